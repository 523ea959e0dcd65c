"""The regularizers of an MSE / NCC evaluation of 2^18 samples or more run on
a second stream beside the measure's passes, and the gather's finish adds
their derivative terms; with pass timing on (and in smaller evaluations) they
run after the gather and add the terms themselves.  Both orders must give
the same bits: the measure, every regularizer value and the derivative, for
the on-the-fly passes and the kernel matrix, and for a value-only evaluation.
"""
import numpy as np
import pytest

from helpers import Case
from skmimageregistration_amd import MEASURE

pytestmark = pytest.mark.gpu


def _evaluate(case, measure, kernel_matrix, timing):
    ev = case.gpu()
    ev.set_tuning(kernel_matrix=kernel_matrix)
    ev.SetRegularizerRD(1e-3)
    ev.SetRegularizerBilateral(1e-3, 12.0, 0.1)
    ev.enable_timing(timing)
    res, der = ev.getValueAndDerivative(measure, case.cp, case.cpw)
    val = ev.getValue(measure, case.cp, case.cpw)
    ev.close()
    return res, der, val


def _fields(r):
    return (r.measure, r.reg_value_rkhs, r.reg_value_rd, r.reg_value_pg, r.reg_value_bilateral,
            r.numPixelsCountedGpu)


@pytest.mark.parametrize("measure", [MEASURE.MSE, MEASURE.NCC])
@pytest.mark.parametrize("mode,kernel_matrix", [(0, 1), (2, 2)])
def test_regularizers_beside_equal_in_sequence(measure, mode, kernel_matrix):
    # 66 x 72 x 64 samples at subsample 1: above the 2^18 of the second stream
    case = Case(dim=3, shape=(64, 72, 66), cp_spacing=8.0, sigma=20.0, mode=mode, seed=61, coef_scale=0.5)
    print(f"measure={int(measure)} mode={mode} kernel_matrix={kernel_matrix}", flush=True)
    beside = _evaluate(case, measure, kernel_matrix, False)
    in_sequence = _evaluate(case, measure, kernel_matrix, True)
    assert beside[0].reg_value_rd > 0 and beside[0].reg_value_bilateral > 0
    assert _fields(beside[0]) == _fields(in_sequence[0])
    np.testing.assert_array_equal(beside[1], in_sequence[1])
    assert _fields(beside[2]) == _fields(in_sequence[2])
