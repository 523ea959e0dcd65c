"""The FP32 CPU baseline (oracle/cpu_fp32.c) against the FP64 oracle on
small 3D NCC problems (stationary and anisotropic kernels, RD): value 1e-5,
RD value 1e-5, derivative 1e-4 of its largest entry -- the baseline computes
what the GPU path computes."""
import numpy as np
import pytest

import cpu_fp32
import oracle as orc
from helpers import Case, rel_err, scalar_rel


@pytest.mark.parametrize("mode", [0, 2])
def test_cpu_fp32_equals_oracle(mode):
    case = Case(dim=3, shape=(18, 20, 22), cp_spacing=5.0, sigma=12.0, mode=mode, seed=61)
    case.cpw[:] = 1.0
    ores, oder = case.oracle().evaluate(1, True, 0.0, 1e-3, 0.0)
    coeffs = orc.bspline_coefficients(case.moving.astype(np.float64))
    cov = case.weight.reshape(-1, 9) if mode == 2 else None
    cp = cpu_fp32.CpuProblem(case.fixed, coeffs, (case.cp_size, case.cp_origin, case.cp_spacing), case.cp, case.cpw,
                             case.sigma_c4, cov=cov, max_eig_cov=case.max_eig if mode == 2 else None)
    cpu_fp32.load().skmf_set_threads(4)
    m, rd, der = cp.evaluate(1e-3)
    assert scalar_rel(m, ores["measure"]) < 1e-5, (m, ores["measure"])
    assert scalar_rel(rd, ores["reg_rd"]) < 1e-5, (rd, ores["reg_rd"])
    assert rel_err(der, oder) < 1e-4, rel_err(der, oder)


def test_cpu_fp32_bilateral_equals_oracle():
    """The baseline's bilateral term against the oracle's (value 1e-5,
    weighted derivative added to NCC's within 1e-4 of the largest entry)."""
    case = Case(dim=3, shape=(18, 20, 22), cp_spacing=5.0, sigma=12.0, seed=62, coef_scale=0.8)
    case.cpw[:] = 1.0
    pr = case.oracle()
    ores, oder = pr.evaluate(1, True)
    bval, bder = pr.bilateral(0.3, 8.0, 0.2)
    coeffs = orc.bspline_coefficients(case.moving.astype(np.float64))
    cp = cpu_fp32.CpuProblem(case.fixed, coeffs, (case.cp_size, case.cp_origin, case.cp_spacing), case.cp, case.cpw,
                             case.sigma_c4)
    cpu_fp32.load().skmf_set_threads(4)
    m, _, bv, der = cp.evaluate(0.0, bilateral=(0.3, 8.0, 0.2))
    assert bval > 0
    assert scalar_rel(bv, bval) < 1e-5, (bv, bval)
    assert scalar_rel(m, ores["measure"]) < 1e-5
    assert rel_err(der, oder + bder) < 1e-4, rel_err(der, oder + bder)
