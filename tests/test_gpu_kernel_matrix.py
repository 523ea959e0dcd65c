"""The kernel matrix (subsample 1): the MSE / NCC value and derivative read
from the assembled k(x_m, p_j) against the same evaluator's on-the-fly
kernels (skm_tuning.kernel_matrix = 1) and against the FP64 oracle.

The automatic policy assembles it for the tensor kernel only.

Tolerances: against the oracle, the parity bar (value 1e-5 relative,
derivative 1e-4 of its largest entry); between the two GPU forms, which sum
the same FP32 pair terms in different orders, value 1e-6 and derivative 1e-5.
Repeated evaluations from the matrix are the same bits.
"""
import os

import numpy as np
import pytest

from helpers import Case, rel_err, scalar_rel
from skmimageregistration_amd import MEASURE
from skmimageregistration_amd._capi import SKM_OPT_FIXED

pytestmark = pytest.mark.gpu

VALUE_TOL, DERIV_TOL = 1e-5, 1e-4
PAIR_VALUE_TOL, PAIR_DERIV_TOL = 1e-6, 1e-5


def _eval(case, measure, kernel_matrix, gather_split=0, regs=None):
    ev = case.gpu()
    ev.set_tuning(kernel_matrix=kernel_matrix, gather_split=gather_split)
    if regs:
        ev.SetRegularizerRD(regs)
    res, der = ev.getValueAndDerivative(measure, case.cp, case.cpw, case.df_image())
    info = ev.kernel_matrix_info()
    return ev, res, der, info


def _check(case, measure, gather_split=0):
    print(f"case dim={case.dim} shape={case.shape} mode={case.mode} measure={int(measure)} "
          f"exact={case.exact} df={case.df is not None} split={gather_split}", flush=True)
    ev_km, r_km, d_km, info = _eval(case, measure, 2, gather_split)
    assert info["state"] == 1 and info["pairs_value"] > 0 and info["pairs_gather"] > 0, info
    # a second evaluation from the matrix: the same bits
    r2, d2 = ev_km.getValueAndDerivative(measure, case.cp, case.cpw, case.df_image())
    assert r2.measure == r_km.measure
    np.testing.assert_array_equal(d2, d_km)
    ev_km.close()
    ev_f, r_f, d_f, info_f = _eval(case, measure, 1, gather_split)
    assert info_f["state"] == 0
    ev_f.close()
    assert r_km.numPixelsCountedGpu == r_f.numPixelsCountedGpu
    assert scalar_rel(r_km.measure, r_f.measure) < PAIR_VALUE_TOL, (r_km.measure, r_f.measure)
    assert rel_err(d_km, d_f) < PAIR_DERIV_TOL, rel_err(d_km, d_f)
    ores, oder = case.oracle().evaluate(int(measure), True, 0.0, 0.0, 0.0)
    assert r_km.numPixelsCountedGpu == ores["pixels_counted"]
    assert scalar_rel(r_km.measure, ores["measure"]) < VALUE_TOL
    assert rel_err(d_km, oder) < DERIV_TOL, rel_err(d_km, oder)


@pytest.mark.parametrize("measure", [MEASURE.MSE, MEASURE.NCC])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_kernel_matrix_2d(measure, mode):
    _check(Case(dim=2, mode=mode, seed=41), measure)


@pytest.mark.parametrize("measure", [MEASURE.MSE, MEASURE.NCC])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_kernel_matrix_3d(measure, mode):
    _check(Case(dim=3, shape=(18, 20, 22), cp_spacing=5.0, sigma=12.0, mode=mode, seed=43), measure)


def test_kernel_matrix_unequal_supports():
    # C4 and C0 supports differ: the reach is the smaller one
    _check(Case(dim=3, shape=(16, 18, 20), cp_spacing=4.0, sigma=11.0, sigma_c0=8.0, seed=45), MEASURE.NCC)


@pytest.mark.parametrize("dim", [2, 3])
def test_kernel_matrix_exact_windows_and_previous_field(dim):
    shape = (40, 44) if dim == 2 else (16, 18, 20)
    _check(Case(dim=dim, shape=shape, cp_spacing=5.0, sigma=12.0, exact=True, with_df=True, mode=2, seed=47),
           MEASURE.NCC)


@pytest.mark.parametrize("split", [0, 8])
def test_kernel_matrix_geometry_and_gather_slab(split):
    # the kernel-matrix gather takes 4-plane slabs unless 8 is asked for
    case = Case(dim=3, shape=(17, 21, 23), cp_spacing=6.0, sigma=14.0, mode=1, origin=(-3.5, 2.25, 1.0),
                spacing=(1.25, 0.8, 1.5), seed=49)
    _check(case, MEASURE.MSE, gather_split=split)


def test_kernel_matrix_refused_above_the_limit():
    case = Case(dim=3, shape=(16, 18, 20), cp_spacing=5.0, sigma=12.0, mode=2, seed=51)
    os.environ["SKM_KERNEL_MATRIX_MAX_BYTES"] = "1000"
    try:
        ev, res, der, info = _eval(case, MEASURE.NCC, 2)
    finally:
        del os.environ["SKM_KERNEL_MATRIX_MAX_BYTES"]
    ev.close()
    assert info["state"] == 2 and info["bytes"] > 1000, info
    ores, oder = case.oracle().evaluate(int(MEASURE.NCC), True, 0.0, 0.0, 0.0)
    assert scalar_rel(res.measure, ores["measure"]) < VALUE_TOL
    assert rel_err(der, oder) < DERIV_TOL


def test_kernel_matrix_release_and_reassembly():
    case = Case(dim=3, shape=(16, 18, 20), cp_spacing=5.0, sigma=12.0, mode=2, seed=53)
    ev, r1, d1, info = _eval(case, MEASURE.NCC, 2)
    ev.release_kernel_matrix()
    assert ev.kernel_matrix_info()["state"] == 0
    r2, d2 = ev.getValueAndDerivative(MEASURE.NCC, case.cp, case.cpw)
    assert ev.kernel_matrix_info()["state"] == 1
    ev.close()
    assert r1.measure == r2.measure
    np.testing.assert_array_equal(d1, d2)


@pytest.mark.parametrize("graph", [False, True])
def test_kernel_matrix_optimizer(graph):
    """The device optimizer reads the matrix from its first iteration; the
    graph replay gives the same bits as eager launches, and a fixed-step loop
    of 20 iterations stays close to the on-the-fly loop's."""
    case = Case(dim=3, shape=(18, 20, 22), cp_spacing=5.0, sigma=12.0, mode=2, seed=55)
    start = np.zeros(case.n_cp * 3, np.float32)
    opts = dict(metric=MEASURE.NCC, iterations=20, kind=SKM_OPT_FIXED, a=50.0)
    out = {}
    for km in (2, 1):
        ev = case.gpu()
        ev.set_tuning(kernel_matrix=km)
        p, rec = ev.optimize(start, graph=graph, **opts)
        assert ev.kernel_matrix_info()["state"] == (1 if km == 2 else 0)
        if km == 2 and graph:
            p_eager, rec_eager = ev.optimize(start, graph=False, **opts)
            np.testing.assert_array_equal(p, p_eager)
            assert [r[0] for r in rec] == [r[0] for r in rec_eager]
        ev.close()
        out[km] = (p, rec)
    (p0, r0), (p1, r1) = out[2], out[1]
    assert np.abs(p0).max() > 0
    assert rel_err(p0, p1) < 1e-4, rel_err(p0, p1)
    assert abs(r0[-1][1] - r1[-1][1]) < 1e-5 * abs(r1[-1][1])


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_kernel_matrix_automatic_policy(mode):
    """kernel_matrix 0: assembled for the tensor kernel, on the fly otherwise."""
    case = Case(dim=3, shape=(16, 18, 20), cp_spacing=5.0, sigma=12.0, mode=mode, seed=57)
    ev, res, der, info = _eval(case, MEASURE.NCC, 0)
    ev.close()
    assert info["state"] == (1 if mode == 2 else 0)
    ores, oder = case.oracle().evaluate(int(MEASURE.NCC), True, 0.0, 0.0, 0.0)
    assert scalar_rel(res.measure, ores["measure"]) < VALUE_TOL
    assert rel_err(der, oder) < DERIV_TOL
