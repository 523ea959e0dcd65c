"""Device-resident optimisation loop (skm_optimize) against the same loop run
on the FP64 oracle: GetValueAndDerivative, host regularizers L1/L21/L2 with
the sign flip, and the gradient-descent update with the standard, adaptive or
fixed step (itkStandardGradientDescentOptimizerv4.hxx:116-166,
itkAdaptiveStepGradientDescentOptimizerv4.hxx:71-128).

Tolerance: parameters after K steps within 2e-4 of the largest parameter
(FP32 device loop against FP64), recorded values relative 1e-5.
"""
import ctypes

import numpy as np
import pytest

import oracle as orc
from helpers import Case, rel_err
from skmimageregistration_amd import MEASURE, _capi

pytestmark = pytest.mark.gpu


def oracle_loop(case, measure, iters, kind, a, A, alpha, fmax, fmin, omega, orthant, l1=0.0, l21=0.0, l2=0.0,
                rd=0.0):
    L = orc.load()
    n, dim = case.n_cp, case.dim
    theta = case.cp.astype(np.float64).T.reshape(-1).copy()  # SoA
    prev_g = np.zeros_like(theta)
    t = 0.0
    values = []
    for k in range(iters):
        cp = theta.reshape(dim, n).T.copy()
        res, der = case.oracle(cp=cp).evaluate(int(measure), True, 0.0, rd, 0.0)
        v1, v21, v2 = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        reg = L.skmo_host_regularizers(dim, n, orc.dp(theta), l1, l21, l2, 1, orc.dp(der), ctypes.byref(v1),
                                       ctypes.byref(v21), ctypes.byref(v2))
        values.append(res["measure"] + reg + rd * res["reg_rd"])
        if kind == _capi.SKM_OPT_STANDARD:
            rate = L.skmo_standard_rate(a, A, alpha, k)
        elif kind == _capi.SKM_OPT_ADAPTIVE:
            rate = L.skmo_adaptive_rate(a, A, alpha, t)
        else:
            rate = a
        dot = float(np.dot(der, prev_g))
        out = np.zeros_like(theta)
        L.skmo_gd_update(theta.size, orc.dp(theta), orc.dp(der), rate, int(orthant), orc.dp(out))
        theta = out
        if kind == _capi.SKM_OPT_ADAPTIVE:
            t = L.skmo_adaptive_time(t, dot, fmax, fmin, omega)
        prev_g = der
    return theta, values


@pytest.mark.parametrize("kind", [_capi.SKM_OPT_STANDARD, _capi.SKM_OPT_ADAPTIVE, _capi.SKM_OPT_FIXED])
def test_device_loop_matches_oracle_mse_l1(kind):
    case = Case(dim=2, shape=(40, 44), cp_spacing=5.0, sigma=12.0, seed=51, coef_scale=0.3)
    case.cpw[:] = 1.0
    opts = dict(a=50.0, A=20.0, alpha=1.0, fmax=1e7, fmin=-1e6, omega=1.0)
    ev = case.gpu()
    params0 = case.cp.T.reshape(-1).astype(np.float32)
    params, recs = ev.optimize(params0, metric=MEASURE.MSE, reg_l1=1e-4, kind=kind, orthant_projection=True,
                               iterations=5, cpwData=case.cpw, **opts)
    ev.close()
    theta, values = oracle_loop(case, MEASURE.MSE, 5, kind, orthant=True, l1=1e-4, **opts)
    assert rel_err(params, theta) < 2e-4
    for (v, *_), ov in zip(recs, values):
        assert abs(v - ov) / abs(ov) < 1e-5


def test_device_loop_matches_oracle_ncc_rd_l21_l2():
    case = Case(dim=3, shape=(16, 18, 20), cp_spacing=5.0, sigma=12.0, seed=53, coef_scale=0.3)
    case.cpw[:] = 1.0
    opts = dict(a=2.0, A=20.0, alpha=1.0, fmax=1e7, fmin=-1e6, omega=1.0)
    ev = case.gpu()
    ev.SetRegularizerRD(0.01)
    params0 = case.cp.T.reshape(-1).astype(np.float32)
    params, recs = ev.optimize(params0, metric=MEASURE.NCC, reg_l21=1e-4, reg_l2=1e-3,
                               kind=_capi.SKM_OPT_ADAPTIVE, iterations=4, cpwData=case.cpw, **opts)
    ev.close()
    theta, values = oracle_loop(case, MEASURE.NCC, 4, _capi.SKM_OPT_ADAPTIVE, orthant=False, l21=1e-4, l2=1e-3,
                                rd=0.01, **opts)
    assert rel_err(params, theta) < 2e-4
    for (v, *_), ov in zip(recs, values):
        assert abs(v - ov) / abs(ov) < 1e-5


def test_registration_recovers_blob_warp():
    """2D blobs (128x128 version of the 256^2 workload), MSE + L1 1e-3 with
    orthant projection, fixed step: the MSE drops by more than 3x in 100 steps."""
    case = Case(dim=2, shape=(128, 128), cp_spacing=8.0, sigma=24.0, seed=57, coef_scale=0.0)
    case.cpw[:] = 1.0
    ev = case.gpu()
    params0 = np.zeros(case.n_cp * 2, np.float32)
    params, recs = ev.optimize(params0, metric=MEASURE.MSE, reg_l1=1e-5, kind=_capi.SKM_OPT_FIXED, a=400.0,
                               orthant_projection=True, iterations=100, cpwData=case.cpw)
    ev.close()
    first, last = recs[0][1], recs[-1][1]
    assert np.isfinite(params).all()
    assert last < first / 3.0, (first, last)


def test_return_best_is_the_lowest_value_position():
    """ReturnBestParametersAndValue: the result is the evaluated position with
    the lowest value, not the position after the last step."""
    case = Case(dim=2, shape=(40, 44), cp_spacing=5.0, sigma=12.0, seed=53, coef_scale=0.3)
    case.cpw[:] = 1.0
    ev = case.gpu()
    p0 = case.cp.T.reshape(-1).astype(np.float32)
    opts = dict(metric=MEASURE.MSE, kind=_capi.SKM_OPT_FIXED, a=4e5, iterations=10)
    p_last, recs = ev.optimize(p0, **opts)
    p_best, recs_b = ev.optimize(p0, return_best=True, **opts)
    values = [r[0] for r in recs_b]
    assert values == [r[0] for r in recs]          # same trajectory
    k = int(np.argmin(values))
    p_k, _ = ev.optimize(p0, **dict(opts, iterations=k))   # the position evaluated at step k
    ev.close()
    np.testing.assert_array_equal(p_best, p_k if k > 0 else p0)
    if k < len(values) - 1:
        assert not np.array_equal(p_best, p_last)
