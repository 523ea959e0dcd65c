"""skm_optimize replays its iterations as one captured HIP graph when asked
to; replay and kernel-by-kernel launches must give the same parameters and
per-iteration records, bit for bit (the same kernels in the same order on the same data:
every replay reads the parameters, optimizer state, iteration count and
random-stream state from device memory).

Each case checks every replay (one record per iteration) against the eager
run, and repeats the pair of runs after eager work on the same evaluators
(an evaluation through the host entry point and a synchronisation), so that
the second capture starts from state the first left behind.
"""
import numpy as np
import pytest

from helpers import Case
from skmimageregistration_amd import MEASURE, _capi

pytestmark = pytest.mark.gpu

CASES = [
    # (id, Case kwargs, optimize kwargs, regularizer set-up)
    ("3d_tensor_ncc_sub2_rd", dict(dim=3, shape=(20, 22, 18), cp_spacing=5.0, sigma=10.0, mode=2, subsample=2,
                                   seed=61, coef_scale=0.3),
     dict(metric=MEASURE.NCC, kind=_capi.SKM_OPT_ADAPTIVE, a=20.0, do_resampling=True), ("rd", 1e-3)),
    ("2d_stationary_mse_l1", dict(dim=2, shape=(40, 44), cp_spacing=5.0, sigma=12.0, seed=62, coef_scale=0.3),
     dict(metric=MEASURE.MSE, kind=_capi.SKM_OPT_STANDARD, a=50.0, reg_l1=1e-4, orthant_projection=True), None),
    ("2d_weight_image_lcc_sub2", dict(dim=2, shape=(36, 40), cp_spacing=6.0, sigma=10.0, mode=1, subsample=2,
                                      subsample_neighborhood=2, seed=63, coef_scale=0.3),
     dict(metric=MEASURE.LCC, kind=_capi.SKM_OPT_ADAPTIVE, a=5.0, do_resampling=True, return_best=True),
     ("rkhs", 1e-3)),
]


def _evaluator(case, reg):
    ev = case.gpu()
    if reg is not None:
        name, w = reg
        if name == "rd":
            ev.SetRegularizerRD(w)
        else:
            ev.SetRegularizerRKHS(w)
    return ev


@pytest.mark.parametrize("cid,ckw,okw,reg", CASES, ids=[c[0] for c in CASES])
def test_graph_replay_equals_eager_launches(cid, ckw, okw, reg):
    case = Case(**ckw)
    params0 = case.cp.T.reshape(-1).astype(np.float32)
    iters = 12
    ev_e, ev_g = _evaluator(case, reg), _evaluator(case, reg)
    for round_ in range(2):
        print(f"{cid}: round {round_}, eager", flush=True)
        p_e, r_e = ev_e.optimize(params0, iterations=iters, graph=False, **okw)
        print(f"{cid}: round {round_}, graph", flush=True)
        p_g, r_g = ev_g.optimize(params0, iterations=iters, graph=True, **okw)
        assert len(r_e) == len(r_g) == iters
        for k, (a, b) in enumerate(zip(r_e, r_g)):
            assert a == b, f"iteration {k}: eager {a} vs graph {b}"
        np.testing.assert_array_equal(p_e, p_g)
        # eager work on both evaluators between the captures
        cp = np.ascontiguousarray(p_e.reshape(case.dim, -1).T)
        res_e, der_e = ev_e.getValueAndDerivative(okw["metric"], cp)
        res_g, der_g = ev_g.getValueAndDerivative(okw["metric"], cp)
        ev_g.synchronize()
        assert res_e.measure == res_g.measure
        np.testing.assert_array_equal(der_e, der_g)
        params0 = p_e
