"""skm_optimize_batch: several evaluators (different image and control-point
grids, kernel modes and regularizers) optimised together, kernel by kernel
and as one captured graph per batch step (one branch per evaluator).  Every
evaluator's parameters and per-iteration records must equal those of
skm_optimize on it alone, bit for bit: the branches share no device state
and each runs the same kernels in the same order as the single run.
"""
import numpy as np
import pytest

from helpers import Case
from skmimageregistration_amd import MEASURE, _capi, optimize_batch

pytestmark = pytest.mark.gpu

MEMBERS = [
    # (Case kwargs, regularizer)
    (dict(dim=3, shape=(20, 22, 18), cp_spacing=5.0, sigma=10.0, mode=2, subsample=2, seed=71, coef_scale=0.3),
     ("rd", 1e-3)),
    (dict(dim=3, shape=(24, 16, 20), cp_spacing=6.0, sigma=12.0, subsample=2, seed=72, coef_scale=0.3), None),
    (dict(dim=3, shape=(16, 16, 16), cp_spacing=4.0, sigma=8.0, mode=1, seed=73, coef_scale=0.3), ("rkhs", 1e-3)),
]

OPTIONS = {
    "ncc_adaptive": dict(metric=MEASURE.NCC, kind=_capi.SKM_OPT_ADAPTIVE, a=20.0, do_resampling=True),
    "mse_standard_l1": dict(metric=MEASURE.MSE, kind=_capi.SKM_OPT_STANDARD, a=50.0, reg_l1=1e-4,
                            orthant_projection=True, do_resampling=True),
    "lcc_best": dict(metric=MEASURE.LCC, kind=_capi.SKM_OPT_ADAPTIVE, a=5.0, do_resampling=True, return_best=True),
}


def _members(extra):
    out = []
    for ckw, reg in MEMBERS:
        case = Case(**ckw, **extra)
        ev = case.gpu()
        if reg is not None:
            name, w = reg
            (ev.SetRegularizerRD if name == "rd" else ev.SetRegularizerRKHS)(w)
        out.append((case, ev))
    return out


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("opt", sorted(OPTIONS))
def test_batch_equals_each_alone(opt, graph):
    okw = OPTIONS[opt]
    extra = {"subsample_neighborhood": 2} if opt == "lcc_best" else {}  # small LCC windows
    iters = 10
    alone = _members(extra)
    batch = _members(extra)
    starts = [c.cp.T.reshape(-1).astype(np.float32) for c, _ in alone]
    cpws = [c.cpw for c, _ in alone]
    want = []
    for (case, ev), p0, w in zip(alone, starts, cpws):
        print(f"{opt}: alone {case.shape}", flush=True)
        want.append(ev.optimize(p0, iterations=iters, graph=False, cpwData=w, **okw))
    print(f"{opt}: batch of {len(batch)}, graph={graph}", flush=True)
    got = optimize_batch([ev for _, ev in batch], starts, iterations=iters, graph=graph, cpwData=cpws, **okw)
    for i, ((p_w, r_w), (p_g, r_g)) in enumerate(zip(want, got)):
        assert len(r_g) == iters
        for k, (a, b) in enumerate(zip(r_w, r_g)):
            assert a == b, f"member {i} iteration {k}: alone {a} vs batch {b}"
        np.testing.assert_array_equal(p_w, p_g, err_msg=f"member {i}")
    for _, ev in alone + batch:
        ev.close()


def test_batch_rejects_a_repeated_evaluator():
    case = Case(**MEMBERS[1][0])
    ev = case.gpu()
    p0 = case.cp.T.reshape(-1).astype(np.float32)
    with pytest.raises(Exception, match="twice"):
        optimize_batch([ev, ev], [p0, p0], iterations=2, graph=True, metric=MEASURE.NCC)
    ev.close()
