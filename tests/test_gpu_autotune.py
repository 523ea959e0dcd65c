"""skm_autotune and the launch-shape table.

The autotuner times every launch shape skm_set_tuning offers and keeps the
fastest; whatever it picks, an evaluation must still agree with the oracle
(value 1e-5 relative, derivative 1e-4 of its largest entry, the parity bar
of the other metric tests) and with the evaluation under the library's
automatic shapes to FP32 rounding of per-workgroup partial sums (value 1e-6
relative, derivative 1e-5).  The chosen shapes are recorded for the
evaluator's shape class: a new evaluator of the same class starts with them,
a saved table loads back, and a line out of range is refused.  The table
is process-wide, so the cases use grids no other test uses.
"""
import numpy as np
import pytest
import torch

from helpers import Case, rel_err, scalar_rel
from skmimageregistration_amd import MEASURE, _capi, tuning_table_load, tuning_table_save

pytestmark = pytest.mark.gpu

CASES = [
    ("mse_tensor_sub2", dict(dim=3, shape=(23, 22, 21), cp_spacing=5.0, sigma=10.0, mode=2, subsample=2, seed=81),
     MEASURE.MSE),
    ("ncc_stationary", dict(dim=3, shape=(27, 25, 26), cp_spacing=6.0, sigma=12.0, seed=82), MEASURE.NCC),
    ("lcc_reference_windows", dict(dim=3, shape=(25, 23, 24), cp_spacing=5.0, sigma=10.0, subsample=2,
                                   subsample_neighborhood=2, seed=83), MEASURE.LCC),
    ("lcc_exact", dict(dim=3, shape=(21, 19, 20), cp_spacing=5.0, sigma=10.0, subsample_neighborhood=2, seed=84,
                       exact=True), MEASURE.LCC),
]


def _device(case):
    cp = torch.from_numpy(np.ascontiguousarray(case.cp.reshape(-1))).cuda()
    cpw = torch.from_numpy(case.cpw).cuda()
    torch.cuda.synchronize()
    return cp, cpw


@pytest.mark.parametrize("cid,ckw,metric", CASES, ids=[c[0] for c in CASES])
def test_autotuned_shapes_keep_results(cid, ckw, metric):
    case = Case(coef_scale=0.5, **ckw)
    # the first evaluation draws the offsets; none of the later ones resamples
    ores, oder = case.oracle(offsets=case.offsets(1) if case.sub > 1 else None).evaluate(int(metric), True)
    ev = case.gpu()
    base_res, base_der = ev.getValueAndDerivative(metric, case.cp, case.cpw, do_resampling=False)
    assert scalar_rel(base_res.measure, ores["measure"]) < 1e-5
    cp, cpw = _device(case)
    rep = ev.autotune_device(metric, cp.data_ptr(), cpw.data_ptr(), repeats=3)
    print(f"{cid}: {rep}", flush=True)
    assert rep["candidates"] >= 2 and rep["best_ms"] <= rep["first_ms"] and rep["best_ms"] > 0
    tuned = ev.get_tuning()
    keys = ("lcc_ref_chunks", "gather_split", "value_split", "gather_regions", "lcc_ref_walk", "kernel_matrix")
    assert {k: tuned[k] for k in keys} == {k: rep[k] for k in keys}
    res, der = ev.getValueAndDerivative(metric, case.cp, case.cpw, do_resampling=False)
    assert res.numPixelsCountedGpu == base_res.numPixelsCountedGpu == ores["pixels_counted"]
    assert scalar_rel(res.measure, ores["measure"]) < 1e-5
    assert rel_err(der, oder) < 1e-4
    assert scalar_rel(res.measure, base_res.measure) < 1e-6
    assert rel_err(der, base_der) < 1e-5
    # a second evaluator of the same shape class starts with the tuned shapes
    ev2 = case.gpu()
    assert ev2.get_tuning() == tuned
    ev2.close()
    ev.close()


def test_table_save_and_load(tmp_path):
    case = Case(dim=3, shape=(19, 20, 23), cp_spacing=5.0, sigma=10.0, subsample=2, seed=85)
    ev = case.gpu()
    cp, cpw = _device(case)
    ev.autotune_device(MEASURE.NCC, cp.data_ptr(), cpw.data_ptr(), repeats=1)
    ev.close()
    path = tmp_path / "shapes.txt"
    tuning_table_save(path)
    lines = [ln for ln in path.read_text().splitlines() if ln and not ln.startswith("#")]
    assert lines
    # rewrite every saved class with a fixed valid choice, load it back
    fixed = [ln.split()[0] + " 0 4 8 1 0" for ln in lines]
    path.write_text("\n".join(fixed) + "\n")
    tuning_table_load(path)
    ev = case.gpu()
    assert ev.get_tuning() == dict(lcc_ref_chunks=0, gather_split=4, value_split=8, gather_regions=1,
                                   lcc_ref_walk=0, kernel_matrix=0)
    ev.close()
    bad = tmp_path / "bad.txt"
    bad.write_text(lines[0].split()[0] + " 0 5 8 1 0\n")
    with pytest.raises(_capi.SkmError, match="out of range"):
        tuning_table_load(bad)
    with pytest.raises(_capi.SkmError):
        tuning_table_load(tmp_path / "missing.txt")
