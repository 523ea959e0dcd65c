"""Device memory comes from a block cache: an evaluator created after another
one of the same sizes receives the blocks that held the other's data.  No
kernel may read memory it has not written, so an evaluation repeated after a
different problem of the same sizes gives the same bits (MSE, NCC with
jittered subsampling, LCC with the reference's windows; the exact-gradient
LCC mode adds with atomics and is left out)."""
import numpy as np
import pytest

from skmimageregistration_amd import MEASURE, GpuKernelEvaluator, Grid, Image, Kernel, synthetic

pytestmark = pytest.mark.gpu

SHAPE, H, S = (24, 28, 32), 4.0, 10.0


def _evaluate(mov, fix, measure, sub, subn):
    size = tuple(reversed(SHAPE))
    cs = tuple(int(np.ceil((s - 1) / H)) for s in size)
    co = tuple(abs((s - 1) - H * (n - 1)) / 2.0 for s, n in zip(size, cs))
    n = int(np.prod(cs))
    cp = (np.random.default_rng(3).uniform(-1, 1, (n, 3))).astype(np.float32)
    ev = GpuKernelEvaluator(n * 3, sub, subn, Kernel(S, S), Image(mov), Image(fix), Grid(cs, co, (H,) * 3))
    ev.SetRegularizerRD(1e-3)
    r, d = ev.getValueAndDerivative(measure, cp)
    ev.close()
    return r.measure, r.reg_value_rd, np.array(d)


@pytest.mark.parametrize("measure,sub,subn", [(MEASURE.MSE, 1, 1), (MEASURE.NCC, 2, 1), (MEASURE.LCC, 2, 3)])
def test_results_do_not_depend_on_recycled_memory(measure, sub, subn):
    mov_a, fix_a, _ = synthetic.ellipsoids_3d(SHAPE, max_disp=2.0)
    mov_b, fix_b, _ = synthetic.ellipsoids_3d(SHAPE, max_disp=3.0, seeds=(41, 42, 43))
    print(f"measure {measure}, subsample {sub}", flush=True)
    first = _evaluate(mov_a, fix_a, measure, sub, subn)
    other = _evaluate(mov_b, fix_b, measure, sub, subn)
    again = _evaluate(mov_a, fix_a, measure, sub, subn)
    assert other[0] != first[0]
    assert again[0] == first[0] and again[1] == first[1]
    np.testing.assert_array_equal(again[2], first[2])
