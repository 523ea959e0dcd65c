"""The warm-start network's HIP kernels (skm_ws_*) against the FP64 numpy
oracle with the same BF16 operand rounding (oracle/warmstart_oracle.py), at
a small shape and at the workload's shape (input 64 x 64 x 48, control
points 22 x 22 x 16), and a training loop that fits a fixed batch.

Layer by layer: each conv's output and the head against the oracle fed the
GPU's own input activations (max-norm relative 1e-5; both sides round
identical values to BF16, measured <= 6e-7), and every parameter gradient
against the oracle's backward through the GPU's activations (5e-3: the data
gradients, FP32 here and FP64 there, are rounded to BF16 at each of up to
nine layers; measured at the first layer 3.1e-4 small, 1.3e-3 at the
workload's shape).  End to end (forward 5e-3, gradients 0.1 per tensor,
measured 4.4e-2): the oracle's own forward, where an activation
computed in FP32 on the GPU and in FP64 by the oracle can round to
different BF16 neighbours (~2^-8 each, rare) and move a few ReLU masks.
"""
import numpy as np
import pytest

import warmstart_oracle as wo

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FWD_TOL, GRAD_TOL = 5e-3, 0.1
LAYER_TOL, LAYER_GRAD_TOL = 1e-5, 5e-3


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / max(np.max(np.abs(b)), 1e-30))


def _tensor_errors(grad, gref):
    off, errs = 0, {}
    for name, s in wo.param_layout():
        n = int(np.prod(s))
        errs[name] = _rel(grad[off:off + n], gref[off:off + n])
        off += n
    return errs


def _net(in_size, cp, batch):
    from skmimageregistration_amd.warmstart import WarmStartNet

    return WarmStartNet(in_size, *cp, max_batch=batch)


def _params(seed):
    rng = np.random.default_rng(seed)
    p = wo.init_params(seed)
    return (p + rng.normal(0, 0.02, size=p.shape)).astype(np.float32)


@pytest.mark.parametrize("shape", ["small", "workload"])
def test_forward_and_gradients_against_oracle(shape):
    if shape == "small":
        in_size, cp, batch = (24, 16, 16), ((6, 5, 4), (0.5, 1.0, 0.75), (3.5, 3.0, 3.5)), 2
    else:
        # 128 x 128 x 96 volumes downsampled by 2; control points every 6 voxels from 1.5
        in_size, cp, batch = (64, 64, 48), ((22, 22, 16), (0.5, 0.5, 0.5), (3.0, 3.0, 3.0)), 1
    print(f"warm-start network {shape}: input {in_size}, cp {cp[0]}, batch {batch}", flush=True)
    rng = np.random.default_rng(5)
    params = _params(7)
    x = rng.uniform(0, 1, size=(batch, in_size[2], in_size[1], in_size[0], 2)).astype(np.float32)
    net = _net(in_size, cp, batch)
    net.set_params(params)
    assert np.array_equal(net.get_params(), params)
    xt = torch.from_numpy(x).cuda()
    out = net.forward(xt).cpu().numpy()
    ref = wo.Network(params.astype(np.float64), *cp)
    oref = ref.forward(x.astype(np.float64))
    err = _rel(out, oref)
    print(f"forward rel {err:.2e}", flush=True)
    assert np.isfinite(out).all() and err < FWD_TOL
    dout = rng.normal(size=out.shape).astype(np.float32)
    grad = net.backward(torch.from_numpy(dout).cuda()).cpu().numpy()
    gref = ref.backward(dout.astype(np.float64))
    errs = _tensor_errors(grad, gref)
    print("end-to-end gradients rel: " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
    assert max(errs.values()) < GRAD_TOL, errs
    # layer by layer, from the GPU's own activations
    acts = [net.activation(l, batch).cpu().numpy() for l in range(10)]
    head = ref.use_activations(x, acts)
    e_head = _rel(out, head)
    layer_errs = [_rel(a, b) for a, b in zip(acts, ref.layer_outputs())]
    print(f"layers rel: {', '.join(f'{e:.1e}' for e in layer_errs)}; head {e_head:.1e}", flush=True)
    assert max(layer_errs) < LAYER_TOL and e_head < LAYER_TOL
    gl = _tensor_errors(grad, ref.backward(dout.astype(np.float64)))
    print("layer-by-layer gradients rel: " + ", ".join(f"{k} {v:.1e}" for k, v in gl.items()), flush=True)
    assert max(gl.values()) < LAYER_GRAD_TOL, gl
    net.close()


def test_training_fits_a_fixed_batch():
    in_size, cp, batch = (24, 16, 16), ((6, 5, 4), (0.5, 1.0, 0.75), (3.5, 3.0, 3.5)), 4
    rng = np.random.default_rng(9)
    x = torch.from_numpy(rng.uniform(0, 1, size=(batch, 16, 16, 24, 2)).astype(np.float32)).cuda()
    target = torch.from_numpy(rng.normal(0, 0.5, size=(batch, 4, 5, 6, 3)).astype(np.float32)).cuda()
    net = _net(in_size, cp, batch)
    net.set_params(wo.init_params(11).astype(np.float32))
    losses = [net.train_step(x, target, 3e-3) for _ in range(150)]
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.1 * losses[0]
    # the step's loss is the MSE of the forward before the update
    out = net.forward(x)
    mse = float(((out - target) ** 2).mean())
    assert abs(net.train_step(x, target, 0.0) - mse) < 1e-6 * max(1.0, mse)
    net.close()
