"""Pins of the warm-start network oracle (oracle/warmstart_oracle.py):
BF16 rounding against known values, a single-channel convolution against
scipy's correlate, and the backward pass against central differences of the
plain-FP64 network (bf16 off)."""
import numpy as np
import pytest
from scipy import ndimage

import warmstart_oracle as wo


def test_bf16_rounding_known_values():
    one = np.float32(1.0)
    assert wo.bf16(one) == 1.0
    assert wo.bf16(1.0 + 2.0 ** -8) == 1.0                      # tie: to even
    assert wo.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6      # tie: to even (up)
    assert wo.bf16(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    assert wo.bf16(-3.0) == -3.0 and wo.bf16(0.0) == 0.0
    x = np.random.default_rng(0).normal(size=1000)
    r = wo.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 + 1e-300)


@pytest.mark.parametrize("stride", [1, 2])
def test_conv_against_scipy_correlate(stride):
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 8, 10, 12, 1))
    k = rng.normal(size=(3, 3, 3))
    w = k.reshape(1, 27, 1)
    y, z = wo.conv_forward(x, w, np.array([0.25]), stride, lambda a: np.asarray(a, np.float64))
    ref = ndimage.correlate(x[0, ..., 0], k, mode="constant", cval=0.0)[::stride, ::stride, ::stride] + 0.25
    np.testing.assert_allclose(z[0, ..., 0], ref, rtol=0, atol=1e-12)
    np.testing.assert_allclose(y, np.maximum(z, 0))


def test_network_backward_central_differences():
    rng = np.random.default_rng(2)
    params = wo.init_params(3)
    params += rng.normal(0, 0.05, size=params.shape)  # non-zero biases
    x = rng.uniform(0, 1, size=(2, 8, 8, 16, 2))
    cp = ((5, 4, 3), (0.5, 0.25, 1.0), (3.0, 2.0, 2.5))
    net = wo.Network(params, *cp, bf16_operands=False)
    out = net.forward(x)
    assert out.shape == (2, 3, 4, 5, 3)
    dout = rng.normal(size=out.shape)
    grad = net.backward(dout)
    layout = wo.param_layout()
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in layout])
    h = 1e-6
    checked = 0
    for li, (name, _) in enumerate(layout):
        for idx in rng.choice(offs[li + 1] - offs[li], 2, replace=False):
            i = offs[li] + idx
            vals = []
            for sgn in (1, -1):
                p = params.copy()
                p[i] += sgn * h
                vals.append(float(np.sum(dout * wo.Network(p, *cp, bf16_operands=False).forward(x))))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - grad[i]) < 1e-6 * max(1.0, abs(fd)), (name, idx, fd, grad[i])
            checked += 1
    assert checked == 2 * len(layout)
