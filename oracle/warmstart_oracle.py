"""FP64 numpy restatement of the warm-start network (forward and backward).

TEST INFRASTRUCTURE ONLY: the checker of the library's warm-start kernels
(skm_ws_* in libskm.so).  The network has no counterpart in the reference
(which starts every registration from zero coefficients,
src/SKMReg.cxx:375-430).

Layers, activations channels-last [batch][z][y][x][channel]:
  conv0 2->16, conv1 16->16                   (level 0, the input grid)
  conv2 16->32 stride 2, conv3 32->32         (level 1)
  conv4 32->48 stride 2, conv5 48->48         (level 2)
  conv6 48->64 stride 2, conv7 64->64         (level 3)
  conv8 [up2(conv7), conv5] 112->48           (level 2)
  conv9 [up2(conv8), conv3] 80->32            (level 1)
  trilinear sample of conv9 at the control points, head 32->3 (linear)
every conv 3x3x3, zero padding 1, ReLU.  Weights of conv l: [cout][27][cin]
(tap = 9 dz + 3 dy + dx) then the bias; then the head [3][32] and its bias.

With bf16=True the operands of every matrix product are rounded to BF16
(round to nearest even of their float32 value) as the GPU's matrix cores
take them: conv inputs and weights (forward), output gradient and weights
(data gradient), output gradient and inputs (weight gradient), the head
weights.  With bf16=False the network is plain FP64, which the finite
difference checks use.
"""
import numpy as np

CIN = (2, 16, 16, 32, 32, 48, 48, 64, 112, 80)
COUT = (16, 16, 32, 32, 48, 48, 64, 64, 48, 32)
STRIDE = (1, 1, 2, 1, 2, 1, 2, 1, 1, 1)
HEAD_IN, HEAD_OUT = 32, 3


def bf16(a):
    """Round to BF16 (nearest even) through float32; returns float64."""
    f = np.ascontiguousarray(np.asarray(a, np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def param_layout():
    """[(name, shape)] in the library's parameter order."""
    out = []
    for l in range(10):
        out.append((f"w{l}", (COUT[l], 27, CIN[l])))
        out.append((f"b{l}", (COUT[l],)))
    out.append(("wh", (HEAD_OUT, HEAD_IN)))
    out.append(("bh", (HEAD_OUT,)))
    return out


def num_params():
    return sum(int(np.prod(s)) for _, s in param_layout())


def unpack(flat):
    flat = np.asarray(flat, np.float64)
    out, off = {}, 0
    for name, shape in param_layout():
        n = int(np.prod(shape))
        out[name] = flat[off:off + n].reshape(shape)
        off += n
    return out


def pack(params):
    return np.concatenate([np.asarray(params[name], np.float64).reshape(-1) for name, _ in param_layout()])


def _taps():
    return [(dz, dy, dx) for dz in range(3) for dy in range(3) for dx in range(3)]


def conv_forward(x, w, b, stride, rnd):
    """x [B,D,H,W,Ci], w [Co,27,Ci] -> relu(conv + b) and the pre-activation."""
    B, D, H, W, _ = x.shape
    Do, Ho, Wo = D // stride, H // stride, W // stride
    xp = np.pad(rnd(x), ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    wr = rnd(w)
    z = np.zeros((B, Do, Ho, Wo, w.shape[0]))
    for t, (dz, dy, dx) in enumerate(_taps()):
        sl = xp[:, dz:dz + stride * Do:stride, dy:dy + stride * Ho:stride, dx:dx + stride * Wo:stride, :]
        z += sl @ wr[:, t, :].T
    z += b
    return np.maximum(z, 0.0), z


def conv_backward(x, w, dz, stride, rnd, need_dx=True):
    """Gradients of sum(dz * conv(x)) w.r.t. x, w and b (dz already masked by the ReLU)."""
    B, D, H, W, _ = x.shape
    Do, Ho, Wo = dz.shape[1:4]
    xp = np.pad(rnd(x), ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    dzr = rnd(dz)
    wr = rnd(w)
    dw = np.zeros_like(w, dtype=np.float64)
    dxp = np.zeros(xp.shape) if need_dx else None
    for t, (dzt, dyt, dxt) in enumerate(_taps()):
        sl = (slice(None), slice(dzt, dzt + stride * Do, stride), slice(dyt, dyt + stride * Ho, stride),
              slice(dxt, dxt + stride * Wo, stride), slice(None))
        dw[:, t, :] = np.einsum("bzyxo,bzyxi->oi", dzr, xp[sl])
        if need_dx:
            dxp[sl] += dzr @ wr[:, t, :]
    db = dz.sum(axis=(0, 1, 2, 3))
    dx = dxp[:, 1:-1, 1:-1, 1:-1, :] if need_dx else None
    return dx, dw, db


def upcat(coarse, skip):
    up = coarse.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)
    return np.concatenate([up, skip], axis=-1)


def upcat_backward(dout, ca):
    dup = dout[..., :ca]
    B, D, H, W, _ = dup.shape
    dcoarse = dup.reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, ca).sum(axis=(2, 4, 6))
    return dcoarse, dout[..., ca:]


def _sample_weights(n_cp, origin, spacing, n_feat):
    """[n_cp, n_feat] trilinear weights along one axis: max(0, 1 - |q_j - f|)
    with q_j = clamp(origin + spacing j, 0, n_feat - 1)."""
    q = np.clip(np.float32(origin) + np.float32(spacing) * np.arange(n_cp, dtype=np.float32), 0, n_feat - 1)
    f = np.arange(n_feat)
    return np.maximum(0.0, 1.0 - np.abs(q.astype(np.float64)[:, None] - f[None, :]))


def sample_matrices(feat_shape, cp_size, cp_origin, cp_spacing):
    """Per-axis weight matrices (z, y, x) of the control points in feature voxels."""
    D, H, W = feat_shape
    wx = _sample_weights(cp_size[0], cp_origin[0], cp_spacing[0], W)
    wy = _sample_weights(cp_size[1], cp_origin[1], cp_spacing[1], H)
    wz = _sample_weights(cp_size[2], cp_origin[2], cp_spacing[2], D)
    return wz, wy, wx


def sample(feat, mats):
    wz, wy, wx = mats
    return np.einsum("az,by,cx,nzyxk->nabck", wz, wy, wx, feat)


def sample_backward(dout, mats):
    wz, wy, wx = mats
    return np.einsum("az,by,cx,nabck->nzyxk", wz, wy, wx, dout)


class Network:
    """cp_origin / cp_spacing in input voxel units, as skm_ws_desc takes them."""

    def __init__(self, params, cp_size, cp_origin, cp_spacing, bf16_operands=True):
        self.p = unpack(params)
        self.cp_size = tuple(cp_size)
        self.cp_origin = tuple(float(v) * 0.5 for v in cp_origin)   # level-1 feature voxels
        self.cp_spacing = tuple(float(v) * 0.5 for v in cp_spacing)
        self.rnd = bf16 if bf16_operands else (lambda a: np.asarray(a, np.float64))

    def forward(self, x):
        x = np.asarray(x, np.float64)
        p, rnd = self.p, self.rnd
        acts, srcs = [], []
        src = x
        for l in range(8):
            srcs.append(src)
            y, _ = conv_forward(src, p[f"w{l}"], p[f"b{l}"], STRIDE[l], rnd)
            acts.append(y)
            src = y
        c3 = upcat(acts[7], acts[5])
        srcs.append(c3)
        y8, _ = conv_forward(c3, p["w8"], p["b8"], 1, rnd)
        acts.append(y8)
        c2 = upcat(y8, acts[3])
        srcs.append(c2)
        y9, _ = conv_forward(c2, p["w9"], p["b9"], 1, rnd)
        acts.append(y9)
        mats = sample_matrices(y9.shape[1:4], self.cp_size, self.cp_origin, self.cp_spacing)
        s = sample(y9, mats)
        out = s @ rnd(p["wh"]).T + p["bh"]
        self._cache = (srcs, acts, mats, s)
        return out

    def use_activations(self, x, acts):
        """Take the ten conv outputs `acts` (e.g. the GPU's) as the forward's,
        so that backward() and layer_outputs() start from them; returns the
        head output from acts[9]."""
        x = np.asarray(x, np.float64)
        acts = [np.asarray(a, np.float64) for a in acts]
        srcs = [x] + acts[:7] + [upcat(acts[7], acts[5]), upcat(acts[8], acts[3])]
        mats = sample_matrices(acts[9].shape[1:4], self.cp_size, self.cp_origin, self.cp_spacing)
        s = sample(acts[9], mats)
        self._cache = (srcs, acts, mats, s)
        return s @ self.rnd(self.p["wh"]).T + self.p["bh"]

    def layer_outputs(self):
        """Each conv's output recomputed from the inputs of use_activations."""
        srcs, _, _, _ = self._cache
        return [conv_forward(srcs[l], self.p[f"w{l}"], self.p[f"b{l}"], STRIDE[l], self.rnd)[0] for l in range(10)]

    def backward(self, dout):
        """Gradients of sum(dout * forward(x)) for the last forward, packed."""
        p, rnd = self.p, self.rnd
        srcs, acts, mats, s = self._cache
        g = {}
        g["wh"] = np.einsum("nabck,nabcj->kj", dout, s)
        g["bh"] = dout.sum(axis=(0, 1, 2, 3))
        ds = dout @ rnd(p["wh"])
        d = sample_backward(ds, mats) * (acts[9] > 0)
        dc2, g["w9"], g["b9"] = conv_backward(srcs[9], p["w9"], d, 1, rnd)
        d8, dskip2 = upcat_backward(dc2, 48)
        d = d8 * (acts[8] > 0)
        dc3, g["w8"], g["b8"] = conv_backward(srcs[8], p["w8"], d, 1, rnd)
        d7, dskip3 = upcat_backward(dc3, 64)
        d = d7
        for l in range(7, -1, -1):
            d = d * (acts[l] > 0)
            dx, g[f"w{l}"], g[f"b{l}"] = conv_backward(srcs[l], p[f"w{l}"], d, STRIDE[l], rnd, need_dx=l > 0)
            if l == 0:
                break
            d = dx
            if l - 1 == 5:
                d = d + dskip3
            if l - 1 == 3:
                d = d + dskip2
        return pack(g)


def init_params(seed=0):
    """He-normal conv weights, zero biases, small head (library-independent)."""
    rng = np.random.default_rng(seed)
    p = {}
    for l in range(10):
        fan_in = 27 * CIN[l]
        p[f"w{l}"] = rng.normal(0.0, np.sqrt(2.0 / fan_in), size=(COUT[l], 27, CIN[l]))
        p[f"b{l}"] = np.zeros(COUT[l])
    p["wh"] = rng.normal(0.0, np.sqrt(1.0 / HEAD_IN), size=(HEAD_OUT, HEAD_IN))
    p["bh"] = np.zeros(HEAD_OUT)
    return pack(p)
