"""Learned warm start of SKM registrations (no reference counterpart: the
reference starts every registration from zero coefficients,
src/SKMReg.cxx:375-430).

WarmStartNet wraps the library's network (skm_ws_*: a 3D conv
encoder-decoder with BF16 weights and FP32 accumulation, trained and run by
libskm.so's own kernels).  Device arrays are torch tensors on the GPU (torch
is used for device memory and streams only; every computation is a library
kernel).
"""
import ctypes

import numpy as np

from . import _capi
from ._capi import check


class WarmStartNet:
    """in_size (x, y, z) of the network input (multiples of 8); control points
    in input voxel coordinates (cp_size, cp_origin, cp_spacing, each x, y, z)."""

    def __init__(self, in_size, cp_size, cp_origin, cp_spacing, max_batch=1):
        import torch

        self._torch = torch
        self._lib = _capi.load()
        d = _capi.SkmWsDesc()
        for a in range(3):
            d.in_size[a] = int(in_size[a])
            d.cp_size[a] = int(cp_size[a])
            d.cp_origin[a] = float(cp_origin[a])
            d.cp_spacing[a] = float(cp_spacing[a])
        d.max_batch = int(max_batch)
        h = ctypes.c_void_p()
        check(self._lib.skm_ws_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self.in_size = tuple(int(v) for v in in_size)
        self.cp_size = tuple(int(v) for v in cp_size)
        self.max_batch = int(max_batch)
        self.n_params = int(self._lib.skm_ws_num_params(h))
        check(self._lib.skm_ws_set_stream(h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def close(self):
        if self._h:
            self._lib.skm_ws_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 -- interpreter shutdown
            pass

    # -- parameters (host float32, the library's layout)
    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        check(self._lib.skm_ws_set_params(self._h, _capi.fptr(p)))

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        check(self._lib.skm_ws_get_params(self._h, _capi.fptr(p)))
        return p

    def save(self, path):
        self.get_params().tofile(path)

    def load(self, path):
        self.set_params(np.fromfile(path, np.float32))

    # -- device calls (torch cuda float32 tensors)
    def _shape_in(self, batch):
        return (batch, self.in_size[2], self.in_size[1], self.in_size[0], 2)

    def _shape_out(self, batch):
        return (batch, self.cp_size[2], self.cp_size[1], self.cp_size[0], 3)

    def _check(self, t, shape):
        if tuple(t.shape) != tuple(shape) or t.dtype != self._torch.float32 or not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"expected a contiguous float32 cuda tensor of shape {shape}, got {tuple(t.shape)}")

    def forward(self, x):
        batch = x.shape[0]
        self._check(x, self._shape_in(batch))
        out = self._torch.empty(self._shape_out(batch), device=x.device, dtype=self._torch.float32)
        check(self._lib.skm_ws_forward(self._h, ctypes.c_void_p(x.data_ptr()), batch, ctypes.c_void_p(out.data_ptr())))
        return out

    def backward(self, dout):
        """Parameter gradients for the output gradient of the last forward."""
        grad = self._torch.empty(self.n_params, device=dout.device, dtype=self._torch.float32)
        check(self._lib.skm_ws_backward(self._h, ctypes.c_void_p(dout.data_ptr()), ctypes.c_void_p(grad.data_ptr())))
        return grad

    LAYER_CHANNELS = (16, 16, 32, 32, 48, 48, 64, 64, 48, 32)
    LAYER_LEVELS = (0, 0, 1, 1, 2, 2, 3, 3, 2, 1)

    def activation(self, layer, batch):
        """Output of conv `layer` in the last forward (torch tensor)."""
        lev = self.LAYER_LEVELS[layer]
        shape = (batch, self.in_size[2] >> lev, self.in_size[1] >> lev, self.in_size[0] >> lev,
                 self.LAYER_CHANNELS[layer])
        out = self._torch.empty(shape, device="cuda", dtype=self._torch.float32)
        check(self._lib.skm_ws_get_activation(self._h, int(layer), ctypes.c_void_p(out.data_ptr())))
        return out

    def train_step(self, x, target, learning_rate):
        batch = x.shape[0]
        self._check(x, self._shape_in(batch))
        self._check(target, self._shape_out(batch))
        loss = ctypes.c_double()
        check(self._lib.skm_ws_train_step(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(target.data_ptr()),
                                          batch, float(learning_rate), ctypes.byref(loss)))
        return loss.value


# ---------------------------------------------------------------- registration workload

WORKLOAD_SHAPE = (96, 128, 128)  # z, y, x


def synthetic_pair(seed, shape=WORKLOAD_SHAPE, n_ellipsoids=20, smoothing_sigma=1.5, n_rbf=8, rbf_sigma=None,
                   max_displacement=6.0, displacement_scale=1.0, noise_sigma=0.02):
    """skm_synthetic_pair: (reference, target, displacement) as numpy arrays,
    shapes (z, y, x) and (z, y, x, 3); the displacement RBFs' standard
    deviation defaults to a quarter of the smallest extent."""
    lib = _capi.load()
    d = _capi.SkmSyntheticDesc()
    for a, n in enumerate(reversed(shape)):
        d.size[a] = int(n)
    d.n_ellipsoids = int(n_ellipsoids)
    d.smoothing_sigma = float(smoothing_sigma)
    d.n_rbf = int(n_rbf)
    d.rbf_sigma = float(rbf_sigma if rbf_sigma is not None else 0.25 * min(shape))
    d.max_displacement = float(max_displacement)
    d.displacement_scale = float(displacement_scale)
    d.noise_sigma = float(noise_sigma)
    d.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    ref = np.empty(shape, np.float32)
    tar = np.empty(shape, np.float32)
    disp = np.empty(tuple(shape) + (3,), np.float32)
    check(lib.skm_synthetic_pair(ctypes.byref(d), _capi.fptr(ref), _capi.fptr(tar), _capi.fptr(disp)))
    return ref, tar, disp


def plain_config(iterations=200, bilateral=1e-5):
    """The plain 3-level registration of the 128x128x96 workload: support
    96 / 48 / 24, control points every support / 4, tensors from the target
    (weights 400 / 200 / 100), NCC + bilateral (spatial sigma 12, intensity
    sigma 0.1), sampling rates 5 / 4 / 3, adaptive steps from 1000."""
    from .evaluator import RegistrationConfig

    return RegistrationConfig(sigma_C4=[96.0, 48.0, 24.0], sigma_C0=[96.0, 48.0, 24.0], tensor_weights=[400, 200, 100],
                              num_scales=3, num_function_evaluations=[iterations], initial_step_size=[1000.0],
                              sampling_rate=[5, 4, 3], neighborhood_sampling=[1], smoothing_variance=[2, 1, 0],
                              reg_bilateral=[bilateral], reg_bilateral_spatial_sigma=[12.0],
                              reg_bilateral_intensity_sigma=[0.1], metric="ncc")


def fine_config(iterations=50, bilateral=1e-5):
    """The last level of plain_config alone (support 24, control points every
    6 voxels, tensor weight 100, sampling rate 3, no smoothing)."""
    from .evaluator import RegistrationConfig

    return RegistrationConfig(sigma_C4=[24.0], sigma_C0=[24.0], tensor_weights=[100], num_scales=1,
                              num_function_evaluations=[iterations], initial_step_size=[1000.0], sampling_rate=[3],
                              neighborhood_sampling=[1], smoothing_variance=[0], reg_bilateral=[bilateral],
                              reg_bilateral_spatial_sigma=[12.0], reg_bilateral_intensity_sigma=[0.1], metric="ncc")


class WarmStart:
    """The warm start of one volume shape.  The network predicts
    coefficients on the fine level's control points (every 6 voxels) in a
    stationary Wendland basis of support 24; their dense field is the
    previous displacement of a fine-level registration (SKM's own way of
    composing levels, a SummationTransform), refined by <= `iterations`
    plain steps and kept only when it evaluates better than none."""

    COEF_SCALE = 10.0   # network outputs are the coefficients times this
    # outputs -> coefficients: one FP32 product by this, on the host and on
    # the device alike (a division by COEF_SCALE rounds differently from
    # torch's, which multiplies by the reciprocal)
    INV_COEF_SCALE = np.float32(1.0 / COEF_SCALE)
    SUPPORT = 24.0
    FIT_ITERATIONS, FIT_RIDGE = 60, 300.0

    def __init__(self, shape=WORKLOAD_SHAPE, max_batch=1, weights=None):
        import torch

        from .evaluator import GpuKernelEvaluator, Image, Kernel, level_control_points

        self._torch = torch
        self.shape = tuple(shape)
        if any(n % 16 for n in self.shape):
            raise ValueError("volume sizes must be multiples of 16 (network input = half, multiple of 8)")
        zero = Image(np.zeros(self.shape, np.float32))
        self.cp = level_control_points(zero, fine_config(), 0)
        self.n_cp = self.cp.count()
        # the stationary basis the coefficients live in
        self.basis = GpuKernelEvaluator(self.n_cp * 3, 1, 1, Kernel(self.SUPPORT, self.SUPPORT), zero, zero, self.cp)
        in_size = tuple(n // 2 for n in reversed(self.shape))
        # input voxel i covers image voxels 2i, 2i + 1: image coordinate p -> (p - 0.5) / 2
        origin = tuple((o - 0.5) / 2.0 for o in self.cp.origin)
        spacing = tuple(h / 2.0 for h in self.cp.spacing)
        self.net = WarmStartNet(in_size, self.cp.size(), origin, spacing, max_batch=max_batch)
        self.in_shape = (in_size[2], in_size[1], in_size[0], 2)
        if weights is not None:
            self.net.load(weights)

    def close(self):
        self.basis.close()
        self.net.close()

    def prepare(self, reference, target, d_input, index=0):
        """Network input sample `index` from host or device images."""
        t = self._torch
        mov = reference if isinstance(reference, t.Tensor) else t.from_numpy(np.ascontiguousarray(reference)).cuda()
        fix = target if isinstance(target, t.Tensor) else t.from_numpy(np.ascontiguousarray(target)).cuda()
        check(self.net._lib.skm_ws_prepare_input(self.net._h, ctypes.c_void_p(mov.data_ptr()),
                                                 ctypes.c_void_p(fix.data_ptr()), int(index),
                                                 ctypes.c_void_p(d_input.data_ptr())))

    def fit(self, displacement):
        """Least-squares coefficients (SoA) of a dense field (shape + (3,)) in the basis."""
        t = self._torch
        d_u = t.from_numpy(np.ascontiguousarray(displacement, np.float32).reshape(-1)).cuda()
        d_c = t.empty(self.n_cp * 3, device="cuda", dtype=t.float32)
        t.cuda.synchronize()  # the fit runs on the evaluator's own stream
        check(self.basis._lib.skm_fit_coefficients_device(self.basis.handle, ctypes.c_void_p(d_u.data_ptr()),
                                                          self.FIT_ITERATIONS, self.FIT_RIDGE,
                                                          ctypes.c_void_p(d_c.data_ptr())))
        return d_c.cpu().numpy()

    def field(self, coefficients_soa):
        """The basis' dense field (shape + (3,)) of SoA coefficients."""
        c = np.asarray(coefficients_soa, np.float32).reshape(3, self.n_cp)
        u = self.basis.evaluateDisplacementField(np.ascontiguousarray(c.T), np.ones(self.n_cp, np.float32))
        return u.reshape(self.shape + (3,))

    def label(self, coefficients_soa):
        """SoA coefficients -> the network's [z][y][x][3] target."""
        nx, ny, nz = self.cp.size()
        c = np.asarray(coefficients_soa, np.float32).reshape(3, nz, ny, nx)
        return np.ascontiguousarray(c.transpose(1, 2, 3, 0)) * self.COEF_SCALE

    def coefficients(self, out):
        """The network's [z][y][x][3] output -> SoA coefficients."""
        nx, ny, nz = self.cp.size()
        o = np.asarray(out, np.float32).reshape(nz, ny, nx, 3)
        return np.ascontiguousarray(o.transpose(3, 0, 1, 2)).reshape(-1) * self.INV_COEF_SCALE

    def predict(self, reference, target):
        x = self._torch.empty((1,) + self.in_shape, device="cuda", dtype=self._torch.float32)
        self.prepare(reference, target, x)
        return self.coefficients(self.net.forward(x).cpu().numpy())

    def predict_field_device(self, reference, target):
        """The predicted coefficients' dense field as a float32 torch tensor
        in device memory (shape + (3,), complete on return), without a host
        round trip: the network's [z][y][x][3] output is the coefficients in
        the control points' AoS order times COEF_SCALE.  For
        register_multiscale / register_batch's initial displacement."""
        t = self._torch
        x = t.empty((1,) + self.in_shape, device="cuda", dtype=t.float32)
        self.prepare(reference, target, x)
        cp = (self.net.forward(x).reshape(-1) * float(self.INV_COEF_SCALE)).contiguous()
        cpw = t.ones(self.n_cp, device="cuda", dtype=t.float32)
        u = t.empty(self.shape + (3,), device="cuda", dtype=t.float32)
        t.cuda.current_stream().synchronize()  # the basis evaluator has a stream of its own
        check(self.basis._lib.skm_evaluate_displacement_field_device(
            self.basis.handle, ctypes.c_void_p(cp.data_ptr()), ctypes.c_void_p(cpw.data_ptr()),
            ctypes.c_void_p(u.data_ptr())))
        check(self.basis._lib.skm_synchronize(self.basis.handle))
        return u

    def register(self, reference, target, iterations=50, bilateral=1e-5, guard=True, coefficients=None):
        """Warm-started registration: the predicted (or given) coefficients'
        field refined by `iterations` steps of the fine level.  Returns
        (displacement, report dict of the level)."""
        from .evaluator import Image, register_multiscale

        u = self.predict_field_device(reference, target) if coefficients is None else self.field(coefficients)
        df, reps = register_multiscale(Image(reference), Image(target), fine_config(iterations, bilateral),
                                       initial_displacement=u, initial_guard=guard)
        return df, reps[0]


def build_dataset(ws, seeds, concurrency=4, chunk=32, plain_iterations=200, bilateral=1e-5, log=print):
    """Network inputs (device tensor [n, z, y, x, 2]) and labels (device
    tensor [n, cz, cy, cx, 3]) for the synthetic pairs of `seeds`: each pair
    registered by the plain 3-level method, its final field fitted onto the
    warm start's basis.  Also returns the plain runs' final NCC."""
    import time

    from .evaluator import Image, register_batch

    torch = ws._torch
    n = len(seeds)
    x = torch.empty((n,) + ws.in_shape, device="cuda", dtype=torch.float32)
    nx, ny, nz = ws.cp.size()
    y = torch.empty((n, nz, ny, nx, 3), device="cuda", dtype=torch.float32)
    nccs = []
    t0 = time.perf_counter()
    for c0 in range(0, n, chunk):
        pairs = []
        for i in range(c0, min(n, c0 + chunk)):
            ref, tar, _ = synthetic_pair(seeds[i], ws.shape)
            pairs.append((Image(ref), Image(tar)))
            ws.prepare(ref, tar, x[i:i + 1])
        res = register_batch(pairs, plain_config(plain_iterations, bilateral), concurrency=concurrency)
        for k, (df, reps) in enumerate(res):
            y[c0 + k] = torch.from_numpy(ws.label(ws.fit(df))).cuda()
            nccs.append(reps[-1]["measure"])
        log(f"labels {min(n, c0 + chunk)}/{n}: {time.perf_counter() - t0:.1f} s, "
            f"mean plain NCC {np.mean(nccs):.4f}")
    return x, y, np.array(nccs)


def train(ws, x, y, epochs=20, batch=8, lr=1e-3, val_fraction=0.05, seed=0, log=print):
    """Adam on the mean squared error between the network's output and the
    labels, cosine learning-rate decay, the last val_fraction held out.
    Returns the per-epoch (train loss, validation loss)."""
    import math

    torch = ws._torch
    n = x.shape[0]
    n_val = max(1, int(round(n * val_fraction))) if n > 1 else 0
    n_tr = n - n_val
    rng = np.random.default_rng(seed)
    steps = epochs * max(1, n_tr // batch)
    step = 0
    hist = []
    for ep in range(epochs):
        perm = rng.permutation(n_tr)
        losses = []
        for b0 in range(0, n_tr - batch + 1, batch):
            idx = torch.from_numpy(perm[b0:b0 + batch]).cuda()
            xb = x.index_select(0, idx).contiguous()
            yb = y.index_select(0, idx).contiguous()
            rate = lr * 0.5 * (1.0 + math.cos(math.pi * step / max(1, steps)))
            losses.append(ws.net.train_step(xb, yb, rate))
            step += 1
        val = float("nan")
        if n_val:
            vl = []
            for b0 in range(n_tr, n, batch):
                xb = x[b0:min(n, b0 + batch)].contiguous()
                out = ws.net.forward(xb).cpu().numpy()
                yt = y[b0:min(n, b0 + batch)].cpu().numpy()
                vl.append(float(np.mean((out.astype(np.float64) - yt) ** 2)) * xb.shape[0])
            val = sum(vl) / n_val
        hist.append((float(np.mean(losses)) if losses else float("nan"), val))
        log(f"epoch {ep + 1}/{epochs}: train {hist[-1][0]:.5f} validation {val:.5f}")
    return hist


def evaluate(ws, seeds, iterations=50, plain_iterations=200, bilateral=1e-5, log=print):
    """Per held-out pair: the plain 3-level registration, the warm-started
    one (prediction + `iterations` fine-level steps, guarded) and the same
    fine-level steps from zero (what the prediction adds).  Wall times
    include the host images in and the dense field out.  Returns a list of
    dicts."""
    import time

    from .evaluator import Image, register_multiscale

    inner = (slice(10, -10),) * 3
    rows = []
    for seed in seeds:
        ref, tar, disp = synthetic_pair(seed, ws.shape)
        t0 = time.perf_counter()
        dfp, rp = register_multiscale(Image(ref), Image(tar), plain_config(plain_iterations, bilateral))
        t_plain = time.perf_counter() - t0
        t0 = time.perf_counter()
        dfw, rw = ws.register(ref, tar, iterations, bilateral)
        t_warm = time.perf_counter() - t0
        dfz, rz = register_multiscale(Image(ref), Image(tar), fine_config(iterations, bilateral))
        row = {"seed": seed, "plain_ms": t_plain * 1e3, "warm_ms": t_warm * 1e3,
               "plain_ncc": rp[-1]["measure"], "warm_ncc": rw["measure"], "zero_ncc": rz[0]["measure"],
               "warm_used": rw["initial_used"], "warm_initial": rw["initial_value"], "zero_initial": rw["zero_value"],
               "plain_err": float(np.abs(dfp[inner] - disp[inner]).mean()),
               "warm_err": float(np.abs(dfw[inner] - disp[inner]).mean()),
               "zero_err": float(np.abs(dfz[inner] - disp[inner]).mean())}
        rows.append(row)
        log(" ".join(f"{k} {v:.4f}" if isinstance(v, float) else f"{k} {v}" for k, v in row.items()))
    return rows
