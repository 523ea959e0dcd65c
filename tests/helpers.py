"""Builds identical inputs for the HIP evaluator (libskm.so) and the FP64 oracle."""
import numpy as np

import oracle as orc
from skmimageregistration_amd import GpuKernelEvaluator, Grid, Image, Kernel, MEASURE, synthetic

REF_STATE = 0xBA5EBA11DEADBEEF


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def scalar_rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def tensor_field(shape, dim, seed, strength=0.5, scale=20.0):
    """Smooth SPD tensor field (D*D per voxel, row-major): scale*(I + strength*n n^T)."""
    rng = np.random.default_rng(seed)
    n = np.prod(shape)
    v = rng.normal(size=(n, dim))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    t = np.einsum("ni,nj->nij", v, v) * strength + np.eye(dim)[None]
    return (t * scale).reshape(tuple(shape) + (dim, dim)).astype(np.float32)


class Case:
    """One registration problem, its cp grid and random coefficients."""

    def __init__(self, dim=2, shape=None, cp_spacing=8.0, sigma=24.0, sigma_c0=None, mode=0, subsample=1,
                 subsample_neighborhood=1, seed=0, coef_scale=2.0, with_df=False, exact=False,
                 origin=None, spacing=None):
        if shape is None:
            shape = (48, 52) if dim == 2 else (20, 22, 24)
        self.dim = dim
        self.shape = tuple(shape)
        rng = np.random.default_rng(seed)
        if dim == 2:
            mov, fix, _ = synthetic.blobs_2d(self.shape, n_blobs=6, max_disp=3.0, seeds=(seed + 1, seed + 2, seed + 3))
        else:
            mov, fix, _ = synthetic.ellipsoids_3d(self.shape, n=8, smooth=1.5, max_disp=2.0,
                                                  seeds=(seed + 1, seed + 2, seed + 3))
        self.moving = mov
        self.fixed = fix
        self.origin = tuple(origin) if origin is not None else (0.0,) * dim
        self.spacing = tuple(spacing) if spacing is not None else (1.0,) * dim
        size = tuple(reversed(self.shape))
        # control-point grid of TransformAdapter::ComputeControlPointImage
        cs, co = [], []
        for d in range(dim):
            lower = self.origin[d]
            upper = self.origin[d] + self.spacing[d] * (size[d] - 1)
            n = int(np.ceil(abs((upper - lower) / cp_spacing)))
            cs.append(n)
            co.append(self.origin[d] + abs(upper - (self.origin[d] + cp_spacing * (n - 1))) / 2.0)
        self.cp_size = tuple(cs)
        self.cp_origin = tuple(co)
        self.cp_spacing = (float(cp_spacing),) * dim
        self.n_cp = int(np.prod(cs))
        self.cp = (rng.uniform(-1, 1, size=(self.n_cp, dim)) * coef_scale).astype(np.float32)
        self.cpw = rng.uniform(0.5, 1.5, size=self.n_cp).astype(np.float32)
        self.sigma_c4 = float(sigma)
        self.sigma_c0 = float(sigma if sigma_c0 is None else sigma_c0)
        self.mode = mode
        self.sub = subsample
        self.subn = subsample_neighborhood
        self.exact = exact
        self.weight = None
        self.max_weight = 1.0
        self.max_eig = np.eye(dim)
        if mode == 1:
            w = rng.uniform(0.6, 1.4, size=self.shape).astype(np.float32)
            w = (w + np.roll(w, 1, 0) + np.roll(w, 1, 1)) / 3.0
            self.weight = w.astype(np.float32)
            self.max_weight = float(self.weight.max())
        elif mode == 2:
            self.weight = tensor_field(self.shape, dim, seed + 7, scale=1.0)
            eig = np.linalg.eigvalsh(self.weight.reshape(-1, dim, dim).astype(np.float64))
            k = int(np.argmax(np.abs(eig[:, -1])))
            self.max_eig = self.weight.reshape(-1, dim, dim)[k].astype(np.float64)
        self.df = None
        if with_df:
            self.df = (rng.normal(size=self.shape + (dim,)) * 0.7).astype(np.float32)

    # ---------------------------------------------------------------- GPU
    def gpu(self, **kw):
        kernel = Kernel(self.sigma_c0, self.sigma_c4,
                        useWeightImage=self.mode == 1, maximumWeight=self.max_weight,
                        wImage=Image(self.weight, self.origin, self.spacing) if self.mode == 1 else None,
                        useWeightTensor=self.mode == 2, maximumEigenvalueCovariance=self.max_eig,
                        wTensor=(Grid(tuple(reversed(self.shape)), self.origin, self.spacing), self.weight)
                        if self.mode == 2 else None)
        ev = GpuKernelEvaluator(self.n_cp * self.dim, self.sub, self.subn, kernel,
                                Image(self.moving, self.origin, self.spacing),
                                Image(self.fixed, self.origin, self.spacing),
                                Grid(self.cp_size, self.cp_origin, self.cp_spacing), exact_gradient=self.exact, **kw)
        return ev

    def df_image(self):
        if self.df is None:
            return None
        return Image(self.df, self.origin, self.spacing, dim=self.dim)

    # ---------------------------------------------------------------- oracle
    def oracle(self, offsets=None, cp=None, parallel=False):
        weight = None
        if self.mode != 0:
            weight = (tuple(reversed(self.shape)), self.origin, self.spacing, self.weight)
        df = None
        if self.df is not None:
            df = (tuple(reversed(self.shape)), self.origin, self.spacing, self.df.reshape(-1, self.dim))
        return orc.OracleProblem(
            self.fixed, self.moving, (self.cp_size, self.cp_origin, self.cp_spacing),
            self.cp if cp is None else cp, self.cpw, sigma_c0=self.sigma_c0, sigma_c4=self.sigma_c4,
            kernel_mode=self.mode, max_weight=self.max_weight, max_eig_cov=self.max_eig, weight=weight, df=df,
            subsample=self.sub, subsample_neighborhood=self.subn, offsets=offsets, exact=self.exact,
            fixed_geom=(self.origin, self.spacing), moving_geom=(self.origin, self.spacing), parallel=parallel)

    def n_cells(self):
        return int(np.prod([s // self.sub for s in self.shape]))

    def offsets(self, calls=1):
        """Offsets of the `calls`-th resampled evaluation from the serial PCG32 stream."""
        state = REF_STATE
        off = None
        for _ in range(calls):
            off, state = orc.cube_offsets(self.n_cells(), self.dim, self.sub, state)
        return off
