"""ctypes wrapper of the FP64 oracle (oracle/_build/libskm_oracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker of the HIP library.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libskm_oracle.so")
# the same source built with OpenMP: the value and derivative loops run in
# parallel (bench.py's CPU baseline); results equal the serial build's
LIB_OMP = os.path.join(HERE, "_build", "libskm_oracle_omp.so")

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int_p = ctypes.POINTER(ctypes.c_int)


class Grid(ctypes.Structure):
    _fields_ = [("size", ctypes.c_int * 3), ("origin", ctypes.c_double * 3), ("spacing", ctypes.c_double * 3)]


class Problem(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("fixed", Grid), ("fixed_data", c_double_p),
        ("moving", Grid), ("moving_coeffs", c_double_p),
        ("cp", Grid), ("cp_aos", c_double_p),
        ("cpw", c_double_p),
        ("sigma_c0", ctypes.c_double), ("sigma_c4", ctypes.c_double),
        ("kernel_mode", ctypes.c_int),
        ("max_weight", ctypes.c_double),
        ("max_eig_cov", ctypes.c_double * 9),
        ("weight", Grid), ("weight_data", c_double_p),
        ("has_df", ctypes.c_int), ("df", Grid), ("df_data", c_double_p),
        ("subsample", ctypes.c_int), ("subsample_neighborhood", ctypes.c_int),
        ("offsets", c_int_p),
        ("exact", ctypes.c_int),
    ]


class Result(ctypes.Structure):
    _fields_ = [("measure", ctypes.c_double), ("reg_rkhs", ctypes.c_double), ("reg_rd", ctypes.c_double),
                ("reg_pg", ctypes.c_double), ("pixels_counted", ctypes.c_longlong)]


_lib = None
_lib_omp = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load_path(path):
    """An oracle build at `path` (a copy of skm_oracle.c compiled elsewhere)."""
    return _declare(ctypes.CDLL(path))


def load(parallel=False):
    """The serial oracle; parallel=True the OpenMP build of the same source."""
    global _lib, _lib_omp
    if parallel:
        if _lib_omp is None:
            if not os.path.exists(LIB_OMP):
                build()
            _lib_omp = _declare(ctypes.CDLL(LIB_OMP))
        return _lib_omp
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _declare(ctypes.CDLL(LIB))
    return _lib


def _declare(lib):
    """Declares the signatures of the oracle's entry points."""
    lib.skmo_wendland_c4.restype = ctypes.c_double
    lib.skmo_wendland_c4.argtypes = [ctypes.c_double]
    lib.skmo_wendland_c0.restype = ctypes.c_double
    lib.skmo_wendland_c0.argtypes = [ctypes.c_double]
    lib.skmo_kernel_eval.restype = ctypes.c_double
    lib.skmo_kernel_eval.argtypes = [ctypes.POINTER(Problem), c_double_p, c_double_p, ctypes.c_double, c_double_p]
    lib.skmo_region_support.restype = ctypes.c_double
    lib.skmo_region_support.argtypes = [ctypes.POINTER(Problem), c_double_p]
    lib.skmo_max_eigenvalue.restype = ctypes.c_double
    lib.skmo_max_eigenvalue.argtypes = [ctypes.c_int, c_double_p]
    lib.skmo_determinant.restype = ctypes.c_double
    lib.skmo_determinant.argtypes = [ctypes.c_int, c_double_p]
    lib.skmo_bspline_coefficients.restype = None
    lib.skmo_bspline_coefficients.argtypes = [ctypes.c_int, c_int_p, c_double_p, c_double_p]
    lib.skmo_bspline_sample.restype = ctypes.c_int
    lib.skmo_bspline_sample.argtypes = [ctypes.c_int, ctypes.POINTER(Grid), c_double_p, c_double_p,
                                        c_double_p, c_double_p]
    lib.skmo_pcg32_next.restype = ctypes.c_uint32
    lib.skmo_pcg32_next.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64]
    lib.skmo_cube_offsets.restype = None
    lib.skmo_cube_offsets.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64, ctypes.c_longlong,
                                      ctypes.c_int, ctypes.c_int, c_int_p]
    lib.skmo_displacement.restype = None
    lib.skmo_displacement.argtypes = [ctypes.POINTER(Problem), c_double_p]
    lib.skmo_evaluate.restype = ctypes.c_int
    lib.skmo_evaluate.argtypes = [ctypes.POINTER(Problem), ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_double, c_double_p, c_double_p,
                                  ctypes.POINTER(Result)]
    lib.skmo_regularizer.restype = None
    lib.skmo_regularizer.argtypes = [ctypes.POINTER(Problem), ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     c_double_p, c_double_p]
    lib.skmo_regularizer_bilateral.restype = None
    lib.skmo_regularizer_bilateral.argtypes = [ctypes.POINTER(Problem), ctypes.c_double, ctypes.c_double,
                                               ctypes.c_double, c_double_p, c_double_p]
    lib.skmo_host_regularizers.restype = ctypes.c_double
    lib.skmo_host_regularizers.argtypes = [ctypes.c_int, ctypes.c_longlong, c_double_p, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_int, c_double_p,
                                           c_double_p, c_double_p, c_double_p]
    lib.skmo_standard_rate.restype = ctypes.c_double
    lib.skmo_standard_rate.argtypes = [ctypes.c_double] * 3 + [ctypes.c_longlong]
    lib.skmo_adaptive_rate.restype = ctypes.c_double
    lib.skmo_adaptive_rate.argtypes = [ctypes.c_double] * 4
    lib.skmo_adaptive_time.restype = ctypes.c_double
    lib.skmo_adaptive_time.argtypes = [ctypes.c_double] * 5
    lib.skmo_gd_update.restype = None
    lib.skmo_gd_update.argtypes = [ctypes.c_longlong, c_double_p, c_double_p, ctypes.c_double, ctypes.c_int,
                                   c_double_p]
    lib.skmo_set_threads.restype = ctypes.c_int
    lib.skmo_set_threads.argtypes = [ctypes.c_int]
    lib.skmo_control_point_grid.restype = None
    lib.skmo_control_point_grid.argtypes = [ctypes.c_int, ctypes.POINTER(Grid), c_double_p, ctypes.POINTER(Grid)]
    return lib


def dp(a):
    return a.ctypes.data_as(c_double_p)


def make_grid(size, origin=None, spacing=None):
    g = Grid()
    dim = len(size)
    for d in range(3):
        g.size[d] = int(size[d]) if d < dim else 1
        g.origin[d] = float(origin[d]) if origin is not None and d < dim else 0.0
        g.spacing[d] = float(spacing[d]) if spacing is not None and d < dim else 1.0
    return g


def bspline_coefficients(image):
    """Padded (n+3)^D coefficient array of a numpy image (x fastest)."""
    lib = load()
    img = np.ascontiguousarray(image, dtype=np.float64)
    dim = img.ndim
    size = (ctypes.c_int * 3)(*(list(reversed(img.shape)) + [1] * (3 - dim)))
    out = np.zeros(tuple(s + 3 for s in img.shape), np.float64)
    lib.skmo_bspline_coefficients(dim, size, dp(img), dp(out))
    return out


def cube_offsets(n_cells, dim, subsample, state=0xBA5EBA11DEADBEEF, sequence=0):
    """Serial PCG32 cube offsets and the advanced state."""
    lib = load()
    st = ctypes.c_uint64(state)
    out = np.zeros((n_cells, dim), np.int32)
    lib.skmo_cube_offsets(ctypes.byref(st), sequence, n_cells, dim, subsample, out.ctypes.data_as(c_int_p))
    return out, st.value


class OracleProblem:
    """Builds an skmo_problem from the same inputs the HIP evaluator takes.

    fixed, moving: numpy images (x fastest) with (origin, spacing); cp_grid:
    (size, origin, spacing); cp_aos (N, D); weight: (size, origin, spacing,
    data) for the weighted kernel modes; df: (size, origin, spacing, data AoS).
    """

    def __init__(self, fixed, moving, cp_grid, cp_aos, cpw=None, sigma_c0=1.0, sigma_c4=1.0, kernel_mode=0,
                 max_weight=1.0, max_eig_cov=None, weight=None, df=None, subsample=1, subsample_neighborhood=1,
                 offsets=None, exact=False, fixed_geom=None, moving_geom=None, parallel=False, lib=None):
        # `lib`: another build of the oracle source (the mutation tests load
        # deliberately broken copies to show that the pins catch them)
        lib = load(parallel) if lib is None else lib
        self.lib = lib
        fixed = np.asarray(fixed)
        dim = fixed.ndim
        self.dim = dim
        fo, fs = fixed_geom if fixed_geom else ((0.0,) * dim, (1.0,) * dim)
        mo, ms = moving_geom if moving_geom else ((0.0,) * dim, (1.0,) * dim)
        self._fixed = np.ascontiguousarray(fixed, np.float64)
        self._coeffs = bspline_coefficients(np.asarray(moving, np.float64))
        csize, corig, cspac = cp_grid
        self.n_cp = int(np.prod(csize))
        self._cp = np.ascontiguousarray(np.asarray(cp_aos, np.float64).reshape(self.n_cp, dim))
        self._cpw = np.ascontiguousarray(np.ones(self.n_cp) if cpw is None else np.asarray(cpw, np.float64))
        p = Problem()
        p.dim = dim
        p.fixed = make_grid(tuple(reversed(fixed.shape)), fo, fs)
        p.fixed_data = dp(self._fixed)
        p.moving = make_grid(tuple(reversed(np.asarray(moving).shape)), mo, ms)
        p.moving_coeffs = dp(self._coeffs)
        p.cp = make_grid(csize, corig, cspac)
        p.cp_aos = dp(self._cp)
        p.cpw = dp(self._cpw)
        p.sigma_c0 = sigma_c0
        p.sigma_c4 = sigma_c4
        p.kernel_mode = kernel_mode
        p.max_weight = max_weight
        m = np.eye(dim) if max_eig_cov is None else np.asarray(max_eig_cov, np.float64).reshape(dim, dim)
        for i in range(dim * dim):
            p.max_eig_cov[i] = m.reshape(-1)[i]
        self._weight = None
        if weight is not None:
            wsize, worig, wspac, wdata = weight
            self._weight = np.ascontiguousarray(np.asarray(wdata, np.float64))
            p.weight = make_grid(wsize, worig, wspac)
            p.weight_data = dp(self._weight)
        self._df = None
        if df is not None:
            dsize, dorig, dspac, ddata = df
            self._df = np.ascontiguousarray(np.asarray(ddata, np.float64))
            p.has_df = 1
            p.df = make_grid(dsize, dorig, dspac)
            p.df_data = dp(self._df)
        p.subsample = subsample
        p.subsample_neighborhood = subsample_neighborhood
        self._offsets = None
        if subsample > 1:
            self._offsets = np.ascontiguousarray(np.asarray(offsets, np.int32))
            p.offsets = self._offsets.ctypes.data_as(c_int_p)
        p.exact = 1 if exact else 0
        self.p = p

    def evaluate(self, measure, with_derivative=True, reg_rkhs=0.0, reg_rd=0.0, reg_pg=0.0, pg_scaling=1.0,
                 per_sample=False):
        """(result dict, derivative SoA); with per_sample also the per-sample
        value array (MSE: M - F, NCC: F, LCC: cc; 0 outside)."""
        der = np.zeros(self.n_cp * self.dim, np.float64)
        n_sub = int(np.prod([s // self.p.subsample for s in self._fixed.shape]))
        samples = np.zeros(n_sub, np.float64)
        r = Result()
        rc = self.lib.skmo_evaluate(ctypes.byref(self.p), int(measure), int(with_derivative), reg_rkhs, reg_rd,
                                    reg_pg, pg_scaling, dp(der), dp(samples), ctypes.byref(r))
        if rc != 0:
            raise RuntimeError("oracle evaluation failed")
        res = dict(measure=r.measure, reg_rkhs=r.reg_rkhs, reg_rd=r.reg_rd, reg_pg=r.reg_pg,
                   pixels_counted=r.pixels_counted)
        if per_sample:
            return res, der, samples
        return res, der

    def bilateral(self, weight, sigma_s, sigma_i, with_derivative=True):
        """(sum of the per-cp bilateral values, weight x derivative SoA)."""
        vals = np.zeros(self.n_cp, np.float64)
        der = np.zeros(self.n_cp * self.dim, np.float64)
        self.lib.skmo_regularizer_bilateral(ctypes.byref(self.p), weight, sigma_s, sigma_i, dp(vals),
                                            dp(der) if with_derivative else None)
        return float(vals.sum()), der

    def displacement(self):
        n = int(np.prod(self._fixed.shape))
        out = np.zeros((n, self.dim), np.float64)
        self.lib.skmo_displacement(ctypes.byref(self.p), dp(out))
        return out

    def kernel(self, p, q, sigma_p=0.0, cov_p=None):
        pa = np.ascontiguousarray(p, np.float64)
        qa = np.ascontiguousarray(q, np.float64)
        cv = np.ascontiguousarray(np.eye(3) if cov_p is None else cov_p, np.float64)
        return self.lib.skmo_kernel_eval(ctypes.byref(self.p), dp(pa), dp(qa), sigma_p, dp(cv))
