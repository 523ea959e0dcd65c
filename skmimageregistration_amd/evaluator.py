"""Host-side mirror of the reference's GpuKernelEvaluator interface over libskm.so.

Names, argument meaning and order follow include/cuda/GpuEvaluator.h:138-178
and include/cuda/Kernel.h:134-146 of the reference; arrays are numpy.  Images
are numpy arrays in C order with x fastest (shape (ny, nx) or (nz, ny, nx)),
geometry tuples are in (x, y[, z]) order.
"""
import ctypes
import enum
import json
import os
import warnings

import numpy as np

from . import _capi
from ._capi import check, f32, fptr, make_grid


class MEASURE(enum.IntEnum):
    """GpuKernelEvaluator::MEASURE (GpuEvaluator.h:160)."""
    MSE = _capi.SKM_MEASURE_MSE
    NCC = _capi.SKM_MEASURE_NCC
    LCC = _capi.SKM_MEASURE_LCC


class Image:
    """An image with geometry: origin + spacing * index (ImageBase.h:25-57)."""

    def __init__(self, data, origin=None, spacing=None, dim=None):
        self.data = np.asarray(data)
        # vector-valued images (a displacement field) pass dim; the trailing
        # axis then holds the components
        self.dim = int(dim) if dim is not None else self.data.ndim
        self.origin = tuple(float(v) for v in (origin if origin is not None else (0.0,) * self.dim))
        self.spacing = tuple(float(v) for v in (spacing if spacing is not None else (1.0,) * self.dim))

    def size(self):
        return tuple(reversed(self.data.shape[: self.dim]))

    def grid(self):
        return make_grid(self.size(), self.origin, self.spacing)


class Grid:
    """Geometry without data (the control-point image of the reference)."""

    def __init__(self, size, origin, spacing):
        self.size_ = tuple(int(v) for v in size)
        self.dim = len(self.size_)
        self.origin = tuple(float(v) for v in origin)
        self.spacing = tuple(float(v) for v in spacing)

    def size(self):
        return self.size_

    def grid(self):
        return make_grid(self.size_, self.origin, self.spacing)

    def count(self):
        return int(np.prod(self.size_))


def image_geometry(size, origin, spacing):
    return Grid(size, origin, spacing)


class Kernel:
    """GKernel construction arguments (Kernel.h:134-146).

    wImage: Image of scalars (useWeightImage); wTensor: (grid Grid, data
    array of shape (..., D, D)) for useWeightTensor.
    """

    def __init__(self, sigmaC0, sigmaC4, useWeightImage=False, maximumWeight=1.0, wImage=None,
                 useWeightTensor=False, maximumEigenvalueCovariance=None, wTensor=None):
        self.sigmaC0 = float(sigmaC0)
        self.sigmaC4 = float(sigmaC4)
        self.useWeightImage = bool(useWeightImage)
        self.maximumWeight = float(maximumWeight)
        self.wImage = wImage
        self.useWeightTensor = bool(useWeightTensor)
        self.maximumEigenvalueCovariance = maximumEigenvalueCovariance
        self.wTensor = wTensor


class EvaluationResult:
    """GpuKernelEvaluator::EvaluationResult (GpuEvaluator.h:138-145)."""

    def __init__(self, r):
        self.measure = r.measure
        self.reg_value_rkhs = r.reg_value_rkhs
        self.reg_value_rd = r.reg_value_rd
        self.reg_value_pg = r.reg_value_pg
        self.reg_value_bilateral = r.reg_value_bilateral
        self.numPixelsCountedGpu = r.num_pixels_counted

    def __repr__(self):
        return (f"EvaluationResult(measure={self.measure!r}, rkhs={self.reg_value_rkhs!r}, "
                f"rd={self.reg_value_rd!r}, pg={self.reg_value_pg!r}, "
                f"numPixelsCountedGpu={self.numPixelsCountedGpu})")


class GpuKernelEvaluator:
    """Drop-in for GpuKernelEvaluator (GpuEvaluator.h:69-179) on libskm.so.

    movingImage holds raw intensities; the cubic B-spline prefilter that the
    reference runs in TransformAdapter::Initialize (itkTransformAdapter.hxx:
    435-439) runs inside the library.  cpImage is a Grid (geometry of the
    control-point image).  exact_gradient selects derivative windows that
    cover the whole kernel support.  LCC centres failing the reference's
    derivative guard |u| > 1e-16 (GpuEvaluatorLCC.cu:517) add to the value
    but not to the derivative in either mode.  moving_coefficients, when
    given, are the moving image's prefiltered cubic B-spline coefficients in
    the reference BsplineImage's layout ((n+3)^D, mirrored borders, shifted by
    one; BsplineImage.h:209-221): they are taken as they are
    (SKM_FLAG_MOVING_COEFFICIENTS) and movingImage gives the geometry only.
    """

    def __init__(self, numParameters, subsample, subsampleNeighborhood, kernel, movingImage, fixedImage,
                 cpImage, cpwImage=None, exact_gradient=False, rng_state=None, rng_sequence=None,
                 moving_coefficients=None):
        self._lib = _capi.load()
        dim = fixedImage.dim
        self.dim = dim
        self.cp_grid = cpImage
        self.n_cp = cpImage.count()
        if numParameters != self.n_cp * dim:
            raise ValueError("numParameters does not match the control-point grid")
        desc = _capi.SkmDesc()
        self._lib.skm_desc_init(ctypes.byref(desc))
        desc.dim = dim
        self._fixed = f32(fixedImage.data)
        flags = _capi.SKM_FLAG_EXACT_GRADIENT if exact_gradient else 0
        if moving_coefficients is not None:
            padded = tuple(n + 3 for n in movingImage.data.shape[:dim])
            self._moving = f32(moving_coefficients)
            if self._moving.shape != padded:
                raise ValueError(f"moving_coefficients must have shape {padded} (image size + 3 per axis)")
            flags |= _capi.SKM_FLAG_MOVING_COEFFICIENTS
        else:
            self._moving = f32(movingImage.data)
        desc.fixed = fixedImage.grid()
        desc.fixed_data = fptr(self._fixed)
        desc.moving = movingImage.grid()
        desc.moving_data = fptr(self._moving)
        desc.cp = cpImage.grid()
        desc.sigma_c0 = kernel.sigmaC0
        desc.sigma_c4 = kernel.sigmaC4
        desc.maximum_weight = kernel.maximumWeight
        self._weight = None
        if kernel.useWeightImage:
            desc.kernel_mode = _capi.SKM_KERNEL_WEIGHT_IMAGE
            self._weight = f32(kernel.wImage.data)
            desc.weight = kernel.wImage.grid()
            desc.weight_data = fptr(self._weight)
        elif kernel.useWeightTensor:
            desc.kernel_mode = _capi.SKM_KERNEL_WEIGHT_TENSOR
            tgrid, tdata = kernel.wTensor
            self._weight = f32(tdata)
            desc.weight = tgrid.grid()
            desc.weight_data = fptr(self._weight)
        if kernel.maximumEigenvalueCovariance is not None:
            m = np.asarray(kernel.maximumEigenvalueCovariance, dtype=np.float64).reshape(dim, dim)
            for i in range(9):
                desc.maximum_eigenvalue_covariance[i] = 0.0
            for i in range(dim):
                for k in range(dim):
                    desc.maximum_eigenvalue_covariance[i * 3 + k] = m[i, k]
        desc.subsample = int(subsample)
        desc.subsample_neighborhood = int(subsampleNeighborhood)
        if rng_state is not None:
            desc.rng_state = int(rng_state)
        if rng_sequence is not None:
            desc.rng_sequence = int(rng_sequence)
        desc.flags = flags
        handle = ctypes.c_void_p()
        check(self._lib.skm_create(ctypes.byref(desc), ctypes.byref(handle)))
        self._h = handle
        self._df_key = None

    def close(self):
        if getattr(self, "_h", None):
            self._lib.skm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # -- GpuEvaluator.h:156-158
    def SetRegularizerRKHS(self, weight):
        check(self._lib.skm_set_regularizer_rkhs(self._h, float(weight)))

    def SetRegularizerRD(self, weight, scaling=1.0):
        check(self._lib.skm_set_regularizer_rd(self._h, float(weight), float(scaling)))

    def SetRegularizerPG(self, weight, scaling=1.0):
        check(self._lib.skm_set_regularizer_pg(self._h, float(weight), float(scaling)))

    def SetRegularizerBilateral(self, weight, spatial_sigma=12.0, intensity_sigma=0.1):
        """skm_set_regularizer_bilateral (no reference counterpart)."""
        check(self._lib.skm_set_regularizer_bilateral(self._h, float(weight), float(spatial_sigma),
                                                      float(intensity_sigma)))

    def SetDisplacementField(self, displacementField):
        """Previous-scale displacement (skm_set_displacement_field): an Image
        with dim set and D components per voxel, or None to clear it."""
        self._set_df(displacementField)

    def _set_df(self, displacementField):
        """The previous-scale field is uploaded when it changes (the reference
        keys its device copy on the host pointer, GpuEvaluator.cu:120-126)."""
        key = None if displacementField is None else id(displacementField.data)
        if key == self._df_key:
            return
        if displacementField is None:
            check(self._lib.skm_set_displacement_field(self._h, None, None))
        else:
            self._df = f32(displacementField.data)
            g = displacementField.grid()
            check(self._lib.skm_set_displacement_field(self._h, ctypes.byref(g), fptr(self._df)))
        self._df_key = key

    def _cp(self, cpData, cpwData):
        cp = f32(cpData).reshape(self.n_cp, self.dim)
        cpw = f32(cpwData if cpwData is not None else np.ones(self.n_cp, np.float32)).reshape(self.n_cp)
        return cp, cpw

    # -- GpuEvaluator.h:162-166
    def getValue(self, metric, cpData, cpwData=None, displacementField=None, do_resampling=True):
        cp, cpw = self._cp(cpData, cpwData)
        self._set_df(displacementField)
        r = _capi.SkmResult()
        check(self._lib.skm_get_value(self._h, int(metric), fptr(cp), fptr(cpw), int(bool(do_resampling)),
                                      ctypes.byref(r)))
        return EvaluationResult(r)

    # -- GpuEvaluator.h:168-173; returns (result, derivative SoA)
    def getValueAndDerivative(self, metric, cpData, cpwData=None, displacementField=None, derivatives=None,
                              do_resampling=True):
        cp, cpw = self._cp(cpData, cpwData)
        self._set_df(displacementField)
        out = derivatives if derivatives is not None else np.empty(self.n_cp * self.dim, np.float32)
        if out.dtype != np.float32 or not out.flags["C_CONTIGUOUS"] or out.size != self.n_cp * self.dim:
            raise ValueError("derivatives must be a contiguous float32 array of numParameters values")
        r = _capi.SkmResult()
        check(self._lib.skm_get_value_and_derivative(self._h, int(metric), fptr(cp), fptr(cpw), fptr(out),
                                                     int(bool(do_resampling)), ctypes.byref(r)))
        return EvaluationResult(r), out

    # -- GpuEvaluator.h:175-178; returns the AoS field (N_vox, D)
    def evaluateDisplacementField(self, cpData, cpwData=None, displacementField=None):
        cp, cpw = self._cp(cpData, cpwData)
        self._set_df(displacementField)
        n = int(np.prod(self._fixed.shape))
        dst = np.empty((n, self.dim), np.float32)
        check(self._lib.skm_evaluate_displacement_field(self._h, fptr(cp), fptr(cpw), fptr(dst)))
        return dst

    def enable_timing(self, enable=True):
        check(self._lib.skm_enable_timing(self._h, int(bool(enable))))

    def read_timing(self):
        """{pass name: (ms, launches)} accumulated since the last read."""
        t = _capi.SkmPassTimes()
        check(self._lib.skm_read_timing(self._h, ctypes.byref(t)))
        return {name: (t.ms[i], t.launches[i]) for i, name in enumerate(_capi.PASS_NAMES)}

    def set_stream(self, stream_handle):
        check(self._lib.skm_set_stream(self._h, ctypes.c_void_p(stream_handle)))

    def get_value_and_derivative_device(self, metric, d_cp_aos, d_cpw, d_derivative, do_resampling=True,
                                        d_result=None):
        """Device pointers (ints); enqueued on the evaluator's stream."""
        check(self._lib.skm_get_value_and_derivative_device(
            self._h, int(metric), ctypes.c_void_p(d_cp_aos), ctypes.c_void_p(d_cpw), ctypes.c_void_p(d_derivative),
            int(bool(do_resampling)), ctypes.c_void_p(d_result) if d_result else None))

    def synchronize(self):
        check(self._lib.skm_synchronize(self._h))

    def set_tuning(self, lcc_ref_chunks=0, gather_split=0, value_split=0, gather_regions=0, lcc_ref_walk=0,
                   kernel_matrix=0):
        """Launch shapes (skm_set_tuning); 0 keeps the library's choice.
        kernel_matrix at subsample 1: 2 reads the assembled kernel matrix, 1
        evaluates the kernels on the fly, 0 the former for the tensor kernel."""
        t = _capi.SkmTuning(int(lcc_ref_chunks), int(gather_split), int(value_split), int(gather_regions),
                            int(lcc_ref_walk), int(kernel_matrix))
        check(self._lib.skm_set_tuning(self._h, ctypes.byref(t)))

    def kernel_matrix_info(self):
        """skm_get_kernel_matrix_info as a dict (state 0 none, 1 assembled, 2 refused)."""
        i = _capi.SkmKernelMatrixInfo()
        check(self._lib.skm_get_kernel_matrix_info(self._h, ctypes.byref(i)))
        return dict(state=i.state, bytes=i.bytes, pairs_value=i.pairs_value, pairs_gather=i.pairs_gather,
                    read_value=i.read_value, read_gather=i.read_gather, build_ms=i.build_ms)

    def release_kernel_matrix(self):
        check(self._lib.skm_release_kernel_matrix(self._h))

    def autotune_device(self, metric, d_cp_aos, d_cpw, repeats=3):
        """skm_autotune at device control points (ints); returns the report as a dict."""
        r = _capi.SkmAutotuneReport()
        check(self._lib.skm_autotune(self._h, int(metric), ctypes.c_void_p(d_cp_aos), ctypes.c_void_p(d_cpw),
                                     int(repeats), ctypes.byref(r)))
        b = r.best
        return {"lcc_ref_chunks": b.lcc_ref_chunks, "gather_split": b.gather_split, "value_split": b.value_split,
                "gather_regions": b.gather_regions, "lcc_ref_walk": b.lcc_ref_walk,
                "kernel_matrix": b.kernel_matrix, "best_ms": r.best_ms, "first_ms": r.first_ms, "candidates": r.candidates}

    def get_tuning(self):
        t = _capi.SkmTuning()
        check(self._lib.skm_get_tuning(self._h, ctypes.byref(t)))
        return dict(lcc_ref_chunks=t.lcc_ref_chunks, gather_split=t.gather_split, value_split=t.value_split,
                    gather_regions=t.gather_regions, lcc_ref_walk=t.lcc_ref_walk, kernel_matrix=t.kernel_matrix)

    def cube_offsets(self):
        n = ctypes.c_int64()
        check(self._lib.skm_get_cube_offsets(self._h, None, ctypes.byref(n)))
        out = np.empty((n.value, self.dim), np.int32)
        check(self._lib.skm_get_cube_offsets(self._h, out.ctypes.data_as(_capi.c_int32_p), None))
        return out

    # -- device-resident optimisation loop
    def optimize(self, params_soa, metric=MEASURE.MSE, reg_l1=0.0, reg_l21=0.0, reg_l2=0.0,
                 use_negative_gradient=True, do_resampling=True, kind=_capi.SKM_OPT_ADAPTIVE, a=0.1, A=20.0,
                 alpha=1.0, fmax=1e7, fmin=-1e6, omega=1.0, orthant_projection=False, iterations=100,
                 cpwData=None, return_best=False, graph=False):
        params = f32(params_soa).copy()
        cpw = f32(cpwData if cpwData is not None else np.ones(self.n_cp, np.float32))
        mo, oo = _optimizer_options(metric, reg_l1, reg_l21, reg_l2, use_negative_gradient, do_resampling, kind, a,
                                    A, alpha, fmax, fmin, omega, orthant_projection, iterations, return_best, graph)
        recs = (_capi.SkmIterationRecord * max(1, iterations))()
        check(self._lib.skm_optimize(self._h, ctypes.byref(mo), ctypes.byref(oo), fptr(cpw), fptr(params), recs))
        return params, _records(recs, iterations)


def _optimizer_options(metric, reg_l1, reg_l21, reg_l2, use_negative_gradient, do_resampling, kind, a, A, alpha,
                       fmax, fmin, omega, orthant_projection, iterations, return_best, graph):
    mo = _capi.SkmMetricOptions(int(metric), reg_l1, reg_l21, reg_l2, int(bool(use_negative_gradient)),
                                int(bool(do_resampling)))
    oo = _capi.SkmOptimizerOptions(int(kind), a, A, alpha, fmax, fmin, omega, int(bool(orthant_projection)),
                                   int(iterations), int(bool(return_best)), int(bool(graph)))
    return mo, oo


def _records(recs, iterations):
    return [(r.value, r.measure, r.learning_rate, r.num_pixels_counted) for r in recs[:iterations]]


def optimize_batch(evaluators, params_soa, metric=MEASURE.MSE, reg_l1=0.0, reg_l21=0.0, reg_l2=0.0,
                   use_negative_gradient=True, do_resampling=True, kind=_capi.SKM_OPT_ADAPTIVE, a=0.1, A=20.0,
                   alpha=1.0, fmax=1e7, fmin=-1e6, omega=1.0, orthant_projection=False, iterations=100,
                   cpwData=None, return_best=False, graph=False):
    """GpuKernelEvaluator.optimize of several evaluators at once through
    skm_optimize_batch (no counterpart in the reference): with graph=True
    every iteration after the first is one launch of one captured graph
    holding an optimizer step of every evaluator.  params_soa and cpwData
    (optional) are per-evaluator sequences; returns [(params, records)]."""
    evaluators = list(evaluators)
    n = len(evaluators)
    if n == 0:
        return []
    lib = evaluators[0]._lib
    params = [f32(p).copy() for p in params_soa]
    cpws = [f32(c) for c in cpwData] if cpwData is not None else [np.ones(ev.n_cp, np.float32) for ev in evaluators]
    if len(params) != n or len(cpws) != n:
        raise ValueError("one start point (and cpw) per evaluator")
    mo, oo = _optimizer_options(metric, reg_l1, reg_l21, reg_l2, use_negative_gradient, do_resampling, kind, a, A,
                                alpha, fmax, fmin, omega, orthant_projection, iterations, return_best, graph)
    recs = [(_capi.SkmIterationRecord * max(1, iterations))() for _ in range(n)]
    hs = (ctypes.c_void_p * n)(*[ev._h for ev in evaluators])
    cps = (_capi.c_float_p * n)(*[fptr(c) for c in cpws])
    pps = (_capi.c_float_p * n)(*[fptr(p) for p in params])
    rps = (ctypes.POINTER(_capi.SkmIterationRecord) * n)(
        *[ctypes.cast(r, ctypes.POINTER(_capi.SkmIterationRecord)) for r in recs])
    check(lib.skm_optimize_batch(hs, n, ctypes.byref(mo), ctypes.byref(oo), cps, pps, rps))
    return [(p, _records(r, iterations)) for p, r in zip(params, recs)]


def tuning_table_save(path):
    """Writes the process's launch-shape table (skm_autotune results) to path."""
    check(_capi.load().skm_tuning_table_save(str(path).encode()))


def tuning_table_load(path):
    """Merges a saved launch-shape table into the process's; evaluators
    created afterwards with a listed shape class start with its shapes."""
    check(_capi.load().skm_tuning_table_load(str(path).encode()))


def set_cache_limit(idle_bytes_per_device):
    """Idle device bytes the library keeps per device for later evaluators
    (skm_set_cache_limit; 0 disables the cache)."""
    check(_capi.load().skm_set_cache_limit(int(idle_bytes_per_device)))


def cache_info():
    """{"idle": bytes kept idle, "live": bytes in use, "limit": idle bytes per device}."""
    idle, live, limit = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    check(_capi.load().skm_get_cache_info(ctypes.byref(idle), ctypes.byref(live), ctypes.byref(limit)))
    return {"idle": idle.value, "live": live.value, "limit": limit.value}


def release_cached_memory():
    """Frees every idle block of the library's device cache (skm_release_cached_memory)."""
    check(_capi.load().skm_release_cached_memory())


def control_point_grid(image, spacing):
    """TransformAdapter::ComputeControlPointImage geometry (itkTransformAdapter.hxx:246-308)."""
    lib = _capi.load()
    dim = image.dim
    g_in = image.grid()
    sp = (ctypes.c_double * 3)(*([float(s) for s in spacing] + [1.0] * (3 - len(spacing))))
    out = _capi.SkmGrid()
    check(lib.skm_control_point_grid(dim, ctypes.byref(g_in), sp, ctypes.byref(out)))
    return Grid(tuple(out.size[:dim]), tuple(out.origin[:dim]), tuple(out.spacing[:dim]))


def GaussianSmoothing(image, variance=1.0):
    """GaussianSmoothing(image, variance) (include/itkUtils.h:140-151) on the GPU:
    discrete Gaussian, variance in physical units, edge replication.  Returns
    an Image with the same geometry."""
    lib = _capi.load()
    src = f32(image.data)
    out = np.empty_like(src)
    g = image.grid()
    check(lib.skm_gaussian_smooth(image.dim, ctypes.byref(g), fptr(src), float(variance), fptr(out)))
    return Image(out, image.origin, image.spacing)


def ComputeAnisotropicTensor(image, sigma, weight, alignment_direction=2, anisotropic_filtering=False):
    """ComputeAnisotropicTensor (include/AnisotropicTensorUtils.h:37-355) with the
    maximum-tilt handling of TransformAdapter::SetWeightingTensor
    (itkTransformAdapter.hxx:335-366), on the GPU.

    Returns (tensor, maximum_tilt, maximum_weight, maximum_eigenvalue_covariance):
    tensor has shape image.data.shape + (D, D) on the image grid (the wTensor
    data of Kernel), maximum_tilt is the value the reference returns (the
    lowest double when no voxel is re-weighted), and the last two are what
    Kernel(maximumWeight=..., maximumEigenvalueCovariance=...) takes."""
    lib = _capi.load()
    dim = image.dim
    src = f32(image.data)
    out = np.empty(src.shape + (dim, dim), dtype=np.float32)
    g = image.grid()
    opt = _capi.SkmTensorOptions(float(sigma), float(weight), int(alignment_direction), int(bool(anisotropic_filtering)))
    tilt, mw = ctypes.c_double(), ctypes.c_double()
    cov = (ctypes.c_double * 9)()
    check(lib.skm_anisotropic_tensor(dim, ctypes.byref(g), fptr(src), ctypes.byref(opt), fptr(out),
                                     ctypes.byref(tilt), ctypes.byref(mw), cov))
    mcov = np.array(cov[:], dtype=np.float64).reshape(3, 3)[:dim, :dim].copy()
    return out, tilt.value, mw.value, mcov


# ------------------------------------------------------------ files and warped images

def read_image(path):
    """itk::ImageFileReader (include/IOUtils.h:27-37) for VTK legacy
    (.vtk) and MetaImage (.mhd, .mha) files.  Returns an Image; vector and
    tensor images carry their components on trailing axes ((D,) or (D, D))."""
    lib = _capi.load()
    h = _capi.SkmImageHeader()
    check(lib.skm_read_image_header(os.fsencode(path), ctypes.byref(h)))
    dim, nc = h.dim, h.components
    size = [h.grid.size[d] for d in range(dim)]
    n = int(np.prod(size)) * nc
    buf = np.empty(max(n, 1), np.float32)
    check(lib.skm_read_image(os.fsencode(path), fptr(buf), n, None))
    trailing = () if nc == 1 else ((dim, dim) if nc == dim * dim and dim != 1 else (nc,))
    data = buf[:n].reshape(tuple(reversed(size)) + trailing)
    return Image(data, tuple(h.grid.origin[:dim]), tuple(h.grid.spacing[:dim]), dim=dim)


def write_image(path, image):
    """itk::ImageFileWriter (include/IOUtils.h:39-52): .vtk (big-endian
    binary float), .mhd (+ .raw) or .mha."""
    lib = _capi.load()
    data = f32(image.data)
    h = _capi.SkmImageHeader()
    h.dim = image.dim
    h.grid = image.grid()
    h.components = int(np.prod(data.shape[image.dim:])) if data.ndim > image.dim else 1
    check(lib.skm_write_image(os.fsencode(path), ctypes.byref(h), fptr(data)))


def warp_image(image, displacement, interpolation="linear", out_grid=None, default_value=0.0):
    """image(x + u(x)) on the GPU for every voxel x of out_grid (default: the
    displacement's grid); default_value outside the image.  "linear": SKMReg's
    final resampling (src/SKMReg.cxx:507-525); "bspline3": WarpImage
    (include/itkUtils.h:163-179), out_grid must be the displacement's grid.
    displacement is an Image with dim=D and a trailing axis of D components."""
    lib = _capi.load()
    kind = {"linear": _capi.SKM_INTERPOLATION_LINEAR, "bspline3": _capi.SKM_INTERPOLATION_BSPLINE3}[interpolation]
    dim = image.dim
    img = f32(image.data)
    df = f32(displacement.data)
    gi, gd = image.grid(), displacement.grid()
    og = out_grid if out_grid is not None else Grid(displacement.size(), displacement.origin, displacement.spacing)
    go = og.grid()
    out = np.empty(tuple(reversed(og.size())), np.float32)
    check(lib.skm_warp_image(dim, ctypes.byref(gi), fptr(img), ctypes.byref(gd), fptr(df), ctypes.byref(go), kind,
                             float(default_value), fptr(out)))
    return Image(out, og.origin, og.spacing)


def read_landmarks(path, dim):
    """ReadLandmarks (include/IOUtils.h:65-107): an (n, dim) float64 array;
    warns, as the reference does, when n differs from the POINTS count."""
    lib = _capi.load()
    n = ctypes.c_int64()
    declared = ctypes.c_int64()
    check(lib.skm_read_landmarks(os.fsencode(path), dim, None, 0, ctypes.byref(n), ctypes.byref(declared)))
    pts = np.zeros((max(n.value, 1), dim), np.float64)
    check(lib.skm_read_landmarks(os.fsencode(path), dim, pts.ctypes.data_as(_capi.c_double_p), n.value,
                                 ctypes.byref(n), ctypes.byref(declared)))
    if n.value != declared.value:
        warnings.warn("number of points read does not correspond to the number of points declared in the header")
    return pts[: n.value]


def landmark_tre(displacement, fixed_points, moving_points):
    """CalculateTRE (include/itkUtils.h:268-293): mean |m + u(m) - f|, -1
    without landmarks."""
    lib = _capi.load()
    dim = displacement.dim
    f = np.ascontiguousarray(fixed_points, np.float64).reshape(-1, dim)
    m = np.ascontiguousarray(moving_points, np.float64).reshape(-1, dim)
    if len(f) != len(m):
        return -1.0
    df = f32(displacement.data)
    g = displacement.grid()
    tre = ctypes.c_double()
    check(lib.skm_landmark_tre(dim, ctypes.byref(g), fptr(df), f.ctypes.data_as(_capi.c_double_p),
                               m.ctypes.data_as(_capi.c_double_p), len(f), ctypes.byref(tre)))
    return tre.value


class RegistrationConfig:
    """The SKMReg parameter set (include/ConfigUtils.h:29-62): per-scale lists
    whose last entry repeats (CURRENT, src/SKMReg.cxx:48-49).  One key has no
    SKMReg counterpart: exact_gradient (1: derivative windows covering the
    whole kernel support at every level; 0, the default: the reference's),
    the bilateral regularizer reg_bilateral with its
    reg_bilateral_spatial_sigma / reg_bilateral_intensity_sigma, and
    autotune (1: each level's launch shapes from skm_autotune, once per
    shape class and process)."""

    STRING_KEYS = ("metric", "reference_filename", "target_filename", "reference_landmarks", "target_landmarks",
                   "weight_filename", "weight_tensor_filename", "temp_directory")

    DEFAULTS = {
        "reference_filename": "/tmp/ref.vtk", "target_filename": "/tmp/tar.vtk", "reference_landmarks": "none",
        "target_landmarks": "none", "weight_filename": "none", "weight_tensor_filename": "none",
        "temp_directory": "/tmp/skm/", "verbosity": [1],
        "sigma_C4": [64, 32, 16], "sigma_C0": [128, 64, 32], "cp_spacing_scaler": [4.0],
        "tensor_weights": [400, 200, 100], "reg_l1": [0, 0, 0], "reg_l21": [0, 0, 0], "reg_l2": [0, 0, 0],
        "reg_rkhs": [0, 0, 0], "reg_rd": [0, 0, 0], "reg_rd_scaling": [1, 1, 1], "reg_pg": [0, 0, 0],
        "reg_pg_scaling": [1, 1, 1], "num_scales": [3], "restart_at": [0],
        "num_function_evaluations": [500, 500, 500], "initial_step_size": [0.1, 0.1, 0.1],
        "sampling_rate": [5, 4, 3], "neighborhood_sampling": [15, 12, 9], "intensity_thresholds": [-1500, 3500],
        "smoothing_variance": [2, 1, 0], "metric": "mse", "exact_gradient": [0],
        "reg_bilateral": [0, 0, 0], "reg_bilateral_spatial_sigma": [12.0], "reg_bilateral_intensity_sigma": [0.1],
        "autotune": [0],
    }

    def __init__(self, **overrides):
        self.values = {k: (list(v) if isinstance(v, list) else v) for k, v in self.DEFAULTS.items()}
        for k, v in overrides.items():
            if k not in self.values:
                raise KeyError(f"unknown option {k}")
            if k in self.STRING_KEYS:
                self.values[k] = str(v)
            else:
                self.values[k] = list(v) if isinstance(v, (list, tuple)) else [v]

    @classmethod
    def from_json(cls, path):
        """load_json_config(filename) over the defaults (include/ConfigUtils.h:64-75,
        src/SKMReg.cxx:528-560): keys of the file replace the defaults."""
        with open(path) as f:
            return cls(**json.load(f))

    def CURRENT(self, key, scale):
        v = self.values[key]
        return v[min(scale, len(v) - 1)]

    def __getitem__(self, key):
        return self.values[key]


def _registration_desc(lib, reference, target, config, weight_image, weight_tensor, guidance_image,
                       restart_displacement, temp_directory, reference_landmarks, target_landmarks):
    """skm_registration_desc of one registration, the arrays it points to
    (kept alive by the caller), the displacement output and the reports."""
    cfg = config or RegistrationConfig()
    dim = target.dim
    d = _capi.SkmRegistrationDesc()
    lib.skm_registration_desc_init(ctypes.byref(d))
    d.dim = dim
    ref = f32(reference.data)
    tar = f32(target.data)
    d.reference = reference.grid()
    d.reference_data = fptr(ref)
    d.target = target.grid()
    d.target_data = fptr(tar)
    n = int(cfg.CURRENT("num_scales", 0))
    if not 1 <= n <= _capi.SKM_MAX_SCALES:
        raise ValueError("num_scales out of range")
    d.num_scales = n
    for s in range(n):
        p = d.scale[s]
        p.sigma_c4 = cfg.CURRENT("sigma_C4", s)
        p.sigma_c0 = cfg.CURRENT("sigma_C0", s)
        p.tensor_weight = cfg.CURRENT("tensor_weights", s)
        for key in ("reg_l1", "reg_l21", "reg_l2", "reg_rkhs", "reg_rd", "reg_rd_scaling", "reg_pg",
                    "reg_pg_scaling", "reg_bilateral", "reg_bilateral_spatial_sigma",
                    "reg_bilateral_intensity_sigma"):
            setattr(p, key, float(cfg.CURRENT(key, s)))
        p.iterations = int(cfg.CURRENT("num_function_evaluations", s))
        p.initial_step_size = float(cfg.CURRENT("initial_step_size", s))
        p.sampling_rate = int(cfg.CURRENT("sampling_rate", s))
        p.neighborhood_sampling = int(cfg.CURRENT("neighborhood_sampling", s))
        p.smoothing_variance = float(cfg.CURRENT("smoothing_variance", s))
    d.cp_spacing_scaler = float(cfg.CURRENT("cp_spacing_scaler", 0))
    d.exact_gradient = int(bool(cfg.CURRENT("exact_gradient", 0)))
    d.autotune = int(bool(cfg.CURRENT("autotune", 0)))
    d.metric = {"mse": MEASURE.MSE, "ncc": MEASURE.NCC, "lcc": MEASURE.LCC}[cfg["metric"]]
    th = cfg["intensity_thresholds"]
    d.use_intensity_thresholds = int(len(th) == 2)
    if len(th) == 2:
        d.threshold_min, d.threshold_max = float(th[0]), float(th[1])
    keep = []
    for name, img in (("weight_image", weight_image), ("weight_tensor", weight_tensor),
                      ("guidance_image", guidance_image)):
        if img is not None:
            arr = f32(img.data)
            keep.append(arr)
            d.weight = img.grid()
            setattr(d, name, fptr(arr))
    d.restart_at = int(cfg.CURRENT("restart_at", 0))
    if restart_displacement is not None:
        rd = f32(restart_displacement)
        keep.append(rd)
        d.restart_displacement = fptr(rd)
    if temp_directory is not None:
        d.temp_directory = os.fsencode(temp_directory)
    d.verbosity = int(cfg.CURRENT("verbosity", 0))
    if reference_landmarks is not None and target_landmarks is not None:
        rl = np.ascontiguousarray(reference_landmarks, np.float64).reshape(-1, dim)
        tl = np.ascontiguousarray(target_landmarks, np.float64).reshape(-1, dim)
        if len(rl) == len(tl) and len(rl) > 0:
            keep += [rl, tl]
            d.num_landmarks = len(rl)
            d.reference_landmarks = rl.ctypes.data_as(_capi.c_double_p)
            d.target_landmarks = tl.ctypes.data_as(_capi.c_double_p)
    keep += [ref, tar]
    out = np.empty(tar.shape + (dim,), np.float32)
    reps = (_capi.SkmScaleReport * n)()
    return d, keep, out, reps


def _reports(reps, dim):
    return [{"value": r.value, "measure": r.measure, "first_value": r.first_value, "iterations": r.iterations,
             "cp_size": tuple(r.cp_size[:dim]), "maximum_tilt": r.maximum_tilt, "ms_setup": r.ms_setup,
             "ms_optimize": r.ms_optimize, "ms_displacement": r.ms_displacement, "tre": r.tre,
             "zero_parameters": r.zero_parameters, "initial_used": bool(r.initial_used),
             "initial_value": r.initial_value, "zero_value": r.zero_value} for r in reps]


def level_control_points(reference, config, scale):
    """The control-point grid skm_register builds for a level
    (SKMReg.cxx:219-232): spacing min(sigma_C4, sigma_C0) / cp_spacing_scaler."""
    cfg = config or RegistrationConfig()
    smin = min(cfg.CURRENT("sigma_C4", scale), cfg.CURRENT("sigma_C0", scale))
    scaler = float(cfg.CURRENT("cp_spacing_scaler", 0))
    h = smin if scaler < 1 else smin / scaler
    return control_point_grid(reference, (h,) * reference.dim)


def _attach_warm_start(d, keep, reference, target, config, initial_coefficients, initial_displacement,
                       initial_guard):
    """Warm-start fields of a registration descriptor (no SKMReg counterpart)."""
    cfg = config or RegistrationConfig()
    if initial_coefficients is not None:
        grid = level_control_points(reference, cfg, max(0, int(cfg.CURRENT("restart_at", 0)) - 1))
        c = f32(initial_coefficients).reshape(-1)
        if c.size != grid.count() * reference.dim:
            raise ValueError(f"initial coefficients: expected {grid.count() * reference.dim}, got {c.size}")
        keep.append(c)
        d.initial_coefficients = fptr(c)
    if initial_displacement is not None and getattr(initial_displacement, "is_cuda", False):
        # a torch tensor in device memory: handed over without a host round trip
        u = initial_displacement
        if str(u.dtype) != "torch.float32" or not u.is_contiguous():
            raise ValueError("initial displacement: a contiguous float32 tensor expected")
        if u.numel() != target.data.size * target.dim:
            raise ValueError("initial displacement: D floats per target voxel expected")
        keep.append(u)
        d.initial_displacement = ctypes.cast(ctypes.c_void_p(u.data_ptr()), _capi.c_float_p)
        d.initial_displacement_on_device = 1
    elif initial_displacement is not None:
        u = f32(initial_displacement).reshape(-1)
        if u.size != target.data.size * target.dim:
            raise ValueError("initial displacement: D floats per target voxel expected")
        keep.append(u)
        d.initial_displacement = fptr(u)
    d.initial_guard = int(bool(initial_guard))


def register_multiscale(reference, target, config=None, weight_image=None, weight_tensor=None, guidance_image=None,
                        restart_displacement=None, temp_directory=None, reference_landmarks=None,
                        target_landmarks=None, initial_coefficients=None, initial_displacement=None,
                        initial_guard=False):
    """The scale-level loop of SKMReg (src/SKMReg.cxx:84-525) through
    skm_register.  reference (moving) and target (fixed) are Images; the
    optional weight image / tensor / guidance image are Images on one common
    weight grid.  With temp_directory, SKMReg's result files are written
    there and restart_at without restart_displacement resumes from its
    df-full file.  Landmarks ((n, D) arrays) give the TRE per level.
    Returns (displacement on the target grid, shape target.data.shape + (D,),
    list of per-level report dicts).

    Warm start (no SKMReg counterpart): initial_coefficients (SoA on the
    first level's control points) start that level's optimizer;
    initial_displacement (the target grid's shape + (D,), a host array or a
    float32 torch tensor in device memory, which is copied on the device) is
    the first level run's previous displacement; with initial_guard either is
    used only when it evaluates better than none."""
    lib = _capi.load()
    d, keep, out, reps = _registration_desc(lib, reference, target, config, weight_image, weight_tensor,
                                            guidance_image, restart_displacement, temp_directory,
                                            reference_landmarks, target_landmarks)
    _attach_warm_start(d, keep, reference, target, config, initial_coefficients, initial_displacement,
                       initial_guard)
    check(lib.skm_register(ctypes.byref(d), fptr(out), reps))
    return out, _reports(reps, target.dim)


def register_batch(pairs, config=None, concurrency=4, initial_displacements=None, initial_guard=False,
                   lockstep=False):
    """Independent registrations through skm_register_batch (no counterpart
    in the reference): pairs is a sequence of (reference, target) Images, all
    registered with one configuration, up to `concurrency` at once on the
    current device.  Each result equals register_multiscale's for its pair.
    Returns a list of (displacement, per-level report dicts);
    initial_displacements is an optional per-pair sequence (see
    register_multiscale).  lockstep=True runs skm_register_batch_lockstep
    instead: all pairs in flight, each optimizer step of the whole batch one
    captured-graph launch (concurrency is then ignored)."""
    lib = _capi.load()
    jobs = [_registration_desc(lib, r, t, config, None, None, None, None, None, None, None) for r, t in pairs]
    if initial_displacements is not None:
        for (d, keep, _, _), (r, t), u in zip(jobs, pairs, initial_displacements):
            _attach_warm_start(d, keep, r, t, config, None, u, initial_guard)
    n = len(jobs)
    descs = (_capi.SkmRegistrationDesc * max(n, 1))()
    for i, (d, _, _, _) in enumerate(jobs):
        descs[i] = d
    outs = (_capi.c_float_p * max(n, 1))(*[fptr(o) for _, _, o, _ in jobs])
    reps = (ctypes.POINTER(_capi.SkmScaleReport) * max(n, 1))(
        *[ctypes.cast(r, ctypes.POINTER(_capi.SkmScaleReport)) for _, _, _, r in jobs])
    if lockstep:
        check(lib.skm_register_batch_lockstep(descs, n, outs, reps, None))
    else:
        check(lib.skm_register_batch(descs, n, int(concurrency), outs, reps, None))
    return [(o, _reports(r, t.dim)) for (_, _, o, r), (_, t) in zip(jobs, pairs)]


def register_files(config):
    """SKMReg's run on files (src/SKMReg.cxx:55-526): reads reference_filename,
    target_filename, the optional weight image / weight tensor (or guidance
    image when tensor_weights > 0) and landmark files named in the config,
    registers, and writes the result files into temp_directory (created if
    missing).  Returns what register_multiscale returns."""
    cfg = config if isinstance(config, RegistrationConfig) else RegistrationConfig(**dict(config))
    reference = read_image(cfg["reference_filename"])
    target = read_image(cfg["target_filename"])
    dim = target.dim
    weight_image = weight_tensor = guidance = None
    if cfg["weight_filename"] != "none":
        weight_image = read_image(cfg["weight_filename"])
    if cfg["weight_tensor_filename"] != "none":
        img = read_image(cfg["weight_tensor_filename"])
        if img.data.ndim > img.dim:
            weight_tensor = img
        else:
            guidance = img
    rl = tl = None
    if cfg["reference_landmarks"] != "none" and cfg["target_landmarks"] != "none":
        rl = read_landmarks(cfg["reference_landmarks"], dim)
        tl = read_landmarks(cfg["target_landmarks"], dim)
    os.makedirs(cfg["temp_directory"], exist_ok=True)
    return register_multiscale(reference, target, cfg, weight_image=weight_image, weight_tensor=weight_tensor,
                               guidance_image=guidance, temp_directory=cfg["temp_directory"],
                               reference_landmarks=rl, target_landmarks=tl)
