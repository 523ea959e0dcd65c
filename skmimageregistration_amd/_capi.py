"""ctypes declarations of the C ABI in include/skm.h (libskm.so).

Loading fails loudly when the in-tree library is missing: there is no CPU or
PyTorch fallback for any entry point.
"""
import ctypes
import os

import numpy as np

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG_DIR, "libskm.so")

SKM_SUCCESS = 0
SKM_ERROR_INVALID_ARGUMENT, SKM_ERROR_HIP, SKM_ERROR_OUT_OF_MEMORY, SKM_ERROR_NOT_SUPPORTED, SKM_ERROR_IO = 1, 2, 3, 4, 5
SKM_INTERPOLATION_LINEAR, SKM_INTERPOLATION_BSPLINE3 = 1, 3
SKM_MEASURE_MSE, SKM_MEASURE_NCC, SKM_MEASURE_LCC = 0, 1, 2
SKM_KERNEL_STATIONARY, SKM_KERNEL_WEIGHT_IMAGE, SKM_KERNEL_WEIGHT_TENSOR = 0, 1, 2
SKM_FLAG_EXACT_GRADIENT = 0x1
SKM_FLAG_MOVING_COEFFICIENTS = 0x2
SKM_OPT_STANDARD, SKM_OPT_ADAPTIVE, SKM_OPT_FIXED = 0, 1, 2

c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_int32_p = ctypes.POINTER(ctypes.c_int32)


class SkmGrid(ctypes.Structure):
    _fields_ = [("size", ctypes.c_int32 * 3), ("origin", ctypes.c_double * 3), ("spacing", ctypes.c_double * 3)]


class SkmDesc(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("fixed", SkmGrid), ("fixed_data", c_float_p),
        ("moving", SkmGrid), ("moving_data", c_float_p),
        ("cp", SkmGrid),
        ("sigma_c0", ctypes.c_double), ("sigma_c4", ctypes.c_double),
        ("kernel_mode", ctypes.c_int32),
        ("maximum_weight", ctypes.c_double),
        ("maximum_eigenvalue_covariance", ctypes.c_double * 9),
        ("weight", SkmGrid), ("weight_data", c_float_p),
        ("subsample", ctypes.c_int32), ("subsample_neighborhood", ctypes.c_int32),
        ("rng_state", ctypes.c_uint64), ("rng_sequence", ctypes.c_uint64),
        ("flags", ctypes.c_uint32),
    ]


class SkmResult(ctypes.Structure):
    _fields_ = [("measure", ctypes.c_double), ("reg_value_rkhs", ctypes.c_double),
                ("reg_value_rd", ctypes.c_double), ("reg_value_pg", ctypes.c_double),
                ("num_pixels_counted", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("reg_value_bilateral", ctypes.c_double)]


class SkmMetricOptions(ctypes.Structure):
    _fields_ = [("measure", ctypes.c_int32), ("reg_l1", ctypes.c_double), ("reg_l21", ctypes.c_double),
                ("reg_l2", ctypes.c_double), ("use_negative_gradient", ctypes.c_int32),
                ("do_resampling", ctypes.c_int32)]


class SkmOptimizerOptions(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("a", ctypes.c_double), ("A", ctypes.c_double),
                ("alpha", ctypes.c_double), ("fmax", ctypes.c_double), ("fmin", ctypes.c_double),
                ("omega", ctypes.c_double), ("orthant_projection", ctypes.c_int32),
                ("iterations", ctypes.c_int32), ("return_best", ctypes.c_int32), ("graph", ctypes.c_int32)]


SKM_NUM_PASSES = 11
PASS_NAMES = ("offsets", "value", "reduce", "prepare", "gather", "regularizers", "lcc_dense", "lcc_centres",
              "lcc_derivative", "displacement", "kernel_matrix")


class SkmPassTimes(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double * SKM_NUM_PASSES), ("launches", ctypes.c_uint32 * SKM_NUM_PASSES)]


class SkmTuning(ctypes.Structure):
    _fields_ = [("lcc_ref_chunks", ctypes.c_int32), ("gather_split", ctypes.c_int32),
                ("value_split", ctypes.c_int32), ("gather_regions", ctypes.c_int32),
                ("lcc_ref_walk", ctypes.c_int32), ("kernel_matrix", ctypes.c_int32), ("reserved", ctypes.c_int32 * 2)]


class SkmKernelMatrixInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("reserved", ctypes.c_int32), ("bytes", ctypes.c_uint64),
                ("pairs_value", ctypes.c_uint64), ("pairs_gather", ctypes.c_uint64), ("read_value", ctypes.c_uint64),
                ("read_gather", ctypes.c_uint64), ("build_ms", ctypes.c_double)]


class SkmAutotuneReport(ctypes.Structure):
    _fields_ = [("best", SkmTuning), ("best_ms", ctypes.c_double), ("first_ms", ctypes.c_double),
                ("candidates", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class SkmWsDesc(ctypes.Structure):
    _fields_ = [("in_size", ctypes.c_int32 * 3), ("cp_size", ctypes.c_int32 * 3), ("cp_origin", ctypes.c_double * 3),
                ("cp_spacing", ctypes.c_double * 3), ("max_batch", ctypes.c_int32)]


class SkmSyntheticDesc(ctypes.Structure):
    _fields_ = [("size", ctypes.c_int32 * 3), ("n_ellipsoids", ctypes.c_int32), ("smoothing_sigma", ctypes.c_double),
                ("n_rbf", ctypes.c_int32), ("rbf_sigma", ctypes.c_double), ("max_displacement", ctypes.c_double),
                ("displacement_scale", ctypes.c_double), ("noise_sigma", ctypes.c_double), ("seed", ctypes.c_uint64)]


class SkmIterationRecord(ctypes.Structure):
    _fields_ = [("value", ctypes.c_double), ("measure", ctypes.c_double), ("learning_rate", ctypes.c_double),
                ("num_pixels_counted", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class SkmTensorOptions(ctypes.Structure):
    _fields_ = [("sigma", ctypes.c_double), ("weight", ctypes.c_double), ("alignment_direction", ctypes.c_int32),
                ("anisotropic_filtering", ctypes.c_int32)]


SKM_MAX_SCALES = 8


class SkmScaleParams(ctypes.Structure):
    _fields_ = [("sigma_c4", ctypes.c_double), ("sigma_c0", ctypes.c_double), ("tensor_weight", ctypes.c_double),
                ("reg_l1", ctypes.c_double), ("reg_l21", ctypes.c_double), ("reg_l2", ctypes.c_double),
                ("reg_rkhs", ctypes.c_double), ("reg_rd", ctypes.c_double), ("reg_rd_scaling", ctypes.c_double),
                ("reg_pg", ctypes.c_double), ("reg_pg_scaling", ctypes.c_double), ("iterations", ctypes.c_int32),
                ("initial_step_size", ctypes.c_double), ("sampling_rate", ctypes.c_int32),
                ("neighborhood_sampling", ctypes.c_int32), ("smoothing_variance", ctypes.c_double),
                ("reg_bilateral", ctypes.c_double), ("reg_bilateral_spatial_sigma", ctypes.c_double),
                ("reg_bilateral_intensity_sigma", ctypes.c_double)]


class SkmRegistrationDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32),
                ("reference", SkmGrid), ("reference_data", c_float_p),
                ("target", SkmGrid), ("target_data", c_float_p),
                ("num_scales", ctypes.c_int32), ("scale", SkmScaleParams * SKM_MAX_SCALES),
                ("cp_spacing_scaler", ctypes.c_double), ("metric", ctypes.c_int32),
                ("use_intensity_thresholds", ctypes.c_int32), ("threshold_min", ctypes.c_double),
                ("threshold_max", ctypes.c_double),
                ("weight", SkmGrid), ("weight_image", c_float_p), ("weight_tensor", c_float_p),
                ("guidance_image", c_float_p), ("restart_at", ctypes.c_int32),
                ("restart_displacement", c_float_p), ("temp_directory", ctypes.c_char_p),
                ("verbosity", ctypes.c_int32), ("num_landmarks", ctypes.c_int64),
                ("reference_landmarks", c_double_p), ("target_landmarks", c_double_p),
                ("exact_gradient", ctypes.c_int32), ("initial_coefficients", c_float_p),
                ("initial_guard", ctypes.c_int32), ("autotune", ctypes.c_int32),
                ("initial_displacement", c_float_p), ("initial_displacement_on_device", ctypes.c_int32)]


class SkmScaleReport(ctypes.Structure):
    _fields_ = [("value", ctypes.c_double), ("measure", ctypes.c_double), ("first_value", ctypes.c_double),
                ("iterations", ctypes.c_int32), ("cp_size", ctypes.c_int32 * 3), ("maximum_tilt", ctypes.c_double),
                ("ms_setup", ctypes.c_double), ("ms_optimize", ctypes.c_double),
                ("ms_displacement", ctypes.c_double), ("tre", ctypes.c_double),
                ("zero_parameters", ctypes.c_double), ("initial_used", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("initial_value", ctypes.c_double), ("zero_value", ctypes.c_double)]


class SkmImageHeader(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("grid", SkmGrid), ("components", ctypes.c_int32)]


# Every symbol include/skm.h declares, with (restype, argtypes).
_SIGNATURES = {
    "skm_desc_init": (None, [ctypes.POINTER(SkmDesc)]),
    "skm_create": (ctypes.c_int, [ctypes.POINTER(SkmDesc), ctypes.POINTER(ctypes.c_void_p)]),
    "skm_destroy": (None, [ctypes.c_void_p]),
    "skm_last_error": (ctypes.c_char_p, []),
    "skm_set_regularizer_rkhs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "skm_set_regularizer_rd": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    "skm_set_regularizer_pg": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    "skm_set_regularizer_bilateral": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                                     ctypes.c_double]),
    "skm_set_displacement_field": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SkmGrid), c_float_p]),
    "skm_get_value": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_float_p, c_float_p, ctypes.c_int32,
                                     ctypes.POINTER(SkmResult)]),
    "skm_get_value_and_derivative": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_float_p, c_float_p,
                                                    c_float_p, ctypes.c_int32, ctypes.POINTER(SkmResult)]),
    "skm_evaluate_displacement_field": (ctypes.c_int, [ctypes.c_void_p, c_float_p, c_float_p, c_float_p]),
    "skm_set_stream": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "skm_get_value_and_derivative_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                                           ctypes.c_void_p]),
    "skm_evaluate_displacement_field_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                              ctypes.c_void_p]),
    "skm_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "skm_enable_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "skm_read_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SkmPassTimes)]),
    "skm_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SkmTuning)]),
    "skm_get_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SkmTuning)]),
    "skm_get_kernel_matrix_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SkmKernelMatrixInfo)]),
    "skm_release_kernel_matrix": (ctypes.c_int, [ctypes.c_void_p]),
    "skm_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                    ctypes.POINTER(SkmAutotuneReport)]),
    "skm_tuning_table_save": (ctypes.c_int, [ctypes.c_char_p]),
    "skm_set_cache_limit": (ctypes.c_int, [ctypes.c_uint64]),
    "skm_get_cache_info": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                          ctypes.POINTER(ctypes.c_uint64)]),
    "skm_release_cached_memory": (ctypes.c_int, []),
    "skm_tuning_table_load": (ctypes.c_int, [ctypes.c_char_p]),
    "skm_get_cube_offsets": (ctypes.c_int, [ctypes.c_void_p, c_int32_p, ctypes.POINTER(ctypes.c_int64)]),
    "skm_control_point_grid": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(SkmGrid), c_double_p,
                                              ctypes.POINTER(SkmGrid)]),
    "skm_optimize": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SkmMetricOptions),
                                    ctypes.POINTER(SkmOptimizerOptions), c_float_p, c_float_p,
                                    ctypes.POINTER(SkmIterationRecord)]),
    "skm_gaussian_smooth": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(SkmGrid), c_float_p, ctypes.c_double,
                                           c_float_p]),
    "skm_gaussian_smooth_device_noalloc": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(SkmGrid), ctypes.c_void_p,
                                                          ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]),
    "skm_anisotropic_tensor": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(SkmGrid), c_float_p,
                                              ctypes.POINTER(SkmTensorOptions), c_float_p, c_double_p, c_double_p,
                                              c_double_p]),
    "skm_registration_desc_init": (None, [ctypes.POINTER(SkmRegistrationDesc)]),
    "skm_register": (ctypes.c_int, [ctypes.POINTER(SkmRegistrationDesc), c_float_p, ctypes.POINTER(SkmScaleReport)]),
    "skm_register_batch": (ctypes.c_int, [ctypes.POINTER(SkmRegistrationDesc), ctypes.c_int32, ctypes.c_int32,
                                          ctypes.POINTER(c_float_p), ctypes.POINTER(ctypes.POINTER(SkmScaleReport)),
                                          ctypes.POINTER(ctypes.c_int)]),
    "skm_register_batch_lockstep": (ctypes.c_int, [ctypes.POINTER(SkmRegistrationDesc), ctypes.c_int32,
                                                   ctypes.POINTER(c_float_p),
                                                   ctypes.POINTER(ctypes.POINTER(SkmScaleReport)),
                                                   ctypes.POINTER(ctypes.c_int)]),
    "skm_optimize_batch": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32,
                                          ctypes.POINTER(SkmMetricOptions), ctypes.POINTER(SkmOptimizerOptions),
                                          ctypes.POINTER(c_float_p), ctypes.POINTER(c_float_p),
                                          ctypes.POINTER(ctypes.POINTER(SkmIterationRecord))]),
    "skm_warp_image": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(SkmGrid), c_float_p, ctypes.POINTER(SkmGrid),
                                      c_float_p, ctypes.POINTER(SkmGrid), ctypes.c_int32, ctypes.c_double,
                                      c_float_p]),
    "skm_read_image_header": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(SkmImageHeader)]),
    "skm_read_image": (ctypes.c_int, [ctypes.c_char_p, c_float_p, ctypes.c_int64, ctypes.POINTER(SkmImageHeader)]),
    "skm_write_image": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(SkmImageHeader), c_float_p]),
    "skm_read_landmarks": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, c_double_p, ctypes.c_int64,
                                          ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "skm_landmark_tre": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(SkmGrid), c_float_p, c_double_p, c_double_p,
                                        ctypes.c_int64, c_double_p]),
    "skm_host_regularizers": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, c_double_p, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_int32, c_double_p,
                                             c_double_p, c_double_p, c_double_p, c_double_p]),
    "skm_ws_create": (ctypes.c_int, [ctypes.POINTER(SkmWsDesc), ctypes.POINTER(ctypes.c_void_p)]),
    "skm_ws_destroy": (None, [ctypes.c_void_p]),
    "skm_ws_num_params": (ctypes.c_int64, [ctypes.c_void_p]),
    "skm_ws_set_stream": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "skm_ws_set_params": (ctypes.c_int, [ctypes.c_void_p, c_float_p]),
    "skm_ws_get_params": (ctypes.c_int, [ctypes.c_void_p, c_float_p]),
    "skm_ws_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    "skm_ws_backward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "skm_ws_prepare_input": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                            ctypes.c_void_p]),
    "skm_fit_coefficients_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_double,
                                                   ctypes.c_void_p]),
    "skm_synthetic_pair": (ctypes.c_int, [ctypes.POINTER(SkmSyntheticDesc), c_float_p, c_float_p, c_float_p]),
    "skm_ws_get_activation": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    "skm_ws_train_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                         ctypes.c_double, ctypes.POINTER(ctypes.c_double)]),
}

EXPORTED_SYMBOLS = tuple(_SIGNATURES)

_lib = None


def load():
    """Returns the loaded libskm.so; raises if the in-tree build is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libskm.so not found at {LIB_PATH}; build it with "
            "`python -m skmimageregistration_amd.build` (there is no fallback path)")
    # One HIP runtime per process: PyTorch's ROCm wheel bundles its own
    # libamdhip64.so.7, and whichever copy is loaded first serves every
    # library that needs that soname.  Loading torch first makes libskm use
    # torch's runtime, so that torch tensors and streams handed to libskm and
    # torch's own later GPU use both work (the other order leaves torch on a
    # runtime it was not built for, and it then finds no GPU).
    try:
        import torch  # noqa: F401
    except Exception:  # noqa: BLE001 -- without a usable torch the system HIP runtime serves libskm alone
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class SkmError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"skm status {status}: {message}")
        self.status = status


def check(status):
    if status != SKM_SUCCESS:
        msg = load().skm_last_error()
        raise SkmError(status, msg.decode() if msg else "")


def fptr(a):
    return a.ctypes.data_as(c_float_p)


def dptr(a):
    return a.ctypes.data_as(c_double_p)


def make_grid(size, origin=None, spacing=None):
    g = SkmGrid()
    dim = len(size)
    for d in range(3):
        g.size[d] = int(size[d]) if d < dim else 1
        g.origin[d] = float(origin[d]) if origin is not None and d < dim else 0.0
        g.spacing[d] = float(spacing[d]) if spacing is not None and d < dim else 1.0
    return g


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)
