"""MI355X (gfx950) implementation of the SKM kernel-sum / metric / regularizer path.

The compute path is libskm.so (HIP kernels + C ABI, include/skm.h); this
package is the host-side mirror of the reference's GpuKernelEvaluator
interface over that ABI.
"""
from . import _capi
from .evaluator import (MEASURE, ComputeAnisotropicTensor, EvaluationResult, GaussianSmoothing,
                        GpuKernelEvaluator, Grid, Image, Kernel, RegistrationConfig, control_point_grid,
                        landmark_tre, optimize_batch, read_image, read_landmarks, register_batch, register_files,
                        register_multiscale, release_cached_memory, set_cache_limit, cache_info, tuning_table_load,
                        tuning_table_save, warp_image, write_image)

__all__ = ["MEASURE", "ComputeAnisotropicTensor", "EvaluationResult", "GaussianSmoothing", "GpuKernelEvaluator",
           "Grid", "Image", "Kernel", "RegistrationConfig", "control_point_grid", "landmark_tre", "load_library",
           "optimize_batch", "read_image", "read_landmarks", "register_batch", "register_files", "register_multiscale",
           "tuning_table_load", "tuning_table_save", "warp_image", "write_image", "release_cached_memory",
           "set_cache_limit", "cache_info"]


def load_library():
    """Loads the in-tree libskm.so (raises if it was not built)."""
    return _capi.load()
