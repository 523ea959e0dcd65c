"""skm_gaussian_smooth_device_noalloc (temporaries in the library's fixed
device array) against skm_gaussian_smooth (host arrays, temporaries
allocated per call): the same bits, on consecutive calls of decreasing size
(the second call reuses the array the first one left behind), 2D and 3D,
in place and out of place; and a clear error beyond the size limit.
"""
import ctypes

import numpy as np
import pytest

from skmimageregistration_amd import _capi

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

LIMIT = 8388608


def _grid(shape, spacing):
    size = tuple(reversed(shape))
    return _capi.make_grid(size, (0.0,) * len(size), tuple(reversed(spacing)))


def _host_smooth(lib, img, grid, variance):
    out = np.empty_like(img)
    _capi.check(lib.skm_gaussian_smooth(img.ndim, ctypes.byref(grid), _capi.fptr(img), variance, _capi.fptr(out)))
    return out


def _noalloc(lib, d_in, d_out, grid, dim, variance):
    stream = torch.cuda.current_stream().cuda_stream
    _capi.check(lib.skm_gaussian_smooth_device_noalloc(dim, ctypes.byref(grid), ctypes.c_void_p(d_in.data_ptr()),
                                                       variance, ctypes.c_void_p(d_out.data_ptr()),
                                                       ctypes.c_void_p(stream)))
    torch.cuda.synchronize()
    return d_out.cpu().numpy()


CALLS = [
    # shape (z, y, x), spacing (z, y, x), variance: the second call is smaller than the first
    ((48, 52, 60), (1.5, 1.0, 0.8), 3.0),
    ((20, 18, 22), (1.0, 1.0, 1.0), 1.2),
]


def test_noalloc_smoothing_equals_allocating_entry_point():
    lib = _capi.load()
    rng = np.random.default_rng(7)
    for shape, spacing, variance in CALLS:
        print(f"noalloc smoothing 3D {shape} variance {variance}", flush=True)
        img = rng.normal(size=shape).astype(np.float32)
        grid = _grid(shape, spacing)
        expected = _host_smooth(lib, img, grid, variance)
        d_in = torch.from_numpy(img).cuda()
        d_out = torch.full_like(d_in, float("nan"))
        got = _noalloc(lib, d_in, d_out, grid, 3, variance)
        assert np.array_equal(got, expected)
        got = _noalloc(lib, d_in, d_in, grid, 3, variance)  # in place
        assert np.array_equal(got, expected)
    for shape, variance in (((70, 90), 4.0), ((30, 26), 0.7)):
        print(f"noalloc smoothing 2D {shape} variance {variance}", flush=True)
        img = rng.normal(size=shape).astype(np.float32)
        grid = _grid(shape, (1.0, 1.25))
        expected = _host_smooth(lib, img, grid, variance)
        d_in = torch.from_numpy(img).cuda()
        got = _noalloc(lib, d_in, torch.empty_like(d_in), grid, 2, variance)
        assert np.array_equal(got, expected)


def test_noalloc_smoothing_rejects_images_beyond_the_limit():
    lib = _capi.load()
    small = torch.zeros(8, device="cuda")
    grid = _grid((129, 256, 256), (1.0, 1.0, 1.0))  # 8,454,144 voxels > the limit
    st = lib.skm_gaussian_smooth_device_noalloc(3, ctypes.byref(grid), ctypes.c_void_p(small.data_ptr()), 1.0,
                                                ctypes.c_void_p(small.data_ptr()), None)
    assert st == _capi.SKM_ERROR_INVALID_ARGUMENT
    assert "allocation-free limit" in lib.skm_last_error().decode()
    assert 129 * 256 * 256 > LIMIT
