"""CPU checks of the C ABI: the library loads, exports every declared symbol,
and its host-only entry points agree with the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as orc
from skmimageregistration_amd import _capi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(REPO, "include", "skm.h")).read()
    return sorted(set(re.findall(r"SKM_API\s+[\w\s\*]+?\b(skm_\w+)\s*\(", text)))


def test_header_symbols_bound():
    assert _declared_symbols() == sorted(_capi.EXPORTED_SYMBOLS)


def test_library_exports_every_symbol():
    lib = _capi.load()
    for name in _declared_symbols():
        assert hasattr(lib, name), name


def test_desc_defaults():
    lib = _capi.load()
    d = _capi.SkmDesc()
    lib.skm_desc_init(ctypes.byref(d))
    assert d.rng_state == 0xBA5EBA11DEADBEEF and d.rng_sequence == 0
    assert d.subsample == 1 and d.subsample_neighborhood == 1
    assert list(d.maximum_eigenvalue_covariance) == [1, 0, 0, 0, 1, 0, 0, 0, 1]


def test_create_rejects_bad_arguments():
    lib = _capi.load()
    d = _capi.SkmDesc()
    lib.skm_desc_init(ctypes.byref(d))
    h = ctypes.c_void_p()
    d.dim = 4
    assert lib.skm_create(ctypes.byref(d), ctypes.byref(h)) == 1
    d.dim = 2
    assert lib.skm_create(ctypes.byref(d), ctypes.byref(h)) == 1  # images missing
    img = np.zeros((8, 8), np.float32)
    d.fixed = _capi.make_grid((8, 8))
    d.moving = _capi.make_grid((8, 8))
    d.cp = _capi.make_grid((2, 2), spacing=(4, 4))
    d.fixed_data = _capi.fptr(img)
    d.moving_data = _capi.fptr(img)
    d.sigma_c0 = -1.0
    assert lib.skm_create(ctypes.byref(d), ctypes.byref(h)) == 1
    assert b"Kernel parameters" in lib.skm_last_error()
    d.sigma_c0 = 4.0
    d.subsample = 101
    assert lib.skm_create(ctypes.byref(d), ctypes.byref(h)) == 1
    assert b"Sampling rate" in lib.skm_last_error()
    d.subsample = 1
    d.kernel_mode = 1
    d.weight = _capi.make_grid((8, 8))
    assert lib.skm_create(ctypes.byref(d), ctypes.byref(h)) == 1
    assert b"Weight image not provided" in lib.skm_last_error()


@pytest.mark.parametrize("dim", [2, 3])
def test_control_point_grid_matches_oracle(dim):
    lib = _capi.load()
    olib = orc.load()
    size = (255, 200, 96)[:dim]
    origin = (-12.5, 3.0, 7.25)[:dim]
    spacing = (1.0, 0.75, 2.0)[:dim]
    sp = (8.0, 6.0, 5.0)[:dim]
    out = _capi.SkmGrid()
    cs = (ctypes.c_double * 3)(*(list(sp) + [1.0] * (3 - dim)))
    assert lib.skm_control_point_grid(dim, ctypes.byref(_capi.make_grid(size, origin, spacing)), cs,
                                      ctypes.byref(out)) == 0
    og = orc.Grid()
    olib.skmo_control_point_grid(dim, ctypes.byref(orc.make_grid(size, origin, spacing)), cs, ctypes.byref(og))
    for d in range(dim):
        assert out.size[d] == og.size[d]
        assert out.origin[d] == og.origin[d]
        # closed form: ceil(extent / h) points, centred in the image extent
        extent = spacing[d] * (size[d] - 1)
        n = int(np.ceil(extent / sp[d]))
        assert out.size[d] == n
        assert abs(out.origin[d] - (origin[d] + (extent - sp[d] * (n - 1)) / 2)) < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_host_regularizers_match_oracle(dim):
    lib = _capi.load()
    olib = orc.load()
    rng = np.random.default_rng(5)
    ncp = 50
    th = rng.normal(size=ncp * dim)
    th[rng.integers(0, ncp * dim, 20)] = 0.0
    th[:dim] = 0.0
    g0 = rng.normal(size=ncp * dim) * 0.01
    for l1, l21, l2 in ((0.005, 0, 0), (0, 0.004, 0), (0, 0, 0.3), (0.002, 0.001, 0.1)):
        a = g0.copy()
        b = g0.copy()
        rv = ctypes.c_double()
        v = [ctypes.c_double() for _ in range(3)]
        assert lib.skm_host_regularizers(dim, ncp, _capi.dptr(th), l1, l21, l2, 1, _capi.dptr(a),
                                         ctypes.byref(rv), *[ctypes.byref(x) for x in v]) == 0
        ov = [ctypes.c_double() for _ in range(3)]
        oreg = olib.skmo_host_regularizers(dim, ncp, orc.dp(th), l1, l21, l2, 1, orc.dp(b),
                                           *[ctypes.byref(x) for x in ov])
        np.testing.assert_array_equal(a, b)
        assert rv.value == oreg


def test_register_batch_rejects_bad_arguments():
    """skm_register_batch validates before touching a device."""
    lib = _capi.load()
    descs = (_capi.SkmRegistrationDesc * 1)()
    lib.skm_registration_desc_init(ctypes.byref(descs[0]))
    buf = np.empty(8, np.float32)
    outs = (_capi.c_float_p * 1)(buf.ctypes.data_as(_capi.c_float_p))
    assert lib.skm_register_batch(descs, 0, 1, None, None, None) == 0  # nothing to do
    assert lib.skm_register_batch(descs, -1, 1, outs, None, None) == _capi.SKM_ERROR_INVALID_ARGUMENT
    assert lib.skm_register_batch(None, 1, 1, outs, None, None) == _capi.SKM_ERROR_INVALID_ARGUMENT
    assert lib.skm_register_batch(descs, 1, 1, None, None, None) == _capi.SKM_ERROR_INVALID_ARGUMENT
    assert lib.skm_register_batch(descs, 1, 0, outs, None, None) == _capi.SKM_ERROR_INVALID_ARGUMENT
    assert b"concurrency" in lib.skm_last_error()
    nulls = (_capi.c_float_p * 1)()
    assert lib.skm_register_batch(descs, 1, 1, nulls, None, None) == _capi.SKM_ERROR_INVALID_ARGUMENT
    assert b"displacements_out[0]" in lib.skm_last_error()


def test_create_rejects_unknown_flags():
    lib = _capi.load()
    d = _capi.SkmDesc()
    lib.skm_desc_init(ctypes.byref(d))
    img = np.zeros((8, 8), np.float32)
    d.dim = 2
    d.fixed = _capi.make_grid((8, 8))
    d.moving = _capi.make_grid((8, 8))
    d.cp = _capi.make_grid((2, 2), spacing=(4, 4))
    d.fixed_data = _capi.fptr(img)
    d.moving_data = _capi.fptr(img)
    d.sigma_c0 = d.sigma_c4 = 4.0
    d.flags = 0x80
    h = ctypes.c_void_p()
    assert lib.skm_create(ctypes.byref(d), ctypes.byref(h)) == _capi.SKM_ERROR_INVALID_ARGUMENT
    assert b"unknown flags" in lib.skm_last_error()


def test_cache_limit_and_release_without_device_blocks():
    """The cache's host bookkeeping: the limit round-trips, releasing with
    nothing idle succeeds, and no block is live before any evaluator."""
    from skmimageregistration_amd import cache_info, release_cached_memory, set_cache_limit
    before = cache_info()
    try:
        set_cache_limit(123 << 20)
        info = cache_info()
        assert info["limit"] == 123 << 20
        release_cached_memory()
        assert cache_info()["idle"] == 0
        set_cache_limit(0)
        assert cache_info()["limit"] == 0
    finally:
        set_cache_limit(before["limit"])
