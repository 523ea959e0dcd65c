"""The library's device-memory cache: blocks an evaluator gives back stay
idle for the next one up to the per-device limit, skm_release_cached_memory
frees them, and a limit of 0 keeps nothing idle (results unchanged)."""
import numpy as np
import pytest

from helpers import Case
from skmimageregistration_amd import MEASURE, cache_info, release_cached_memory, set_cache_limit

pytestmark = pytest.mark.gpu


def test_release_and_limit():
    before = cache_info()
    case = Case(dim=3, shape=(20, 22, 24), cp_spacing=5.0, sigma=10.0, seed=91)
    try:
        ev = case.gpu()
        r1, d1 = ev.getValueAndDerivative(MEASURE.NCC, case.cp, case.cpw)
        assert cache_info()["live"] > 0
        ev.close()
        assert cache_info()["idle"] > 0
        release_cached_memory()
        assert cache_info()["idle"] == 0
        set_cache_limit(0)
        ev = case.gpu()
        r2, d2 = ev.getValueAndDerivative(MEASURE.NCC, case.cp, case.cpw)
        ev.close()
        assert cache_info()["idle"] == 0
        assert r1.measure == r2.measure and np.array_equal(d1, d2)
    finally:
        set_cache_limit(before["limit"])
