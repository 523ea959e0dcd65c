"""Pins of the FP64 oracle against facts fixed outside the oracle's own code:
published algorithms (PCG32 demo stream, Unser's cubic B-spline prefilter as
implemented by scipy.ndimage), the defining properties of the Wendland
functions, numpy linear algebra, closed forms and finite differences.

The reference cannot be compiled here (ITK 4.10 and nvcc are absent), so no
reference-built golden vectors exist; every oracle function is checked by the
properties below instead.
"""
import ctypes
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle as orc
from helpers import Case


def lib():
    return orc.load()


# ------------------------------------------------------------------ a1 kernel

def _poly_fit(fn, deg):
    t = np.linspace(0.0, 0.999, 400)
    y = np.array([fn(v) for v in t])
    return np.polynomial.polynomial.polyfit(t, y, deg)


def test_wendland_c4_defining_properties():
    """psi_{3,2} (Wendland 1995) normalised to 1 at 0: a degree-8 polynomial on
    [0,1] with a 6-fold zero at t=1 (C4 at the support edge), value 1 and
    vanishing odd derivatives of order 1 and 3 at t=0 (C4 at the centre).
    These five facts fix the polynomial uniquely."""
    c = _poly_fit(lib().skmo_wendland_c4, 10)
    assert np.all(np.abs(c[9:]) < 1e-8)                  # degree 8
    assert abs(c[0] - 1.0) < 1e-12                       # k(0) = 1
    assert abs(c[1]) < 1e-9 and abs(c[3]) < 1e-8         # even up to order 4
    p = np.polynomial.Polynomial(c[:9])
    for k in range(6):                                   # (1-t)^6 factor
        assert abs(p.deriv(k)(1.0)) < 1e-6 * math.factorial(k + 2), k
    assert lib().skmo_wendland_c4(1.0) == 0.0 and lib().skmo_wendland_c4(1.7) == 0.0


def test_wendland_c0_defining_properties():
    """(1 - t)_+^2: degree 2, double zero at 1, value 1 at 0."""
    c = _poly_fit(lib().skmo_wendland_c0, 4)
    assert np.all(np.abs(c[3:]) < 1e-10)
    assert abs(c[0] - 1) < 1e-12 and abs(c[1] + 2) < 1e-10 and abs(c[2] - 1) < 1e-10
    assert lib().skmo_wendland_c0(2.0) == 0.0


def _stationary_problem(dim=3, sigma=24.0, sigma_c0=None, mode=0, weight=None, max_weight=1.0, max_eig=None):
    fixed = np.zeros((4,) * dim)
    return orc.OracleProblem(fixed, fixed, ((1,) * dim, (0.0,) * dim, (1.0,) * dim), np.zeros((1, dim)),
                             sigma_c0=sigma if sigma_c0 is None else sigma_c0, sigma_c4=sigma, kernel_mode=mode,
                             weight=weight, max_weight=max_weight, max_eig_cov=max_eig)


def test_kernel_closed_form_value():
    """k(6; sigma_C4 = sigma_C0 = 24) = 0.75^6 (3 + 4.5 + 2.1875)/3 * 0.75^2."""
    pr = _stationary_problem(3, 24.0)
    k = pr.kernel([0, 0, 0], [6, 0, 0])
    assert abs(k - 0.75 ** 6 * (3 + 18 * 0.25 + 35 * 0.0625) / 3 * 0.75 ** 2) < 1e-15
    assert abs(k - 0.323281288) < 1e-9
    assert pr.kernel([1, 2, 3], [1, 2, 3]) == 1.0
    assert pr.kernel([0, 0, 0], [24, 0, 0]) == 0.0


def test_kernel_symmetric_and_positive_definite():
    """Products of Wendland functions positive definite on R^3 give a positive
    definite Gram matrix (Schur product theorem)."""
    pr = _stationary_problem(3, 20.0, 30.0)
    rng = np.random.default_rng(0)
    pts = rng.uniform(0, 30, size=(60, 3))
    G = np.array([[pr.kernel(p, q) for q in pts] for p in pts])
    assert np.allclose(G, G.T, atol=0)
    assert np.linalg.eigvalsh(G).min() > 0


def test_nonstationary_reduces_to_scaled_stationary():
    """sigma_P = sigma_Q = s: F^2 = 1 and the radius is scaled by sqrt(1/s) (Kernel.h:55-61, 79-85, 105-117)."""
    s = 1.7
    w = np.full((4, 4, 4), s)
    pr = _stationary_problem(3, 20.0, weight=((4, 4, 4), (0, 0, 0), (1, 1, 1), w), mode=1)
    st = _stationary_problem(3, 20.0)
    p, q = np.array([1.0, 2.0, 1.5]), np.array([9.0, -3.0, 4.0])
    r = np.linalg.norm(p - q)
    k_ns = pr.kernel(p, np.array([1.0, 1.0, 1.0]) + 0 * q, sigma_p=s) if False else None
    k_ns = pr.kernel(p, np.array([2.0, 2.0, 2.0]), sigma_p=s)
    r2 = np.linalg.norm(p - np.array([2.0, 2.0, 2.0]))
    assert abs(k_ns - st.kernel([0, 0, 0], [r2 / math.sqrt(s), 0, 0])) < 1e-14
    # outside the weight image sigma_Q = 0 (ImageNearest default), so F = 0
    assert pr.kernel(p, np.array([50.0, 0, 0]), sigma_p=s) == 0.0
    del r


def test_anisotropic_reduces_and_is_rotation_invariant():
    """cov = c I gives the stationary kernel of r / sqrt(c); k(Rp, Rq; R C R^T) = k(p, q; C)."""
    c = 2.3
    rng = np.random.default_rng(1)
    w = np.broadcast_to(np.eye(3) * c, (4, 4, 4, 3, 3)).copy()
    pr = _stationary_problem(3, 15.0, mode=2, weight=((4, 4, 4), (0, 0, 0), (1, 1, 1), w.reshape(-1, 9)))
    st = _stationary_problem(3, 15.0)
    p, q = np.array([0.5, 1.0, 2.0]), np.array([1.0, 1.0, 1.0])
    r = np.linalg.norm(p - q)
    assert abs(pr.kernel(p, q, cov_p=np.eye(3) * c) - st.kernel([0, 0, 0], [r / math.sqrt(c), 0, 0])) < 1e-14
    A = rng.normal(size=(3, 3))
    C = A @ A.T + np.eye(3)
    R, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    w1 = np.broadcast_to(C, (4, 4, 4, 3, 3)).reshape(-1, 9).copy()
    w2 = np.broadcast_to(R @ C @ R.T, (4, 4, 4, 3, 3)).reshape(-1, 9).copy()
    pr1 = _stationary_problem(3, 15.0, mode=2, weight=((4, 4, 4), (0, 0, 0), (1, 1, 1), w1))
    pr2 = _stationary_problem(3, 15.0, mode=2, weight=((4, 4, 4), (-50, -50, -50), (100, 100, 100), w2))
    d = np.array([3.0, -2.0, 1.0])
    k1 = pr1.kernel(q + d, q, cov_p=C)
    k2 = pr2.kernel(R @ d, np.zeros(3), cov_p=R @ C @ R.T)
    assert abs(k1 - k2) < 1e-13
    # F^2 factor: det(C)^(1/2) det(C)^(1/2) / det(C) = 1 for equal tensors; quadratic form v C^-1 v
    assert abs(k1 - st.kernel([0, 0, 0], [math.sqrt(d @ np.linalg.solve(C, d)), 0, 0])) < 1e-13


@pytest.mark.parametrize("dim", [2, 3])
def test_determinant_and_max_eigenvalue(dim):
    rng = np.random.default_rng(2)
    for _ in range(50):
        A = rng.normal(size=(dim, dim))
        S = A @ A.T + 0.1 * np.eye(dim)
        m = np.ascontiguousarray(S.reshape(-1))
        assert abs(lib().skmo_determinant(dim, orc.dp(m)) - np.linalg.det(S)) < 1e-10 * max(1, abs(np.linalg.det(S)))
        assert abs(lib().skmo_max_eigenvalue(dim, orc.dp(m)) - np.linalg.eigvalsh(S)[-1]) < 1e-9
    D = np.ascontiguousarray(np.diag([1.0, 4.0, 2.0][:dim]).reshape(-1))
    assert lib().skmo_max_eigenvalue(dim, orc.dp(D)) == 4.0


def test_region_support():
    """getRegionSupport (Kernel.h:158-171): min(sigmas), /sqrt(2/(maxW+sigma)), /sqrt(2/lambda_max)."""
    pr = _stationary_problem(3, 24.0, 30.0)
    assert pr.lib.skmo_region_support(ctypes.byref(pr.p), orc.dp(np.zeros(3))) == 24.0
    w = np.full((4, 4, 4), 1.5)
    pr = _stationary_problem(3, 24.0, 30.0, mode=1, max_weight=2.5, weight=((4, 4, 4), (0, 0, 0), (1, 1, 1), w))
    assert abs(pr.lib.skmo_region_support(ctypes.byref(pr.p), orc.dp(np.ones(3))) - 24.0 / math.sqrt(2 / 4.0)) < 1e-12


# ------------------------------------------------------------------ a3 B-spline

@pytest.mark.parametrize("shape", [(17, 23), (9, 11, 13)])
def test_prefilter_matches_scipy_mirror(shape):
    """Unser's cubic prefilter with mirror boundary = scipy.ndimage.spline_filter(mode='mirror');
    the coefficient grid is shifted by one and padded by mirroring (BsplineImage.h:72-108)."""
    rng = np.random.default_rng(3)
    img = rng.normal(size=shape)
    coef = orc.bspline_coefficients(img)
    ref = ndimage.spline_filter(img, order=3, mode="mirror")
    inner = coef[tuple(slice(1, s + 1) for s in shape)]
    assert np.max(np.abs(inner - ref)) < 1e-12
    # mirrored borders: index -1 -> 1, n -> n-2, n+1 -> n-3
    ax = img.ndim - 1
    c0 = np.take(coef, 0, axis=ax)
    c2 = np.take(coef, 2, axis=ax)
    np.testing.assert_array_equal(c0, c2)


@pytest.mark.parametrize("dim", [2, 3])
def test_bspline_sampling_matches_scipy_and_fd(dim):
    rng = np.random.default_rng(4)
    shape = (14, 15) if dim == 2 else (9, 10, 11)
    img = rng.normal(size=shape)
    coef = orc.bspline_coefficients(img)
    spacing = (1.5, 0.5, 2.0)[:dim]
    origin = (3.0, -1.0, 0.5)[:dim]
    g = orc.make_grid(tuple(reversed(shape)), origin, spacing)
    val = ctypes.c_double()
    grad = np.zeros(3)
    for _ in range(40):
        idx = np.array([rng.uniform(0, s - 1) for s in reversed(shape)])
        p = np.ascontiguousarray(np.array(origin) + np.array(spacing) * idx)
        assert lib().skmo_bspline_sample(dim, ctypes.byref(g), orc.dp(coef), orc.dp(p), ctypes.byref(val), orc.dp(grad))
        ref = ndimage.map_coordinates(img, [[v] for v in reversed(idx)], order=3, mode="mirror")[0]
        assert abs(val.value - ref) < 1e-10
        for d in range(dim):
            h = 1e-5
            pp, pm = p.copy(), p.copy()
            pp[d] += h
            pm[d] -= h
            vp, vm = ctypes.c_double(), ctypes.c_double()
            lib().skmo_bspline_sample(dim, ctypes.byref(g), orc.dp(coef), orc.dp(pp), ctypes.byref(vp), orc.dp(np.zeros(3)))
            lib().skmo_bspline_sample(dim, ctypes.byref(g), orc.dp(coef), orc.dp(pm), ctypes.byref(vm), orc.dp(np.zeros(3)))
            assert abs((vp.value - vm.value) / (2 * h) - grad[d]) < 1e-6
    # node interpolation
    for _ in range(10):
        i = np.array([rng.integers(0, s) for s in reversed(shape)], dtype=float)
        p = np.ascontiguousarray(np.array(origin) + np.array(spacing) * i)
        lib().skmo_bspline_sample(dim, ctypes.byref(g), orc.dp(coef), orc.dp(p), ctypes.byref(val), orc.dp(grad))
        assert abs(val.value - img[tuple(int(v) for v in reversed(i))]) < 1e-12


def test_bspline_inside_and_truncation():
    """inside iff -0.5 <= idx < n - 0.5; for idx in (-0.5, 0) the cubic piece of
    [0, 1) is extrapolated (int() truncation, BsplineImage.h:235-236)."""
    rng = np.random.default_rng(5)
    img = rng.normal(size=(1, 12))[0]
    img2 = np.tile(img, (8, 1))  # constant along y
    coef = orc.bspline_coefficients(img2)
    g = orc.make_grid((12, 8))
    val = ctypes.c_double()
    gr = np.zeros(3)

    def sample(x):
        p = np.ascontiguousarray([x, 3.0])
        ok = lib().skmo_bspline_sample(2, ctypes.byref(g), orc.dp(coef), orc.dp(p), ctypes.byref(val), orc.dp(gr))
        return ok, val.value

    assert sample(-0.5)[0] and not sample(-0.5000001)[0]
    assert sample(11.4999)[0] and not sample(11.5)[0]
    ts = np.linspace(0.0, 0.99, 30)
    piece = np.polynomial.polynomial.polyfit(ts, [sample(t)[1] for t in ts], 3)
    for t in (-0.1, -0.3, -0.49):
        assert abs(sample(t)[1] - np.polynomial.polynomial.polyval(t, piece)) < 1e-9


# ------------------------------------------------------------------ a5 PCG32

def test_pcg32_published_demo_stream():
    """pcg32_srandom_r(42, 54) then pcg32_random_r (O'Neill, pcg-c-basic demo)."""
    L = lib()
    inc = (54 << 1) | 1
    st = ctypes.c_uint64(0)
    L.skmo_pcg32_next(ctypes.byref(st), inc)
    st.value = (st.value + 42) & 0xFFFFFFFFFFFFFFFF
    L.skmo_pcg32_next(ctypes.byref(st), inc)
    out = [L.skmo_pcg32_next(ctypes.byref(st), inc) for _ in range(6)]
    assert out == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_cube_offsets_fixed_seed_properties():
    """offsets = int(nextI/UINT_MAX * sub) in [0, sub]; the stream is the reference
    seed 0xBA5EBA11DEADBEEF, sequence 0 (UniformRng.h:32)."""
    off, state = orc.cube_offsets(5000, 3, 4)
    assert off.min() >= 0 and off.max() <= 4
    counts = np.bincount(off.reshape(-1), minlength=5)
    assert counts[4] == 0 and np.all(np.abs(counts[:4] / off.size - 0.25) < 0.02)
    st = ctypes.c_uint64(0xBA5EBA11DEADBEEF)
    first = lib().skmo_pcg32_next(ctypes.byref(st), 0)
    assert off[0, 0] == int(first / 4294967295.0 * 4)
    assert state != 0xBA5EBA11DEADBEEF


# ------------------------------------------------------------------ a6 displacement

@pytest.mark.parametrize("dim", [2, 3])
def test_displacement_single_coefficient_is_kernel_response(dim):
    """One non-zero coefficient c at cp j: u(x) = w_j k(x, p_j) c (SKMReg.cxx:303-323)."""
    case = Case(dim=dim, shape=(20, 22) if dim == 2 else (10, 11, 12), cp_spacing=4.0, sigma=9.0, seed=1)
    cp = np.zeros_like(case.cp)
    j = case.n_cp // 2
    cp[j] = (0.7, -1.2, 0.4)[:dim]
    pr = case.oracle(cp=cp)
    u = pr.displacement()
    jc = np.unravel_index(j, tuple(reversed(case.cp_size)))[::-1]
    pj = np.array(case.cp_origin) + np.array(case.cp_spacing) * np.array(jc)
    c4, c0 = lib().skmo_wendland_c4, lib().skmo_wendland_c0
    for v in range(0, u.shape[0], 7):
        x = np.array(np.unravel_index(v, case.shape)[::-1], float)
        r = np.linalg.norm(x - pj)
        k = c4(r / case.sigma_c4) * c0(r / case.sigma_c0)
        want = float(case.cpw[j]) * k * cp[j].astype(np.float64)
        np.testing.assert_allclose(u[v], want, rtol=1e-12, atol=1e-18)


# ------------------------------------------------------------------ a7/a8 metrics

def _fd_check(case, measure, n_checks=6, h=1e-3, tol=2e-6):
    pr = case.oracle(case.offsets(1) if case.sub > 1 else None)
    res, der = pr.evaluate(measure, True)
    rng = np.random.default_rng(9)
    scale = np.abs(der).max()
    for idx in rng.choice(der.size, n_checks, replace=False):
        j, d = idx % case.n_cp, idx // case.n_cp
        cpp = case.cp.astype(np.float64).copy()
        cpm = cpp.copy()
        cpp[j, d] += h
        cpm[j, d] -= h
        off = case.offsets(1) if case.sub > 1 else None
        vp = case.oracle(off, cp=cpp).evaluate(measure, False)[0]["measure"]
        vm = case.oracle(off, cp=cpm).evaluate(measure, False)[0]["measure"]
        fd = (vp - vm) / (2 * h)
        assert abs(fd - der[idx]) < tol * max(scale, 1e-12) + 1e-9, (idx, fd, der[idx])
    return res


@pytest.mark.parametrize("measure", [0, 1])
@pytest.mark.parametrize("dim,sub", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_metric_derivative_fd_exact_mode(measure, dim, sub):
    shape = (24, 26) if dim == 2 else (10, 11, 12)
    case = Case(dim=dim, shape=shape, cp_spacing=4.0, sigma=9.0, subsample=sub, exact=True, seed=2, coef_scale=0.6)
    _fd_check(case, measure)


@pytest.mark.parametrize("mode", [1, 2])
def test_metric_derivative_fd_weighted_modes(mode):
    case = Case(dim=3, shape=(10, 11, 12), cp_spacing=4.0, sigma=8.0, mode=mode, exact=True, seed=4, coef_scale=0.6)
    _fd_check(case, 1)


def test_reference_windows_equal_exact_in_3d_and_2d_dense():
    """3D derivative windows are a superset (GpuEvaluatorMSE.cu:126-132), and the
    2D windows are exact at subsample 1, so reference mode equals the gradient there."""
    for dim, sub in ((3, 2), (2, 1)):
        shape = (24, 26) if dim == 2 else (10, 11, 12)
        a = Case(dim=dim, shape=shape, cp_spacing=4.0, sigma=9.0, subsample=sub, seed=3)
        b = Case(dim=dim, shape=shape, cp_spacing=4.0, sigma=9.0, subsample=sub, seed=3, exact=True)
        off = a.offsets(1) if sub > 1 else None
        _, da = a.oracle(off).evaluate(0, True)
        _, db = b.oracle(off).evaluate(0, True)
        assert np.max(np.abs(da - db)) <= 1e-12 * np.abs(db).max()


def test_reference_2d_subsampled_window_drops_cells():
    """With subsample > 1 the 2D lower bound int((idx - s/h)/sub + 1) drops a cell of
    in-support samples (GpuEvaluatorMSE.cu:267-273): reference != gradient."""
    a = Case(dim=2, shape=(24, 26), cp_spacing=4.0, sigma=9.0, subsample=4, seed=3)
    b = Case(dim=2, shape=(24, 26), cp_spacing=4.0, sigma=9.0, subsample=4, seed=3, exact=True)
    off = a.offsets(1)
    _, da = a.oracle(off).evaluate(0, True)
    _, db = b.oracle(off).evaluate(0, True)
    # the dropped samples sit near the edge of the support, where k ~ (1-t)^8 is small
    diff = np.max(np.abs(da - db))
    assert 1e-9 * np.abs(db).max() < diff < 0.1 * np.abs(db).max()


def test_identity_registration():
    """Moving = fixed, zero coefficients: MSE = 0 with zero derivative, NCC = -1."""
    case = Case(dim=2, shape=(20, 22), cp_spacing=4.0, sigma=9.0, seed=5)
    case.moving = case.fixed.copy()
    case.cp[:] = 0
    pr = case.oracle()
    r, d = pr.evaluate(0, True)
    assert r["measure"] < 1e-28 and np.abs(d).max() < 1e-15
    r, _ = pr.evaluate(1, False)
    assert abs(r["measure"] + 1) < 1e-14


def test_ncc_invariant_to_affine_intensity():
    case = Case(dim=2, shape=(20, 22), cp_spacing=4.0, sigma=9.0, seed=6)
    r1, d1 = case.oracle().evaluate(1, True)
    case.fixed = case.fixed.astype(np.float64) * 3.5 + 2.0
    r2, d2 = case.oracle().evaluate(1, True)
    # one-pass centred sums (GpuEvaluator.cu:239-254) lose ~eps * mean^2 / var
    assert abs(r1["measure"] - r2["measure"]) < 1e-10
    assert np.max(np.abs(d1 - d2)) < 1e-8 * np.abs(d1).max()


def test_mse_matches_independent_warp():
    """Per-sample M(x+u) - T(x) against R sampled by scipy at the oracle's
    displacement, on samples whose warped point is interior; MSE = mean of the
    squared differences over the inside samples."""
    case = Case(dim=2, shape=(30, 32), cp_spacing=5.0, sigma=10.0, seed=7, coef_scale=0.5)
    pr = case.oracle()
    u = pr.displacement().reshape(case.shape + (2,))
    yy, xx = np.meshgrid(np.arange(30.0), np.arange(32.0), indexing="ij")
    wx, wy = xx + u[..., 0], yy + u[..., 1]
    inside = (wx >= -0.5) & (wx < 31.5) & (wy >= -0.5) & (wy < 29.5)
    interior = (wx >= 0) & (wx <= 31) & (wy >= 0) & (wy <= 29)
    m = ndimage.map_coordinates(case.moving.astype(np.float64), [wy[interior], wx[interior]], order=3, mode="mirror")
    r, _, diffs = pr.evaluate(0, False, per_sample=True)
    diffs = diffs.reshape(case.shape)
    assert np.max(np.abs(diffs[interior] - (m - case.fixed[interior]))) < 1e-12
    assert np.all(diffs[~inside] == 0.0)
    assert r["pixels_counted"] == inside.sum()
    assert abs(r["measure"] - np.sum(diffs ** 2) / inside.sum()) < 1e-15


# ------------------------------------------------------------------ a9 LCC

@pytest.mark.parametrize("dim", [2, 3])
def test_lcc_exact_derivative_fd(dim):
    shape = (14, 15) if dim == 2 else (8, 9, 10)
    case = Case(dim=dim, shape=shape, cp_spacing=3.0, sigma=5.0, subsample=2, subsample_neighborhood=2,
                exact=True, seed=8, coef_scale=0.5)
    _fd_check(case, 2, n_checks=4, h=1e-4, tol=1e-5)


def test_lcc_reference_equals_gradient_without_truncation():
    """When every centre lies in every cp window (support covers the image) and
    w = 1, the reference LCC derivative is the exact gradient."""
    case = Case(dim=2, shape=(9, 10), cp_spacing=4.0, sigma=30.0, seed=9, coef_scale=0.3)
    case.cpw[:] = 1.0
    _fd_check(case, 2, n_checks=4, h=1e-4, tol=1e-5)


def test_lcc_identity_is_minus_one():
    case = Case(dim=2, shape=(16, 17), cp_spacing=4.0, sigma=6.0, seed=10)
    case.moving = case.fixed.copy()
    case.cp[:] = 0
    r, _ = case.oracle().evaluate(2, False)
    assert abs(r["measure"] + 1) < 1e-12


# ------------------------------------------------------------------ a10 regularizers

def _two_cp_problem(dim, c, w=(1.0, 1.0), h=5.0, sigma=8.0):
    size = (2,) + (1,) * (dim - 1)
    fixed = np.zeros((4,) * dim)
    return orc.OracleProblem(fixed, fixed, (size, (0.0,) * dim, (h,) * dim), np.asarray(c, float),
                             cpw=np.asarray(w), sigma_c0=sigma, sigma_c4=sigma)


@pytest.mark.parametrize("dim", [2, 3])
def test_regularizer_closed_forms_two_points(dim):
    rng = np.random.default_rng(11)
    c = rng.normal(size=(2, dim))
    h, s = 5.0, 8.0
    pr = _two_cp_problem(dim, c, h=h, sigma=s)
    k = lib().skmo_wendland_c4(h / s) * lib().skmo_wendland_c0(h / s)
    vals = np.zeros(2)
    pr.lib.skmo_regularizer(ctypes.byref(pr.p), 0, 1.0, 1.0, orc.dp(vals), None)
    assert abs(vals.sum() - ((c ** 2).sum() + 2 * k * c[0] @ c[1])) < 1e-12
    pr.lib.skmo_regularizer(ctypes.byref(pr.p), 1, 1.0, 1.0, orc.dp(vals), None)
    assert abs(vals.sum() - 2 * k * ((c[0] - c[1]) ** 2).sum()) < 1e-12
    scale = 0.7
    pr.lib.skmo_regularizer(ctypes.byref(pr.p), 2, 1.0, scale, orc.dp(vals), None)
    cr2 = np.sum(np.cross(c[0], c[1]) ** 2) if dim == 3 else (c[0, 0] * c[1, 1] - c[0, 1] * c[1, 0]) ** 2
    assert abs(vals.sum() - 2 * k * scale ** 2 / 2 * math.log(1 + cr2 / scale)) < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("which", [0, 1, 2])
def test_regularizer_derivative_fd(dim, which):
    """With uniform cp weights the GPU-side regularizer gradients are exact."""
    case = Case(dim=dim, shape=(14, 15) if dim == 2 else (8, 9, 10), cp_spacing=3.0, sigma=7.0, seed=12,
                coef_scale=0.8)
    case.cpw[:] = 1.0
    pr = case.oracle()
    n = case.n_cp
    der = np.zeros(n * dim)
    vals = np.zeros(n)
    pr.lib.skmo_regularizer(ctypes.byref(pr.p), which, 1.0, 0.9, orc.dp(vals), orc.dp(der))
    rng = np.random.default_rng(13)
    h = 1e-5
    for idx in rng.choice(n * dim, 6, replace=False):
        j, d = idx % n, idx // n
        out = []
        for sgn in (1, -1):
            cp = case.cp.astype(np.float64).copy()
            cp[j, d] += sgn * h
            p2 = case.oracle(cp=cp)
            v = np.zeros(n)
            p2.lib.skmo_regularizer(ctypes.byref(p2.p), which, 1.0, 0.9, orc.dp(v), None)
            out.append(v.sum())
        fd = (out[0] - out[1]) / (2 * h)
        assert abs(fd - der[idx]) < 1e-6 * max(1.0, abs(fd)), (idx, fd, der[idx])


# ------------------------------------------------------------------ a11 host regularizers

def test_host_regularizers_l1_subgradient_cases():
    """itkRegularizedImageToImageMetricv4.hxx:444-462: sign for non-zero params,
    soft threshold at zero; negation at the end (:528-529)."""
    th = np.array([0.5, -0.2, 0.0, 0.0, 0.0, 0.0])
    g = np.array([0.1, 0.1, -0.3, 0.3, 0.05, -0.05])
    l1 = ctypes.c_double()
    l21 = ctypes.c_double()
    l2 = ctypes.c_double()
    reg = lib().skmo_host_regularizers(2, 3, orc.dp(th), 0.1, 0.0, 0.0, 0, orc.dp(g), ctypes.byref(l1),
                                       ctypes.byref(l21), ctypes.byref(l2))
    np.testing.assert_allclose(g, [0.2, 0.0, -0.2, 0.2, 0.0, 0.0], atol=1e-15)
    assert abs(l1.value - 0.7) < 1e-15 and abs(reg - 0.07) < 1e-15
    th = np.array([3.0, 0.0, 4.0, 0.0])
    g = np.zeros(4)
    reg = lib().skmo_host_regularizers(2, 2, orc.dp(th), 0.0, 1.0, 0.5, 1, orc.dp(g), ctypes.byref(l1),
                                       ctypes.byref(l21), ctypes.byref(l2))
    # L21: ||(3,4)|| = 5, ||(0,0)|| = 0; L2: 25; derivative -(0.5*2*theta + theta/5)
    assert abs(l21.value - 5.0) < 1e-15 and abs(l2.value - 25.0) < 1e-15 and abs(reg - 17.5) < 1e-12
    np.testing.assert_allclose(g, -np.array([3.0 + 0.6, 0.0, 4.0 + 0.8, 0.0]), atol=1e-14)


# ------------------------------------------------------------------ a13 optimizers

def test_optimizer_schedules():
    L = lib()
    assert abs(L.skmo_standard_rate(2.0, 20.0, 1.0, 4) - 2.0 / 25.0) < 1e-16
    assert abs(L.skmo_standard_rate(2.0, 20.0, 0.5, 4) - math.sqrt(2.0 / 25.0)) < 1e-16
    assert abs(L.skmo_adaptive_rate(0.1, 20.0, 1.0, 5.5) - 0.1 / 26.5) < 1e-16
    fmax, fmin, omega = 1e7, -1e6, 1.0
    # f(0) = fmin + (fmax - fmin) / (1 - fmax/fmin); large positive products -> fmin, negative -> fmax
    f0 = fmin + (fmax - fmin) / (1 - fmax / fmin)
    assert abs(L.skmo_adaptive_time(100.0, 0.0, fmax, fmin, omega) - max(0.0, 100.0 + f0)) < 1e-6
    assert L.skmo_adaptive_time(5.0, 1e3, fmax, fmin, omega) == 0.0
    assert abs(L.skmo_adaptive_time(0.0, -1e3, fmax, fmin, omega) - fmax) < 1e-3


def test_gd_update_orthant_projection_and_quadratic_bowl():
    L = lib()
    prev = np.array([1.0, -1.0, 0.5, 0.0])
    grad = np.array([-3.0, 0.5, 0.1, 2.0])
    out = np.zeros(4)
    L.skmo_gd_update(4, orc.dp(prev), orc.dp(grad), 0.5, 1, orc.dp(out))
    np.testing.assert_array_equal(out, [0.0, -0.75, 0.55, 1.0])
    # on 1/2 |theta|^2 with eta_k = 1/(k+2): theta_K = theta_0 prod (k+1)/(k+2) = theta_0 / (K+1)
    th0 = np.array([4.0, -2.0, 1.0])
    th = th0.copy()
    for k in range(200):
        g = -th
        rate = L.skmo_standard_rate(1.0, 1.0, 1.0, k)
        L.skmo_gd_update(3, orc.dp(th.copy()), orc.dp(g), rate, 0, orc.dp(th))
    np.testing.assert_allclose(th, th0 / 201.0, rtol=1e-12)


@pytest.mark.parametrize("mode,measure", [(0, 1), (2, 1), (1, 0), (2, 0)])
def test_openmp_build_equals_serial_build(mode, measure):
    """The OpenMP build of the oracle (bench.py's CPU baseline) runs the
    per-sample and per-control-point loops in parallel with the reductions
    left serial, so its results equal the serial build's exactly."""
    case = Case(dim=3, mode=mode, cp_spacing=5.0, sigma=10.0, subsample=1, seed=41)
    serial = case.oracle()
    assert orc.load(parallel=True).skmo_set_threads(4) == 4
    assert orc.load().skmo_set_threads(4) == 1
    par = case.oracle(parallel=True)
    rs, ds = serial.evaluate(measure, True, 0.0, 1e-3, 0.0)
    rp, dpar = par.evaluate(measure, True, 0.0, 1e-3, 0.0)
    assert rs == rp
    np.testing.assert_array_equal(ds, dpar)
