"""Pins of the tensor-preprocessing oracle (oracle/tensor_oracle.py) against
facts fixed outside its own code: scipy's modified Bessel functions and the
error bounds Abramowitz & Stegun publish for 9.8.1-9.8.4, the exact discrete
Gaussian exp(-t) I_n(t), scipy.ndimage.correlate1d with edge replication,
central differences, numpy's symmetric eigensolver and the closed form of the
covariance the reference assembles.

The reference holds no test or fixture for this path (its test/ directory
only has itkFiniteDifferenceTest.cpp), so these properties are the pins.
"""
import math

import numpy as np
import pytest
from scipy import ndimage, special

import tensor_oracle as tor


# ------------------------------------------------------------ Gaussian kernel

@pytest.mark.parametrize("lo,hi,bound0,bound1", [
    (0.01, 3.74, 1.6e-7, 8e-9 * 2.0),       # 9.8.1 |eps| < 1.6e-7 on I0 >= 1; 9.8.3 |eps| < 8e-9 on I1/x >= 0.5
    (3.75, 300.0, 1.9e-7 / 0.39, 2.2e-7 / 0.39),   # 9.8.2 / 9.8.4 on sqrt(x) e^-x I(x) >= 0.39
])
def test_bessel_within_published_bounds(lo, hi, bound0, bound1):
    xs = np.linspace(lo, hi, 997)
    e0 = max(abs(tor.bessel_i0(x) / special.i0(x) - 1.0) for x in xs)
    e1 = max(abs(tor.bessel_i1(x) / special.i1(x) - 1.0) for x in xs)
    assert e0 < bound0 and e1 < bound1
    assert tor.bessel_i1(-2.0) == -tor.bessel_i1(2.0)


def _exact_radius(t, cap=0.99):
    r = 1
    while special.ive(0, t) + 2 * sum(special.ive(n, t) for n in range(1, r + 1)) < cap:
        r += 1
    return r


@pytest.mark.parametrize("t", [0.3, 1.0, 1.78, 4.0, 7.11, 16.0, 64.0])
def test_gaussian_coefficients_match_exact_discrete_gaussian(t):
    c = tor.gaussian_coefficients(t)
    r = (len(c) - 1) // 2
    assert len(c) % 2 == 1
    np.testing.assert_array_equal(c, c[::-1])
    assert abs(c.sum() - 1.0) < 1e-12
    assert r == _exact_radius(t)
    exact = special.ive(np.abs(np.arange(-r, r + 1)), t)
    np.testing.assert_allclose(c, exact / exact.sum(), rtol=0, atol=2e-6)
    # discrete Gaussian of variance t, truncated at 99 % of its mass: the
    # second moment is below t but close to it
    m2 = float(np.sum(np.arange(-r, r + 1) ** 2 * c))
    assert 0.85 * t < m2 <= t * 1.0000001


def test_gaussian_coefficients_truncated_at_maximum_width():
    c = tor.gaussian_coefficients(650.0)
    assert len(c) == 2 * 64 + 1                 # half kernel of 65 entries ends the loop
    assert abs(c.sum() - 1.0) < 1e-12
    exact = special.ive(np.arange(0, 65), 650.0)
    assert exact[0] + 2 * exact[1:].sum() < 0.99  # the width limit, not the error cap, ended it
    exact = exact / (exact[0] + 2 * exact[1:].sum())
    np.testing.assert_allclose(c[64:], exact, rtol=0, atol=2e-6)
    # exp(t) overflows beyond t ~ 709.8: exp(-t) * inf is NaN, as in C
    assert np.isnan(tor.gaussian_coefficients(800.0)).all()


def test_gaussian_smooth_matches_correlate1d_nearest():
    rng = np.random.default_rng(3)
    img = rng.normal(size=(9, 13, 17)).astype(np.float32)
    spacing = (1.5, 1.0, 2.0)                   # x, y, z
    out = tor.gaussian_smooth(img, spacing, 6.0)
    ref = img.astype(np.float64)
    for d in (2, 1, 0):
        c = tor.gaussian_coefficients(6.0 / spacing[d] ** 2)
        ref = ndimage.correlate1d(ref, c, axis=2 - d, mode="nearest").astype(np.float32).astype(np.float64)
    np.testing.assert_allclose(out, ref, rtol=2e-6, atol=2e-6)


def test_gaussian_smooth_impulse_and_constant():
    spacing = (1.0, 2.0, 0.5)
    img = np.zeros((41, 41, 41), np.float32)
    img[20, 20, 20] = 1.0
    out = tor.gaussian_smooth(img, spacing, 2.0)
    cx, cy, cz = (tor.gaussian_coefficients(2.0 / s ** 2) for s in spacing)
    expect = np.zeros_like(out, dtype=np.float64)
    rz, ry, rx = (len(cz) - 1) // 2, (len(cy) - 1) // 2, (len(cx) - 1) // 2
    expect[20 - rz:21 + rz, 20 - ry:21 + ry, 20 - rx:21 + rx] = np.einsum("k,j,i->kji", cz, cy, cx)
    np.testing.assert_allclose(out, expect, rtol=0, atol=1e-7)
    const = np.full((7, 8, 9), 3.25, np.float32)
    np.testing.assert_allclose(tor.gaussian_smooth(const, spacing, 5.0), const, rtol=1e-6)


def test_gradient_central_differences_with_edge_replication():
    rng = np.random.default_rng(4)
    img = rng.normal(size=(6, 7, 8)).astype(np.float32)
    spacing = (0.5, 2.0, 1.25)
    g = tor.gradient(img, spacing)
    f = img.astype(np.float64)
    for d in range(3):
        axis = 2 - d
        num = np.gradient(f, spacing[d], axis=axis)        # central in the interior
        sl = [slice(None)] * 3
        sl[axis] = slice(1, -1)
        np.testing.assert_allclose(g[d][tuple(sl)], num[tuple(sl)], rtol=1e-12, atol=1e-12)
        first = [slice(None)] * 3
        first[axis] = 0
        second = [slice(None)] * 3
        second[axis] = 1
        np.testing.assert_allclose(g[d][tuple(first)], (f[tuple(second)] - f[tuple(first)]) / (2 * spacing[d]),
                                   rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------ eigen / tensors

def test_eigen_rows_ordered_by_magnitude():
    t = np.diag([3.0, -5.0, 1.0])[None]
    rows, lam = tor.eigen_rows_by_magnitude(t)
    np.testing.assert_allclose(lam[0], [1.0, 3.0, -5.0])
    np.testing.assert_allclose(np.abs(rows[0]), np.abs(np.array([[0, 0, 1.0], [3.0, 0, 0], [0, 5.0, 0]])))


def _closed_form(t, alignment):
    """cov = a e1 e1^T + c (I - e1 e1^T): the reference's U S S^T U^T has the
    columns e1/(1+s), c v_mid, c v_min, i.e. eigenvalues a = 1/(1+s) and c = 2 - a."""
    dim = t.shape[-1]
    lam, vec = np.linalg.eigh(t)
    k = int(np.argmax(np.abs(lam)))
    e1 = vec[:, k]
    factor = 1.0
    if 0 <= alignment < dim:
        md = np.eye(dim)[alignment]
        cr = np.linalg.norm(np.cross(e1, md)) if dim == 3 else abs(e1[0] * md[1] - e1[1] * md[0])
        factor = (cr + 1e-10) ** 2 / math.sqrt(dim)
    rs = abs(lam[k]) * factor
    s = rs ** 1.5
    a = 1.0 / (1.0 + s)
    c = 1.0 + s / (1.0 + s)
    return c * np.eye(dim) + (a - c) * np.outer(e1, e1), rs


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("alignment", [-1, 0, 1, 2])
def test_covariance_closed_form(dim, alignment):
    if alignment >= dim:
        pytest.skip("alignment outside the dimension")
    rng = np.random.default_rng(10 + dim + alignment)
    ts = []
    for _ in range(40):
        q, _ = np.linalg.qr(rng.normal(size=(dim, dim)))
        lam = rng.uniform(0.05, 8.0, size=dim) ** 2
        ts.append(q @ np.diag(lam) @ q.T)
    t = np.array(ts)
    dirs, maximum = tor.directions(t, alignment)
    cov, ok = tor.covariance(dirs)
    assert ok
    worst = 0.0
    tilts = []
    for v in range(len(t)):
        expect, rs = _closed_form(t[v], alignment)
        tilts.append(rs)
        worst = max(worst, np.max(np.abs(cov[v] - expect)))
        ev = np.linalg.eigvalsh(cov[v])
        assert ev.min() > 0
    assert worst < 1e-12
    assert abs(maximum - max(tilts)) < 1e-12 * max(tilts)


def test_degenerate_tensors_give_identity():
    zero = np.zeros((3, 3))
    rank1 = np.outer([1.0, 2.0, 3.0], [1.0, 2.0, 3.0])
    tiny = np.diag([5.0, 2.0, 1e-9])
    huge = np.diag([1e12, 2.0, 1.0])   # a = 1/(1+s) < 1e-8: covariance branch falls back to identity
    t = np.array([zero, rank1, tiny, huge])
    dirs, maximum = tor.directions(t, -1)
    cov, ok = tor.covariance(dirs)
    assert ok
    for v in range(4):
        np.testing.assert_allclose(cov[v], np.eye(3), atol=1e-12)
    assert maximum == pytest.approx(1e12)        # the huge voxel took the re-weighting branch


def test_constant_and_ramp_images_give_identity():
    const = np.full((6, 7, 8), 2.0, np.float32)
    ten, tilt, mw, mcov = tor.anisotropic_tensor(const, (1.0, 1.0, 1.0), 4.0, 400.0, 2)
    np.testing.assert_array_equal(ten, np.broadcast_to(np.eye(3, dtype=np.float32), ten.shape))
    assert tilt == tor.LOWEST and mw == 1.0
    np.testing.assert_array_equal(mcov, np.eye(3))
    ramp = np.broadcast_to(np.arange(8, dtype=np.float32) * 0.3, (6, 7, 8)).copy()
    ten, tilt, _, _ = tor.anisotropic_tensor(ramp, (1.0, 1.0, 1.0), 4.0, 400.0, 2)
    np.testing.assert_array_equal(ten, np.broadcast_to(np.eye(3, dtype=np.float32), ten.shape))
    assert tilt == tor.LOWEST


def test_set_weighting_tensor():
    ten = np.broadcast_to(np.eye(3), (4, 5, 6, 3, 3)).copy()
    ten[2, 3, 1] = np.diag([1.0, 2.5, 1.0])
    mw, mcov = tor.set_weighting_tensor(ten, tor.LOWEST)
    assert mw == pytest.approx(2.5)
    np.testing.assert_allclose(mcov, np.diag([1.0, 2.5, 1.0]))
    mw, mcov = tor.set_weighting_tensor(ten, 7.5)
    assert mw == 7.5
    np.testing.assert_array_equal(mcov, np.eye(3))


def test_anisotropic_tensor_structure_on_blobs():
    """End to end on a small 3D image: symmetric positive definite output with
    eigenvalues {a, 2 - a, 2 - a} per voxel and the tilt equal to the
    largest |lambda_max| * factor over re-weighted voxels."""
    rng = np.random.default_rng(5)
    z, y, x = np.meshgrid(np.arange(10), np.arange(12), np.arange(14), indexing="ij")
    img = (np.exp(-((x - 6.0) ** 2 + (y - 5.0) ** 2 + (z - 4.0) ** 2) / 18.0)
           + 0.02 * rng.normal(size=x.shape)).astype(np.float32)
    spacing = (1.0, 1.5, 2.0)
    ten, tilt, mw, mcov = tor.anisotropic_tensor(img, spacing, 8.0, 50.0, 2)
    assert tilt > 0 and mw == tilt
    np.testing.assert_array_equal(mcov, np.eye(3))
    flat = ten.reshape(-1, 3, 3).astype(np.float64)
    np.testing.assert_allclose(flat, np.swapaxes(flat, 1, 2), atol=1e-7)
    ev = np.linalg.eigvalsh(flat)
    assert ev.min() > 0
    # two equal eigenvalues c and one a with a + c = 2
    np.testing.assert_allclose(ev[:, 0] + ev[:, 2], 2.0, atol=1e-5)
    st = tor.structure_tensor(img, spacing, 8.0, 50.0).reshape(-1, 3, 3)
    expect = max(_closed_form(t.astype(np.float64), 2)[1] for t in st)
    assert tilt == pytest.approx(expect, rel=1e-12)
