"""Pins of the warp / Jacobian / TRE oracle (oracle/warp_oracle.py) against
facts fixed outside its own code: scipy.ndimage.map_coordinates (linear with
edge replication, cubic B-spline with the mirror boundary), the
interpolation property of the cubic B-spline at the nodes, exact
reproduction of affine functions by linear interpolation, closed forms of
det(I + J) for linear fields and of the TRE for constant fields.

The reference holds no test or fixture for these outputs (its test/
directory only has itkFiniteDifferenceTest.cpp), so these properties are
the pins.
"""
import numpy as np
import pytest
from scipy import ndimage

import warp_oracle as wo


def _smooth(shape, seed):
    rng = np.random.default_rng(seed)
    return ndimage.gaussian_filter(rng.normal(size=shape), 2.0)


@pytest.mark.parametrize("shape", [(9, 11, 13), (17, 12)])
def test_zero_field_returns_the_image(shape):
    img = _smooth(shape, 1)
    dim = len(shape)
    geom = ((0.5,) * dim, (1.5,) * dim)
    df = np.zeros(shape + (dim,))
    out = wo.warp_linear(img, geom, df, geom, shape, geom)
    np.testing.assert_allclose(out, img, rtol=0, atol=1e-12)
    out3 = wo.warp_bspline3(img, geom, df, geom)
    np.testing.assert_allclose(out3, img, rtol=0, atol=1e-10)   # B-spline interpolation property


def test_integer_shift_and_half_voxel_buffer():
    img = _smooth((6, 7, 8), 2)
    geom = ((0.0,) * 3, (1.0,) * 3)
    df = np.zeros((6, 7, 8, 3))
    df[..., 0] = 2.0                       # x + 2
    out = wo.warp_linear(img, geom, df, geom, (6, 7, 8), geom)
    np.testing.assert_allclose(out[:, :, :6], img[:, :, 2:], atol=1e-12)
    assert np.all(out[:, :, 6:] == 0.0)    # index 8, 9 >= n - 0.5
    df[..., 0] = 0.4                       # last column moves to index 7.4 < 7.5: clamped, inside
    out = wo.warp_linear(img, geom, df, geom, (6, 7, 8), geom)
    np.testing.assert_allclose(out[:, :, 7], img[:, :, 7], atol=1e-12)
    df[..., 0] = -0.6                      # first column at -0.6 < -0.5: outside
    out = wo.warp_linear(img, geom, df, geom, (6, 7, 8), geom)
    assert np.all(out[:, :, 0] == 0.0)


def test_linear_warp_matches_map_coordinates_nearest():
    rng = np.random.default_rng(3)
    shape = (10, 12, 14)
    img = _smooth(shape, 3)
    geom = ((0.0,) * 3, (1.0,) * 3)
    df = rng.uniform(-2.0, 2.0, size=shape + (3,))
    out = wo.warp_linear(img, geom, df, geom, shape, geom)
    pts = wo.grid_points(shape, *geom) + df
    inside = np.all((pts >= -0.5) & (pts < np.array(shape[::-1]) - 0.5), axis=-1)
    coords = np.stack([pts[..., 2], pts[..., 1], pts[..., 0]])
    ref = ndimage.map_coordinates(img, coords, order=1, mode="nearest")
    np.testing.assert_allclose(out[inside], ref[inside], rtol=0, atol=1e-12)
    assert np.all(out[~inside] == 0.0)


def test_linear_warp_reproduces_affine_functions_and_interpolates_the_field():
    shape = (12, 13, 14)
    geom_img = ((1.0, -2.0, 0.5), (0.9, 1.1, 1.3))
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    px, py, pz = 1.0 + 0.9 * x, -2.0 + 1.1 * y, 0.5 + 1.3 * z
    img = 0.3 * px - 0.2 * py + 0.7 * pz + 1.0
    # field on a coarser grid, output on the image grid: u interpolated, affine in x
    df_shape = (7, 8, 9)
    geom_df = ((0.0, -3.0, -1.0), (2.5, 2.0, 2.4))
    pd = wo.grid_points(df_shape, *geom_df)
    df = np.stack([0.1 * pd[..., 1], -0.05 * pd[..., 2], 0.02 * pd[..., 0]], axis=-1)
    out = wo.warp_linear(img, geom_img, df, geom_df, shape, geom_img)
    p = wo.grid_points(shape, *geom_img)
    u = np.stack([0.1 * p[..., 1], -0.05 * p[..., 2], 0.02 * p[..., 0]], axis=-1)
    q = p + u
    expect = 0.3 * q[..., 0] - 0.2 * q[..., 1] + 0.7 * q[..., 2] + 1.0
    # interior: the field covers the point and the moved point stays >= 1 voxel inside the image
    ci_df = (p - np.array(geom_df[0])) / np.array(geom_df[1])
    ci_img = (q - np.array(geom_img[0])) / np.array(geom_img[1])
    ok = np.all((ci_df >= 0) & (ci_df <= np.array(df_shape[::-1]) - 1), axis=-1) & \
        np.all((ci_img >= 0) & (ci_img <= np.array(shape[::-1]) - 1), axis=-1)
    assert ok.sum() > 500
    np.testing.assert_allclose(out[ok], expect[ok], rtol=0, atol=1e-10)


@pytest.mark.parametrize("shape", [(8, 9, 10), (15, 11)])
def test_bspline3_matches_map_coordinates_mirror(shape):
    rng = np.random.default_rng(4)
    img = _smooth(shape, 4)
    dim = len(shape)
    ci = np.stack([rng.uniform(-0.5, n - 0.5 - 1e-9, size=500) for n in shape[::-1]], axis=-1)
    got = wo.bspline3_at(wo.bspline3_coefficients(img), ci)
    coords = np.stack([ci[:, d] for d in reversed(range(dim))])
    ref = ndimage.map_coordinates(img, coords, order=3, mode="mirror", prefilter=True)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10)


def test_cubic_weights_partition_of_unity_and_known_values():
    t = np.linspace(0.0, 1.0, 11)
    w = wo.cubic_weights(t)
    np.testing.assert_allclose(w.sum(axis=-1), 1.0, atol=1e-15)
    np.testing.assert_allclose(w[0], [1 / 6, 4 / 6, 1 / 6, 0.0], atol=1e-15)
    np.testing.assert_allclose(w[5], [1 / 48, 23 / 48, 23 / 48, 1 / 48], atol=1e-15)
    # first moment reproduces the position (linear precision): sum w_k (k - 1) = t
    np.testing.assert_allclose(w @ np.array([-1.0, 0.0, 1.0, 2.0]), t, atol=1e-15)


def test_jacobian_determinant_of_linear_fields():
    shape = (6, 7, 8)
    sp = (0.5, 1.0, 2.0)
    A = np.array([[0.1, 0.02, -0.03], [0.0, -0.2, 0.05], [0.04, 0.01, 0.3]])
    p = wo.grid_points(shape, (0.0,) * 3, sp)
    v = p @ A.T                            # v_c = sum_a A[c, a] x_a
    det = wo.jacobian_determinant(v, sp)
    np.testing.assert_allclose(det[1:-1, 1:-1, 1:-1], np.linalg.det(np.eye(3) + A), rtol=1e-12)
    # on a face the clamped neighbour halves that axis' derivative
    B = A.copy()
    B[:, 0] *= 0.5
    np.testing.assert_allclose(det[3, 3, 0], np.linalg.det(np.eye(3) + B), rtol=1e-12)
    assert np.all(wo.jacobian_determinant(np.zeros(shape + (3,)), sp) == 1.0)


def test_tre_closed_forms():
    shape = (5, 6, 7)
    geom = ((0.0,) * 3, (1.0,) * 3)
    df = np.zeros(shape + (3,))
    df[...] = [0.5, -1.0, 0.25]
    f = np.array([[1.0, 1.0, 1.0], [2.0, 3.0, 1.0]])
    m = np.array([[1.0, 2.0, 1.0], [2.5, 4.0, 0.5]])
    expect = np.mean(np.linalg.norm(m + [0.5, -1.0, 0.25] - f, axis=1))
    assert wo.tre(df, geom, f, m) == pytest.approx(expect, rel=1e-14)
    # a point outside the field's buffer stays where it is
    m_out = np.array([[100.0, 0.0, 0.0]])
    assert wo.tre(df, geom, np.zeros((1, 3)), m_out) == pytest.approx(100.0)
    assert wo.tre(df, geom, np.zeros((0, 3)), np.zeros((0, 3))) == -1.0
