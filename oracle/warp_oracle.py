"""FP64 numpy restatement of the outputs the registration driver writes:
the warped images, the Jacobian determinant of the control-point image and
the target registration error.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
library; never part of the product path.

The reference takes these from an external ITK (not in the reference tree;
its build pins ITK 4.x).  They are restated from ITK's published behaviour:

* warp_linear -- itk::ResampleImageFilter with an
  itk::DisplacementFieldTransform and the default
  itk::LinearInterpolateImageFunction, default value 0, output on the target
  grid (src/SKMReg.cxx:507-525).  The transform adds the displacement
  interpolated linearly in the field (VectorLinearInterpolateImageFunction)
  where the point lies inside the field's buffer and leaves the point
  unchanged elsewhere; the image is sampled by linear interpolation where the
  moved point lies inside the image's buffer.  Both interpolators take the
  neighbours of floor(index) and floor(index) + 1 clamped to the buffer.
  "Inside the buffer" is a continuous index in [-0.5, n - 0.5) on every axis
  (ImageFunction::IsInsideBuffer).
* warp_bspline3 -- WarpImage (include/itkUtils.h:163-179):
  itk::WarpImageFilter on the field's grid with an
  itk::BSplineInterpolateImageFunction of order 3: coefficients of the
  mirror-boundary cubic B-spline prefilter (scipy.ndimage.spline_filter,
  mode 'mirror'), taps floor(x) - 1 .. floor(x) + 2 with the uniform cubic
  B-spline weights, tap indices reflected at the edges (-i below 0,
  2(n-1) - i above n-1).
* jacobian_determinant -- itk::DisplacementFieldJacobianDeterminantFilter
  (include/itkUtils.h:153-161): det(I + J), J_ij = (v_j[x + e_i] -
  v_j[x - e_i]) * 0.5 / spacing_i with the neighbours clamped to the image.
* tre -- CalculateTRE (include/itkUtils.h:268-293): mean |m + u(m) - f| with
  u from the DisplacementFieldTransform above.

Arrays are indexed [z, y, x] (3D) or [y, x] (2D); geometry tuples are
(origin, spacing) in (x, y[, z]) order.  Vector fields carry D components on
a trailing axis.
"""
import numpy as np
from scipy import ndimage


def _cindex(points, origin, spacing):
    """Physical points (..., D) in x, y, z order -> continuous indices."""
    return (points - np.asarray(origin, np.float64)) / np.asarray(spacing, np.float64)


def _inside(ci, shape_xyz):
    n = np.asarray(shape_xyz, np.float64)
    return np.all((ci >= -0.5) & (ci < n - 0.5), axis=-1)


def _linear(data, ci):
    """Clamped-neighbour linear interpolation of data ([z, y, x] + trailing
    components) at continuous indices ci (..., D) in x, y, z order."""
    dim = ci.shape[-1]
    shape_xyz = data.shape[:dim][::-1]
    base = np.floor(ci).astype(np.int64)
    frac = ci - base
    comps = data.shape[dim:]
    out = np.zeros(ci.shape[:-1] + comps, np.float64)
    for corner in range(1 << dim):
        w = np.ones(ci.shape[:-1], np.float64)
        idx = []
        for d in range(dim):
            up = (corner >> d) & 1
            i = np.clip(base[..., d] + up, 0, shape_xyz[d] - 1)
            w = w * (frac[..., d] if up else 1.0 - frac[..., d])
            idx.append(i)
        vals = data[tuple(reversed(idx))].astype(np.float64)
        out += (w.reshape(w.shape + (1,) * len(comps))) * vals
    return out


def grid_points(shape, origin, spacing):
    """Physical points of every voxel of a grid, shape + (D,), x, y, z order."""
    dim = len(shape)
    axes = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    idx_xyz = np.stack(list(reversed(axes)), axis=-1)
    return np.asarray(origin, np.float64) + idx_xyz * np.asarray(spacing, np.float64)


def transform_points(points, df, df_geom):
    """DisplacementFieldTransform::TransformPoint for points (..., D)."""
    dim = points.shape[-1]
    ci = _cindex(points, *df_geom)
    inside = _inside(ci, df.shape[:dim][::-1])
    u = _linear(df, np.where(inside[..., None], ci, 0.0))
    return points + np.where(inside[..., None], u, 0.0)


def warp_linear(image, image_geom, df, df_geom, out_shape, out_geom, default=0.0):
    pts = transform_points(grid_points(out_shape, *out_geom), df, df_geom)
    dim = len(out_shape)
    ci = _cindex(pts, *image_geom)
    inside = _inside(ci, image.shape[::-1])
    vals = _linear(image, np.where(inside[..., None], ci, 0.0))
    return np.where(inside, vals, default)


def _mirror(i, n):
    if n == 1:
        return np.zeros_like(i)
    i = np.where(i < 0, -i, i)
    i = np.where(i > n - 1, 2 * (n - 1) - i, i)
    return np.clip(i, 0, n - 1)


def cubic_weights(t):
    """Uniform cubic B-spline weights of the taps floor(x)-1 .. floor(x)+2."""
    s = 1.0 - t
    return np.stack([s ** 3 / 6.0, (3 * t ** 3 - 6 * t ** 2 + 4) / 6.0, (-3 * t ** 3 + 3 * t ** 2 + 3 * t + 1) / 6.0,
                     t ** 3 / 6.0], axis=-1)


def bspline3_at(coeffs, ci):
    """Cubic B-spline value at continuous indices ci (..., D) from
    mirror-boundary coefficients."""
    dim = ci.shape[-1]
    shape_xyz = coeffs.shape[::-1]
    fl = np.floor(ci)
    w = cubic_weights(ci - fl)
    base = fl.astype(np.int64) - 1
    out = np.zeros(ci.shape[:-1], np.float64)
    for taps in np.ndindex(*(4,) * dim):
        wt = np.ones(ci.shape[:-1], np.float64)
        idx = []
        for d in range(dim):
            wt = wt * w[..., d, taps[d]]
            idx.append(_mirror(base[..., d] + taps[d], shape_xyz[d]))
        out += wt * coeffs[tuple(reversed(idx))]
    return out


def bspline3_coefficients(image):
    return ndimage.spline_filter(np.asarray(image, np.float64), order=3, mode="mirror")


def warp_bspline3(image, image_geom, df, df_geom, default=0.0):
    dim = image.ndim
    pts = grid_points(df.shape[:dim], *df_geom) + df.astype(np.float64)
    ci = _cindex(pts, *image_geom)
    inside = _inside(ci, image.shape[::-1])
    vals = bspline3_at(bspline3_coefficients(image), np.where(inside[..., None], ci, 0.0))
    return np.where(inside, vals, default)


def jacobian_determinant(v, spacing):
    """det(I + J) of a vector image v ([z, y, x] + (D,))."""
    dim = v.shape[-1]
    J = np.zeros(v.shape[:-1] + (dim, dim), np.float64)
    vv = v.astype(np.float64)
    for a in range(dim):
        ax = dim - 1 - a                     # array axis of x, y, z
        n = v.shape[ax]
        hi = np.take(vv, np.minimum(np.arange(n) + 1, n - 1), axis=ax)
        lo = np.take(vv, np.maximum(np.arange(n) - 1, 0), axis=ax)
        J[..., a, :] = 0.5 / spacing[a] * (hi - lo)
        J[..., a, a] += 1.0
    return np.linalg.det(J)


def tre(df, df_geom, fixed_points, moving_points):
    f = np.asarray(fixed_points, np.float64)
    m = np.asarray(moving_points, np.float64)
    if len(f) == 0 or len(f) != len(m):
        return -1.0
    moved = transform_points(m, df, df_geom)
    return float(np.mean(np.linalg.norm(moved - f, axis=1)))
