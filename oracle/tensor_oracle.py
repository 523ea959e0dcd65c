"""FP64 numpy restatement of the reference's image preprocessing for the
anisotropic kernel: Gaussian smoothing (include/itkUtils.h:140-151) and
ComputeAnisotropicTensor (include/AnisotropicTensorUtils.h:37-355), followed by
the maximum-tilt logic of TransformAdapter::SetWeightingTensor
(include/itkTransformAdapter.hxx:335-366).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
library; never part of the product path.

The reference takes the filters from an external ITK (not in the reference
tree; its own build pins ITK 4.x).  They are restated here from ITK's
published algorithms:

* itk::GaussianOperator::GenerateCoefficients -- discrete Gaussian kernel
  T(n, t) = exp(-t) I_n(t) for the variance t in pixels.  I_0 and I_1 come
  from the polynomial approximations of Abramowitz & Stegun 9.8.1-9.8.4,
  higher orders from the recurrence I_{n+1} = I_{n-1} - (2n / t) I_n.  The
  half kernel grows until the two-sided sum reaches 1 - maximum_error (0.01)
  or a coefficient is not positive or the half kernel holds more than
  maximum_kernel_width (64) entries; it is then normalised by that sum.
* itk::DiscreteGaussianImageFilter -- separable passes, the last axis first,
  per-axis variance = variance / spacing^2 (UseImageSpacing), zero-flux
  Neumann boundary (edge replication), sums in double, images stored in the
  pixel type (float) between passes.
* itk::GradientImageFilter -- (0.5 f[i+1] - 0.5 f[i-1]) / spacing with edge
  replication, in double.
* itk::SymmetricEigenAnalysis ordered by magnitude -- eigenvalues sorted by
  |lambda| ascending, eigenvectors as rows (numpy.linalg.eigh here).

Arrays are numpy arrays indexed [z, y, x] (3D) or [y, x] (2D), x fastest,
matching the float images of the C ABI.  The per-voxel tensor output has
shape image.shape + (D, D).
"""
import math

import numpy as np

MAXIMUM_ERROR = 0.01       # itk::DiscreteGaussianImageFilter default
MAXIMUM_KERNEL_WIDTH = 64  # GaussianSmoothing sets it (itkUtils.h:148)
LOWEST = -np.finfo(np.float64).max  # std::numeric_limits<double>::lowest()


def _exp(x):
    """exp with C overflow semantics (inf instead of an exception)."""
    return math.exp(x) if x < 709.78 else math.inf


def bessel_i0(y):
    """Modified Bessel I_0, Abramowitz & Stegun 9.8.1 / 9.8.2."""
    d = abs(y)
    if d < 3.75:
        m = (y / 3.75) ** 2
        return 1.0 + m * (3.5156229 + m * (3.0899424 + m * (1.2067492 + m * (0.2659732 + m * (0.360768e-1
                                                                                            + m * 0.45813e-2)))))
    m = 3.75 / d
    poly = 0.39894228 + m * (0.1328592e-1 + m * (0.225319e-2 + m * (-0.157565e-2 + m * (0.916281e-2 + m * (
        -0.2057706e-1 + m * (0.2635537e-1 + m * (-0.1647633e-1 + m * 0.392377e-2)))))))
    return _exp(d) / math.sqrt(d) * poly


def bessel_i1(y):
    """Modified Bessel I_1, Abramowitz & Stegun 9.8.3 / 9.8.4."""
    d = abs(y)
    if d < 3.75:
        m = (y / 3.75) ** 2
        acc = d * (0.5 + m * (0.87890594 + m * (0.51498869 + m * (0.15084934 + m * (0.2658733e-1 + m * (
            0.301532e-2 + m * 0.32411e-3))))))
    else:
        m = 3.75 / d
        acc = 0.2282967e-1 + m * (-0.2895312e-1 + m * (0.1787654e-1 - m * 0.420059e-2))
        acc = 0.39894228 + m * (-0.3988024e-1 + m * (-0.362018e-2 + m * (0.163801e-2 + m * (-0.1031555e-1
                                                                                           + m * acc))))
        acc *= _exp(d) / math.sqrt(d)
    return -acc if y < 0 else acc


def gaussian_coefficients(variance, maximum_error=MAXIMUM_ERROR, maximum_width=MAXIMUM_KERNEL_WIDTH):
    """Symmetric kernel (length 2r+1) of itk::GaussianOperator for a variance in pixels."""
    et = math.exp(-variance)
    cap = 1.0 - maximum_error
    half = [et * bessel_i0(variance), et * bessel_i1(variance)]
    total = half[0] + 2.0 * half[1]
    i = 2
    while total < cap:
        half.append(half[i - 2] - 2.0 * (i - 1) * half[i - 1] / variance)
        total += 2.0 * half[i]
        if half[i] <= 0.0:
            break
        if len(half) > maximum_width:
            break
        i += 1
    half = [c / total for c in half]
    return np.array(half[:0:-1] + half, dtype=np.float64)


def _smooth_axis(img, axis, coeffs):
    """One NeighborhoodOperatorImageFilter pass: double sums, float result."""
    r = (len(coeffs) - 1) // 2
    n = img.shape[axis]
    src = img.astype(np.float64)
    acc = np.zeros(img.shape, dtype=np.float64)
    base = np.arange(n)
    for k in range(2 * r + 1):
        idx = np.clip(base + (k - r), 0, n - 1)
        acc += coeffs[k] * np.take(src, idx, axis=axis)
    return acc.astype(np.float32)


def gaussian_smooth(image, spacing, variance):
    """GaussianSmoothing(image, variance) (itkUtils.h:140-151); variance in physical units."""
    img = np.asarray(image, dtype=np.float32)
    dim = img.ndim
    for d in reversed(range(dim)):          # z (or y in 2D) first, x last
        c = gaussian_coefficients(variance / (spacing[d] * spacing[d]))
        img = _smooth_axis(img, dim - 1 - d, c)
    return img


def gradient(image, spacing):
    """itk::GradientImageFilter with image spacing: list of D double arrays (x first)."""
    f = np.asarray(image, dtype=np.float32).astype(np.float64)
    dim = f.ndim
    out = []
    for d in range(dim):
        axis = dim - 1 - d
        n = f.shape[axis]
        up = np.take(f, np.clip(np.arange(n) + 1, 0, n - 1), axis=axis)
        dn = np.take(f, np.clip(np.arange(n) - 1, 0, n - 1), axis=axis)
        out.append((0.5 * up - 0.5 * dn) * (1.0 / spacing[d]))
    return out


def structure_tensor(image, spacing, sigma, weight):
    """Smoothed structure tensor of ComputeAnisotropicTensor
    (AnisotropicTensorUtils.h:54-150): float array image.shape + (D, D)."""
    img = np.asarray(image, dtype=np.float32)
    dim = img.ndim
    scaled = (img * np.float32(weight)).astype(np.float32)         # MultiplyConstant, float
    smoothed = gaussian_smooth(scaled, spacing, sigma)              # variance = sigma
    g = gradient(smoothed, spacing)
    t = np.zeros(img.shape + (dim, dim), dtype=np.float32)
    for i in range(dim):
        for j in range(dim):
            comp = (g[i] * g[j]).astype(np.float32)                # outer product stored as float
            t[..., i, j] = gaussian_smooth(comp, spacing, sigma / 4.0)
    return t


def eigen_rows_by_magnitude(t):
    """SymmetricEigenDecompositionImageFilter ordered by magnitude: rows are
    eigenvectors times eigenvalues, |lambda| ascending
    (itkSymmetricEigenDecompositionImageFilter.h:54-69)."""
    lam, vec = np.linalg.eigh(t.astype(np.float64))
    order = np.argsort(np.abs(lam), axis=-1, kind="stable")
    lam = np.take_along_axis(lam, order, axis=-1)
    vec = np.take_along_axis(vec, order[..., None, :], axis=-1)
    rows = np.swapaxes(vec, -1, -2) * lam[..., :, None]
    return rows, lam


def directions(t, alignment_direction):
    """Re-weighted eigen-directions and the maximum tilt
    (AnisotropicTensorUtils.h:177-300).  Returns (M, max_reverse_scale or None)."""
    dim = t.shape[-1]
    rows, _ = eigen_rows_by_magnitude(t)
    flat = rows.reshape(-1, dim, dim).copy()
    md = np.zeros(dim)
    if 0 <= alignment_direction < dim:
        md[alignment_direction] = 1.0
    out = np.zeros_like(flat)
    maximum = None
    for v in range(flat.shape[0]):
        m = flat[v]
        mags = [np.linalg.norm(m[r]) for r in range(dim)]
        if all(mg > 1e-8 for mg in mags):
            top = dim - 1
            normalized = m[top] / mags[top]
            if 0 <= alignment_direction < dim:
                if dim == 3:
                    c = np.linalg.norm(np.cross(normalized, md))
                else:
                    c = abs(normalized[0] * md[1] - normalized[1] * md[0])
                factor = (abs(c) + 1e-10) ** 2 / math.sqrt(dim)
                m[top] = m[top] * factor
            reverse_scale = np.linalg.norm(m[top])
            scaler = reverse_scale ** 1.5
            if maximum is None or maximum < reverse_scale:
                maximum = reverse_scale
            m[top] = normalized / (1.0 + scaler)
            for r in range(top):
                m[r] = m[r] / mags[r] * (1.0 + reverse_scale ** 1.5 / (1.0 + reverse_scale ** 1.5))
            out[v] = np.flipud(m).T
        else:
            out[v] = np.eye(dim)
    return out.reshape(t.shape), maximum


def covariance(directions_field):
    """Covariance U S S U^T from the direction columns (AnisotropicTensorUtils.h:311-355).
    Returns (cov, ok) where ok is False if a determinant is not positive."""
    dim = directions_field.shape[-1]
    flat = directions_field.reshape(-1, dim, dim)
    cov = np.zeros_like(flat)
    ok = True
    for v in range(flat.shape[0]):
        m = flat[v]
        if np.linalg.norm(m[:, 0]) < 1e-8:
            c = np.eye(dim)
        else:
            u = np.zeros((dim, dim))
            s = np.zeros((dim, dim))
            for d in range(dim):
                nrm = np.linalg.norm(m[:, d])
                u[:, d] = m[:, d] / nrm
                s[d, d] = math.sqrt(nrm)
            c = u @ s @ s @ u.T
        if np.linalg.det(c) <= 0:
            ok = False
        cov[v] = c
    return cov.reshape(directions_field.shape), ok


def set_weighting_tensor(tensor, maximum_tilt):
    """TransformAdapter::SetWeightingTensor (itkTransformAdapter.hxx:335-366):
    returns (maximum_weight, maximum_eigenvalue_covariance)."""
    dim = tensor.shape[-1]
    if maximum_tilt >= 0:
        return float(maximum_tilt), np.eye(dim)
    maximum = -np.inf
    best = np.eye(dim)
    for m in tensor.reshape(-1, dim, dim).astype(np.float64):
        top = abs(np.linalg.eigvalsh(m)[-1])
        if top > maximum:
            maximum, best = top, m
    return float(maximum), best


def anisotropic_tensor(image, spacing, sigma, weight, alignment_direction=2):
    """ComputeAnisotropicTensor(image, sigma, weight, alignment, tilt, false)
    (AnisotropicTensorUtils.h:37-355) as SKMReg calls it (src/SKMReg.cxx:271-272).

    Returns (tensor float32 image.shape + (D, D), maximum_tilt, maximum_weight,
    maximum_eigenvalue_covariance).  maximum_tilt is the lowest double when no
    voxel has D non-vanishing eigenvalues, as in the reference."""
    t = structure_tensor(image, spacing, sigma, weight)
    dirs, maximum = directions(t, alignment_direction)
    cov, ok = covariance(dirs)
    if not ok:
        raise ValueError("Matrix is not positive definite")
    tilt = LOWEST if maximum is None else float(maximum)
    cov32 = cov.astype(np.float32)
    mw, mcov = set_weighting_tensor(cov32, tilt)
    return cov32, tilt, mw, mcov
