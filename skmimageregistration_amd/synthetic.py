"""Seeded synthetic registration workloads (numpy/scipy, host side).

Moving = reference image R, fixed = target T = R(x + d(x)) + noise, so a
registration recovers u ~ d.  Arrays are float32, C order (x fastest).
"""
import numpy as np
from scipy import ndimage


def _warp(image, disp):
    """R(x + d(x)) with cubic interpolation; disp has shape image.shape + (D,) in (x, y[, z]) order."""
    dim = image.ndim
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in image.shape], indexing="ij")
    coords = [grids[a] + disp[..., dim - 1 - a] for a in range(dim)]
    return ndimage.map_coordinates(image.astype(np.float64), coords, order=3, mode="nearest")


def _rbf_field(shape, n, sigma, max_disp, rng):
    dim = len(shape)
    grids = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    field = np.zeros(tuple(shape) + (dim,))
    for _ in range(n):
        c = [rng.uniform(0, s - 1) for s in shape]
        w = rng.uniform(-1, 1, size=dim)
        r2 = sum((g - ci) ** 2 for g, ci in zip(grids, c))
        field += np.exp(-r2 / (2 * sigma * sigma))[..., None] * w
    mag = np.sqrt((field ** 2).sum(-1)).max()
    return field * (max_disp / max(mag, 1e-12))


def blobs_2d(shape=(256, 256), n_blobs=12, max_disp=5.0, noise=0.01, seeds=(1, 2, 3)):
    """2D reference: ramp 0.2*x/(nx-1) plus Gaussian blobs (sigma U[8,20], amplitude U[0.3,1])."""
    ny, nx = shape
    r1 = np.random.default_rng(seeds[0])
    yy, xx = np.meshgrid(np.arange(ny, dtype=np.float64), np.arange(nx, dtype=np.float64), indexing="ij")
    img = 0.2 * xx / max(nx - 1, 1)
    for _ in range(n_blobs):
        s = r1.uniform(8, 20) * min(nx, ny) / 256.0
        a = r1.uniform(0.3, 1.0)
        cx, cy = r1.uniform(0, nx - 1), r1.uniform(0, ny - 1)
        img += a * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * s * s))
    disp = _rbf_field(shape, 8, 30.0 * min(nx, ny) / 256.0, max_disp, np.random.default_rng(seeds[1]))
    target = _warp(img, disp) + np.random.default_rng(seeds[2]).normal(0.0, noise, size=shape)
    return img.astype(np.float32), target.astype(np.float32), disp


def ellipsoids_3d(shape=(96, 128, 128), n=20, smooth=1.5, max_disp=6.0, noise=0.02, seeds=(11, 12, 13)):
    """3D reference: random ellipsoids (intensity U[0,1]) smoothed with a Gaussian."""
    nz, ny, nx = shape
    r1 = np.random.default_rng(seeds[0])
    zz, yy, xx = np.meshgrid(*[np.arange(s, dtype=np.float32) for s in shape], indexing="ij")
    img = np.zeros(shape, np.float32)
    for _ in range(n):
        c = [r1.uniform(0.15, 0.85) * s for s in (nx, ny, nz)]
        ax = [r1.uniform(0.05, 0.2) * s for s in (nx, ny, nz)]
        val = r1.uniform(0.0, 1.0)
        inside = ((xx - c[0]) / ax[0]) ** 2 + ((yy - c[1]) / ax[1]) ** 2 + ((zz - c[2]) / ax[2]) ** 2 <= 1.0
        img[inside] = val
    img = ndimage.gaussian_filter(img, smooth)
    disp = _rbf_field(shape, 8, 0.25 * min(shape), max_disp, np.random.default_rng(seeds[1]))
    target = _warp(img, disp) + np.random.default_rng(seeds[2]).normal(0.0, noise, size=shape)
    return img.astype(np.float32), target.astype(np.float32), disp


def sliding_3d(shape=(128, 256, 256), shift=4.0, breathing=8.0, seeds=(21, 22, 23)):
    """3D sliding motion: two halves split by the plane x = nx/2 move +-shift along y,
    plus a smooth breathing-like field along z (max `breathing`)."""
    nz, ny, nx = shape
    img, _, _ = ellipsoids_3d(shape, n=30, smooth=1.5, max_disp=0.0, noise=0.0, seeds=(seeds[0], seeds[1], seeds[2]))
    zz, yy, xx = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    disp = np.zeros(shape + (3,))
    disp[..., 1] = np.where(xx < nx / 2, shift, -shift)
    disp[..., 2] = breathing * np.sin(np.pi * zz / max(nz - 1, 1)) * np.exp(-((xx - nx / 2) ** 2 + (yy - ny / 2) ** 2) / (2 * (0.3 * nx) ** 2))
    target = _warp(img, disp) + np.random.default_rng(seeds[2]).normal(0.0, 0.02, size=shape)
    return img.astype(np.float32), target.astype(np.float32), disp


# frame sizes of the time series, (z, y, x): 96^3, 128x128x64 and 160x160x80
SERIES_SIZES = ((96, 96, 96), (64, 128, 128), (80, 160, 160))


def time_series_3d(n_frames=32, n=20, smooth=1.5, max_disp=7.0, noise=0.02, seed=31, sizes=SERIES_SIZES):
    """Time series of uneven-sized 3D frames.  One phantom of `n` random
    ellipsoids (intensity U[0,1], smoothed with a Gaussian) fills the box
    that holds every size; frame t, of a size drawn from `sizes`, samples it
    on its own grid (origin 0, spacing 1) at x + sin(2 pi t / n_frames) d(x),
    with d a smooth field of maximum length max_disp, plus N(0, noise).
    Frame 0 (phase 0) is the reference the others register to.
    Returns (frames, displacements), each frame's displacement in (x, y, z)
    order on its grid."""
    rng = np.random.default_rng(seed)
    box = tuple(max(s[a] for s in sizes) for a in range(3))
    phantom, _, _ = ellipsoids_3d(box, n=n, smooth=smooth, max_disp=0.0, noise=0.0,
                                  seeds=(seed + 1, seed + 2, seed + 3))
    coeffs = ndimage.spline_filter(phantom.astype(np.float64), order=3, mode="nearest")
    field = _rbf_field(box, 8, 0.25 * min(box), max_disp, np.random.default_rng(seed + 4))
    picks = rng.integers(0, len(sizes), size=n_frames)
    frames, disps = [], []
    for t in range(n_frames):
        shape = sizes[picks[t]]
        d = np.sin(2.0 * np.pi * t / n_frames) * field[:shape[0], :shape[1], :shape[2]]
        grids = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
        coords = [grids[a] + d[..., 2 - a] for a in range(3)]
        img = ndimage.map_coordinates(coeffs, coords, order=3, mode="nearest", prefilter=False)
        img += np.random.default_rng(seed + 100 + t).normal(0.0, noise, size=shape)
        frames.append(img.astype(np.float32))
        disps.append(d)
    return frames, disps
