"""Independent brute-force LCC value in numpy (test infrastructure).

Written from the mathematics of the reference's LCC value
(include/cuda/GpuEvaluatorLCC.cu:28-35 neighborhoodWeight, :38-204
evaluateValueLCC, host sum include/cuda/GpuEvaluator.cu:263-298), not from the
oracle's loops:

* the warp u(y) is the kernel sum over ALL control points (the reference's
  truncating cp window is a superset of the kernel's support whenever
  max_weight bounds the weight image, so both sums are equal);
* a centre x's neighbours are the fixed-grid indices i with
  idx(x) - s_x/h < i <= idx(x) + s_x/h on every axis (the truncating box of
  the reference, clamped to the image), taken on the lattice that starts at
  the box's LOWER bound with stride `subsample_neighborhood`;
* kw(x, y) = (1 - (|x - y| / s_y)^10)_+^2 with the NEIGHBOUR's support s_y
  (the shadowed `support` local of the reference);
* cc(x) = -cov^2 / (var_F var_M) over the included neighbours (kw > 0 and the
  warped neighbour inside the moving image), only when both variances are
  positive, else 0; here the sums are centred two-pass sums;
* E = sum_x cc(x) / N_in, N_in = centres whose own warped point is inside.

The cubic B-spline moving image uses scipy's mirror-boundary prefilter and
the standard cubic B-spline weights with the reference's truncating base
index and [-0.5, n-0.5) inside test (include/cuda/BsplineImage.h:229-291).
"""
import math

import numpy as np
from scipy import ndimage


def wendland_c4(t):
    f = max(0.0, 1.0 - t)
    return f ** 6 * (3.0 + 18.0 * t + 35.0 * t * t) / 3.0


def wendland_c0(t):
    f = max(0.0, 1.0 - t)
    return f * f


class BruteLCC:
    """fixed, moving: numpy images (C order, x fastest), spacing 1, origin 0.
    cp: (size, origin, spacing) of the control-point grid, coefficients (N, D)
    and weights (N,). weight: None (stationary) or a weight image of the fixed
    grid's shape (non-stationary kernel)."""

    def __init__(self, fixed, moving, cp_grid, coefs, cpw, sigma_c4, sigma_c0, weight=None, max_weight=1.0,
                 subsample=1, subsample_neighborhood=1, offsets=None):
        self.F = np.asarray(fixed, np.float64)
        self.dim = self.F.ndim
        self.size = tuple(reversed(self.F.shape))  # (nx, ny[, nz])
        self.coef = ndimage.spline_filter(np.asarray(moving, np.float64), order=3, mode="mirror")
        csize, corig, cspac = cp_grid
        # flat order x fastest: index tuples over (.., ny, nx), reversed to (x, y[, z])
        idx = np.stack(np.meshgrid(*[np.arange(n) for n in reversed(csize)], indexing="ij"), -1)
        idx = idx.reshape(-1, self.dim)[:, ::-1]
        self.cp_pos = np.asarray(corig, np.float64) + np.asarray(cspac, np.float64) * idx
        self.c = np.asarray(coefs, np.float64).reshape(-1, self.dim)
        self.w = np.asarray(cpw, np.float64)
        self.s4, self.s0 = float(sigma_c4), float(sigma_c0)
        self.smin = min(self.s4, self.s0)
        self.weight = None if weight is None else np.asarray(weight, np.float64)
        self.max_weight = float(max_weight)
        self.sub = subsample
        self.sn = subsample_neighborhood
        self.offsets = offsets

    # ------------------------------------------------------------ pieces
    def _nn_weight(self, p):
        """Nearest-neighbour weight lookup (0 outside the image)."""
        i = [int(p[d] + 0.5) for d in range(self.dim)]
        if any(i[d] < 0 or i[d] >= self.size[d] for d in range(self.dim)):
            return 0.0
        return self.weight[tuple(reversed(i))]

    def support(self, p):
        if self.weight is None:
            return self.smin
        return self.smin * math.sqrt((self.max_weight + self._nn_weight(p)) / 2.0)

    def kernel(self, p, q):
        r = float(np.linalg.norm(np.asarray(p) - np.asarray(q)))
        if self.weight is None:
            return wendland_c4(r / self.s4) * wendland_c0(r / self.s0)
        sp, sq = self._nn_weight(p), self._nn_weight(q)
        if sp + sq <= 0.0:
            return 0.0
        ff = math.sqrt(sp * sq) / ((sp + sq) / 2.0)  # F^2: geometric over arithmetic mean
        rr = r * math.sqrt(2.0 / (sp + sq))
        return ff * wendland_c4(rr / self.s4) * wendland_c0(rr / self.s0)

    def warp(self, y):
        u = np.zeros(self.dim)
        for j in range(len(self.c)):
            k = self.kernel(y, self.cp_pos[j])
            if k != 0.0:
                u += self.w[j] * k * self.c[j]
        return np.asarray(y, np.float64) + u

    def sample(self, p):
        """(inside, value) of the cubic B-spline moving image at index position p."""
        taps = []
        for d in range(self.dim):
            n = self.size[d]
            if not (-0.5 <= p[d] < n - 0.5):
                return False, 0.0
            i0 = int(p[d])  # truncation toward zero
            t = p[d] - i0
            wts = [(1 - t) ** 3 / 6.0, (3 * t ** 3 - 6 * t * t + 4) / 6.0,
                   (-3 * t ** 3 + 3 * t * t + 3 * t + 1) / 6.0, t ** 3 / 6.0]
            ids = [(n - 1) - abs((n - 1) - abs(i0 + k)) for k in (-1, 0, 1, 2)]
            taps.append(list(zip(ids, wts)))
        val = 0.0
        if self.dim == 2:
            for iy, wy in taps[1]:
                for ix, wx in taps[0]:
                    val += wy * wx * self.coef[iy, ix]
        else:
            for iz, wz in taps[2]:
                for iy, wy in taps[1]:
                    for ix, wx in taps[0]:
                        val += wz * wy * wx * self.coef[iz, iy, ix]
        return True, val

    def neighbours(self, x, s):
        """Fixed-grid neighbours of centre x: idx - s < i <= idx + s per axis,
        clamped to the image, on the lattice from the lower bound."""
        axes = []
        for d in range(self.dim):
            lo = max(0, math.floor(x[d] - s) + 1)
            hi = min(self.size[d] - 1, math.floor(x[d] + s))
            axes.append(range(lo, hi + 1, self.sn))
        if self.dim == 2:
            return [(a, b) for b in axes[1] for a in axes[0]]
        return [(a, b, c) for c in axes[2] for b in axes[1] for a in axes[0]]

    def centres(self):
        ns = [n // self.sub for n in self.size]
        out = []
        flat = 0
        rng = [range(n) for n in ns]
        grid = [(a, b) for b in rng[1] for a in rng[0]] if self.dim == 2 else \
            [(a, b, c) for c in rng[2] for b in rng[1] for a in rng[0]]
        for m in grid:
            if self.sub > 1:
                pos = tuple(m[d] * self.sub + int(self.offsets[flat][d]) for d in range(self.dim))
            else:
                pos = m
            out.append(pos)
            flat += 1
        return out

    # ------------------------------------------------------------ LCC
    def centre_cc(self, x, inside, mval):
        s = self.support(x)
        fs, ms = [], []
        for y in self.neighbours(x, s):
            sy = self.support(y)
            r = float(np.linalg.norm(np.asarray(x, np.float64) - np.asarray(y, np.float64)))
            kw = max(0.0, 1.0 - (r / sy) ** 10) ** 2
            if kw > 0.0 and inside[y]:
                fs.append(kw * self.F[tuple(reversed(y))])
                ms.append(kw * mval[y])
        if not fs:
            return 0.0
        f = np.array(fs)
        m = np.array(ms)
        vf = float(((f - f.mean()) ** 2).sum())
        vm = float(((m - m.mean()) ** 2).sum())
        if vf > 0 and vm > 0:
            cov = float(((f - f.mean()) * (m - m.mean())).sum())
            return -cov * cov / (vf * vm)
        return 0.0

    def evaluate(self):
        """(E, per-centre cc list, N_in)."""
        inside, mval = {}, {}
        rng = [range(n) for n in self.size]
        voxels = [(a, b) for b in rng[1] for a in rng[0]] if self.dim == 2 else \
            [(a, b, c) for c in rng[2] for b in rng[1] for a in rng[0]]
        for y in voxels:
            ok, v = self.sample(self.warp(y))
            inside[y], mval[y] = ok, v
        ccs = []
        n_in = 0
        for x in self.centres():
            ccs.append(self.centre_cc(x, inside, mval))
            n_in += int(inside[x])
        return sum(ccs) / n_in, ccs, n_in
