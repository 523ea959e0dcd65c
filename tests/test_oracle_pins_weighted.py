"""Pins of the oracle's weighted kernels and LCC value that do not rest on the
oracle's own formulas.

* Non-stationary and anisotropic kernels with UNEQUAL weights (sigma_P !=
  sigma_Q, Sigma_P != Sigma_Q): values computed by hand from the definition
  in include/cuda/Kernel.h:79-131, with the normalisation F applied in both
  the C4 and the C0 factor (Kernel.h:112-113, 126-127), so k = F^2 C4 C0:
    non-stationary  F^2 = sqrt(sigma_P sigma_Q) / ((sigma_P + sigma_Q) / 2),
                    r   = |p - q| sqrt(2 / (sigma_P + sigma_Q))
    anisotropic     F^2 = sqrt(det S_P det S_Q) / det((S_P + S_Q) / 2),
                    r   = sqrt(v^T ((S_P + S_Q) / 2)^-1 v)
  The radial factors use the Wendland functions pinned by their defining
  properties in test_oracle_pins.py.
* Symmetry k(p, q; P, Q) = k(q, p; Q, P) and a positive-definite Gram matrix
  under smoothly varying weight fields.
* The LCC value against the independent numpy brute force of
  tests/lcc_bruteforce.py on tiny 2D/3D problems, per centre.

Every check is a function of the oracle library so that
test_oracle_mutants.py can run it against deliberately broken copies of the
oracle source and show that each plausible mistake fails at least one pin.
"""
import math

import numpy as np
import pytest

import oracle as orc
from helpers import Case
from lcc_bruteforce import BruteLCC


def _lib():
    return orc.load()


def _problem(lib, dim, weight, mode, sigma_c4=10.0, sigma_c0=None, max_weight=1.0, max_eig=None):
    """A kernel-only problem whose weight image / tensor field covers [0, 7]^D."""
    fixed = np.zeros((8,) * dim)
    size = (8,) * dim
    return orc.OracleProblem(fixed, fixed, ((1,) * dim, (0.0,) * dim, (1.0,) * dim), np.zeros((1, dim)),
                             sigma_c0=sigma_c4 if sigma_c0 is None else sigma_c0, sigma_c4=sigma_c4,
                             kernel_mode=mode, weight=(size, (0.0,) * dim, (1.0,) * dim, weight),
                             max_weight=max_weight, max_eig_cov=max_eig, lib=lib)


def _w(lib, r, s4, s0):
    return lib.skmo_wendland_c4(r / s4) * lib.skmo_wendland_c0(r / s0)


# ------------------------------------------------------------ closed forms

def check_nonstationary_closed_forms(lib):
    # 3D: sigma_P = 1 (given), sigma_Q = 4 (weight image): F^2 = sqrt(4) / 2.5 = 0.8,
    # r = 3 sqrt(2/5); sigma_C4 = 10, sigma_C0 = 14 (the two radial factors differ)
    pr = _problem(lib, 3, np.full(8 ** 3, 4.0), 1, 10.0, 14.0, max_weight=4.0)
    k = pr.kernel([5.0, 3.0, 2.0], [2.0, 3.0, 2.0], sigma_p=1.0)
    r = 3.0 * math.sqrt(0.4)
    assert abs(k - 0.8 * _w(lib, r, 10.0, 14.0)) < 1e-14, k
    # 2D: sigma_P = 9, sigma_Q = 1: F^2 = 3 / 5 = 0.6, |v| = 5, r = 5 sqrt(0.2) = sqrt(5)
    pr = _problem(lib, 2, np.full(64, 1.0), 1, 10.0, max_weight=1.0)
    k = pr.kernel([4.0, 6.0], [1.0, 2.0], sigma_p=9.0)
    assert abs(k - 0.6 * _w(lib, math.sqrt(5.0), 10.0, 10.0)) < 1e-14, k


def check_anisotropic_closed_forms(lib):
    # 3D: S_P = [[2,1,0],[1,2,0],[0,0,1]], S_Q = [[2,-1,0],[-1,2,0],[0,0,3]]:
    # (S_P + S_Q)/2 = 2 I, det S_P = 3, det S_Q = 9, F^2 = sqrt(27) / 8;
    # v = (2, 1, 2): v^T (2I)^-1 v = 9/2
    sp = np.array([[2.0, 1, 0], [1, 2, 0], [0, 0, 1]])
    sq = np.array([[2.0, -1, 0], [-1, 2, 0], [0, 0, 3]])
    pr = _problem(lib, 3, np.broadcast_to(sq.reshape(-1), (8 ** 3, 9)).copy(), 2, 10.0, 12.0)
    k = pr.kernel([5.0, 4.0, 6.0], [3.0, 3.0, 4.0], cov_p=sp)
    assert abs(k - math.sqrt(27.0) / 8.0 * _w(lib, math.sqrt(4.5), 10.0, 12.0)) < 1e-14, k
    # 2D: S_P = [[1,.5],[.5,1]], S_Q = [[3,-.5],[-.5,1]]: mean diag(2, 1), det S_P = 0.75,
    # det S_Q = 2.75, F^2 = sqrt(2.0625) / 2; v = (2, 1): v^T diag(1/2, 1) v = 3
    sp = np.array([[1.0, 0.5], [0.5, 1.0]])
    sq = np.array([[3.0, -0.5], [-0.5, 1.0]])
    pr = _problem(lib, 2, np.broadcast_to(sq.reshape(-1), (64, 4)).copy(), 2, 10.0)
    # the 2D oracle reads the first D*D entries of cov_p row-major
    k = pr.kernel([5.0, 4.0], [3.0, 3.0], cov_p=sp.reshape(-1).tolist() + [0.0] * 5)
    assert abs(k - math.sqrt(2.0625) / 2.0 * _w(lib, math.sqrt(3.0), 10.0, 10.0)) < 1e-14, k


def _smooth_weight(shape, phase):
    g = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    w = 1.5 + 0.6 * np.sin(g[-1] / 2.5 + phase) * np.cos(g[-2] / 3.0)
    if len(shape) == 3:
        w += 0.3 * np.sin(g[0] / 2.0)
    return w


def _smooth_tensor(shape, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=tuple(shape) + (3, 3)) * 0.4
    g = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    t = np.einsum("...ij,...kj->...ik", a, a) + np.eye(3) * (1.0 + 0.5 * np.sin(g[-1] / 3.0))[..., None, None]
    return t


def _nn(field, p):
    i = tuple(int(v + 0.5) for v in reversed(p))
    return field[i]


def check_symmetry_and_gram(lib):
    rng = np.random.default_rng(5)
    w = _smooth_weight((8, 8, 8), 0.3)
    pr = _problem(lib, 3, w.reshape(-1), 1, 6.0, max_weight=w.max())
    pts = rng.permutation(np.stack(np.meshgrid(*[np.arange(8.0)] * 3, indexing="ij"), -1).reshape(-1, 3))[:70]
    G = np.array([[pr.kernel(p, q, sigma_p=_nn(w, p)) for q in pts] for p in pts])
    assert np.array_equal(G, G.T)
    assert np.linalg.eigvalsh(G).min() > 1e-6
    t = _smooth_tensor((8, 8, 8), 6)
    pr = _problem(lib, 3, t.reshape(-1, 9), 2, 6.0)
    G = np.array([[pr.kernel(p, q, cov_p=_nn(t, p)) for q in pts] for p in pts])
    assert np.max(np.abs(G - G.T)) < 1e-15
    assert np.linalg.eigvalsh(G).min() > 1e-6


def test_nonstationary_unequal_sigma_closed_forms():
    check_nonstationary_closed_forms(_lib())


def test_anisotropic_unequal_covariance_closed_forms():
    check_anisotropic_closed_forms(_lib())


def test_weighted_kernels_symmetric_positive_definite():
    """Both weighted kernels are symmetric under exchanging the points with
    their weights; their Gram matrices on 70 grid points of a smoothly varying
    field are positive definite (an empirical property here: the compactly
    supported profile is not a Gaussian scale mixture, so Paciorek's theorem
    does not apply)."""
    check_symmetry_and_gram(_lib())


# ------------------------------------------------------------ LCC brute force

LCC_CASES = [
    # dim, shape, sigma, cp spacing, mode, subsample, neighbourhood stride, seed
    (2, (11, 12), 4.3, 2.5, 0, 1, 3, 1),
    (2, (10, 11), 3.6, 2.0, 1, 1, 2, 2),
    (3, (7, 8, 9), 3.7, 2.5, 0, 2, 2, 3),
    (3, (7, 7, 8), 3.3, 2.0, 1, 1, 3, 4),
]


def _lcc_pair(dim, shape, sigma, h, mode, sub, sn, seed, lib=None, zero_moving=False):
    case = Case(dim=dim, shape=shape, cp_spacing=h, sigma=sigma, mode=mode, subsample=sub,
                subsample_neighborhood=sn, seed=seed, coef_scale=0.6)
    if zero_moving:
        case.moving = np.zeros_like(case.moving)
    offsets = case.offsets(1) if sub > 1 else None
    pr = case.oracle(offsets)
    if lib is not None:
        pr = orc.OracleProblem(case.fixed, case.moving, (case.cp_size, case.cp_origin, case.cp_spacing), case.cp,
                               case.cpw, sigma_c0=case.sigma_c0, sigma_c4=case.sigma_c4, kernel_mode=mode,
                               max_weight=case.max_weight,
                               weight=None if mode == 0 else (tuple(reversed(case.shape)), case.origin,
                                                              case.spacing, case.weight),
                               subsample=sub, subsample_neighborhood=sn, offsets=offsets, lib=lib)
    brute = BruteLCC(case.fixed, case.moving, (case.cp_size, case.cp_origin, case.cp_spacing), case.cp, case.cpw,
                     case.sigma_c4, case.sigma_c0, weight=case.weight if mode == 1 else None,
                     max_weight=case.max_weight, subsample=sub, subsample_neighborhood=sn, offsets=offsets)
    return pr, brute


def check_lcc_against_brute_force(lib, cases=LCC_CASES):
    for c in cases:
        pr, brute = _lcc_pair(*c, lib=lib)
        res, _, cc = pr.evaluate(2, False, per_sample=True)
        e, bcc, n_in = brute.evaluate()
        assert res["pixels_counted"] == n_in, c
        assert np.max(np.abs(cc - np.array(bcc))) < 1e-10, (c, np.max(np.abs(cc - np.array(bcc))))
        assert abs(res["measure"] - e) < 1e-11, (c, res["measure"], e)
        assert np.count_nonzero(bcc) > len(bcc) // 2, c  # the pin is not vacuous


def check_lcc_zero_variance(lib):
    """A zero moving image has zero moving variance at every centre: cc = 0
    (the reference sets cc only when both variances are positive)."""
    pr, brute = _lcc_pair(*LCC_CASES[0], lib=lib, zero_moving=True)
    res, _, cc = pr.evaluate(2, False, per_sample=True)
    assert np.all(cc == 0.0) and res["measure"] == 0.0
    e, bcc, _ = brute.evaluate()
    assert e == 0.0 and not any(bcc)


@pytest.mark.parametrize("case", LCC_CASES, ids=lambda c: f"{c[0]}d-mode{c[4]}-sub{c[5]}-sn{c[6]}")
def test_lcc_value_against_brute_force(case):
    check_lcc_against_brute_force(_lib(), [case])


def test_lcc_zero_moving_variance():
    check_lcc_zero_variance(_lib())


# ------------------------------------------------------------ bilateral regularizer

def _bilateral_problem(lib, dim, size, h, cp, cpw, fixed):
    return orc.OracleProblem(fixed, fixed, (size, (0.0,) * dim, (h,) * dim), np.asarray(cp, float),
                             cpw=np.asarray(cpw, float), sigma_c0=5.0, sigma_c4=5.0, lib=lib)


def check_bilateral_closed_forms(lib):
    """Two control points 5 apart with fixed-image values 0.3 and 0.55 at
    their voxels: B = exp(-25 / (2 * 4^2)) exp(-0.25^2 / (2 * 0.2^2)) =
    exp(-1.5625), value = (w_0 + w_1) B |c_0 - c_1|^2, derivative
    +-2 B (w_0 + w_1)(c_0 - c_1); a pair beyond 3 sigma_s contributes nothing."""
    fixed = np.zeros((4, 4, 8))
    fixed[0, 0, 0], fixed[0, 0, 5] = 0.3, 0.55
    c = np.array([[0.4, -1.0, 0.5], [-0.2, 0.3, 1.1]])
    w = np.array([0.7, 1.6])
    pr = _bilateral_problem(lib, 3, (2, 1, 1), 5.0, c, w, fixed)
    val, der = pr.bilateral(2.0, 4.0, 0.2)
    b = math.exp(-1.5625)
    diff = c[0] - c[1]
    assert abs(val - (w[0] + w[1]) * b * (diff @ diff)) < 1e-14, val
    g0 = 2.0 * b * (w[0] + w[1]) * diff * 2.0  # times the weight 2
    np.testing.assert_allclose(der.reshape(3, 2).T, np.stack([g0, -g0]), rtol=0, atol=1e-14)
    # 13 > 3 sigma_s = 12: outside the cut
    fixed2 = np.zeros((4, 4, 16))
    pr = _bilateral_problem(lib, 3, (2, 1, 1), 13.0, c, w, fixed2)
    val, der = pr.bilateral(1.0, 4.0, 0.2)
    assert val == 0.0 and not der.any()


def check_bilateral_fd(lib):
    """The derivative is the exact gradient of the summed value for
    non-uniform control-point weights (B is symmetric)."""
    rng = np.random.default_rng(3)
    fixed = rng.uniform(0, 1, size=(12, 13, 14))
    size, h = (5, 4, 4), 3.0
    n = int(np.prod(size))
    c = rng.normal(size=(n, 3))
    w = rng.uniform(0.5, 1.5, size=n)
    pr = _bilateral_problem(lib, 3, size, h, c, w, fixed)
    _, der = pr.bilateral(1.0, 2.5, 0.3)
    eps = 1e-6
    for idx in rng.choice(n * 3, 8, replace=False):
        j, d = idx % n, idx // n
        vals = []
        for sgn in (1, -1):
            cc = c.copy()
            cc[j, d] += sgn * eps
            vals.append(_bilateral_problem(lib, 3, size, h, cc, w, fixed).bilateral(1.0, 2.5, 0.3, False)[0])
        fd = (vals[0] - vals[1]) / (2 * eps)
        assert abs(fd - der[idx]) < 1e-7 * max(1.0, abs(fd)), (idx, fd, der[idx])


def test_bilateral_closed_forms():
    check_bilateral_closed_forms(_lib())


def test_bilateral_derivative_fd():
    check_bilateral_fd(_lib())
