"""Benchmark of one GetValueAndDerivative pass of the SKM path on one MI355X.

Workload (default): the 3D registration of 128x128x96 float32 volumes
(random smoothed ellipsoids, target warped by a smooth field, noise), control
points every 6 voxels (22x22x16 = 7,744, 23,232 parameters), Wendland support
24 voxels, anisotropic kernel tensors computed from the target's structure
(ComputeAnisotropicTensor with tensor weight 400, alignment z, as SKMReg does
by default), NCC metric, bilateral regularizer (weight 1e-3, spatial sigma
12, intensity sigma 0.1), dense sampling.  One step is one
value-and-derivative evaluation with every input resident in HBM.
`--kernel stationary` runs the isotropic kernel instead.

metric: voxel x control-point pair evaluations per second (GVox*ctrl/s),
counting the pairs enumerated by the reference windows in the value and the
derivative pass.  One JSON line on stdout.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
# algorithmic FP32 flops per (voxel, control point) pair: Wendland C4*C0 and
# the weighted accumulate (32); the anisotropic kernel adds the Mahalanobis
# distance and the F factor (75 more)
FLOPS_PER_PAIR = {"stationary": 32, "anisotropic": 32 + 75}
# LCC reference-window derivative: per (control point, centre, neighbour)
# triple the isotropic Wendland ka(p_j, y) and the weighted accumulate
FLOPS_PER_LCC_TRIPLE = 32
# the same derivative regrouped per voxel (k_lcc_ref_table): per candidate
# centre of a voxel's table its weight kw (distance, (r/s)^10, squares) and
# its term kw (kw (F A + M B) + C) added to the row's prefix; per (control
# point, voxel) pair with ka != 0 the kernel ka (32), the 8-corner sum of the
# table (7) and the weighted accumulate (7)
FLOPS_PER_TABLE_ENTRY = 20
FLOPS_PER_TABLE_PAIR = 46


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200,
                    help="timed evaluations (200: one pyramid level's optimizer iterations)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--shape", default="96,128,128", help="z,y,x voxels")
    ap.add_argument("--cp-spacing", type=float, default=6.0)
    ap.add_argument("--sigma", type=float, default=24.0)
    ap.add_argument("--kernel", default="anisotropic", choices=["anisotropic", "stationary"])
    ap.add_argument("--gather-split", type=int, default=0, choices=[0, 4, 8],
                    help="sample planes per derivative-gather slab (skm_tuning.gather_split; 0: automatic)")
    ap.add_argument("--tensor-weight", type=float, default=400.0)
    ap.add_argument("--subsample", type=int, default=1)
    ap.add_argument("--measure", default="ncc", choices=["mse", "ncc", "lcc"])
    ap.add_argument("--reg-rd", type=float, default=0.0)
    ap.add_argument("--reg-bilateral", type=float, default=1e-3)
    ap.add_argument("--bilateral-spatial-sigma", type=float, default=12.0)
    ap.add_argument("--bilateral-intensity-sigma", type=float, default=0.1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-registration", action="store_true", help="skip the multi-scale registration timing")
    ap.add_argument("--no-lcc", action="store_true", help="skip the LCC timing on the sliding-motion volumes")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched time-series registration")
    ap.add_argument("--no-lcc-registration", action="store_true",
                    help="skip the LCC registration of the sliding-motion volumes")
    ap.add_argument("--lcc-iterations", type=int, default=50, help="iterations per level of the LCC registration")
    ap.add_argument("--batch-frames", type=int, default=32, help="frames of the time series (frame 0 = reference)")
    ap.add_argument("--batch-concurrency", type=int, default=4, help="registrations in flight in the batched run")
    ap.add_argument("--no-warm-start", action="store_true", help="skip the learned warm-start comparison")
    ap.add_argument("--warm-pairs", type=int, default=8, help="held-out pairs of the warm-start comparison")
    ap.add_argument("--no-lockstep", action="store_true",
                    help="skip the lockstep (graph-replay) runs of the batch and warm-start legs")
    ap.add_argument("--registration-repeats", type=int, default=3)
    ap.add_argument("--registration-autotune", type=int, default=1, choices=[0, 1],
                    help="1: the registration's levels take their launch shapes from the autotuner "
                         "(tuned in the untimed first run; 0: the automatic shapes)")
    ap.add_argument("--registration-bilateral", type=float, default=1e-5,
                    help="bilateral weight of the registrations (spatial / intensity sigma as above)")
    ap.add_argument("--no-pair-count", action="store_true", help="skip the exact k != 0 pair count")
    ap.add_argument("--cpu-crop", default=None, help="z,y,x voxels of the CPU baseline sample")
    ap.add_argument("--cpu-repeats", type=int, default=5,
                    help="timed CPU baseline evaluations after one untimed one (median)")
    ap.add_argument("--calibration-copy", type=int, default=0, metavar="MIB",
                    help="run one plain device copy of MIB MiB first (reference dispatch for HBM counters)")
    return ap.parse_args()


# ------------------------------------------------------------ geometry

def geometry(shape, cp_spacing):
    """Control-point grid of TransformAdapter::ComputeControlPointImage for a
    unit-spacing image at the origin: (image size xyz, cp size xyz, cp origin xyz)."""
    size = tuple(reversed(shape))
    cs, co = [], []
    for d in range(3):
        upper = size[d] - 1.0
        n = int(np.ceil(upper / cp_spacing))
        cs.append(n)
        co.append(abs(upper - cp_spacing * (n - 1)) / 2.0)
    return size, tuple(cs), tuple(co)


def value_windows(size, cs, co, h, support, sub=1):
    """Reference cp window of every sample (GpuEvaluatorMSE.cu:61-80), samples
    at the cell origins: per-axis lo/hi arrays (N, 3); support per sample (N,)."""
    nx, ny, nz = size
    z, y, x = np.meshgrid(*[np.arange(0, (n // sub) * sub, sub) for n in (nz, ny, nx)], indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.float64)
    ci = (pos - np.asarray(co)) / h
    r = (np.asarray(support, np.float64) / h).reshape(-1, 1) * np.ones((1, 3))
    lo = np.clip(np.trunc(ci - r + 1.0), 0, np.asarray(cs) - 1).astype(np.int64)
    hi = np.clip(np.trunc(ci + r), 0, np.asarray(cs) - 1).astype(np.int64)
    return lo, hi


def derivative_windows(size, cs, co, h, support, sub=1):
    """Sample-cell window of every control point (3D derivative kernels,
    GpuEvaluatorMSE.cu:239-309): lo/hi arrays (N_cp, 3)."""
    ix, iy, iz = (np.arange(c) for c in cs)
    z, y, x = np.meshgrid(iz, iy, ix, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1) * h + np.asarray(co)
    r = np.asarray(support, np.float64).reshape(-1, 1) * np.ones((1, 3))
    lim = np.asarray(size) // sub - 1
    lo = np.clip(np.trunc((p - r) / sub), 0, lim).astype(np.int64)
    hi = np.clip(np.trunc((p + r) / sub + 1.0), 0, lim).astype(np.int64)
    return lo, hi, p


def box_pairs(lo, hi):
    return int(np.prod(hi - lo + 1, axis=1).sum())


def support_pairs(sample_axes, cp_axes, support):
    """Exact number of (sample, cp) pairs with |x - p| < support (stationary k != 0)."""
    d2 = [(s[:, None] - c[None, :]) ** 2 for s, c in zip(sample_axes, cp_axes)]
    s2 = support * support
    dx = np.sort(d2[0].ravel())
    rest = (d2[1].ravel()[:, None] + d2[2].ravel()[None, :]).ravel()
    t = s2 - rest
    t = t[t > 0]
    return int(np.searchsorted(dx, t, side="left").sum())


# ------------------------------------------------------------ anisotropic pair counts

def upper(t):
    """Symmetric part of (..., 3, 3) as its upper triangle (..., 6): xx xy xz yy yz zz."""
    s = 0.5 * (t + np.swapaxes(t, -1, -2))
    return np.stack([s[..., 0, 0], s[..., 0, 1], s[..., 0, 2], s[..., 1, 1], s[..., 1, 2], s[..., 2, 2]], axis=-1)


def tensor_supports(ten, mcov, smin):
    """Region support smin / sqrt(2 / lambda_max(maxEigCov + Sigma)) per voxel (Kernel.h:166-171)."""
    lam = np.linalg.eigvalsh(ten.reshape(-1, 3, 3).astype(np.float64) + mcov[None])[:, -1]
    return smin / np.sqrt(2.0 / lam)


def _nonzero(torch, s, v, smin):
    """k != 0  <=>  r = sqrt(2 v^T adj(S) v / det S) < smin for S = S_p + S_q."""
    s00, s01, s02, s11, s12, s22 = (s[..., i] for i in range(6))
    c00 = s11 * s22 - s12 * s12
    c01 = s02 * s12 - s01 * s22
    c02 = s01 * s12 - s02 * s11
    c11 = s00 * s22 - s02 * s02
    c12 = s01 * s02 - s00 * s12
    c22 = s00 * s11 - s01 * s01
    det = s00 * c00 + s01 * c01 + s02 * c02
    vx, vy, vz = v
    quad = (c00 * vx * vx + c11 * vy * vy + c22 * vz * vz
            + 2.0 * (c01 * vx * vy + c02 * vx * vz + c12 * vy * vz))
    return 2.0 * quad < smin * smin * det


def count_nonzero_pairs(ten, size, cs, co, h, smin, vlo, vhi, dlo, dhi, dev):
    """Exact numbers of pairs with k != 0 inside the value windows and inside
    the derivative windows, counted with torch on the GPU (float64)."""
    import torch
    nx, ny, nz = size
    up = torch.tensor(upper(ten.astype(np.float64)).reshape(-1, 6), device=dev)
    # control-point states: the tensor at the nearest voxel
    ix, iy, iz = (np.arange(c) for c in cs)
    z, y, x = np.meshgrid(iz, iy, ix, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1) * h + np.asarray(co)
    nv = np.clip((p + 0.5).astype(np.int64), 0, np.asarray(size) - 1)
    cp_up = up[torch.tensor(nv[:, 0] + nx * (nv[:, 1] + ny * nv[:, 2]), device=dev)]
    cp_p = torch.tensor(p, device=dev)

    # derivative windows: per cp, voxels of its box
    dlo_t, dhi_t = torch.tensor(dlo, device=dev), torch.tensor(dhi, device=dev)
    W = (dhi - dlo + 1).max(axis=0)
    off = [torch.arange(int(W[d]), device=dev) for d in range(3)]
    der = 0
    B = 8
    for j0 in range(0, len(p), B):
        lo, hi = dlo_t[j0:j0 + B], dhi_t[j0:j0 + B]
        ax = [lo[:, d, None] + off[d][None, :] for d in range(3)]               # (B, W_d)
        ok = [a <= hi[:, d, None] for d, a in enumerate(ax)]
        axc = [torch.clamp(a, max=n - 1) for a, n in zip(ax, (nx, ny, nz))]
        flat = (axc[2][:, :, None, None] * ny + axc[1][:, None, :, None]) * nx + axc[0][:, None, None, :]
        s = up[flat] + cp_up[j0:j0 + B][:, None, None, None, :]
        pp = cp_p[j0:j0 + B]
        v = (pp[:, 0, None, None, None] - ax[0][:, None, None, :].double(),
             pp[:, 1, None, None, None] - ax[1][:, None, :, None].double(),
             pp[:, 2, None, None, None] - ax[2][:, :, None, None].double())
        mask = ok[2][:, :, None, None] & ok[1][:, None, :, None] & ok[0][:, None, None, :]
        der += int((_nonzero(torch, s, v, smin) & mask).sum())

    # value windows: per sample, control points of its box
    vlo_t, vhi_t = torch.tensor(vlo, device=dev), torch.tensor(vhi, device=dev)
    Wc = (vhi - vlo + 1).max(axis=0)
    offc = [torch.arange(int(Wc[d]), device=dev) for d in range(3)]
    ncx, ncy, ncz = cs
    val = 0
    S = 4096
    n = len(vlo)
    for i0 in range(0, n, S):
        idx = torch.arange(i0, min(i0 + S, n), device=dev)
        xs = torch.stack([idx % nx, (idx // nx) % ny, idx // (nx * ny)], dim=1).double()
        lo, hi = vlo_t[i0:i0 + S], vhi_t[i0:i0 + S]
        ax = [lo[:, d, None] + offc[d][None, :] for d in range(3)]
        ok = [a <= hi[:, d, None] for d, a in enumerate(ax)]
        axc = [torch.clamp(a, max=m - 1) for a, m in zip(ax, (ncx, ncy, ncz))]
        flat = (axc[2][:, :, None, None] * ncy + axc[1][:, None, :, None]) * ncx + axc[0][:, None, None, :]
        s = cp_up[flat] + up[idx][:, None, None, None, :]
        v = (xs[:, 0, None, None, None] - (co[0] + h * ax[0][:, None, None, :].double()),
             xs[:, 1, None, None, None] - (co[1] + h * ax[1][:, None, :, None].double()),
             xs[:, 2, None, None, None] - (co[2] + h * ax[2][:, :, None, None].double()))
        mask = ok[2][:, :, None, None] & ok[1][:, None, :, None] & ok[0][:, None, None, :]
        val += int((_nonzero(torch, s, v, smin) & mask).sum())
    return val, der


# ------------------------------------------------------------ workload

def workload(args):
    from skmimageregistration_amd import synthetic
    shape = tuple(int(v) for v in args.shape.split(","))
    return synthetic.ellipsoids_3d(shape, n=20, smooth=1.5, max_disp=6.0, noise=0.02)


def kernel_for(args, fix):
    """Kernel of the workload; anisotropic: ComputeAnisotropicTensor on the
    target (SKMReg.cxx:268-273) and SetWeightingTensor's maximum values."""
    from skmimageregistration_amd import ComputeAnisotropicTensor, Grid, Image, Kernel
    if args.kernel == "stationary":
        return Kernel(args.sigma, args.sigma), None, None, None
    ComputeAnisotropicTensor(Image(fix), args.sigma, args.tensor_weight, 2)  # first HIP use, untimed
    t0 = time.perf_counter()
    ten, tilt, mw, mcov = ComputeAnisotropicTensor(Image(fix), args.sigma, args.tensor_weight, 2)
    tensor_ms = (time.perf_counter() - t0) * 1e3
    size = tuple(reversed(fix.shape))
    k = Kernel(args.sigma, args.sigma, useWeightTensor=True, maximumWeight=mw, maximumEigenvalueCovariance=mcov,
               wTensor=(Grid(size, (0.0,) * 3, (1.0,) * 3), ten))
    return k, ten, mcov, tensor_ms


def pair_counts(args, shape, ten, mcov, dev, exact_nonzero):
    """(value box pairs, derivative box pairs, value k!=0 pairs, derivative k!=0 pairs)."""
    size, cs, co = geometry(shape, args.cp_spacing)
    h = args.cp_spacing
    sub = args.subsample
    n_cells = int(np.prod([n // sub for n in size]))
    if ten is None:
        sv = np.full(n_cells, args.sigma)
        sd = np.full(int(np.prod(cs)), args.sigma)
    else:
        sup = tensor_supports(ten, mcov, args.sigma).reshape(tuple(reversed(size)))
        sv = sup[::sub, ::sub, ::sub][: size[2] // sub, : size[1] // sub, : size[0] // sub].ravel()
        _, _, p = derivative_windows(size, cs, co, h, np.zeros(int(np.prod(cs))))
        nv = np.clip((p + 0.5).astype(np.int64), 0, np.asarray(size) - 1)
        sd = sup.ravel()[nv[:, 0] + size[0] * (nv[:, 1] + size[1] * nv[:, 2])]
    vlo, vhi = value_windows(size, cs, co, h, sv, sub)
    dlo, dhi, _ = derivative_windows(size, cs, co, h, sd, sub)
    vbox, dbox = box_pairs(vlo, vhi), box_pairs(dlo, dhi)
    vnz = dnz = None
    if exact_nonzero:
        if ten is None:
            samples = [np.arange(size[d], dtype=np.float64) for d in range(3)]
            cps = [co[d] + h * np.arange(cs[d]) for d in range(3)]
            vnz = dnz = support_pairs(samples, cps, args.sigma)
        else:
            vnz, dnz = count_nonzero_pairs(ten, size, cs, co, h, args.sigma, vlo, vhi, dlo, dhi, dev)
    return vbox, dbox, vnz, dnz


def workload_key(args, shape):
    z, y, x = shape
    return f"{args.kernel}-{x}x{y}x{z}-h{args.cp_spacing:g}-s{args.sigma:g}-{args.measure}-sub{args.subsample}"


def profiled_traffic(kernel, key):
    """HBM bytes per launch of `kernel` on this workload, as measured by the
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised in profiles/ (the line
    `traffic <kernel> <workload key> <bytes>`); None when not profiled."""
    d = os.path.join(REPO, "profiles")
    if not os.path.isdir(d):
        return None
    for name in sorted(os.listdir(d)):
        if not name.endswith(".txt"):
            continue
        with open(os.path.join(d, name)) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 4 and parts[0] == "traffic" and parts[1] == kernel and parts[2] == key:
                    return float(parts[3])
    return None


def cpu_baseline(args, mov, fix, ten, mcov):
    """The FP32 multi-core CPU restatement of the reference's NCC
    GetValueAndDerivative (oracle/cpu_fp32.c, compiled here with -O3
    -march=native -fopenmp, OpenMP over every logical CPU this process may
    use), one evaluation of the same workload (whole volumes, same kernel,
    sigma, control-point spacing and regularizers; --cpu-crop z,y,x takes a corner crop
    instead), median of --cpu-repeats after one untimed run; reported in the
    same unit.  Also the CPU time of one 3-level registration estimated as
    iterations x pairs per iteration / the measured pair rate."""
    if args.measure != "ncc" or args.subsample != 1:
        return None
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import cpu_fp32
    import oracle as orc
    threads = cpu_fp32.load().skmf_set_threads(cpu_threads())
    crop = tuple(int(v) for v in args.cpu_crop.split(",")) if args.cpu_crop else tuple(fix.shape)
    crop = tuple(min(c, n) for c, n in zip(crop, fix.shape))
    f = np.ascontiguousarray(fix[: crop[0], : crop[1], : crop[2]])
    m = np.ascontiguousarray(mov[: crop[0], : crop[1], : crop[2]])
    size, cs, co = geometry(crop, args.cp_spacing)
    rng = np.random.default_rng(0)
    n_cp = int(np.prod(cs))
    cp = rng.uniform(-1, 1, size=(n_cp, 3)).astype(np.float32)
    tc = None
    if ten is not None:
        tc = np.ascontiguousarray(ten[: crop[0], : crop[1], : crop[2]]).reshape(-1, 9)
    coeffs = orc.bspline_coefficients(m.astype(np.float64))
    pr = cpu_fp32.CpuProblem(f, coeffs, (cs, co, (args.cp_spacing,) * 3), cp, np.ones(n_cp, np.float32), args.sigma,
                             cov=tc, max_eig_cov=mcov)
    bl = bilateral_args(args, args.reg_bilateral) if args.reg_bilateral > 0 else None
    pr.evaluate(args.reg_rd, bilateral=bl)
    times = []
    for _ in range(max(args.cpu_repeats, 1)):
        t0 = time.perf_counter()
        pr.evaluate(args.reg_rd, bilateral=bl)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    v, d, _, _ = pair_counts(args, crop, None if tc is None else tc.reshape(crop + (3, 3)), mcov, None, False)
    rate = (v + d) / dt
    # pairs per iteration of the registration's levels: samples / sub^3 x (2 support / h)^3 box
    # control points, value and derivative pass; support / h = 4 on every level
    n_vox = float(np.prod(fix.shape))
    reg_pairs = sum(200 * 2 * n_vox / sub ** 3 * 8 ** 3 for sub in (5, 4, 3))
    return {"value": rate / 1e9, "unit": "GVox*ctrl/s", "cores": threads, "kind": "port",
            "sample": f"FP32 CPU restatement (oracle/cpu_fp32.c, -O3 -march=native, OpenMP {threads} threads), "
                      f"GetValueAndDerivative on {crop[2]}x{crop[1]}x{crop[0]} voxels (same kernel, sigma, cp "
                      f"spacing, regularizers), median of {len(times)}: {dt:.2f} s; host {host_cpu()}",
            "ms_per_evaluation": round(dt * 1e3, 1),
            "ms_per_registration_estimate": round(reg_pairs / rate * 1e3, 0)}


def regularizer_desc(args):
    """' + RD w' / ' + bilateral w (spatial sigma, intensity sigma)' of the workload."""
    out = ""
    if args.reg_rd > 0:
        out += f" + RD {args.reg_rd:g}"
    if args.reg_bilateral > 0:
        out += (f" + bilateral {args.reg_bilateral:g} (spatial sigma {args.bilateral_spatial_sigma:g}, "
                f"intensity sigma {args.bilateral_intensity_sigma:g})")
    return out


def cpu_threads():
    """Logical CPUs this process may use, capped by OMP_NUM_THREADS when set
    (the host's share of CPUs for this GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def host_cpu():
    """CPU model and the logical CPUs this process may use."""
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return f"{model}, {n} logical CPUs available"


def lcc_triples(shape, cs, co, h, support, sub, subn):
    """The reference-window LCC derivative's enumeration
    (GpuEvaluatorLCC.cu:206-530): for every control point j, every centre in
    j's sample-cell window, every neighbour on the stride-subn lattice of the
    centre's tight support box; centres at their cell origins (the jitter
    changes the count at the volume's faces only).  Separable per axis."""
    size = tuple(reversed(shape))
    lo, hi, _ = derivative_windows(size, cs, co, h, support, sub)
    total = np.ones(len(lo), np.float64)
    for d in range(3):
        n = size[d]
        pos = np.arange(n // sub) * sub
        nlo = np.clip(np.trunc(pos - support + 1.0), 0, n - 1)
        nhi = np.clip(np.trunc(pos + support), 0, n - 1)
        cnt = (nhi - nlo) // subn + 1
        pre = np.concatenate([[0.0], np.cumsum(cnt)])
        total *= pre[hi[:, d] + 1] - pre[lo[:, d]]
    return float(total.sum())


def lcc_reach_cells(n, r, sub, subn):
    """Per voxel index along one axis: the sample cells holding a centre
    position whose neighbour lattice (neighbour_box's arithmetic,
    GpuEvaluatorLCC.cu:95-97) holds the voxel -- the ranks of its table."""
    out = np.zeros(n, np.int64)
    for y in range(n):
        cells = set()
        for p in range(max(0, y - int(np.floor(r)) - 1), min(n - 1, y - int(np.floor(1.0 - r)) + 1) + 1):
            nlo = min(max(int(p - r + 1.0), 0), n - 1)
            nhi = min(max(int(p + r), 0), n - 1)
            if nlo <= y <= nhi and (y - nlo) % subn == 0 and p // sub < n // sub:
                cells.add(p // sub)
        out[y] = len(cells)
    return out


def lcc_table_work(shape, cs, co, h, s, sub, subn):
    """(table entries, (control point, voxel) pairs with ka != 0) of one
    voxel-table derivative: entries = sum over voxels of the product of the
    per-axis reach cells; pairs = exact count of |p - y| < s."""
    size = tuple(reversed(shape))
    entries = float(np.prod([lcc_reach_cells(size[d], s, sub, subn).sum() for d in range(3)]))
    voxels = [np.arange(size[d], dtype=np.float64) for d in range(3)]
    cps = [co[d] + h * np.arange(cs[d]) for d in range(3)]
    return entries, float(support_pairs(voxels, cps, s))


def lcc_cpu_baseline(mov, fix, h, s, sub, subn, triples_full, bilateral, crop=64):
    """The OpenMP build of the FP64 oracle (oracle/skm_oracle.c, the
    reference's LCC loops restated) timed on a crop^3 block at the centre of
    the sliding-motion volumes (the sliding interface runs through it), with
    the same kernel, control-point spacing, subsampling, neighbourhood and
    bilateral term (weight, spatial sigma, intensity sigma); reference-window
    derivative.  Its triple rate scales to the full volume's
    triple count for a per-evaluation estimate."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as orc
    threads = orc.load(parallel=True).skmo_set_threads(cpu_threads())
    c0 = [(n - crop) // 2 for n in fix.shape]
    sl = tuple(slice(a, a + crop) for a in c0)
    f, m = np.ascontiguousarray(fix[sl]), np.ascontiguousarray(mov[sl])
    size, cs, co = geometry(f.shape, h)
    n_cp = int(np.prod(cs))
    cp = np.random.default_rng(0).uniform(-1, 1, (n_cp, 3))
    n_cells = int(np.prod([n // sub for n in f.shape]))
    offsets, _ = orc.cube_offsets(n_cells, 3, sub)
    pr = orc.OracleProblem(f, m, (cs, co, (h,) * 3), cp, sigma_c0=s, sigma_c4=s, subsample=sub,
                           subsample_neighborhood=subn, offsets=offsets, parallel=True)
    t0 = time.perf_counter()
    pr.evaluate(2, True)
    pr.bilateral(*bilateral)
    dt = time.perf_counter() - t0
    tri = lcc_triples(f.shape, cs, co, h, s, sub, subn)
    rate = tri / dt
    return {"value": round(triples_full / rate * 1e3, 1), "unit": "ms per LCC GetValueAndDerivative (estimate)",
            "cores": threads, "kind": "port",
            "sample": f"FP64 oracle (oracle/skm_oracle.c, OpenMP {threads} threads), reference-window LCC value and "
                      f"derivative on the central {crop}^3 crop (cp {cs[0]}x{cs[1]}x{cs[2]}): {dt:.2f} s for "
                      f"{tri:.3g} (cp, centre, neighbour) triples; scaled to the full volume's {triples_full:.3g}; "
                      f"host {host_cpu()}",
            "triples_per_s": round(rate, 1)}


def lcc_timing(args):
    """One LCC GetValueAndDerivative on the sliding-motion workload
    (256x256x128 two halves shifted 4 voxels plus breathing motion,
    control-point spacing 5 -> 51x51x26, isotropic Wendland support 20,
    bilateral 1e-3 (spatial sigma 12, intensity sigma 0.1), subsample 2,
    neighbourhood stride 6 -> 7^3 neighbours per centre):
    the reference's windowed derivative (GpuEvaluatorLCC.cu:206-530), in the
    automatic form (regrouped per voxel at this size) and as the centre walk,
    and the exact gradient (each voxel gathers the terms of the centres that
    reach it, then one kernel-sum gather); median of 3 after one untimed
    evaluation."""
    from skmimageregistration_amd import MEASURE, GpuKernelEvaluator, Grid, Image, Kernel, synthetic
    shape, h, s, sub, subn = (128, 256, 256), 5.0, 20.0, 2, 6
    mov, fix, _ = synthetic.sliding_3d(shape)
    size, cs, co = geometry(shape, h)
    n = int(np.prod(cs))
    cp = np.random.default_rng(0).uniform(-1, 1, (n, 3)).astype(np.float32)
    out = {"workload": "LCC GetValueAndDerivative, 3D 256x256x128 sliding motion, cp 51x51x26 spacing 5, "
                       "isotropic Wendland support 20, bilateral 0.001 (spatial sigma 12, intensity sigma 0.1), "
                       "subsample 2, neighbourhood stride 6"}
    for name, exact, walk in (("exact_gradient", True, 0), ("reference_windows", False, 0),
                              ("reference_windows_walk", False, 1)):
        ev = GpuKernelEvaluator(n * 3, 2, 6, Kernel(s, s), Image(mov), Image(fix), Grid(cs, co, (h,) * 3),
                                exact_gradient=exact)
        ev.SetRegularizerBilateral(1e-3, args.bilateral_spatial_sigma, args.bilateral_intensity_sigma)
        if walk:
            ev.set_tuning(lcc_ref_walk=1)
        ev.getValueAndDerivative(MEASURE.LCC, cp)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            ev.getValueAndDerivative(MEASURE.LCC, cp)
            times.append((time.perf_counter() - t0) * 1e3)
        ev.enable_timing(True)
        r, _ = ev.getValueAndDerivative(MEASURE.LCC, cp)
        passes = {k: round(v[0] / max(v[1], 1), 3) for k, v in ev.read_timing().items() if v[1]}
        ev.close()
        out[name] = {"ms": round(float(np.median(times)), 2), "pass_ms": passes, "measure": round(r.measure, 6)}
    # the reference windows' derivative by voxel tables (the automatic form at
    # this size): its own work, and the reference enumeration's (cp, centre,
    # neighbour) triples per second
    tri = lcc_triples(shape, cs, co, h, s, sub, subn)
    entries, pairs = lcc_table_work(shape, cs, co, h, s, sub, subn)
    key = f"lcc-{size[0]}x{size[1]}x{size[2]}-h{h:g}-s{s:g}-sub{sub}-n{subn}"
    ref_ms = out["reference_windows"]["pass_ms"]["lcc_derivative"]
    flops = entries * FLOPS_PER_TABLE_ENTRY + pairs * FLOPS_PER_TABLE_PAIR
    achieved = flops / (ref_ms * 1e-3) / 1e12
    roof = {"bound": "valu", "kernel": "k_lcc_ref_table", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "launch_ms": ref_ms,
            "table_entries_per_launch": entries, "flops_per_entry": FLOPS_PER_TABLE_ENTRY,
            "pairs_per_launch": pairs, "flops_per_pair": FLOPS_PER_TABLE_PAIR,
            "traffic": profiled_traffic("k_lcc_ref_table", key),
            "reference_triples_per_s": round(tri / (ref_ms * 1e-3), 1)}
    if roof["traffic"]:
        roof["hbm_traffic_frac"] = round(roof["traffic"] / (ref_ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)
    out["roofline"] = roof
    # the centre walk (the reference's enumeration order) on its own work
    walk_ms = out["reference_windows_walk"]["pass_ms"]["lcc_derivative"]
    walk_tf = tri * FLOPS_PER_LCC_TRIPLE / (walk_ms * 1e-3) / 1e12
    out["reference_windows_walk"]["roofline"] = {
        "bound": "valu", "kernel": "k_lcc_gather_ref", "achieved": round(walk_tf, 3), "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s", "frac": round(walk_tf / FP32_PEAK_TFLOPS, 4), "launch_ms": walk_ms,
        "triples_per_launch": tri, "flops_per_triple": FLOPS_PER_LCC_TRIPLE,
        "traffic": profiled_traffic("k_lcc_gather_ref", key)}
    # the exact-gradient path's streaming kernels against HBM
    hbm = {}
    for kname, pname in (("k_lcc_centres", "lcc_centres"), ("k_dense_warp", "lcc_dense")):
        t = profiled_traffic(kname, key)
        ms_k = out["exact_gradient"]["pass_ms"].get(pname) if pname else None
        if t and ms_k:
            hbm[kname] = {"traffic": t, "ms": ms_k, "GB/s": round(t / (ms_k * 1e-3) / 1e9, 1),
                          "frac": round(t / (ms_k * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)}
    out["exact_gradient"]["hbm"] = hbm or None
    out["cpu_baseline"] = None if args.no_cpu_baseline else lcc_cpu_baseline(
        mov, fix, h, s, sub, subn, tri, bilateral_args(args, 1e-3))
    return out


def lcc_registration(args):
    """Wall time of one 3-level LCC registration of the sliding-motion
    volumes (256x256x128, two halves shifted 4 voxels in opposite directions
    plus breathing motion): support 80/40/20, control-point spacing
    support/4 (51x51x26 at the finest level), isotropic Wendland kernel,
    LCC with a 7^3 neighbourhood (strides 23/12/6), the reference's windowed
    derivative, bilateral 1e-3 (spatial sigma 12, intensity sigma 0.1),
    sampling rates 4/3/2, --lcc-iterations
    adaptive-step iterations per level at step 1000.  One timed run, then one
    with the exact-gradient derivative."""
    from skmimageregistration_amd import Image, RegistrationConfig, register_multiscale, synthetic
    mov, fix, disp = synthetic.sliding_3d((128, 256, 256))
    s = 20.0
    cfg = RegistrationConfig(sigma_C4=[4 * s, 2 * s, s], sigma_C0=[4 * s, 2 * s, s], tensor_weights=[0.0],
                             num_scales=3, num_function_evaluations=[args.lcc_iterations],
                             initial_step_size=[1000.0], sampling_rate=[4, 3, 2], neighborhood_sampling=[23, 12, 6],
                             smoothing_variance=[2, 1, 0], metric="lcc", **bilateral_config(args, 1e-3))
    t0 = time.perf_counter()
    df, reports = register_multiscale(Image(mov), Image(fix), cfg)
    ms = (time.perf_counter() - t0) * 1e3
    inner = (slice(10, -10),) * 3
    cfg.values["exact_gradient"] = [1]
    t0 = time.perf_counter()
    df_x, rep_x = register_multiscale(Image(mov), Image(fix), cfg)
    ms_x = (time.perf_counter() - t0) * 1e3
    exact = {"ms_per_registration": round(ms_x, 1), "final_value": [round(r["value"], 6) for r in rep_x],
             "level_ms_optimize": [round(r["ms_optimize"], 1) for r in rep_x],
             "mean_abs_error_vox": round(float(np.abs(df_x[inner] - disp[inner]).mean()), 4)}
    return {"workload": "LCC registration, 3D 256x256x128 sliding motion, 3 levels (support 80/40/20, "
                        f"cp spacing support/4), 7^3 neighbourhood, bilateral 0.001, {args.lcc_iterations} iterations "
                        "per level, reference windowed derivative",
            "ms_per_registration": round(ms, 1),
            "cp_size": [list(r["cp_size"]) for r in reports],
            "first_value": [round(r["first_value"], 6) for r in reports],
            "final_value": [round(r["value"], 6) for r in reports],
            "level_ms": [{k: round(r[k], 1) for k in ("ms_setup", "ms_optimize", "ms_displacement")} for r in reports],
            "mean_abs_error_vox": round(float(np.abs(df[inner] - disp[inner]).mean()), 4),
            "mean_abs_displacement_vox": round(float(np.abs(disp[inner]).mean()), 4),
            "exact_gradient": exact}


def registration_config(args, autotune=0):
    """3 levels with support 4s, 2s, s, control-point spacing support/4,
    tensors from the target with weights 400/200/100, NCC + bilateral, sampling
    rates 5/4/3, 200 adaptive-step iterations per level."""
    from skmimageregistration_amd import RegistrationConfig
    s = args.sigma
    return RegistrationConfig(sigma_C4=[4 * s, 2 * s, s], sigma_C0=[4 * s, 2 * s, s],
                              tensor_weights=[args.tensor_weight, args.tensor_weight / 2, args.tensor_weight / 4]
                              if args.kernel == "anisotropic" else [0.0],
                              num_scales=3, num_function_evaluations=[200], initial_step_size=[1000.0],
                              sampling_rate=[5, 4, 3], neighborhood_sampling=[1], smoothing_variance=[2, 1, 0],
                              metric="ncc", autotune=[int(autotune)],
                              **bilateral_config(args, args.registration_bilateral))


def bilateral_args(args, weight):
    """(weight, spatial sigma, intensity sigma) of the bilateral term."""
    return (weight, args.bilateral_spatial_sigma, args.bilateral_intensity_sigma)


def bilateral_config(args, weight):
    """RegistrationConfig keys of the bilateral term."""
    return {"reg_bilateral": [weight], "reg_bilateral_spatial_sigma": [args.bilateral_spatial_sigma],
            "reg_bilateral_intensity_sigma": [args.bilateral_intensity_sigma]}


def registration(args, mov, fix, disp):
    """Wall time of one multi-scale registration (SKMReg's scale loop,
    SKMReg.cxx:123-500) with registration_config's levels, then the dense
    displacement field.  Host images in, dense field out, timed after one
    untimed run; the median of the repeats."""
    from skmimageregistration_amd import Image, register_multiscale
    s = args.sigma
    # with autotune the untimed first run tunes each level's launch shapes
    cfg = registration_config(args, autotune=args.registration_autotune)
    register_multiscale(Image(mov), Image(fix), cfg)
    times, reports, df = [], None, None
    for _ in range(max(args.registration_repeats, 1)):
        t0 = time.perf_counter()
        df, reports = register_multiscale(Image(mov), Image(fix), cfg)
        times.append((time.perf_counter() - t0) * 1e3)
    inner = (slice(10, -10),) * 3
    return {"ms_per_registration": round(float(np.median(times)), 2),
            "levels": 3, "iterations_per_level": 200, "support": [4 * s, 2 * s, s],
            "cp_size": [list(r["cp_size"]) for r in reports],
            "first_value": round(reports[0]["first_value"], 6), "final_value": round(reports[-1]["value"], 6),
            "final_ncc": round(reports[-1]["measure"], 6),
            "level_ms": [{k: round(r[k], 2) for k in ("ms_setup", "ms_optimize", "ms_displacement")} for r in reports],
            "mean_abs_error_vox": round(float(np.abs(df[inner] - disp[inner]).mean()), 4),
            "mean_abs_displacement_vox": round(float(np.abs(disp[inner]).mean()), 4),
            "repeats_ms": [round(t, 2) for t in times], "reg_bilateral": args.registration_bilateral,
            "autotune": int(getattr(args, "registration_autotune", 0))}


def batch_timing(args):
    """Batched time series: frames of 96^3, 128x128x64 and 160x160x80 voxels
    (sizes drawn with seed 31) sampling one ellipsoid phantom, warped by a
    smooth field (max 7 voxels) times the sine of the frame's phase, plus
    N(0, 0.02); frames 1.. each registered to frame 0 with
    registration_config's levels.  Wall time of the whole series registered
    one at a time, with --batch-concurrency registrations in flight
    (skm_register_batch) and all of them in lockstep (one captured graph per
    optimizer step of the whole batch, skm_register_batch_lockstep), after
    one untimed pair; every run must give the same displacements bit for
    bit."""
    from skmimageregistration_amd import Image, register_batch, synthetic
    frames, disps = synthetic.time_series_3d(n_frames=args.batch_frames)
    cfg = registration_config(args)
    pairs = [(Image(frames[0]), Image(f)) for f in frames[1:]]
    register_batch(pairs[:1], cfg, concurrency=1)
    out = {"workload": f"{len(pairs)} registrations of a {len(frames)}-frame time series to frame 0, frames of "
                       "96^3 / 128x128x64 / 160x160x80 voxels, NCC + bilateral, 3 levels x 200 iterations",
           "frames": {f"{s[2]}x{s[1]}x{s[0]}": sum(f.shape == s for f in frames) for s in synthetic.SERIES_SIZES}}
    results = {}
    for conc in (1, args.batch_concurrency):
        t0 = time.perf_counter()
        results[conc] = register_batch(pairs, cfg, concurrency=conc)
        ms = (time.perf_counter() - t0) * 1e3
        out[f"concurrency_{conc}"] = {"ms": round(ms, 1), "ms_per_registration": round(ms / len(pairs), 2)}
    if not args.no_lockstep:
        t0 = time.perf_counter()
        results["lockstep"] = register_batch(pairs, cfg, lockstep=True)
        ms = (time.perf_counter() - t0) * 1e3
        out["lockstep"] = {"ms": round(ms, 1), "ms_per_registration": round(ms / len(pairs), 2)}
    a, b = results[1], results[args.batch_concurrency]
    out["identical"] = all(np.array_equal(x[0], y[0]) and ("lockstep" not in results or np.array_equal(x[0], z[0]))
                           for x, y, z in zip(a, b, results.get("lockstep", b)))
    out["speedup"] = round(out["concurrency_1"]["ms"] / out[f"concurrency_{args.batch_concurrency}"]["ms"], 3)
    if "lockstep" in out:
        out["speedup_lockstep"] = round(out["concurrency_1"]["ms"] / out["lockstep"]["ms"], 3)
    if not args.no_warm_start:
        out["warm_start"] = series_warm_start(args, pairs, b, out["concurrency_1"]["ms_per_registration"])
    inner = (slice(10, -10),) * 3
    out["final_ncc_mean"] = round(float(np.mean([r[1][-1]["measure"] for r in b])), 6)
    out["first_value_mean"] = round(float(np.mean([r[1][0]["first_value"] for r in b])), 6)
    out["mean_abs_error_vox"] = round(float(np.mean([np.abs(r[0][inner] - d[inner]).mean()
                                                     for r, d in zip(b, disps[1:])])), 4)
    out["mean_abs_displacement_vox"] = round(float(np.mean([np.abs(d[inner]).mean() for d in disps[1:]])), 4)
    return out


def series_warm_start(args, pairs, plain, plain_single_ms):
    """The time series with the learned warm start: per frame the network's
    prediction (one network per frame size; its weights were trained on
    128x128x96 pairs only) and its dense field, both computed and handed to
    the registration in device memory, then the registration leg's
    last level alone (support s, tensor weight / 4, sampling rate 3, bilateral) for
    50 iterations from that field, guarded, --batch-concurrency frames in
    flight; also the same 50 iterations from zero.  Wall time per
    registration includes the predictions; final NCC against the plain
    3-level runs (`plain`, the same frames)."""
    import torch

    from skmimageregistration_amd import RegistrationConfig, register_batch
    from skmimageregistration_amd.warmstart import WarmStart
    weights = os.path.join(REPO, "skmimageregistration_amd", "weights", "warmstart_128x128x96.bin")
    if not os.path.exists(weights):
        return None
    s = args.sigma
    fine = RegistrationConfig(sigma_C4=[s], sigma_C0=[s], tensor_weights=[args.tensor_weight / 4]
                              if args.kernel == "anisotropic" else [0.0], num_scales=1,
                              num_function_evaluations=[50], initial_step_size=[1000.0], sampling_rate=[3],
                              neighborhood_sampling=[1], smoothing_variance=[0], metric="ncc",
                              **bilateral_config(args, args.registration_bilateral))
    nets = {}
    for _, t in pairs:
        if t.data.shape not in nets:
            nets[t.data.shape] = WarmStart(shape=t.data.shape, weights=weights)
            nets[t.data.shape].predict_field_device(t.data, t.data)  # first-call set-up, untimed
    register_batch(pairs[:1], fine, concurrency=1)
    t0 = time.perf_counter()
    ref = torch.from_numpy(pairs[0][0].data).cuda()  # frame 0, the reference of every pair
    fields = [nets[t.data.shape].predict_field_device(ref, t.data) for _, t in pairs]
    warm = register_batch(pairs, fine, concurrency=args.batch_concurrency, initial_displacements=fields,
                          initial_guard=True)
    ms = (time.perf_counter() - t0) * 1e3 / len(pairs)
    zero = register_batch(pairs, fine, concurrency=args.batch_concurrency)
    for ws in nets.values():
        ws.close()
    gaps = [abs(w[1][0]["measure"] - p[1][-1]["measure"]) / abs(p[1][-1]["measure"]) for w, p in zip(warm, plain)]
    used = {}
    for (_, t), w in zip(pairs, warm):
        k = "x".join(str(v) for v in reversed(t.data.shape))
        used.setdefault(k, [0, 0])
        used[k][0] += int(w[1][0]["initial_used"])
        used[k][1] += 1
    return {"ms_per_registration": round(ms, 2), "speedup_vs_plain_single": round(plain_single_ms / ms, 2),
            "final_ncc_mean": round(float(np.mean([w[1][0]["measure"] for w in warm])), 6),
            "zero_start_final_ncc_mean": round(float(np.mean([z[1][0]["measure"] for z in zero])), 6),
            "max_rel_ncc_gap_to_plain": round(float(max(gaps)), 5),
            "predictions_used_per_size": {k: f"{u}/{n}" for k, (u, n) in used.items()}}


BF16_DENSE_PEAK_TFLOPS = 2500.0
WS_CIN = (2, 16, 16, 32, 32, 48, 48, 64, 112, 80)
WS_COUT = (16, 16, 32, 32, 48, 48, 64, 64, 48, 32)
WS_LEVEL = (0, 0, 1, 1, 2, 2, 3, 3, 2, 1)


def network_timing(in_size, cp, steps=10, warmup=3, batch=8):
    """The warm-start network's own kernels at the workload's shape (input
    2 x 64 x 64 x 48, control points 22 x 22 x 16): one training step
    (forward, backward, Adam) of a batch of 8 and one forward of one pair,
    timed with HIP events on the network's stream.  Algorithmic flops of the
    ten 3x3x3 convolutions: forward 2 x voxels x 27 x cin x cout each, a
    training step three times that (data and weight gradients)."""
    import torch

    from skmimageregistration_amd.warmstart import WarmStartNet
    vox0 = in_size[0] * in_size[1] * in_size[2]
    fwd = sum(2.0 * (vox0 >> (3 * lv)) * 27 * ci * co for ci, co, lv in zip(WS_CIN, WS_COUT, WS_LEVEL))
    origin = tuple((o - 0.5) / 2.0 for o in cp.origin)
    spacing = tuple(h / 2.0 for h in cp.spacing)
    net = WarmStartNet(in_size, cp.size(), origin, spacing, max_batch=batch)
    rng = np.random.default_rng(5)
    net.set_params(rng.normal(0.0, 0.05, net.n_params).astype(np.float32))
    x = torch.tensor(rng.uniform(0, 1, (batch, in_size[2], in_size[1], in_size[0], 2)).astype(np.float32),
                     device="cuda")
    t = torch.zeros((batch, cp.size()[2], cp.size()[1], cp.size()[0], 3), device="cuda")
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        net.train_step(x, t, 1e-4)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(steps):
        net.train_step(x, t, 1e-4)
    b.record(stream)
    torch.cuda.synchronize()
    train_ms = a.elapsed_time(b) / steps
    x1 = x[:1].contiguous()
    for _ in range(warmup):
        net.forward(x1)
    a.record(stream)
    for _ in range(steps):
        net.forward(x1)
    b.record(stream)
    torch.cuda.synchronize()
    fwd_ms = a.elapsed_time(b) / steps
    net.close()
    train_tf = 3.0 * fwd * batch / (train_ms * 1e-3) / 1e12
    return {"train_step_ms": round(train_ms, 3), "batch": batch, "forward_ms_per_pair": round(fwd_ms, 3),
            "flops_forward_per_pair": fwd,
            "roofline": {"bound": "mfma", "achieved": round(train_tf, 2), "peak": BF16_DENSE_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(train_tf / BF16_DENSE_PEAK_TFLOPS, 5),
                         "traffic": None, "what": "training step (bf16 operands, fp32 accumulation)"},
            "forward_tflops": round(fwd / (fwd_ms * 1e-3) / 1e12, 2)}


def warm_start_timing(args):
    """The learned warm start on held-out synthetic pairs of the 128x128x96
    workload (seeds 900.., outside the training seeds): per pair the plain
    3-level registration (3 x 200 iterations), the warm-started one (network
    prediction + 50 fine-level iterations, guarded) and the same 50 iterations
    from zero; then the warm-started registrations of --warm-pairs pairs run
    --batch-concurrency at a time (skm_register_batch).  Wall times include
    host images in and the dense field out."""
    from skmimageregistration_amd import Image, register_batch
    from skmimageregistration_amd.warmstart import WarmStart, evaluate, fine_config, synthetic_pair
    weights = os.path.join(REPO, "skmimageregistration_amd", "weights", "warmstart_128x128x96.bin")
    if not os.path.exists(weights):
        return None
    ws = WarmStart(weights=weights)
    ref, tar, _ = synthetic_pair(899)
    ws.register(ref, tar, 50)  # first-call set-up, untimed
    rows = evaluate(ws, list(range(900, 900 + args.warm_pairs)), log=lambda m: None)
    # batched: predictions first (one network pass per pair), then the guarded fine-level runs in flight together
    pairs = [synthetic_pair(900 + i)[:2] for i in range(args.warm_pairs)]
    t0 = time.perf_counter()
    fields = [ws.predict_field_device(r, t) for r, t in pairs]
    res = register_batch([(Image(r), Image(t)) for r, t in pairs], fine_config(50), concurrency=args.batch_concurrency,
                         initial_displacements=fields, initial_guard=True)
    ms_batch = (time.perf_counter() - t0) * 1e3 / len(pairs)
    res_ls, ms_lockstep = None, None
    if not args.no_lockstep:
        t0 = time.perf_counter()
        fields = [ws.predict_field_device(r, t) for r, t in pairs]
        res_ls = register_batch([(Image(r), Image(t)) for r, t in pairs], fine_config(50), lockstep=True,
                                initial_displacements=fields, initial_guard=True)
        ms_lockstep = (time.perf_counter() - t0) * 1e3 / len(pairs)
    net = network_timing(ws.net.in_size, ws.cp)
    ws.close()
    plain_ms = float(np.mean([r["plain_ms"] for r in rows]))
    warm_ms = float(np.mean([r["warm_ms"] for r in rows]))
    return {"workload": f"{len(rows)} held-out 128x128x96 pairs (max displacement 6), NCC + bilateral 1e-5; plain = "
                        "3 levels x 200 iterations, warm = prediction + 50 fine-level iterations",
            "plain_ms_per_registration": round(plain_ms, 2), "warm_ms_per_registration": round(warm_ms, 2),
            "warm_batched_ms_per_registration": round(ms_batch, 2),
            "warm_lockstep_ms_per_registration": None if ms_lockstep is None else round(ms_lockstep, 2),
            "speedup_single": round(plain_ms / warm_ms, 2), "speedup_batched": round(plain_ms / ms_batch, 2),
            "speedup_lockstep": None if ms_lockstep is None else round(plain_ms / ms_lockstep, 2),
            "lockstep_identical_to_batched": None if res_ls is None else
            all(np.array_equal(x[0], y[0]) for x, y in zip(res, res_ls)),
            "plain_final_ncc": round(float(np.mean([r["plain_ncc"] for r in rows])), 5),
            "warm_final_ncc": round(float(np.mean([r["warm_ncc"] for r in rows])), 5),
            "warm_batched_final_ncc": round(float(np.mean([r[1][0]["measure"] for r in res])), 5),
            "zero_start_final_ncc": round(float(np.mean([r["zero_ncc"] for r in rows])), 5),
            "max_rel_ncc_gap": round(float(max(abs(r["warm_ncc"] - r["plain_ncc"]) / abs(r["plain_ncc"])
                                               for r in rows)), 5),
            "predictions_used": int(sum(r["warm_used"] for r in rows)),
            "mean_abs_error_vox": {k: round(float(np.mean([r[k + "_err"] for r in rows])), 4)
                                   for k in ("plain", "warm", "zero")},
            "network": net}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from skmimageregistration_amd import GpuKernelEvaluator, Grid, Image, MEASURE

    mov, fix, disp = workload(args)
    shape = fix.shape
    size, cs, co = geometry(shape, args.cp_spacing)
    n_cp = int(np.prod(cs))
    measure = {"mse": MEASURE.MSE, "ncc": MEASURE.NCC, "lcc": MEASURE.LCC}[args.measure]
    kernel, ten, mcov, tensor_ms = kernel_for(args, fix)
    ev = GpuKernelEvaluator(n_cp * 3, args.subsample, 1, kernel, Image(mov), Image(fix),
                            Grid(cs, co, (args.cp_spacing,) * 3))
    if args.reg_rd > 0:
        ev.SetRegularizerRD(args.reg_rd)
    if args.reg_bilateral > 0:
        ev.SetRegularizerBilateral(*bilateral_args(args, args.reg_bilateral))
    if args.gather_split:
        ev.set_tuning(gather_split=args.gather_split)
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    ev.set_stream(stream.cuda_stream)
    rng = np.random.default_rng(1)
    # two parameter sets, alternated: consecutive evaluations have different inputs
    cps = [torch.tensor(rng.uniform(-1, 1, size=(n_cp, 3)).astype(np.float32), device=dev) for _ in range(2)]
    cpw = torch.ones(n_cp, dtype=torch.float32, device=dev)
    der = torch.empty(n_cp * 3, dtype=torch.float32, device=dev)

    def step(i):
        ev.get_value_and_derivative_device(measure, cps[i & 1].data_ptr(), cpw.data_ptr(), der.data_ptr(), True)

    def timed(n):
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for i in range(n):
            step(i)
        stop.record(stream)
        torch.cuda.synchronize()
        return start.elapsed_time(stop)

    def pass_times(n):
        ev.enable_timing(True)
        ev.read_timing()
        for i in range(n):
            step(i)
        t = ev.read_timing()
        ev.enable_timing(False)
        return {k: v[0] / max(v[1], 1) for k, v in t.items() if v[1] > 0}

    if args.calibration_copy > 0:
        src = torch.full((args.calibration_copy * (1 << 18),), 1.0, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    # the timed steps start without the kernel matrix: its assembly (at the
    # first step) is inside the timed region
    ev.release_kernel_matrix()
    total_ms = timed(args.steps)
    ms_step = total_ms / args.steps
    km = ev.kernel_matrix_info()
    per_kernel = pass_times(min(args.steps, 50))  # per-pass device times, assembled
    steady_ms = timed(min(args.steps, 50)) / min(args.steps, 50)
    # the same evaluations with the kernels evaluated on the fly
    ev.set_tuning(gather_split=args.gather_split, kernel_matrix=1)
    for i in range(2):
        step(i)
    n_fly = min(args.steps, 20)
    fly_ms = timed(n_fly) / n_fly
    fly_kernel = pass_times(n_fly)
    ev.close()

    exact = args.subsample == 1 and not args.no_pair_count
    vbox, dbox, vnz, dnz = pair_counts(args, shape, ten, mcov, dev, exact)
    value = (vbox + dbox) / (ms_step * 1e-3) / 1e9
    key = workload_key(args, shape)
    # the on-the-fly passes: FP32-VALU bound (kernel evaluation per pair)
    fly_roof = None
    fpp = FLOPS_PER_PAIR[args.kernel]
    if vnz is not None:
        dom = max(("value", "gather"), key=lambda k: fly_kernel.get(k, 0.0))
        pairs = vnz if dom == "value" else dnz
        achieved = pairs * fpp / (fly_kernel[dom] * 1e-3) / 1e12
        kname = "k_value" if dom == "value" else "k_gather_tile"
        fly_roof = {"bound": "valu", "kernel": kname, "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                    "traffic": profiled_traffic(kname, key), "launch_ms": round(fly_kernel[dom], 4),
                    "pairs_per_launch": pairs, "flops_per_pair": fpp}
    # the kernel-matrix passes: HBM bound (streaming the stored pairs)
    roof = None
    if km["state"] == 1:
        n_smp = int(np.prod(shape))
        per_smp = {"value": 4 + 4 + 16 + (4 if args.measure == "ncc" else 0),  # F, coefficients, (gradM, a), b
                   "gather": 16}                                              # v
        layout = {"value": km["read_value"], "gather": km["read_gather"]}
        nnz = {"value": km["pairs_value"], "gather": km["pairs_gather"]}
        dom = max(("value", "gather"), key=lambda k: per_kernel.get(k, 0.0))
        t = per_kernel[dom] * 1e-3
        algo = layout[dom] + per_smp[dom] * n_smp
        kname = "k_km_value" if dom == "value" else "k_km_gather"
        roof = {"bound": "hbm", "kernel": kname, "achieved": round(algo / t / 1e9, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(algo / t / 1e9 / HBM_PEAK_GBS, 4),
                "traffic": profiled_traffic(kname, key), "launch_ms": round(per_kernel[dom], 4),
                "bytes_per_launch": algo, "layout_bytes": layout[dom], "pairs_per_launch": nnz[dom],
                # the same launch at 4 bytes per stored non-zero pair (no records, no zero slots)
                "frac_at_4_bytes_per_pair": round((4 * nnz[dom] + per_smp[dom] * n_smp) / t / 1e9 / HBM_PEAK_GBS, 4)}
    else:
        roof = fly_roof  # no kernel matrix (subsampled, or refused): the timed steps ran on the fly
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, mov, fix, ten, mcov)
    kdesc = ("anisotropic Wendland (tensors from the target, weight %g)" % args.tensor_weight
             if args.kernel == "anisotropic" else "stationary Wendland")
    line = {
        "metric": "GVox*ctrl/s (voxel x control-point pair evaluations per second, value + derivative pass)",
        "value": round(value, 3), "unit": "GVox*ctrl/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (FP64 final sums)", "data": "synthetic",
        "config": {"workload": f"SKM GetValueAndDerivative, 3D {size[0]}x{size[1]}x{size[2]}, cp {cs[0]}x{cs[1]}x{cs[2]} "
                               f"spacing {args.cp_spacing:g}, {kdesc} support {args.sigma:g}, "
                               f"{args.measure.upper()}{regularizer_desc(args)}, subsample {args.subsample}",
                   "pairs_value": vbox, "pairs_derivative": dbox,
                   "pairs_nonzero_value": vnz, "pairs_nonzero_derivative": dnz},
        "pass_ms": {k: round(v, 4) for k, v in per_kernel.items()},
        "roofline": roof, "cpu_baseline": cpu,
        "kernel_matrix": {"state": km["state"], "assembly_ms": round(km["build_ms"], 3),
                          "ms_per_step_assembled": round(steady_ms, 4), "bytes": km["bytes"],
                          "read_value": km["read_value"], "read_gather": km["read_gather"],
                          "pairs_value": km["pairs_value"], "pairs_gather": km["pairs_gather"]},
        "on_the_fly": {"ms_per_step": round(fly_ms, 4), "value": round((vbox + dbox) / (fly_ms * 1e-3) / 1e9, 3),
                       "pass_ms": {k: round(v, 4) for k, v in fly_kernel.items()}, "roofline": fly_roof},
    }
    if tensor_ms is not None:
        line["tensor_generation_ms"] = round(tensor_ms, 2)
    if not args.no_registration:
        line["registration"] = registration(args, mov, fix, disp)
    # the series before the 256x256x128 LCC legs: right after them its four
    # registrations in flight took 50.6 ms each instead of 43.0
    if not args.no_batch:
        line["batch"] = batch_timing(args)
    if not args.no_lcc:
        line["lcc"] = lcc_timing(args)
    if not args.no_lcc_registration:
        line["lcc_registration"] = lcc_registration(args)
    if not args.no_warm_start:
        line["warm_start"] = warm_start_timing(args)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
