"""Benchmark of one GetValueAndDerivative pass of the SKM path on one MI355X.

Workload (default): the 3D registration of 128x128x96 float32 volumes
(random smoothed ellipsoids, target warped by a smooth field, noise), control
points every 6 voxels (22x22x16 = 7,744, 23,232 parameters), Wendland support
24 voxels, NCC metric, RD regularizer, dense sampling.  One step is one
value-and-derivative evaluation with every input resident in HBM.

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
FLOPS_PER_PAIR = 32  # Wendland C4*C0 evaluation + weighted accumulate, stationary kernel


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--shape", default="96,128,128", help="z,y,x voxels")
    ap.add_argument("--cp-spacing", type=float, default=6.0)
    ap.add_argument("--sigma", type=float, default=24.0)
    ap.add_argument("--subsample", type=int, default=1)
    ap.add_argument("--measure", default="ncc", choices=["mse", "ncc", "lcc"])
    ap.add_argument("--reg-rd", type=float, default=1e-3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-crop", default="64,96,96", help="z,y,x voxels of the CPU baseline sample")
    return ap.parse_args()


def window_counts(n, origin, spacing, cp_n, cp_origin, cp_spacing, support, sub, tight):
    """Per-axis counts of the reference windows: for value (tight=True) the cp
    window around each sample; for the derivative the sample window per cp."""
    if tight:
        idx = (origin + spacing * np.arange(0, (n // sub) * sub, sub) - cp_origin) / cp_spacing
        r = support / cp_spacing
        lo = np.clip(np.trunc(idx - r + 1.0), 0, cp_n - 1)
        hi = np.clip(np.trunc(idx + r), 0, cp_n - 1)
        return hi - lo + 1
    pc = cp_origin + cp_spacing * np.arange(cp_n)
    idx = (pc - origin) / spacing
    r = support / spacing
    lim = (n // sub) - 1
    lo = np.clip(np.trunc((idx - r) / sub), 0, lim)
    hi = np.clip(np.trunc((idx + r) / sub + 1.0), 0, lim)
    return hi - lo + 1


def support_pairs(sample_axes, cp_axes, support):
    """Exact number of (sample, cp) pairs with |x - p| < support (k != 0)."""
    d2 = [(s[:, None] - c[None, :]) ** 2 for s, c in zip(sample_axes, cp_axes)]
    s2 = support * support
    dx = np.sort(d2[0].ravel())
    rest = d2[1].ravel()
    if len(d2) == 3:
        rest = (d2[1].ravel()[:, None] + d2[2].ravel()[None, :]).ravel()
    t = s2 - rest
    t = t[t > 0]
    return int(np.searchsorted(dx, t, side="left").sum())


def workload(args):
    from skmimageregistration_amd import synthetic
    shape = tuple(int(v) for v in args.shape.split(","))
    mov, fix, _ = synthetic.ellipsoids_3d(shape, n=20, smooth=1.5, max_disp=6.0, noise=0.02)
    return mov, fix


def geometry(shape, cp_spacing):
    size = tuple(reversed(shape))
    cs, co = [], []
    for d in range(3):
        upper = size[d] - 1.0
        n = int(np.ceil(upper / cp_spacing))
        cs.append(n)
        co.append(abs(upper - cp_spacing * (n - 1)) / 2.0)
    return size, tuple(cs), tuple(co)


def pair_counts(shape, cp_spacing, sigma, sub):
    size, cs, co = geometry(shape, cp_spacing)
    val = np.prod([window_counts(size[d], 0.0, 1.0, cs[d], co[d], cp_spacing, sigma, sub, True).sum()
                   for d in range(3)])
    der = np.prod([window_counts(size[d], 0.0, 1.0, cs[d], co[d], cp_spacing, sigma, sub, False).sum()
                   for d in range(3)])
    samples = [np.arange(0, (size[d] // sub) * sub, sub, dtype=np.float64) for d in range(3)]
    cps = [co[d] + cp_spacing * np.arange(cs[d]) for d in range(3)]
    insup = support_pairs(samples, cps, sigma) if sub == 1 else None
    return int(val), int(der), insup


def cpu_baseline(args, mov, fix):
    """FP64 oracle (single thread) on a crop of the same volumes, same kernel
    and control-point spacing; reported in the same unit."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as orc
    crop = tuple(int(v) for v in args.cpu_crop.split(","))
    f = fix[: crop[0], : crop[1], : crop[2]]
    m = mov[: crop[0], : crop[1], : crop[2]]
    size, cs, co = geometry(crop, args.cp_spacing)
    rng = np.random.default_rng(0)
    cp = rng.uniform(-1, 1, size=(int(np.prod(cs)), 3))
    pr = orc.OracleProblem(f, m, (cs, co, (args.cp_spacing,) * 3), cp, sigma_c0=args.sigma, sigma_c4=args.sigma)
    measure = {"mse": 0, "ncc": 1, "lcc": 2}[args.measure]
    t0 = time.perf_counter()
    pr.evaluate(measure, True, 0.0, args.reg_rd, 0.0)
    dt = time.perf_counter() - t0
    v, d, _ = pair_counts(crop, args.cp_spacing, args.sigma, 1)
    return {"value": (v + d) / dt / 1e9, "unit": "GVox*ctrl/s", "cores": 1, "kind": "port",
            "sample": f"FP64 oracle, one GetValueAndDerivative on a {crop[2]}x{crop[1]}x{crop[0]} crop "
                      f"(same sigma and cp spacing), {dt:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from skmimageregistration_amd import GpuKernelEvaluator, Grid, Image, Kernel, MEASURE

    mov, fix = workload(args)
    shape = fix.shape
    size, cs, co = geometry(shape, args.cp_spacing)
    n_cp = int(np.prod(cs))
    measure = {"mse": MEASURE.MSE, "ncc": MEASURE.NCC, "lcc": MEASURE.LCC}[args.measure]
    ev = GpuKernelEvaluator(n_cp * 3, args.subsample, 1, Kernel(args.sigma, args.sigma), Image(mov), Image(fix),
                            Grid(cs, co, (args.cp_spacing,) * 3))
    if args.reg_rd > 0:
        ev.SetRegularizerRD(args.reg_rd)
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    ev.set_stream(stream.cuda_stream)
    rng = np.random.default_rng(1)
    cp = torch.tensor(rng.uniform(-1, 1, size=(n_cp, 3)).astype(np.float32), device=dev)
    cpw = torch.ones(n_cp, dtype=torch.float32, device=dev)
    der = torch.empty(n_cp * 3, dtype=torch.float32, device=dev)

    def step():
        ev.get_value_and_derivative_device(measure, cp.data_ptr(), cpw.data_ptr(), der.data_ptr(), True)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(args.steps):
        step()
    stop.record(stream)
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(stop)
    ms_step = total_ms / args.steps
    # per-pass device times: HIP events around each pass, on the evaluator's stream
    ev.enable_timing(True)
    ev.read_timing()
    for _ in range(args.steps):
        step()
    passes = ev.read_timing()
    ev.enable_timing(False)

    val_pairs, der_pairs, insup = pair_counts(shape, args.cp_spacing, args.sigma, args.subsample)
    value = (val_pairs + der_pairs) / (ms_step * 1e-3) / 1e9
    per_kernel = {k: v[0] / max(v[1], 1) for k, v in passes.items() if v[1] > 0}
    dom = max(("value", "gather"), key=lambda k: per_kernel.get(k, 0.0))
    roof = None
    if insup is not None:
        achieved = insup * FLOPS_PER_PAIR / (per_kernel[dom] * 1e-3) / 1e12
        roof = {"bound": "valu", "kernel": "k_value" if dom == "value" else "k_gather", "achieved": round(achieved, 3),
                "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                "traffic": None, "launch_ms": round(per_kernel[dom], 4),
                "pairs_per_launch": insup, "flops_per_pair": FLOPS_PER_PAIR}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, mov, fix)
    line = {
        "metric": "GVox*ctrl/s (voxel x control-point pair evaluations per second, value + derivative pass)",
        "value": round(value, 3), "unit": "GVox*ctrl/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (FP64 final sums)", "data": "synthetic",
        "config": {"workload": f"SKM GetValueAndDerivative, 3D {size[0]}x{size[1]}x{size[2]}, cp {cs[0]}x{cs[1]}x{cs[2]} "
                               f"spacing {args.cp_spacing:g}, stationary Wendland support {args.sigma:g}, "
                               f"{args.measure.upper()} + RD {args.reg_rd:g}, subsample {args.subsample}",
                   "pairs_value": val_pairs, "pairs_derivative": der_pairs, "pairs_in_support": insup},
        "pass_ms": {k: round(v, 4) for k, v in per_kernel.items()},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ev.close()


if __name__ == "__main__":
    main()
