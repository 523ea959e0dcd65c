"""Trains the warm start of the 128x128x96 workload and writes its weights.

    python -m skmimageregistration_amd.warmstart_train --pairs 4096 --epochs 20 --out weights.bin

Synthetic pairs (seeds first..first+pairs-1) are registered by the plain
3-level method (3 x 200 iterations), each final field is fitted onto the fine
level's kernel basis (the labels), and the network is trained by the
library's own kernels (Adam, BF16 weights, FP32 accumulation).  Progress is
printed as it goes.
"""
import argparse
import json
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=4096)
    ap.add_argument("--first-seed", type=int, default=1000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--concurrency", type=int, default=4)
    ap.add_argument("--init-seed", type=int, default=0)
    ap.add_argument("--out", required=True)
    ap.add_argument("--log", default=None, help="JSON summary written here")
    args = ap.parse_args()

    from .warmstart import WarmStart, build_dataset, train

    def log(msg):
        print(msg, flush=True)

    t0 = time.perf_counter()
    ws = WarmStart(max_batch=args.batch)
    ws.net.set_params(he_init(ws.net.n_params, args.init_seed))
    seeds = list(range(args.first_seed, args.first_seed + args.pairs))
    x, y, nccs = build_dataset(ws, seeds, concurrency=args.concurrency, log=log)
    t_data = time.perf_counter() - t0
    hist = train(ws, x, y, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log)
    t_train = time.perf_counter() - t0 - t_data
    ws.net.save(args.out)
    summary = {"pairs": args.pairs, "first_seed": args.first_seed, "epochs": args.epochs, "batch": args.batch,
               "lr": args.lr, "seconds_data": round(t_data, 1), "seconds_train": round(t_train, 1),
               "plain_ncc_mean": float(np.mean(nccs)), "history": hist}
    log(json.dumps(summary))
    if args.log:
        with open(args.log, "w") as f:
            json.dump(summary, f, indent=1)


def he_init(n_params, seed):
    """He-normal conv weights, zero biases, small head (the layout of skm_ws_get_params)."""
    cin = (2, 16, 16, 32, 32, 48, 48, 64, 112, 80)
    cout = (16, 16, 32, 32, 48, 48, 64, 64, 48, 32)
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co in zip(cin, cout):
        parts.append(rng.normal(0.0, np.sqrt(2.0 / (27 * ci)), size=co * 27 * ci))
        parts.append(np.zeros(co))
    parts.append(rng.normal(0.0, 0.01, size=3 * 32))
    parts.append(np.zeros(3))
    p = np.concatenate(parts).astype(np.float32)
    if p.size != n_params:
        raise RuntimeError("parameter layout mismatch")
    return p


if __name__ == "__main__":
    main()
