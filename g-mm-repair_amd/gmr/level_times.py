"""Per-level times of one sweep under each node-update kernel and block size:
the measurements the schedule proposer learns from, printed as a table.

The matrix is built with the engine's own generator and RePair builder, and
optionally has its top rules inlined; every (block, kernel) pair is timed with
Engine.time_levels (HIP events around every launch on the engine's stream).
Bytes per rule, for the rate column: 8 (the kids) + 3 x 4B (two child rows
read, one row written).

Usage: python -m gmr.level_times [--rows R --cols C --seed S] [--B 64]
           [--inline-top N] [--reps 10] [--blocks 128,256,512,1024]
"""
import argparse
import time

import numpy as np

import gmr


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rows", type=int, default=500000)
    ap.add_argument("--cols", type=int, default=10000)
    ap.add_argument("--seed", type=int, default=20261205)
    ap.add_argument("--B", type=int, default=64)
    ap.add_argument("--inline-top", type=int, default=2, help="inline top rules C uses at most N times (0: none)")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--blocks", default="128,256,512,1024")
    args = ap.parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        raise SystemExit("level_times needs a GPU")

    t = time.time()
    g = gmr.Matrix.genotype(args.seed, args.rows, args.cols).repair()
    if args.inline_top > 0:
        g = g.inline_top(args.inline_top)
    s = g.schedule(args.rows, args.cols, gmr.RESIDENT)
    lo = s.arrays()["level_offsets"]
    eng = s.engine()
    L = eng.info()["max_depth"]
    print(f"genotype {args.rows} x {args.cols} (seed {args.seed}), inline-top {args.inline_top}: "
          f"{int(lo[L])} rules, depth {L}, built in {time.time() - t:.1f} s", flush=True)
    X = torch.randn(args.cols, args.B, device="cuda")
    Y = torch.empty(args.rows, args.B, device="cuda")
    eng.run_torch(0, X, Y, emit=gmr.EMIT_GATHER)  # set-up (gather lists) outside the timings
    res = {}
    for blk in (int(b) for b in args.blocks.split(",")):
        for name, kind in (("group", gmr.EVAL_GROUP), ("wave", gmr.EVAL_WAVE), ("thread", gmr.EVAL_THREAD)):
            lv = eng.time_levels(0, X, Y, reps=args.reps, emit=gmr.EMIT_GATHER, block=blk, eval_kind=kind)
            res[(blk, name)] = lv
            print(f"block {blk:4d} {name:6s}: init {lv[0] * 1e3:8.1f} us  levels {lv[1:-1].sum() * 1e3:8.1f} us  "
                  f"gather {lv[-1] * 1e3:8.1f} us", flush=True)
    base_key = (256, "group") if (256, "group") in res else next(iter(res))
    base = res[base_key]
    per_rule = 8 + 12 * args.B
    print(f"per level at block {base_key[0]}, {base_key[1]} kernel ({per_rule} bytes per rule), and the fastest pair:")
    for k in range(1, L + 1):
        w = int(lo[k] - lo[k - 1])
        us = float(base[k]) * 1e3
        best = min(res, key=lambda key: res[key][k])
        rate = w * per_rule / (us * 1e-6) / 1e12 if us > 0 else 0.0
        print(f"  level {k:2d}: {w:9d} rules {us:8.1f} us {rate:6.2f} TB/s; fastest {best[0]} {best[1]} "
              f"{float(res[best][k]) * 1e3:.1f} us", flush=True)
    return res


if __name__ == "__main__":
    main()
