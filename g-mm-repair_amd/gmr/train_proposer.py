"""Train the schedule proposer on sweep timings and evaluate it on held-out
grammars.

Training set: the haplotype matrix (4,096 x 2,048, founders 64, density 0.08,
segments 128, flips 1%) and genotype matrices (10,000 columns, 256 founders in
16 classes, segments 512; 500,000 rows and smaller row counts), swept at
several vector counts B (1, 4, 16, 64, 256 and a few widths that are not
multiples of 4). Every sample comes from Engine.time_levels / whole-sweep
timings of the engine itself.

Held-out set: a heterogeneous batch of grammars -- genotype matrices and R-MAT
graphs with nnz log-uniform in [nnz_lo, nnz_hi], each with its own B ~
U{1..128}. For each grammar the proposer's choice is scored against the
exhaustive optimum over the same timing table (regret) and against the plain
schedule (per-level launches, emission on the spot, thread-per-rule, tile = B)
(gain). The whole held-out set is then swept in one fused persistent launch
with the proposer's per-grammar choices.

Usage: python -m gmr.train_proposer --out DIR [--quick]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

T0 = time.time()


def log(*a):
    print(f"[proposer {time.time() - T0:6.1f}s]", *a, file=sys.stderr, flush=True)


def training_set(quick, graphs=False, vary=False):
    """Matrices of the haplotype (config 1) and genotype (config 2) families,
    the genotype ones at row counts from 64 to 500,000 so the set spans the
    shapes the decisions depend on (row-parallel work, level widths); with
    `vary`, also members of those two families with their generator parameters
    moved towards weak compression (dense, noisy haplotypes; genotypes with
    short founder segments and many founders); with `graphs`, also R-MAT graphs
    (the config 4 family) at several scales."""
    import gmr
    if vary:
        yield "hap4096x2048_noisy", gmr.Matrix.haplotype(31, 4096, 2048, density=0.2, flip=0.15)
        yield "hap65536x1024_noisy", gmr.Matrix.haplotype(32, 65536, 1024, density=0.1, flip=0.1, seg_mean=16.0)
        yield "hap262144x256_sparse", gmr.Matrix.haplotype(33, 262144, 256, density=0.05, flip=0.05, seg_mean=8.0)
        yield "geno200000x10000_short", gmr.Matrix.genotype(34, 200000, 10000, founders=2048, seg_mean=8.0)
        yield "geno20000x10000_short", gmr.Matrix.genotype(35, 20000, 10000, founders=1024, seg_mean=16.0)
    if graphs:
        for scale, edges, seed in ((12, 60000, 31), (15, 450000, 32), (18, 3500000, 33), (21, 28000000, 34)):
            yield f"rmat{scale}", gmr.Matrix.rmat(seed, scale, edges)
    if quick:
        yield "hap1024x512", gmr.Matrix.haplotype(20261204, 1024, 512, founders=16, seg_mean=32.0)
        yield "hap2048x1024", gmr.Matrix.haplotype(7, 2048, 1024, founders=32, seg_mean=64.0)
        for rows, seed in ((64, 11), (300, 12), (3000, 13)):
            yield f"geno{rows}x10000", gmr.Matrix.genotype(seed, rows, 10000)
        return
    yield "hap4096x2048", gmr.Matrix.haplotype(20261204, 4096, 2048)
    yield "hap4096x2048_s2", gmr.Matrix.haplotype(20261214, 4096, 2048)
    for rows, seed in ((500000, 20261205), (150000, 20261215), (40000, 20261225), (10000, 20261235),
                       (2000, 20261245), (300, 20261255), (64, 20261265)):
        yield f"geno{rows}x10000", gmr.Matrix.genotype(seed, rows, 10000)


def heldout_set(n, seed, nnz_lo, nnz_hi):
    """(name, Matrix, B): genotype and R-MAT, nnz log-uniform, B ~ U{1..128}."""
    import gmr
    rng = np.random.default_rng(seed)
    for i in range(n):
        target = math.exp(rng.uniform(math.log(nnz_lo), math.log(nnz_hi)))
        B = int(rng.integers(1, 129))
        if i % 2 == 0:
            rows = max(64, int(target / 1022.0))  # mean allele count ~1022 at 10,000 columns
            m = gmr.Matrix.genotype(seed * 1000 + i, rows, 10000)
            yield f"geno{rows}", m, B
        else:
            scale = max(8, int(round(math.log2(target / 14.0))))  # ~14 distinct edges per vertex after dedupe
            m = gmr.Matrix.rmat(seed * 1000 + i, scale, int(target))
            yield f"rmat{scale}", m, B


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--quick", action="store_true", help="small grammars (tests)")
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--heldout", type=int, default=32)
    ap.add_argument("--nnz-lo", type=float, default=1e5)
    ap.add_argument("--nnz-hi", type=float, default=1e8)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--graphs", action="store_true", help="also train on R-MAT graphs (config 4 family)")
    ap.add_argument("--vary", action="store_true",
                    help="also train on haplotype/genotype matrices generated with weak-compression parameters")
    ap.add_argument("--weights", default=None,
                    help="score these trained weights (.npz) on the held-out set instead of training")
    ap.add_argument("--save-weights", default=None,
                    help="write the proposer (network, baseline rule and margin) to this .npz after fitting")
    ap.add_argument("--constant-from-training", action="store_true",
                    help="with --weights: also time the training set, to pick the best constant schedule on it")
    ap.add_argument("--inline-top-from", type=int, default=0,
                    help="held-out grammars with B >= this run on Grammar.inline_top(2) (0: never)")
    args = ap.parse_args(argv)
    import torch

    import gmr
    from gmr import proposer as P

    os.makedirs(args.out, exist_ok=True)
    train_B = [1, 3, 4, 16, 29, 64, 100, 256] if not args.quick else [1, 4, 16, 37, 64]
    Xs, Ts, Ms = [], [], []
    train_problems = []  # (timing table, B): where the best constant schedule is chosen
    train_arrs = []      # (schedule arrays, timing table): where the decision margin is fitted
    t_collect = time.time()
    timed = not args.weights or args.constant_from_training
    for name, m in (training_set(args.quick, args.graphs, args.vary) if timed else []):
        g = m.repair()
        s = g.schedule(m.rows, m.cols, "resident")
        arr = s.arrays()
        eng = s.engine()
        table, (X, T, M) = P.collect(eng, arr, train_B, reps=args.reps)
        train_problems += [(table, B) for B in train_B]
        train_arrs.append((arr, table))
        Xs.append(X)
        Ts.append(T)
        Ms.append(M)
        log("collected", name, m.info(), "samples", len(X))
        del eng
        torch.cuda.empty_cache()
    t_collect = time.time() - t_collect
    if args.weights:
        prop = P.Proposer.load(args.weights)
        X, losses, t_train = np.zeros((0, 1)), [float("nan")], 0.0
        log("scoring", args.weights)
    else:
        X, T, M = np.vstack(Xs), np.vstack(Ts), np.vstack(Ms)
        prop = P.Proposer(seed=1)
        t_train = time.time()
        losses = prop.fit(X, T, M, epochs=args.epochs)
        t_train = time.time() - t_train
        log(f"trained on {len(X)} samples: loss {losses[0]:.4f} -> {losses[-1]:.4f} in {t_train:.1f}s")
    # The method without the network, chosen on the training timings, becomes the
    # proposer's fallback: the network's choice is followed only when it is
    # predicted to beat that constant schedule by a margin fitted on the same timings.
    best_train = P.best_constant(train_problems) if train_problems else None
    margins = None
    if best_train is not None:
        prop.baseline = best_train[0]
        margins = prop.fit_margin(train_arrs, train_B, log=log)
        log(f"baseline {P.describe_rule(prop.baseline)}, margin {prop.margin} (training totals {margins})")
    if not args.weights:
        prop.save(os.path.join(args.out, "proposer.npz"))
    if args.save_weights:
        prop.save(args.save_weights)
    del train_arrs

    # ---- held-out evaluation
    rows_out, engines, Xg, opts, held_problems = [], [], [], [], []
    t_tune_total = 0.0
    for name, m, B in heldout_set(args.heldout if not args.quick else 6, 20261203,
                                  args.nnz_lo if not args.quick else 2e4, args.nnz_hi if not args.quick else 3e5):
        g = m.repair()
        inlined = bool(args.inline_top_from) and B >= args.inline_top_from
        if inlined:
            g = g.inline_top(2)
        s = g.schedule(m.rows, m.cols, "resident")
        arr = s.arrays()
        eng = s.engine()
        tiles = P.tile_candidates(B)
        t0 = time.time()
        table, _ = P.collect(eng, arr, tiles, reps=args.reps)
        t_tune = time.time() - t0
        t_tune_total += t_tune
        t0 = time.time()
        cache = {}
        choice = prop.choose(arr, B, tiles=tiles, cache=cache)
        t_choose = time.time() - t0
        net = prop.choose(arr, B, tiles=tiles, use_baseline=False, cache=cache)
        ex = P.exhaustive(table, B)
        he = P.heuristic(table, B)
        mine = P.sweep_cost(table, B, choice[0], choice[1], choice[2])
        net_ms = P.sweep_cost(table, B, net[0], net[1], net[2])
        r = dict(name=name, nnz=m.nnz(), B=B, rules=arr["info"]["num_raw_nt"], depth=arr["info"]["max_depth"],
                 inlined=inlined,
                 exhaustive_ms=ex[0], exhaustive=dict(tile=ex[1], plan=P.PLAN_NAMES[ex[2]]),
                 proposer_ms=mine, proposer=dict(tile=choice[0], plan=P.PLAN_NAMES[choice[1]]),
                 plain_ms=he[0], regret=mine / ex[0] - 1.0, gain_over_plain=he[0] / mine,
                 network_only_ms=net_ms, network_only_regret=net_ms / ex[0] - 1.0,
                 followed_network=bool(choice[:3] == net[:3] and prop.baseline is not None),
                 tune_s=t_tune, choose_s=t_choose)
        rows_out.append(r)
        held_problems.append((table, B))
        log(f"{name:12s} nnz={r['nnz']:>10d} B={B:3d} regret={r['regret'] * 100:6.2f}% "
            f"gain={r['gain_over_plain']:5.2f}x  {r['proposer']} vs {r['exhaustive']}")
        engines.append(eng)
        gen = torch.Generator(device="cuda").manual_seed(len(engines))
        Xg.append(torch.randn(m.cols, B, device="cuda", generator=gen))
        opts.append({"emit": P.PLANS[choice[1]][1], "level_kind": np.array([0] + list(choice[2]), np.int32)})
    # The method without the network: one constant schedule for every grammar
    # and B, the best one on the training set (and, stricter, the best one on
    # the held-out set itself, chosen in hindsight).
    constants = {}
    best_hind = P.best_constant(held_problems)
    for tag, bc in (("training", best_train), ("hindsight", best_hind)):
        if bc is None:
            continue
        costs = [P.constant_cost(t, B, bc[0]) for t, B in held_problems]
        if any(c is None for c in costs):
            continue
        for r, c in zip(rows_out, costs):
            r[f"constant_{tag}_ms"] = c
        creg = np.array([c / r["exhaustive_ms"] - 1.0 for r, c in zip(rows_out, costs)])
        cgain = np.array([c / r["proposer_ms"] for r, c in zip(rows_out, costs)])
        constants[tag] = dict(rule=P.describe_rule(bc[0]), regret_mean=float(creg.mean()),
                              regret_median=float(np.median(creg)), regret_max=float(creg.max()),
                              regret_of_totals=sum(costs) / sum(r["exhaustive_ms"] for r in rows_out) - 1.0,
                              proposer_gain_geomean=float(np.exp(np.log(cgain).mean())),
                              proposer_gain_of_totals=sum(costs) / sum(r["proposer_ms"] for r in rows_out),
                              proposer_wins=int((cgain > 1.0 + 1e-9).sum()),
                              proposer_losses=int((cgain < 1.0 - 1e-9).sum()))
        log(f"best constant schedule ({tag}): {constants[tag]}")
    reg = np.array([r["regret"] for r in rows_out])
    gain = np.array([r["gain_over_plain"] for r in rows_out])
    tot_mine = sum(r["proposer_ms"] for r in rows_out)
    tot_ex = sum(r["exhaustive_ms"] for r in rows_out)
    tot_plain = sum(r["plain_ms"] for r in rows_out)

    # ---- the whole held-out set in one fused launch (and the alternatives)
    batch = gmr.Batch(engines)
    fused = {}
    for vname, v in (("persistent", gmr.VARIANT_PERSISTENT), ("graph", gmr.VARIANT_GRAPH),
                     ("levels", gmr.VARIANT_LEVELS)):
        Ys = batch.run_torch(0, Xg, variant=v, opts=opts)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(3):
            batch.run_torch(0, Xg, Ys, variant=v, opts=opts)
        e1.record()
        torch.cuda.synchronize()
        if v == gmr.VARIANT_PERSISTENT:
            batch.check()
        fused[vname] = e0.elapsed_time(e1) / 3
    nnz_vec = sum(r["nnz"] * r["B"] for r in rows_out)
    report = dict(
        training=dict(samples=int(len(X)), epochs=args.epochs, graphs=bool(args.graphs), vary=bool(args.vary),
                      weights=args.weights, inline_top_from=args.inline_top_from,
                      loss_first=float(losses[0]),
                      loss_last=float(losses[-1]), collect_s=t_collect, train_s=t_train, B=train_B),
        heldout=dict(n=len(rows_out), regret_mean=float(reg.mean()), regret_median=float(np.median(reg)),
                     regret_max=float(reg.max()), regret_of_totals=tot_mine / tot_ex - 1.0,
                     gain_over_plain_geomean=float(np.exp(np.log(gain).mean())),
                     gain_over_plain_of_totals=tot_plain / tot_mine,
                     tune_s_total=t_tune_total, choose_s_total=sum(r["choose_s"] for r in rows_out),
                     best_constant=constants,
                     decision=dict(baseline=P.describe_rule(prop.baseline) if prop.baseline else None,
                                   margin=prop.margin, margin_training_totals=margins,
                                   followed_network=int(sum(r["followed_network"] for r in rows_out))),
                     network_only=dict(
                         regret_mean=float(np.mean([r["network_only_regret"] for r in rows_out])),
                         regret_median=float(np.median([r["network_only_regret"] for r in rows_out])),
                         regret_max=float(np.max([r["network_only_regret"] for r in rows_out])),
                         regret_of_totals=sum(r["network_only_ms"] for r in rows_out) / tot_ex - 1.0)),
        fused_ms=fused, fused_nnz_vec_per_s={k: nnz_vec / (v * 1e-3) for k, v in fused.items()},
        per_grammar=rows_out)
    with open(os.path.join(args.out, "proposer_report.json"), "w") as f:
        json.dump(report, f, indent=1)
    print(json.dumps({k: report[k] for k in ("training", "heldout", "fused_ms", "fused_nnz_vec_per_s")}), flush=True)


if __name__ == "__main__":
    main()
