"""Exhaustive schedule autotuner.

For one engine, semiring and batch width B, every candidate launch schedule
(kernel variant x emission strategy x vector tile x block size) is timed on
the device with HIP events, and the fastest is kept. Schedules never change the
arithmetic, so every candidate computes the same Y (FP32 sums may associate
differently; Boolean and Tropical are identical). Timings are cached on disk
keyed by a hash of the grammar shape, so a repeated run skips the search.
"""
import hashlib
import json
import os

from . import (BOOLEAN, EMIT_GATHER, EMIT_SPOT, EVAL_LDS, RESIDENT, RUN_BITS, RUN_FUSED_TAIL, RUN_LDS_RANGES,
               RUN_PAIRED_LEVELS, RUN_ROW_ORDER, RUN_ROW_SEGMENTS, RUN_SINGLE_LIST, VARIANT_GRAPH, VARIANT_LEVELS, VARIANT_PERSISTENT, GmrError)

VARIANT_NAMES = {VARIANT_LEVELS: "levels", VARIANT_GRAPH: "graph", VARIANT_PERSISTENT: "persistent"}
EMIT_NAMES = {EMIT_SPOT: "spot", EMIT_GATHER: "gather"}


def candidates(B, layout, blocks=(256,), persistent_shapes=((256, 0), (512, 0), (1024, 0), (256, 256)),
               semiring=None):
    """The search space for one engine and batch width: variant x emission x
    vector tile x launch shape (x gather order for row-gather emission of
    several vectors; x XCD-partitioned (1, 2, 4 or automatic passes),
    single-list or row-segment gather for one vector; x XCD-partitioned or
    row-segment gather for tiles of 2-4 vectors; the Boolean semiring on one
    vector adds the bit mode to the levels and graph variants), and the
    level-chunk sizes: for one vector, the chunks per wave of the
    XCD-partitioned gather, the LDS-range gather (levels and graph variants)
    at several range sizes and chunk counts, with the rows in matrix order
    or renumbered so similar rows sit together, the bit mode at several range
    sizes, and the LDS-staged node update (GMR_EVAL_LDS) at several window
    sizes; one vector in the Resident layout repeats the levels and graph
    schedules (bit mode aside) with the top levels fused into one workgroup
    (GMR_RUN_FUSED_TAIL), the row-gather ones also with small levels two to a
    launch (GMR_RUN_PAIRED_LEVELS), and the bit mode repeats its schedules
    with every level pair in one launch. Persistent shapes are (block, grid cap); a cap of 0 means the
    engine's default of 16 waves per CU."""
    emits = [EMIT_SPOT, EMIT_GATHER] if layout == RESIDENT else [EMIT_SPOT]
    tiles = sorted({B} | {t for t in (16, 32, 64, 128) if t < B})
    out = []
    for v in (VARIANT_LEVELS, VARIANT_GRAPH, VARIANT_PERSISTENT):
        for e in emits:
            for t in tiles:
                shapes = persistent_shapes if v == VARIANT_PERSISTENT else [(b, 0) for b in blocks]
                for blk, cap in shapes:
                    for order in ((0, 1) if e == EMIT_GATHER and B > 1 else (0,)):
                        # 2-4 vectors per tile: XCD-partitioned lists (they fix the order;
                        # the default) or row segments (in either order)
                        narrow = e == EMIT_GATHER and 2 <= t <= 4 and layout == RESIDENT
                        if e == EMIT_GATHER and B == 1:
                            flags = (0, RUN_ROW_SEGMENTS, RUN_SINGLE_LIST)
                            if v != VARIANT_PERSISTENT and layout == RESIDENT:
                                flags += (RUN_LDS_RANGES,)
                                if semiring == BOOLEAN:
                                    flags += (RUN_BITS,)
                        elif narrow:
                            flags = (0, RUN_ROW_SEGMENTS) if order == 0 else (RUN_ROW_SEGMENTS,)
                        else:
                            flags = (0,)
                        for fl in flags:
                            base = dict(variant=v, emit=e, b_tile=0 if t == B else t, block=blk, grid_cap=cap,
                                        gather_order=order)
                            if fl:
                                base["flags"] = fl
                            if e == EMIT_GATHER and (B == 1 or narrow) and fl == 0:
                                # XCD-partitioned gather: passes x chunks per wave (one vector)
                                for passes in (0, 1, 2, 4):
                                    for ch in ((2, 4) if B == 1 else (0,)):
                                        c = dict(base, xcd_passes=passes)
                                        if ch:
                                            c["chunks"] = ch
                                        out.append(c)
                            elif fl == RUN_LDS_RANGES:
                                # x rows in matrix order or renumbered (GMR_RUN_ROW_ORDER)
                                for span in (16384, 32768, 40960):
                                    for ch in (2, 4):
                                        for ro in (0, RUN_ROW_ORDER):
                                            out.append(dict(base, flags=fl | ro, lds_values=span, chunks=ch))
                            elif fl == RUN_BITS:
                                for rng in (262144, 524288, 1048576):
                                    for ch in (4, 8):
                                        out.append(dict(base, lds_values=rng, chunks=ch))
                            else:
                                out.append(base)
    if B == 1:
        # the LDS-staged node update at several windows, with the default emission forms
        for v in (VARIANT_LEVELS, VARIANT_GRAPH):
            for e in emits:
                for win in (2048, 4096, 16384):
                    out.append(dict(variant=v, emit=e, b_tile=0, block=blocks[0], grid_cap=0, gather_order=0,
                                    eval_kind=EVAL_LDS, lds_values=win))
        if layout == RESIDENT:
            # every levels / graph schedule (bit mode aside) also with its top levels in one
            # workgroup, and the row-gather ones also with their small levels two to a launch
            base_ = [c for c in out if c["variant"] != VARIANT_PERSISTENT and not c.get("flags", 0) & RUN_BITS]
            bits_ = [c for c in out if c.get("flags", 0) & RUN_BITS]
            out += [dict(c, flags=c.get("flags", 0) | RUN_FUSED_TAIL) for c in base_]
            out += [dict(c, flags=c.get("flags", 0) | RUN_FUSED_TAIL | RUN_PAIRED_LEVELS) for c in base_
                    if c["emit"] == EMIT_GATHER]
            # the bit mode with its levels two to a launch
            out += [dict(c, flags=c["flags"] | RUN_PAIRED_LEVELS) for c in bits_]
    return out


def apply(engine, cand):
    """Set the candidate's engine state; return the run keywords."""
    kw = dict(cand)
    order = kw.pop("gather_order", None)
    if order is not None:
        engine.set_gather_order(order)
    return kw


def describe(c):
    cap = c.get("grid_cap", 0)
    fl = c.get("flags", 0)
    lds = c.get("lds_values", 0)
    return (f"{VARIANT_NAMES[c['variant']]}/{EMIT_NAMES[c['emit']]}/tile={c['b_tile'] or 'B'}/blk={c['block']}"
            + (f"/grid<={cap}" if cap else "") + ("/order=rule" if c.get("gather_order") == 1 else "")
            + ("/segments" if fl & RUN_ROW_SEGMENTS else "")
            + ("/single-list" if fl & RUN_SINGLE_LIST else "")
            + (f"/bits(range={lds or 'default'})" if fl & RUN_BITS else "")
            + (f"/lds-ranges(values={lds or 'default'})" if fl & RUN_LDS_RANGES else "")
            + ("/row-order" if fl & RUN_ROW_ORDER else "")
            + ("/fused-tail" if fl & RUN_FUSED_TAIL else "")
            + ("/paired-levels" if fl & RUN_PAIRED_LEVELS else "")
            + (f"/xcd-passes={c['xcd_passes'] or 'auto'}" if "xcd_passes" in c else "")
            + (f"/chunks={c['chunks']}" if c.get("chunks") else "")
            + (f"/eval=lds(window={lds or 'default'})" if c.get("eval_kind") == EVAL_LDS else ""))


def shape_key(info, B, semiring):
    s = json.dumps({k: info[k] for k in sorted(info)}, sort_keys=True) + f"|B={B}|sr={semiring}"
    return hashlib.sha1(s.encode()).hexdigest()[:16]


def time_candidate(engine, semiring, X, Y, cand, reps=3, warm=1):
    """Mean device milliseconds of one sweep under `cand` (torch events on the
    current stream, which is where the engine launches)."""
    import torch
    cand = apply(engine, cand)
    for _ in range(warm):
        engine.run_torch(semiring, X, Y, **cand)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        engine.run_torch(semiring, X, Y, **cand)
    e1.record()
    torch.cuda.synchronize()
    if cand["variant"] == VARIANT_PERSISTENT:
        engine.check()
    return e0.elapsed_time(e1) / reps


def tune(engine, semiring, X, Y, cands, cache_path=None, reps=3, log=None):
    """Time every candidate; return (best candidate, [(cand, ms), ...])."""
    info = engine.info()
    B = 1 if X.dim() == 1 else X.shape[1]
    key = shape_key(info, B, semiring)
    db = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            db = json.load(f)
    results = []
    if key in db:
        cached = {json.dumps({k: v for k, v in r.items() if k != "ms"}, sort_keys=True): r["ms"] for r in db[key]}
        results = [(c, cached[json.dumps(c, sort_keys=True)]) for c in cands
                   if json.dumps(c, sort_keys=True) in cached]
        if len(results) != len(cands):
            results = []  # a different search space: time it again
        elif log:
            log(f"  tune: {len(results)} timings from {cache_path}")
    if not results:
        for c in cands:
            try:
                ms = time_candidate(engine, semiring, X, Y, c, reps=reps)
            except GmrError as e:
                # a schedule this engine or device refuses (an LDS range larger than a
                # workgroup may hold, a persistent block without residency headroom)
                ms = float("inf")
                if log:
                    log(f"  tune {describe(c):40s} refused: {e}")
            results.append((c, ms))
            if log and ms != float("inf"):
                log(f"  tune {describe(c):40s} {ms * 1e3:9.1f} us")
    best = min(results, key=lambda r: r[1])[0]
    if cache_path:
        db[key] = [dict(c, ms=ms) for c, ms in results]
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(db, f)
    return best, results


def tune_forms(forms, semiring, X, Y, cands, cache_path=None, reps=2, log=None):
    """Tune several engines of the same matrix -- e.g. the grammar as built
    and Grammar.inline_top(2) -- and keep the fastest. forms: [{"label": str,
    "engine": Engine, ...}]. Returns (form, best candidate, [(label, cand, ms),
    ...] over every form)."""
    runs = []
    for f in forms:
        best, res = tune(f["engine"], semiring, X, Y, cands, cache_path=cache_path, reps=reps, log=log)
        runs.append((min(ms for _, ms in res), f, best, res))
    _, form, best, _ = min(runs, key=lambda r: r[0])
    return form, best, [(f["label"], c, ms) for _, f, _, res in runs for c, ms in res]

