"""Benchmark: batched Y = M X over a RePair-compressed genotype matrix on one
MI355X, against the plain per-level sweep and the reference's CPU sweep.

Workload (defaults): the 500,000 x 10,000 genotype-like 0/1 matrix (per-row
allele count k with P(k) ~ 1/k, mosaics of 256 founders in 16 density classes,
segment mean 512, flips to hit k exactly; seed 20261205), RePair-compressed by
the engine's builder, B = 64 FP32 right-hand sides ~N(0, 1), X row-major
cols x B. A step is one sweep of the grammar over the B vectors with X and Y
resident in HBM. The launch schedule is chosen by the exhaustive autotuner
before timing. Prints one JSON line (rank 0).

Also measured, outside the headline value:
  plain      the plain per-level sweep at B = 1 (reference layout, one launch
             per level, atomics on the spot), in nnz/s;
  fp32_err   max_i |y_i - y*_i| / sum_j |M_ij x_j| against the FP64 oracle on a
             sample of columns;
  cpu_baseline  the reference's OpenMP CPU sweep (oracle restatement) on the
             same grammar, on a bounded sample of single-vector sweeps.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "g-mm-repair_amd"))
sys.path.insert(0, ROOT)

METRIC = "effective nnz·vectors/s and HBM GB/s (% of peak), batched vs plain sweep"
HBM_PEAK_GBS = 8000.0
T0 = time.time()


def log(*a):
    print(f"[bench {time.time() - T0:6.1f}s]", *a, file=sys.stderr, flush=True)


def rel_err_cols(y, y_star, den):
    """max |y - y*| / sum_j |M_ij x_j| (exact match required where the sum is 0)."""
    import numpy as np
    d = np.abs(np.asarray(y, np.float64) - np.asarray(y_star, np.float64))
    return float(np.max(np.where(den > 0, d / np.where(den > 0, den, 1), np.where(d == 0, 0, np.inf))))


def dag_bytes(info, B):
    """Algorithmic bytes of one sweep over B vectors, counted on the grammar DAG
    (schedule-independent): per rule 2 child ids + per vector 2 reads, 1 write;
    per emission (slot, row) + per vector 1 read and an RMW of y; per terminal
    (value, column) + per vector an x gather and a T write; y initialised."""
    R, E, Et, U, rows = info["num_raw_nt"], info["emit_total"], info["num_term_emit"], info["U"], info["rows"]
    return 8 * R + 8 * (E + Et) + 8 * (U - 1) + 4 * B * (3 * R + 3 * (E + Et) + 2 * (U - 1) + rows)


def amended_bytes(info, B):
    """The DAG count with the row gather's emission: per emission one 4-byte
    read per vector (no read-modify-write of y), and y written once; rules and
    terminals as in dag_bytes. Counted on the grammar form that runs."""
    R, E, Et, U, rows = info["num_raw_nt"], info["emit_total"], info["num_term_emit"], info["U"], info["rows"]
    return 8 * R + 8 * (E + Et) + 8 * (U - 1) + 4 * B * (3 * R + (E + Et) + 2 * (U - 1) + rows)


def cpu_cores():
    """(threads this process may run on, physical cores from lscpu or None)."""
    avail = len(os.sched_getaffinity(0))
    phys = None
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = {k.strip(): v.strip() for k, v in (ln.split(":", 1) for ln in out.splitlines() if ":" in ln)}
        phys = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except Exception:
        pass
    return avail, phys


def reference_driver_leg(g, rows, cols, x):
    """The reference's own driver (gpu_engine_test.cu, compiled unchanged against
    this engine's drop-in headers into oracle/_ref) on the same grammar and x:
    its OpenMP CPU sweep, its 1-thread sweep, its GPU-vs-CPU verification and
    its own timing of the drop-in (20 vectors) in one run. None if the binary
    was not built."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_ref", "gpu_engine_test")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        base = os.path.join(d, "g")
        g.write(base)
        xf = os.path.join(d, "x.bin")
        x.astype("float32").tofile(xf)
        env = dict(os.environ, SEMIRING="plustimes", XVEC=xf, GMR_ENERGY_SECONDS="0")
        t = time.time()
        r = subprocess.run([exe, base, str(rows), str(cols), "20"], capture_output=True, text=True, env=env,
                           timeout=600)
    out = r.stdout

    def num(tag, field=0):
        if tag not in out:
            return None
        return float(out.split(tag)[1].split()[field].rstrip(","))
    return {"exit_code": r.returncode, "wall_s": round(time.time() - t, 1),
            "cpu_openmp_ms": num("CPU reference (OpenMP):"),
            "cpu_openmp_threads": int(out.split("ms/vector (")[1].split()[0]) if "ms/vector (" in out else None,
            "cpu_sequential_ms": num("CPU reference (sequential):"),
            "gpu_vs_cpu_max_rel": (float(out.split("Verification: ")[1].split("max_rel_diff=")[1].split(",")[0])
                                   if "Verification: " in out else None),
            "dropin_gpu_ms_per_vector": (float(out.split("GPU kernel:")[1].split("ms/vector.")[0].split("=")[-1])
                                         if "GPU kernel:" in out else None),
            "verdict": "SUCCESS" if "SUCCESS: GPU and CPU reference results match" in out else "FAILURE",
            "note": "the reference's gpu_engine_test.cu, unchanged, linked with this engine's drop-in; "
                    "metric |gpu - cpu| / (1 + |cpu|), pass < 1e-4; the drop-in's time is the driver's "
                    "own event timing of its loop over the grammar as built"}


def layered_bytes(info_ref, B):
    """The reference-layered count of one sweep (RefCompat schedule of the
    grammar as built): per layered rule a 12-byte record and per vector 2 reads
    and 1 write, emissions and terminals as in the DAG count, plus the
    double-buffered frontier fills (2 w* per vector)."""
    RL, E, Et, U = info_ref["num_layered"], info_ref["emit_total"], info_ref["num_term_emit"], info_ref["U"]
    rows, w = info_ref["rows"], info_ref["max_width"]
    return 12 * RL + 8 * E + 8 * Et + 8 * (U - 1) + 4 * B * (3 * RL + 3 * E + 3 * Et + 2 * (U - 1) + rows + 2 * w)


def gather_bytes(info, B, n_seg=0):
    """Algorithmic bytes of one k_gather launch over B vectors (Resident layout,
    row gather): the row offsets, the segment table, one symbol id per top-
    sequence symbol, the B-wide value row (rule slot or terminal) of every symbol,
    and Y written once."""
    rows, syms = info["rows"], info["emit_total"] + info["num_term_emit"]
    return 4 * (rows + 1) + 16 * n_seg + 4 * syms + 4 * B * syms + 4 * B * rows


def read_traffic(workload, kernel):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE),
    from profiles/traffic.txt lines `<workload> <kernel> <bytes>`; None if absent."""
    p = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 3 and not line.startswith("#") and parts[0] == workload and parts[1] == kernel:
                return float(parts[2])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--B", type=int, default=64)
    ap.add_argument("--workload", default="genotype", choices=["genotype", "haplotype", "rmat"],
                    help="genotype: 500,000 x 10,000 (the headline); haplotype: 4,096 x 2,048; "
                         "rmat: 2^23-vertex graph, Boolean and min/plus (own JSON line)")
    ap.add_argument("--sweep-B", default=None,
                    help="comma-separated vector counts: throughput per B on this workload (own JSON line)")
    ap.add_argument("--left-B", default=None,
                    help="comma-separated vector counts: the left product X^T M per B on this workload (own JSON line)")
    ap.add_argument("--blocks", type=int, default=0,
                    help="compress the matrix as this many row-block grammars (matrepair -b) and sweep them in one "
                         "batch launch")
    ap.add_argument("--traverse", action="store_true",
                    help="rmat workload: also run the fixpoint traversal (reachability / shortest paths)")
    ap.add_argument("--rows", type=int, default=None)
    ap.add_argument("--cols", type=int, default=None)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bounded CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-vendor", action="store_true", help="skip the rocSPARSE CSR baseline and the energy reads")
    ap.add_argument("--no-tune", action="store_true", help="use the default schedule instead of autotuning")
    ap.add_argument("--semiring", default="all", choices=["all", "boolean", "minplus"],
                    help="rmat workload: which of its two products to run")
    ap.add_argument("--variant", default=None, help="force a variant (levels|graph|persistent)")
    ap.add_argument("--emit", default=None, help="force an emission strategy (spot|gather)")
    ap.add_argument("--b-tile", type=int, default=None)
    ap.add_argument("--schedule", default=None,
                    help="force a schedule given as a JSON object of run keywords (autotune candidate form), e.g. "
                         "'{\"variant\": 1, \"emit\": 1, \"flags\": 16, \"lds_values\": 32768}'")
    ap.add_argument("--profile", action="store_true",
                    help="only the timed batched region (for rocprofv3); no plain sweep, CPU baseline or check")
    ap.add_argument("--calibrate", action="store_true",
                    help="run a 1 GiB device-to-device copy first (known bytes, to scale HBM counters)")
    ap.add_argument("--layout", default="resident", choices=["resident", "wave64", "refcompat"],
                    help="schedule layout of the timed engine")
    ap.add_argument("--inline-top", default="auto",
                    help="grammar form: 0 = as built; N = rules no rule uses and C uses at most N times "
                         "written out in C (Grammar.inline_top); auto = tune over 0 and 2 (untuned runs: "
                         "2 for B >= 32, else 0)")
    ap.add_argument("--grammar-cache", default=None,
                    help="directory to save/load the RePair grammar files between runs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this engine's sweep runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import gmr
    from gmr import autotune

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    if args.workload == "rmat":
        return run_rmat(args)
    dflt = {"genotype": (500000, 10000, 20261205), "haplotype": (4096, 2048, 20261204)}[args.workload]
    args.rows = args.rows or dflt[0]
    args.cols = args.cols or dflt[1]
    args.seed = args.seed or dflt[2]
    B = args.B
    log("generating", args.workload, args.rows, "x", args.cols)
    if args.workload == "genotype":
        m = gmr.Matrix.genotype(args.seed, args.rows, args.cols)
    else:
        m = gmr.Matrix.haplotype(args.seed, args.rows, args.cols)
    nnz = m.nnz()
    if args.blocks:
        return run_blocks(args, m)
    base = None
    if args.grammar_cache:
        os.makedirs(args.grammar_cache, exist_ok=True)
        base = os.path.join(args.grammar_cache, f"{args.workload}_{args.rows}x{args.cols}_{args.seed}")
    if base and os.path.exists(base + ".vc.C"):
        g = gmr.Grammar.read(base)
        log("grammar loaded from", base)
    else:
        log("nnz", nnz, "; RePair ...")
        g = m.repair()
        log(f"RePair {g.build_seconds:.1f}s rounds={g.build_rounds}", g.info())
        if base:
            g.write(base)
    t = time.time()
    s_res = g.schedule(args.rows, args.cols, args.layout)
    s_ref = None if args.profile else g.schedule(args.rows, args.cols, "refcompat")
    sched_s = time.time() - t
    info = s_res.info()
    info_ref = s_ref.info() if s_ref else {"num_passthrough": None}
    log("schedules", f"{sched_s:.1f}s", "resident", info, "refcompat", info_ref)
    eng = s_res.engine()
    log("engine uploaded", eng.info())
    g0 = g
    forms = grammar_forms(args, g, s_res, eng, log)
    if args.left_B:
        return run_left_B(args, m, forms)
    if args.sweep_B:
        return run_sweep_B(args, m, forms)

    gen = torch.Generator(device="cuda").manual_seed(args.seed)
    X = torch.randn(args.cols, B, device="cuda", generator=gen, dtype=torch.float32)
    Y = torch.empty(args.rows, B, device="cuda", dtype=torch.float32)
    if B == 1:
        X, Y = X[:, 0].contiguous(), Y[:, 0].contiguous()

    # ---- schedule choice (grammar form, variant, emission, tile, launch shape)
    if args.no_tune or args.variant or args.schedule:
        form = untuned_form(args, forms, B)
        best = dict(variant=gmr.VARIANTS[args.variant or "persistent"], emit=gmr.EMITS[args.emit or "gather"],
                    b_tile=args.b_tile or 0, block=256)
        if args.schedule:
            best.update(json.loads(args.schedule))
        tuned = []
    else:
        form, best, tuned = tune_forms(forms, X, Y, autotune.candidates(B, gmr.LAYOUTS[args.layout]), reps=2,
                                       log=log)
    g, s_res, eng = form["grammar"], form["schedule"], form["engine"]
    info = s_res.info()
    log("grammar form:", form["label"])
    log("schedule:", autotune.describe(best))
    run_kw = autotune.apply(eng, best)

    if args.calibrate:
        # 1 GiB read + 1 GiB written by one copy kernel: the HBM counters of this
        # dispatch give the bytes-per-count scale for the sweep's dispatches.
        a_ = torch.empty(1 << 28, device="cuda", dtype=torch.float32).fill_(1.0)
        b_ = torch.empty_like(a_)
        torch.cuda.synchronize()
        b_.copy_(a_)
        torch.cuda.synchronize()
        del a_, b_
        log("calibration copy: 1073741824 bytes read, 1073741824 bytes written")

    # ---- timed region: W warm-up steps, then K steps, one event pair per step
    for _ in range(args.warmup):
        eng.run_torch(0, X, Y, **run_kw)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for i in range(args.steps):
        ev[i][0].record()
        eng.run_torch(0, X, Y, **run_kw)
        ev[i][1].record()
    stop.record()
    torch.cuda.synchronize()
    if best["variant"] == gmr.VARIANT_PERSISTENT:
        eng.check()
    total_ms = start.elapsed_time(stop)
    per = [a.elapsed_time(b) for a, b in ev]
    ms_step = total_ms / args.steps
    kern_ms = float(np.mean(per))
    value = nnz * B / (ms_step * 1e-3)
    log(f"batched B={B}: {ms_step:.3f} ms/step, launch avg {kern_ms:.3f} ms, {value / 1e9:.1f} G nnz*vec/s")

    # ---- the same step with X arriving from and Y returning to pinned host
    # memory (what a caller handing over host buffers sees); not the bench value
    pcie_ms = None
    if not args.profile:
        Xh = torch.empty(X.shape, dtype=X.dtype, pin_memory=True).copy_(X)
        Yh = torch.empty(Y.shape, dtype=Y.dtype, pin_memory=True)
        torch.cuda.synchronize()
        h0, h1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        h0.record()
        for _ in range(args.steps):
            X.copy_(Xh, non_blocking=True)
            eng.run_torch(0, X, Y, **run_kw)
            Yh.copy_(Y, non_blocking=True)
        h1.record()
        torch.cuda.synchronize()
        pcie_ms = h0.elapsed_time(h1) / args.steps
        log(f"with host X in / Y out over PCIe: {pcie_ms:.3f} ms/step")
        del Xh, Yh

    # ---- the dominant kernel's own launch duration: HIP events around every
    # phase, on the stream the kernels run on, over as many sweeps as were timed
    gather_ms = None
    if run_kw.get("emit") == gmr.EMIT_GATHER:
        # the same options as the timed steps (time_levels runs the levels variant)
        lv = eng.time_levels(0, X, Y, reps=args.steps,
                             **{k: v for k, v in run_kw.items() if k not in ("variant", "grid_cap", "b_tile")})
        gather_ms = float(lv[info["max_depth"] + 1])
        log(f"phases (ms, mean of {args.steps}): init {lv[0]:.4f}, node updates {float(lv[1:-1].sum()):.4f}, "
            f"k_gather {gather_ms:.4f}")

    if args.profile:
        print(json.dumps({"profile": True, "ms_per_step": ms_step, "launch_ms": kern_ms, "gather_ms": gather_ms,
                          "schedule": autotune.describe(best), "bytes_dag": dag_bytes(info, B)}), flush=True)
        return

    # ---- FP32 error against the FP64 oracle on a sample of columns
    from oracle import oracle as O
    seq, V = m.arrays()
    cols_chk = [0, B // 3, (2 * B) // 3, B - 1]
    Xs = X.reshape(args.cols, -1)[:, cols_chk].contiguous().cpu().numpy()
    Ys, den = O.brute_force_batched(0, args.rows, args.cols, seq, V, Xs)
    Yg = Y.reshape(args.rows, -1)[:, cols_chk].cpu().numpy().astype(np.float64)
    err = rel_err_cols(Yg, Ys, den)
    log(f"fp32 err (4 columns vs FP64 oracle): {err:.3e}")

    # ---- the plain per-level sweep (reference layout, B = 1, spot atomics)
    eng_ref = s_ref.engine()
    x1 = X.reshape(args.cols, -1)[:, 0].contiguous()
    y1 = torch.empty(args.rows, device="cuda", dtype=torch.float32)
    plain = dict(variant=gmr.VARIANT_LEVELS, emit=gmr.EMIT_SPOT, b_tile=0, block=256)
    for _ in range(3):
        eng_ref.run_torch(0, x1, y1, **plain)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(args.steps):
        eng_ref.run_torch(0, x1, y1, **plain)
    b.record()
    torch.cuda.synchronize()
    plain_ms = a.elapsed_time(b) / args.steps
    plain_val = nnz / (plain_ms * 1e-3)
    log(f"plain per-level sweep B=1: {plain_ms:.3f} ms, {plain_val / 1e9:.2f} G nnz/s")
    # best single-vector schedule, for reference
    form1, best1, _ = tune_forms(forms, x1, y1, autotune.candidates(1, gmr.RESIDENT), reps=3, log=log)
    one_ms = autotune.time_candidate(form1["engine"], 0, x1, y1, best1, reps=args.steps)
    # its row gather against HBM peak: algorithmic bytes 4 per top-sequence symbol
    # read + 4 per y row written (+ 4 per value an LDS range fill reads), over the
    # gather launch's own HIP-event time (levels variant, same options)
    one_roof = None
    if best1.get("emit") == gmr.EMIT_GATHER:
        kw1 = {k: v for k, v in autotune.apply(form1["engine"], best1).items() if k not in ("variant", "grid_cap")}
        ph1 = form1["engine"].time_levels(0, x1, y1, reps=args.steps, **kw1)
        fi1 = form1["engine"].info()
        lds1 = bool(best1.get("flags", 0) & gmr.RUN_LDS_RANGES)
        gb1 = (4 * (fi1["emit_total"] + fi1["num_term_emit"]) + 4 * args.rows
               + (4 * (fi1["U"] + fi1["num_layered"]) if lds1 else 0))
        ach1 = gb1 / (float(ph1[-1]) * 1e-3) / 1e9
        key1 = (f"{args.workload}{args.rows}x{args.cols}_B1"
                + (f"_inline{form1['max_emits']}" if form1["max_emits"] else ""))
        one_roof = {"kernel": "k_gather_lds" if lds1 else "k_gather", "bound": "hbm", "achieved": round(ach1, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach1 / HBM_PEAK_GBS, 4),
                    "launch_ms": round(float(ph1[-1]), 4), "bytes_per_launch": int(gb1),
                    "traffic": read_traffic(key1, "k_gather_lds" if lds1 else "k_gather")}
        # the whole one-vector sweep on the amended DAG count of the form that runs,
        # and the traffic rocprofv3 counted for every launch of one such sweep
        sb1 = amended_bytes(form1["schedule"].info(), 1)
        one_roof["sweep"] = {"ms": one_ms, "bytes": int(sb1), "achieved": round(sb1 / (one_ms * 1e-3) / 1e9, 1),
                             "frac": round(sb1 / (one_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                             "traffic": read_traffic(key1, "sweep")}
        log("best single vector", autotune.describe(best1), "gather", one_roof)
    del eng_ref

    # ---- vendor baseline (rocSPARSE CSR product of the uncompressed matrix at
    # the same B, the reference's cuSPARSE comparison) and energy per step
    vendor_line, energy = None, None
    if not args.no_vendor:
        vendor_line, energy = vendor_leg(args, m, X, Y, eng, run_kw, B, nnz, value)

    # ---- roofline of the dominant kernel: k_gather (about 70% of the sweep's
    # device time at B = 64); the whole sweep is reported beside it
    workload = f"{args.workload}{args.rows}x{args.cols}_B{B}"
    sweep_traffic = read_traffic(workload, "sweep")
    amended = amended_bytes(info, B)
    sweep = {"ms": ms_step, "bytes": amended, "achieved": round(amended / (ms_step * 1e-3) / 1e9, 1),
             "frac": round(amended / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
             "bytes_dag_literal": dag_bytes(info, B), "bytes_reference_layered": layered_bytes(info_ref, B),
             "traffic": sweep_traffic,
             "note": "bytes: the DAG count of the grammar form run with the row gather's emission (one 4-byte "
                     "read per emission per vector, y written once), over the whole step's time; "
                     "bytes_dag_literal charges every emission a y read-modify-write that the row gather does not "
                     "do, so it overcounts (it would exceed the chip's bandwidth); the reference-layered count is "
                     "that of the reference's own schedule of the grammar as built"}
    if sweep_traffic:
        sweep["achieved_measured"] = round(sweep_traffic / (kern_ms * 1e-3) / 1e9, 1)
        sweep["frac_measured"] = round(sweep_traffic / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    roofline = None
    if gather_ms:
        gb = gather_bytes(info, B, eng.info().get("n_gseg", 0))
        achieved = gb / (gather_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(workload, "k_gather"),
                    "kernel": "k_gather<PlusTimesOp>", "launch_ms": round(gather_ms, 4), "bytes_per_launch": gb,
                    "sweep": sweep}

    # ---- CPU baseline: the reference's OpenMP FP32 sweep, bounded sample
    cpu = None
    if not args.no_cpu:
        # every core this process may run on, within the OpenMP thread count the
        # machine sets for it (a GPU box gives each GPU a share of its host)
        avail, phys = cpu_cores()
        omp_env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        threads = min(avail, omp_env) if omp_env > 0 else avail
        sa = s_ref.arrays()
        xs = X.reshape(args.cols, -1)[:, 0].cpu().numpy()
        O.set_threads(threads)
        y_cpu = O.cpu_layered_f32(0, sa, args.rows, args.cols, V, xs)  # warm-up; its y is checked below
        times, t0 = [], time.time()
        while not times or time.time() - t0 < args.cpu_seconds:
            t1 = time.time()
            O.cpu_layered_f32(0, sa, args.rows, args.cols, V, xs)
            times.append(time.time() - t1)
        cpu_s = float(np.mean(times))
        O.set_threads(1)
        one = []
        for _ in range(3):
            t1 = time.time()
            O.cpu_layered_f32(0, sa, args.rows, args.cols, V, xs)
            one.append(time.time() - t1)
        # the same run's GPU column 0 against this CPU y, and both against the oracle
        y0 = Yg[:, 0]  # the timed run's column 0, copied before any later leg reused Y
        cpu_vs_gpu = rel_err_cols(y0, np.asarray(y_cpu, np.float64), den[:, 0])
        cpu_err = rel_err_cols(np.asarray(y_cpu, np.float64), Ys[:, 0], den[:, 0])
        log(f"cpu baseline y vs GPU y (column 0): {cpu_vs_gpu:.3e}; vs FP64 oracle {cpu_err:.3e}")
        ref_drv = reference_driver_leg(g0, args.rows, args.cols, xs)
        log("reference driver", ref_drv)
        cpu = {"value": nnz / cpu_s, "unit": "nnz*vectors/s", "cores": threads, "cores_available": avail,
               "cores_physical": phys, "kind": "port", "cpu_vs_gpu_err": cpu_vs_gpu, "cpu_err": cpu_err,
               "reference_driver": ref_drv,
               "sample": f"{len(times)} single-vector sweeps of the same grammar (reference layout, FP32, "
                         f"OpenMP {threads} threads): mean {cpu_s * 1e3:.1f} ms, best {min(times) * 1e3:.1f} ms, "
                         f"median {float(np.median(times)) * 1e3:.1f} ms; schedule build excluded",
               "one_thread": {"value": nnz / min(one), "best_ms": min(one) * 1e3,
                              "median_ms": float(np.median(one)) * 1e3, "sweeps": len(one)}}
        log("cpu baseline", cpu)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "nnz*vectors/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": f"synthetic (seeded {args.workload} mosaic, RePair-compressed by the engine's builder; X ~ N(0,1))",
        "config": {"workload": workload, "rows": args.rows, "cols": args.cols, "nnz": nnz, "B": B,
                   "grammar": {"rules": info["num_raw_nt"], "C": g.info()["clen"], "depth": info["max_depth"],
                               "emissions": info["emit_total"], "term_emissions": info["num_term_emit"],
                               "U": info["U"], "refcompat_passthrough": info_ref["num_passthrough"]},
                   "grammar_form": form["label"],
                   "schedule": autotune.describe(best), "repair_s": round(g0.build_seconds, 1),
                   "schedule_build_s": round(sched_s, 1)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain_sweep": {"value": plain_val, "unit": "nnz/s", "ms": plain_ms,
                        "schedule": "refcompat/levels/spot B=1"},
        "batched_over_plain": value / plain_val,
        "pcie_inclusive": {"ms": pcie_ms, "value": nnz * B / (pcie_ms * 1e-3), "unit": "nnz*vectors/s",
                           "note": "X copied in from and Y copied out to pinned host memory inside every step"},
        "best_single_vector": {"value": nnz / (one_ms * 1e-3), "ms": one_ms, "schedule": autotune.describe(best1),
                               "grammar_form": form1["label"], "roofline": one_roof},
        "fp32_err": err,
        "vendor_csr": vendor_line,
        "energy": energy,
        "tuned": [(lab + " " + autotune.describe(c), round(ms, 4)) for lab, c, ms in tuned],
    }
    print(json.dumps(out), flush=True)


def vendor_leg(args, m, X, Y, eng, run_kw, B, nnz, value):
    """rocSPARSE SpMM (SpMV at B = 1) of the uncompressed matrix in CSR form on
    the same X, every CSR algorithm timed with HIP events over the same number
    of steps (the fastest reported), its agreement with the engine's Y, and the
    energy per step of both from the amdsmi counter over >= 1.5 s loops."""
    import torch

    import gmr
    from gmr import vendor
    t = time.time()
    A = vendor.Csr.from_matrix(m)
    build_s = time.time() - t
    Yv = torch.empty_like(Y)
    per_alg = {}
    for alg in vendor.ALGS:
        try:
            A.run_torch(X, Yv, alg=alg)  # preprocess + warm-up
            torch.cuda.synchronize()
        except gmr.GmrError as e:
            log("rocSPARSE", alg, "unsupported:", e)
            per_alg[alg] = None
            continue
        Yv.zero_()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            A.run_torch(X, Yv, alg=alg)
        e1.record()
        torch.cuda.synchronize()
        # a timing only counts if the timed calls produced the product
        if float((Yv - Y).abs().max() / Y.abs().max().clamp_min(1e-30)) > 1e-4:
            log("rocSPARSE", alg, "timed calls did not produce the product; dropped")
            per_alg[alg] = None
            continue
        per_alg[alg] = e0.elapsed_time(e1) / args.steps
        log(f"rocSPARSE {'SpMM' if B > 1 else 'SpMV'} ({alg}): {per_alg[alg]:.3f} ms")
    ok = [a for a in per_alg if per_alg[a] is not None]
    if not ok:
        return None, None
    best = min(ok, key=per_alg.get)
    Yv.zero_()
    A.run_torch(X, Yv, alg=best)
    torch.cuda.synchronize()
    diff = float((Yv - Y).abs().max() / Y.abs().max().clamp_min(1e-30))
    ms = per_alg[best]
    line = {"library": "rocSPARSE", "op": "SpMM" if B > 1 else "SpMV", "alg": best, "ms": ms,
            "value": nnz * B / (ms * 1e-3), "unit": "nnz*vectors/s",
            "per_alg_ms": {a: (round(v, 4) if v is not None else None) for a, v in per_alg.items()},
            "engine_over_vendor": value / (nnz * B / (ms * 1e-3)),
            "max_abs_diff_vs_engine_over_max_abs_y": diff, "csr_device_bytes": A.device_bytes(),
            "csr_build_s": round(build_s, 1)}
    log("vendor", line)
    energy = None
    if vendor.energy_uj(0) is not None:
        try:
            mj_v, w_v = A.energy(B=B, alg=best, seconds=2.0)
        except gmr.GmrError as e:
            log("vendor energy:", e)
            mj_v = w_v = None

        def batch():
            for _ in range(10):
                eng.run_torch(0, X, Y, **run_kw)
            torch.cuda.synchronize()
            return 10
        r = vendor.energy_between_updates(batch, seconds=2.0)
        if r is not None:
            energy = {"engine_mJ_per_step": r[0], "engine_avg_W": r[1], "vendor_mJ_per_step": mj_v,
                      "vendor_avg_W": w_v,
                      "note": "amdsmi energy counter, between two of its updates >= 2 s apart, steps back to back"}
            log("energy", energy)
    del A
    return line, energy


def grammar_forms(args, g, s_res, eng, log):
    """The grammar forms to run: as built and/or with its top rules written
    out in C (Grammar.inline_top), each with its Resident schedule and engine.
    'auto' tries thresholds 2 (best for many vectors) and 8 (fewer values for
    the one-vector LDS-range gather: fewer ranges, fewer row pieces)."""
    forms = []
    it = str(args.inline_top)
    if it in ("auto", "0"):
        forms.append({"label": "as built", "max_emits": 0, "grammar": g, "schedule": s_res, "engine": eng})
    for n in ((2, 8) if it == "auto" else (() if it == "0" else (int(it),))):
        t = time.time()
        g2 = g.inline_top(n)
        s2 = g2.schedule(args.rows, args.cols, args.layout)
        e2 = s2.engine()
        lab = f"top rules used <= {n} times in C inlined ({g2.n_inlined} of {g.info()['nrules']} rules)"
        log(f"grammar form: {lab}, {time.time() - t:.1f}s", s2.info())
        forms.append({"label": lab, "max_emits": n, "grammar": g2, "schedule": s2, "engine": e2})
    return forms


def untuned_form(args, forms, B):
    """Without tuning, 'auto' takes the inlined form from B = 32 on (where the
    node updates it saves outweigh the longer row gather) and the grammar as
    built below."""
    if len(forms) == 1:
        return forms[0]
    return forms[1] if B >= 32 else forms[0]


def tune_forms(forms, X, Y, cands, reps=2, log=None, semiring=0):
    """autotune.tune_forms with the form labels shortened for the JSON line."""
    from gmr import autotune
    form, best, tuned = autotune.tune_forms(forms, semiring, X, Y, cands, reps=reps, log=log)
    return form, best, [(lab.split(" (")[0], c, ms) for lab, c, ms in tuned]


def _time_steps(eng, semiring, X, Y, cand, steps, warmup):
    import torch

    from gmr import autotune
    cand = autotune.apply(eng, cand)
    for _ in range(warmup):
        eng.run_torch(semiring, X, Y, **cand)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        eng.run_torch(semiring, X, Y, **cand)
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def run_sweep_B(args, m, forms):
    """Throughput per vector count B on one matrix (each B autotuned over the
    grammar forms and schedules), with the FP32 error of two columns against
    the FP64 oracle."""
    import numpy as np
    import torch

    import gmr
    from gmr import autotune
    from oracle import oracle as O
    seq, V = m.arrays()
    nnz = m.nnz()
    info = forms[0]["schedule"].info()
    out = []
    for B in [int(b) for b in args.sweep_B.split(",")]:
        gen = torch.Generator(device="cuda").manual_seed(args.seed + B)
        X = torch.randn(args.cols, B, device="cuda", generator=gen)
        Y = torch.empty(args.rows, B, device="cuda")
        if B == 1:
            X, Y = X[:, 0].contiguous(), Y[:, 0].contiguous()
        form, best, tuned = tune_forms(forms, X, Y, autotune.candidates(B, gmr.RESIDENT), reps=2, log=log)
        eng = form["engine"]
        ms = _time_steps(eng, 0, X, Y, best, args.steps, args.warmup)
        if best["variant"] == gmr.VARIANT_PERSISTENT:
            eng.check()
        cols = [0] if B == 1 else [0, B - 1]
        Xs = X.reshape(args.cols, -1)[:, cols].contiguous().cpu().numpy()
        Ys, den = O.brute_force_batched(0, args.rows, args.cols, seq, V, Xs)
        Yg = Y.reshape(args.rows, -1)[:, cols].cpu().numpy().astype(np.float64)
        d = np.abs(Yg - Ys)
        err = float(np.max(np.where(den > 0, d / np.where(den > 0, den, 1), np.where(d == 0, 0, np.inf))))
        r = {"B": B, "ms": ms, "nnz_vec_per_s": nnz * B / (ms * 1e-3),
             "gbs_dag": dag_bytes(form["schedule"].info(), B) / ms / 1e6,
             "grammar_form": form["label"], "schedule": autotune.describe(best), "fp32_err": err,
             "tuned": [(lab + " " + autotune.describe(c), round(t, 4)) for lab, c, t in tuned]}
        log("B", B, r)
        out.append(r)
        del X, Y
    print(json.dumps({"metric": METRIC, "workload": f"{args.workload}{args.rows}x{args.cols}", "nnz": nnz,
                      "grammar": {"rules": info["num_raw_nt"], "depth": info["max_depth"],
                                  "emissions": info["emit_total"], "term_emissions": info["num_term_emit"]},
                      "per_B": out}), flush=True)


def run_blocks(args, m):
    """The matrix as args.blocks row-block grammars (mm-repair's -b blocks),
    all swept in one batch launch per step (each variant timed, the fastest
    reported), with the FP32 error of four columns against the FP64 oracle."""
    import numpy as np
    import torch

    import gmr
    from gmr.blocks import BlockGrammar
    from oracle import oracle as O
    B, nnz = args.B, m.nnz()
    it = str(args.inline_top)
    n_inline = (2 if B >= 32 else 0) if it == "auto" else int(it)
    bg = BlockGrammar(m, args.blocks, inline_top=n_inline)
    log(f"{args.blocks} row-block grammars built in {bg.build_seconds:.1f}s", bg.info())
    gen = torch.Generator(device="cuda").manual_seed(args.seed)
    X = torch.randn(args.cols, B, device="cuda", generator=gen)
    if B == 1:
        X = X[:, 0].contiguous()
    Y = bg.run_torch(0, X)
    per = {}
    for vname, v in gmr.VARIANTS.items():
        for _ in range(args.warmup):
            bg.run_torch(0, X, Y, variant=v)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            bg.run_torch(0, X, Y, variant=v)
        e1.record()
        torch.cuda.synchronize()
        if v == gmr.VARIANT_PERSISTENT:
            bg.batch.check()
        per[vname] = e0.elapsed_time(e1) / args.steps
        log(f"blocks {vname}: {per[vname]:.3f} ms")
    best = min(per, key=per.get)
    bg.run_torch(0, X, Y, variant=gmr.VARIANTS[best])
    seq, V = m.arrays()
    cols_chk = [0, B // 3, (2 * B) // 3, B - 1]
    Xs = X.reshape(args.cols, -1)[:, cols_chk].contiguous().cpu().numpy()
    Ys, den = O.brute_force_batched(0, args.rows, args.cols, seq, V, Xs)
    d = np.abs(Y.reshape(args.rows, -1)[:, cols_chk].cpu().numpy().astype(np.float64) - Ys)
    err = float(np.max(np.where(den > 0, d / np.where(den > 0, den, 1), np.where(d == 0, 0, np.inf))))
    print(json.dumps({"metric": METRIC, "workload": f"{args.workload}{args.rows}x{args.cols}_B{B}", "nnz": nnz,
                      "blocks": args.blocks, "inline_top": n_inline, "build_s": round(bg.build_seconds, 1),
                      "grammar": bg.info(), "per_variant_ms": {k: round(v, 4) for k, v in per.items()},
                      "variant": best, "ms": per[best], "value": nnz * B / (per[best] * 1e-3),
                      "unit": "nnz*vectors/s", "fp32_err": err}), flush=True)


def run_left_B(args, m, forms):
    """The left product Y^T = X^T M per vector count B (every grammar form in
    the levels and graph variants, the fastest kept), with the FP32 error of
    two columns against the FP64 oracle's left brute force."""
    import numpy as np
    import torch

    import gmr
    from oracle import oracle as O
    seq, V = m.arrays()
    nnz = m.nnz()
    out = []
    for B in [int(b) for b in args.left_B.split(",")]:
        gen = torch.Generator(device="cuda").manual_seed(args.seed + 7 * B)
        X = torch.randn(args.rows, B, device="cuda", generator=gen)
        if B == 1:
            X = X[:, 0].contiguous()
        best = None
        for f in forms:
            eng = f["engine"]
            for vname in ("levels", "graph"):
                kw = dict(variant=gmr.VARIANTS[vname])
                Y = eng.run_left_torch(0, X, **kw)
                for _ in range(args.warmup):
                    eng.run_left_torch(0, X, Y, **kw)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(args.steps):
                    eng.run_left_torch(0, X, Y, **kw)
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / args.steps
                log(f"left B={B} {f['label']} {vname}: {ms:.3f} ms")
                if best is None or ms < best[0]:
                    best = (ms, vname, Y, f["label"])
        ms, vname, Y, label = best
        cols = [0] if B == 1 else [0, B - 1]
        Xs = X.reshape(args.rows, -1)[:, cols].contiguous().cpu().numpy()
        Ys, den = O.brute_force_left(0, args.rows, args.cols, seq, V, Xs)
        Yg = Y.reshape(args.cols, -1)[:, cols].cpu().numpy().astype(np.float64)
        d = np.abs(Yg - Ys)
        err = float(np.max(np.where(den > 0, d / np.where(den > 0, den, 1), np.where(d == 0, 0, np.inf))))
        r = {"B": B, "ms": ms, "nnz_vec_per_s": nnz * B / (ms * 1e-3), "variant": vname, "grammar_form": label,
             "fp32_err": err}
        log("left", r)
        out.append(r)
    print(json.dumps({"metric": "left product Y^T = X^T M, nnz*vectors/s", "workload": f"{args.workload}"
                      f"{args.rows}x{args.cols}", "nnz": nnz, "per_B": out}), flush=True)


def run_rmat(args):
    """R-MAT graph on 2^23 vertices from 2^27 edge draws (a, b, c, d) = (.57, .19,
    .19, .05), duplicates removed: (a) Boolean one-step reachability from a 1%
    frontier; (b) min/plus relaxation with weights U{1..100}, distances 0 on
    1,024 sources and INT32_MAX/2 elsewhere. Both checked bit-exact against the
    oracle; throughput of the tuned single-vector sweep and of the plain sweep."""
    import numpy as np
    import torch

    import gmr
    from gmr import autotune
    from oracle import oracle as O
    scale, edges = 23, 1 << 27
    res = {}
    for tag, nw, sr in (("boolean", 1, gmr.BOOLEAN), ("minplus", 100, gmr.TROPICAL)):
        if args.semiring not in ("all", tag):
            continue
        m = gmr.Matrix.rmat(20261206 + nw, scale, edges, nweights=nw)
        nnz, n = m.nnz(), m.rows
        g = m.repair()
        log(tag, "nnz", nnz, f"RePair {g.build_seconds:.1f}s", g.info())
        s_res = g.schedule(n, n, "resident")
        s_ref = g.schedule(n, n, "refcompat")
        eng = s_res.engine()
        args.rows = args.cols = n
        forms = grammar_forms(args, g, s_res, eng, log)
        rng = np.random.default_rng(2026 + nw)
        if sr == gmr.BOOLEAN:
            x = np.zeros(n, np.float32)
            x[rng.choice(n, n // 100, replace=False)] = 1.0
        else:
            x = np.full(n, np.float32(2 ** 31 - 1) // 2, np.float32)
            x[rng.choice(n, 1024, replace=False)] = 0.0
        xt = torch.from_numpy(x).cuda()
        y = torch.empty(n, dtype=torch.int32, device="cuda")
        if args.no_tune or args.schedule:
            # Boolean: values as bits (the LDS-range gather); min/plus: the XCD-partitioned gather
            # (or the schedule given)
            form, tuned = untuned_form(args, forms, 1), []
            best = dict(variant=gmr.VARIANT_GRAPH, emit=gmr.EMIT_GATHER, b_tile=0, block=256, grid_cap=0,
                        gather_order=0)
            if args.schedule:
                best.update(json.loads(args.schedule))
            elif sr == gmr.BOOLEAN:
                best["flags"] = gmr.RUN_BITS
            else:
                best["xcd_passes"] = 0
        else:
            form, best, tuned = tune_forms(forms, xt, y, autotune.candidates(1, gmr.RESIDENT, semiring=sr), reps=2,
                                           semiring=sr, log=log)
        ms = _time_steps(form["engine"], sr, xt, y, best, args.steps, args.warmup)
        # phases of the chosen schedule (levels variant, HIP events around every launch)
        # and the row gather's rate on its algorithmic bytes: 4 per top-sequence symbol
        # read, 4 per y row written
        kw = autotune.apply(form["engine"], best)
        kw.pop("variant", None)
        ph = form["engine"].time_levels(sr, xt, y, reps=5, **kw)
        finfo = form["engine"].info()
        gbytes = 4 * (finfo["emit_total"] + finfo["num_term_emit"]) + 4 * n
        phases = {"init_ms": round(float(ph[0]), 4), "levels_ms": round(float(ph[1:-1].sum()), 4),
                  "gather_ms": round(float(ph[-1]), 4), "gather_bytes": int(gbytes),
                  "gather_GBps": round(gbytes / (float(ph[-1]) * 1e-3) / 1e9, 1)}
        log(tag, "phases", phases)
        # the gather kernel against HBM peak; traffic as counted by rocprofv3 for the
        # same gather kernel and grammar form (profiles/traffic.txt)
        bits = bool(best.get("flags", 0) & gmr.RUN_BITS) and not (best.get("flags", 0) & gmr.RUN_SINGLE_LIST)
        kernel = ("k_bits_gather_lds" if bits else
                  "k_gather_lds" if best.get("flags", 0) & gmr.RUN_LDS_RANGES else "k_gather")
        roof = {"kernel": kernel, "bound": "hbm", "achieved": phases["gather_GBps"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(phases["gather_GBps"] / HBM_PEAK_GBS, 4),
                "traffic": read_traffic(f"rmat{scale}_{tag}" + (f"_inline{form['max_emits']}" if form["max_emits"] else ""),
                                        kernel)}
        seq, V = m.arrays()
        y_star, _ = O.brute_force(sr, n, n, seq, V, x)
        exact = bool(np.array_equal(y.cpu().numpy().astype(np.int64), y_star.astype(np.int64)))
        eng_ref = s_ref.engine()
        plain_ms = _time_steps(eng_ref, sr, xt, y, dict(variant=gmr.VARIANT_LEVELS, emit=gmr.EMIT_SPOT),
                               args.steps, args.warmup)
        exact_plain = bool(np.array_equal(y.cpu().numpy().astype(np.int64), y_star.astype(np.int64)))
        # fixpoint traversal from the same start (Boolean: the reachable set;
        # min/plus: shortest distances), checked against the oracle's left
        # product iterated on the CPU
        trav = None
        if args.traverse:
            runs = []
            for f in forms:
                xs = torch.from_numpy(x).cuda()
                f["engine"].traverse_torch(sr, xs.clone(), max_iters=2)  # builds the incoming lists
                torch.cuda.synchronize()
                t0 = time.time()
                iters = f["engine"].traverse_torch(sr, xs)
                torch.cuda.synchronize()
                runs.append(((time.time() - t0) * 1e3, iters, xs.cpu().numpy(), f["label"]))
                log(tag, "traversal", f["label"], f"{runs[-1][0]:.2f} ms, {iters} products")
            t_ms, iters, got, t_form = min(runs, key=lambda r: r[0])
            if sr == gmr.TROPICAL:
                got = np.minimum(got, np.float32(1e9))  # distances >= 1e9 mean unreached
            xc = x.copy()
            for it in range(1, 10000):
                yc, _ = O.brute_force_left(sr, n, n, seq, V, xc)
                xn = (np.maximum(xc, yc) if sr == gmr.BOOLEAN else np.minimum(np.minimum(xc, 1e9), yc)).astype(np.float32)
                if np.array_equal(xn, np.minimum(xc, 1e9) if sr == gmr.TROPICAL else xc):
                    break
                xc = xn
            want = np.minimum(xc, 1e9) if sr == gmr.TROPICAL else xc
            clip = (lambda v: np.minimum(v, np.float32(1e9))) if sr == gmr.TROPICAL else (lambda v: v)
            trav = {"products": iters, "ms": t_ms, "ms_per_product": t_ms / iters, "cpu_products": it,
                    "grammar_form": t_form, "bit_exact": all(np.array_equal(clip(r[2]), want) for r in runs),
                    "per_form_ms": {r[3]: round(r[0], 3) for r in runs},
                    "moved": int((got != np.minimum(x, 1e9)).sum())}
            log(tag, "traversal", trav)
            # the same traversal over an engine of M^T: every step a right product
            # (Boolean: bit mode; min/plus: LDS ranges with the fused tail)
            t0 = time.time()
            gt = m.transpose().repair()
            # one-vector right-product forms: Boolean as bits; min/plus, whose x turns
            # finite as the traversal goes (then every row piece emits), also the row
            # segments (each row reduced in registers, stored once)
            T_ = gmr.RUN_TRANSPOSED
            tkws = ([dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_BITS | T_, lds_values=262144, chunks=8)]
                    if sr == gmr.BOOLEAN else
                    [dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_ROW_SEGMENTS | gmr.RUN_FUSED_TAIL | T_),
                     dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_FUSED_TAIL | T_),
                     dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES | gmr.RUN_FUSED_TAIL | T_,
                          lds_values=16384, chunks=2)])
            truns = []
            for lab, gg in (("as built", gt), ("top rules used <= 8 times in C inlined", gt.inline_top(8))):
                et = gg.schedule(n, n, "resident").engine()
                for tkw in tkws:
                    xs = torch.from_numpy(x).cuda()
                    et.traverse_torch(sr, xs.clone(), max_iters=2, **tkw)  # builds the lists
                    torch.cuda.synchronize()
                    t1 = time.time()
                    ti = et.traverse_torch(sr, xs, **tkw)
                    torch.cuda.synchronize()
                    truns.append(((time.time() - t1) * 1e3, ti, xs.cpu().numpy(), lab, autotune.describe(
                        dict(tkw, variant=gmr.VARIANT_LEVELS, b_tile=0, block=256))))
                    log(tag, "transposed traversal", lab, truns[-1][4], f"{truns[-1][0]:.2f} ms, {ti} products")
                del et
            tt_ms, t_it, _, t_lab, t_sched = min(truns, key=lambda r: r[0])
            trav["transposed"] = {"ms": tt_ms, "products": t_it, "ms_per_product": tt_ms / t_it,
                                  "grammar_form": t_lab, "schedule": t_sched,
                                  "bit_exact": all(np.array_equal(clip(r[2]), want) for r in truns),
                                  "per_form_ms": {r[3] + " " + r[4]: round(r[0], 3) for r in truns},
                                  "transpose_and_repair_s": round(time.time() - t0, 1)}
            log(tag, "transposed traversal", trav["transposed"])
            del gt
        info = s_res.info()
        res[tag] = {"nnz": nnz, "ms": ms, "nnz_per_s": nnz / (ms * 1e-3), "schedule": autotune.describe(best),
                    "grammar_form": form["label"], "phases": phases, "roofline": roof,
                    "traversal": trav,
                    "bit_exact": exact, "plain_ms": plain_ms, "plain_nnz_per_s": nnz / (plain_ms * 1e-3),
                    "plain_bit_exact": exact_plain,
                    "grammar": {"rules": info["num_raw_nt"], "depth": info["max_depth"], "C": g.info()["clen"],
                                "U": info["U"], "refcompat_passthrough": s_ref.info()["num_passthrough"]},
                    "repair_s": g.build_seconds,
                    "tuned": [(lab + " " + autotune.describe(c), round(t, 4)) for lab, c, t in tuned]}
        log(tag, res[tag])
        del eng, eng_ref, m, g, s_res, s_ref, seq, forms, form
        torch.cuda.empty_cache()
    print(json.dumps({"metric": METRIC, "workload": f"rmat{scale}_{edges}", "results": res}), flush=True)


if __name__ == "__main__":
    main()
