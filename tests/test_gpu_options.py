"""GPU tests of the run options that change how, not what, a sweep computes:
the level-chunk sizes (LDS window of the EV_LDS node update, LDS range and
chunks per wave of the single-vector gathers), the LDS-range gather of 4-byte
values, the Tropical cap of whole-row stores, and the reporting of a
persistent launch whose grid barrier timed out. Every result is checked
against the oracle (FP64 brute force from the uncompressed matrix): Boolean
and Tropical bit-exact, (+,x) max_i |y_i - y*_i| / sum_j |M_ij x_j| <= 1e-5."""
import sys

import numpy as np
import pytest

import gmr
from gmr import _lib
from helpers import rel_err, x_for
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5

torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.dev()  # the HIP library must load; no fallback exists


def run(eng, sr, X, **kw):
    Xt = torch.from_numpy(np.ascontiguousarray(X, np.float32)).cuda()
    Y = eng.run_torch(sr, Xt, **kw)
    torch.cuda.synchronize()
    if kw.get("variant") == gmr.VARIANT_PERSISTENT:
        eng.check()
    return Y.cpu().numpy()


def check(sr, y, y_star, den, what):
    assert np.shape(y) == np.shape(y_star), what
    if sr == 0:
        e = rel_err(y, y_star, den)
        assert e <= TOL, f"{what}: FP32 error {e} > {TOL}"
    else:
        diff = int((np.asarray(y, np.int64) != np.asarray(y_star, np.int64)).sum())
        assert diff == 0, f"{what}: {diff} rows differ"


@pytest.fixture(scope="module")
def hap():
    m = gmr.Matrix.haplotype(20261204, 4096, 2048)
    g = m.repair()
    seq, V = m.arrays()
    return g.schedule(4096, 2048, "resident").engine(), seq, V


@pytest.fixture(scope="module")
def geno():
    # enough values (rules + terminals) for dozens of LDS ranges
    m = gmr.Matrix.genotype(20261206, 6000, 10000)
    g = m.repair()
    seq, V = m.arrays()
    return g.schedule(6000, 10000, "resident").engine(), seq, V


@pytest.mark.parametrize("sr", [0, 1, 2])
def test_lds_range_gather(hap, geno, sr):
    """GMR_RUN_LDS_RANGES at every range size and chunk count, levels and graph."""
    for name, (eng, seq, V), rows, cols in (("hap", hap, 4096, 2048), ("geno", geno, 6000, 10000)):
        x = x_for(sr, cols, 300 + sr)
        y_star, den = O.brute_force(sr, rows, cols, seq, V, x)
        for span in (4096, 8192, 16384, 32768):
            for ch in (2, 4, 8):
                for var in (gmr.VARIANT_LEVELS, gmr.VARIANT_GRAPH):
                    _say(name, "lds ranges", sr, span, ch, var)
                    y = run(eng, sr, x, variant=var, emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES,
                            lds_values=span, chunks=ch)
                    check(sr, y, y_star, den, f"{name} lds span={span} chunks={ch} var={var}")
        # default range and chunks
        y = run(eng, sr, x, emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES)
        check(sr, y, y_star, den, f"{name} lds default")
        # 160 KiB ranges: only where one workgroup may hold that much LDS
        try:
            y = run(eng, sr, x, emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES, lds_values=40960, chunks=2)
        except gmr.GmrError as e:
            assert e.code == gmr.ERR_INVALID and "fit the LDS" in str(e), str(e)
        else:
            check(sr, y, y_star, den, f"{name} lds span=40960")


@pytest.mark.parametrize("sr", [0, 1, 2])
def test_lds_range_gather_row_order(hap, geno, sr):
    """GMR_RUN_ROW_ORDER: the LDS-range lists with their rows renumbered (rows
    of equal length and similar symbols together), the pieces combined in the
    engine's row-ordered copy of y and written back in matrix row order. The
    ordered and the matrix-order lists live side by side on one engine."""
    for name, (eng, seq, V), rows, cols in (("hap", hap, 4096, 2048), ("geno", geno, 6000, 10000)):
        x = x_for(sr, cols, 310 + sr)
        y_star, den = O.brute_force(sr, rows, cols, seq, V, x)
        for span in (4096, 16384, 32768):
            for ch in (2, 4):
                for var in (gmr.VARIANT_LEVELS, gmr.VARIANT_GRAPH):
                    for fl in (gmr.RUN_LDS_RANGES | gmr.RUN_ROW_ORDER, gmr.RUN_LDS_RANGES):
                        _say(name, "lds ranges", sr, span, ch, var, fl)
                        y = run(eng, sr, x, variant=var, emit=gmr.EMIT_GATHER, flags=fl, lds_values=span, chunks=ch)
                        check(sr, y, y_star, den, f"{name} lds span={span} chunks={ch} var={var} flags={fl}")


def test_lds_range_gather_row_order_empty_rows():
    """Rows with no non-zero get the semiring zero through the row-ordered copy
    too (every third row of a dense 0/1 matrix emptied)."""
    rng = np.random.default_rng(91)
    M = (rng.random((3000, 700)) < 0.05).astype(np.float32)
    M[::3] = 0.0
    m = gmr.Matrix.from_dense(M)
    seq, V = m.arrays()
    eng = m.repair().schedule(3000, 700, "resident").engine()
    for sr in (0, 1, 2):
        x = x_for(sr, 700, 320 + sr)
        y_star, den = O.brute_force(sr, 3000, 700, seq, V, x)
        y = run(eng, sr, x, emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES | gmr.RUN_ROW_ORDER, lds_values=4096)
        check(sr, y, y_star, den, f"empty rows sr={sr}")


@pytest.mark.parametrize("sr", [0, 1, 2])
def test_fused_tail(hap, geno, sr):
    """GMR_RUN_FUSED_TAIL (the top levels with at most 8,192 rules together in
    one workgroup) and GMR_RUN_PAIRED_LEVELS (levels two to a launch), alone
    and together, with spot and gather emission, the levels and graph variants
    and the one-vector gathers; on the smallest grammar every level is in the
    tail."""
    small = gmr.Matrix.haplotype(20261204, 512, 256, founders=16, seg_mean=32.0)
    sseq, sV = small.arrays()
    seng = small.repair().schedule(512, 256, "resident").engine()
    cases = (("small", (seng, sseq, sV), 512, 256), ("hap", hap, 4096, 2048), ("geno", geno, 6000, 10000))
    for name, (eng, seq, V), rows, cols in cases:
        x = x_for(sr, cols, 330 + sr)
        y_star, den = O.brute_force(sr, rows, cols, seq, V, x)
        for var in (gmr.VARIANT_LEVELS, gmr.VARIANT_GRAPH):
            for kw in (dict(emit=gmr.EMIT_SPOT), dict(emit=gmr.EMIT_GATHER),
                       dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_SINGLE_LIST),
                       dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES, lds_values=8192),
                       dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES | gmr.RUN_ROW_ORDER, lds_values=8192),
                       dict(emit=gmr.EMIT_SPOT, eval_kind=gmr.EVAL_LDS)):
                for extra in (gmr.RUN_FUSED_TAIL, gmr.RUN_PAIRED_LEVELS, gmr.RUN_FUSED_TAIL | gmr.RUN_PAIRED_LEVELS):
                    kw2 = dict(kw, flags=kw.get("flags", 0) | extra)
                    _say(name, "fused tail / paired levels", sr, var, kw2)
                    y = run(eng, sr, x, variant=var, **kw2)
                    check(sr, y, y_star, den, f"{name} var={var} {kw2}")


def test_fused_tail_refusals(hap):
    eng, seq, V = hap
    x = torch.zeros(2048, device="cuda")
    for kw in (dict(variant=gmr.VARIANT_PERSISTENT, flags=gmr.RUN_FUSED_TAIL),
               dict(variant=gmr.VARIANT_PERSISTENT, flags=gmr.RUN_PAIRED_LEVELS),
               dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_FUSED_TAIL | gmr.RUN_BITS),
               dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_FUSED_TAIL | gmr.RUN_PAIRED_LEVELS | gmr.RUN_BITS),
               dict(variant=gmr.VARIANT_PERSISTENT, emit=gmr.EMIT_GATHER, flags=gmr.RUN_PAIRED_LEVELS | gmr.RUN_BITS)):
        with pytest.raises(gmr.GmrError) as e:
            eng.run_torch(1, x, **kw)
        assert e.value.code == gmr.ERR_INVALID, kw
    for fl in (gmr.RUN_FUSED_TAIL, gmr.RUN_PAIRED_LEVELS):
        with pytest.raises(gmr.GmrError) as e:
            eng.run_torch(0, torch.zeros(2048, 4, device="cuda"), flags=fl)  # one vector only
        assert e.value.code == gmr.ERR_INVALID
    m = gmr.Matrix.haplotype(20261204, 512, 256, founders=16, seg_mean=32.0)
    weng = m.repair().schedule(512, 256, "wave64").engine()
    with pytest.raises(gmr.GmrError) as e:
        weng.run_torch(0, torch.zeros(256, device="cuda"), flags=gmr.RUN_FUSED_TAIL)  # Resident layout only
    assert e.value.code == gmr.ERR_INVALID
    torch.cuda.synchronize()


def test_fused_tail_in_a_batch(hap, geno):
    """One-vector grammars with GMR_RUN_FUSED_TAIL in a batch, levels variant
    and graph variant (a branch per grammar), against the oracle; the
    persistent batch, whose tiles run every level themselves, refuses it."""
    cases = ((hap, 4096, 2048), (geno, 6000, 10000))
    Xs, stars = [], []
    for i, ((eng, seq, V), rows, cols) in enumerate(cases):
        x = x_for(0, cols, 340 + i)
        stars.append(O.brute_force(0, rows, cols, seq, V, x))
        Xs.append(torch.from_numpy(x).reshape(cols, 1).cuda())
    batch = gmr.Batch([c[0][0] for c in cases])
    opts = [{"emit": gmr.EMIT_GATHER, "flags": gmr.RUN_FUSED_TAIL}, {"emit": gmr.EMIT_SPOT, "flags": gmr.RUN_FUSED_TAIL}]
    for var in ("levels", "graph"):
        Ys = batch.run_torch(0, Xs, variant=gmr.VARIANTS[var], opts=opts)
        torch.cuda.synchronize()
        for i, (Y, (ys, den)) in enumerate(zip(Ys, stars)):
            check(0, Y.cpu().numpy()[:, 0], ys, den, f"batch {var} grammar {i} fused tail")
    with pytest.raises(gmr.GmrError) as e:
        batch.run_torch(0, Xs, variant=gmr.VARIANT_PERSISTENT, opts=opts)
    assert e.value.code == gmr.ERR_INVALID
    torch.cuda.synchronize()


def test_lds_range_gather_sparse_rows():
    """Entries carry their row as a 16-bit offset from their chunk's first row,
    so a row piece that starts 65,536 rows or more after its chunk's first row
    starts a new chunk (here: three non-zeros in 150,000 rows, with and without
    the row order), against the oracle; and the bit-mode gather over the same
    rows (one chunk spanning every row)."""
    rows, cols = 150000, 8
    M = np.zeros((rows, cols))
    M[0, 0], M[100000, 1], M[149999, 2], M[149999, 0] = 2.0, 3.0, 5.0, 7.0
    m = gmr.Matrix.from_dense(M)
    g = m.repair()
    seq, V = m.arrays()
    eng = g.schedule(rows, cols, "resident").engine()
    for sr in (0, 1, 2):
        x = x_for(sr, cols, 990 + sr)
        y_star, den = O.brute_force(sr, rows, cols, seq, V, x)
        for ch in (2, 4):
            for fl in (gmr.RUN_LDS_RANGES, gmr.RUN_LDS_RANGES | gmr.RUN_ROW_ORDER):
                y = run(eng, sr, x, emit=gmr.EMIT_GATHER, flags=fl, lds_values=4096, chunks=ch)
                check(sr, y, y_star, den, f"sparse rows lds chunks={ch} flags={fl}")
    xb = x_for(1, cols, 993)
    yb_star, _ = O.brute_force(1, rows, cols, seq, V, xb)
    for rng in (131072, 1048576):
        y = run(eng, 1, xb, emit=gmr.EMIT_GATHER, flags=gmr.RUN_BITS, lds_values=rng, chunks=2)
        check(1, y, yb_star, None, f"sparse rows bits range={rng}")


def test_lds_range_fills_over_renumbered_terminals():
    """The engine numbers terminals in column order and the range fills compute
    their values from that numbering: terminals spread thinly over many columns
    (one non-zero every 150 of 300,000 columns), more than 65,536 distinct
    values, 60,000 distinct values over dense columns and a small dense matrix,
    at range sizes of two workgroups and of one workgroup per CU, against the
    oracle."""
    rng = np.random.default_rng(77)
    cases = []
    M = np.zeros((6, 300000))
    M[rng.integers(0, 6, 2000), np.arange(2000) * 150] = rng.integers(1, 4, 2000)
    cases.append(("sparse columns", M))
    cases.append(("many values", rng.uniform(1.0, 2.0, (8, 9000)).round(7)))
    cases.append(("60,000 values", rng.integers(1, 60001, (4, 30000)).astype(np.float64)))
    cases.append(("dense", rng.integers(0, 3, (500, 3000)).astype(np.float64)))
    for name, M in cases:
        m = gmr.Matrix.from_dense(M)
        seq, V = m.arrays()
        rows, cols = M.shape
        eng = m.repair().schedule(rows, cols, "resident").engine()
        for sr in (0, 1, 2):
            x = x_for(sr, cols, 1200 + sr)
            y_star, den = O.brute_force(sr, rows, cols, seq, V, x)
            for span in (4096, 40960):
                _say(name, sr, len(V), "values", span)
                y = run(eng, sr, x, emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES, lds_values=span, chunks=4)
                check(sr, y, y_star, den, f"{name} lds ranges span={span}")


def test_graph_cache_keys_on_every_level_chunk_option():
    """A captured graph bakes in the range size, the chunks per wave and the
    EV_LDS window: runs on the same buffers that differ only there (8,192 and
    40,960 values share their lowest set bit) each capture a graph of their
    own and replay it."""
    m = gmr.Matrix.haplotype(20261204, 4096, 2048)
    seq, V = m.arrays()
    eng = m.repair().schedule(4096, 2048, "resident").engine()  # no graphs captured yet
    x = x_for(0, 2048, 77)
    y_star, den = O.brute_force(0, 4096, 2048, seq, V, x)
    xt = torch.from_numpy(x).cuda()
    Y = torch.empty(4096, device="cuda")
    opts = [dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES, lds_values=8192, chunks=2),
            dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES, lds_values=40960, chunks=2),
            dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES, lds_values=8192, chunks=4),
            dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES | gmr.RUN_ROW_ORDER, lds_values=8192, chunks=4),
            dict(eval_kind=gmr.EVAL_LDS, lds_values=2048),
            dict(eval_kind=gmr.EVAL_LDS, lds_values=4096)]
    assert eng.debug()["graphs"] == 0
    ran = 0
    for rep in range(2):
        for kw in opts:
            Y.fill_(12345.0)
            try:
                eng.run_torch(0, xt, Y, variant=gmr.VARIANT_GRAPH, **kw)
            except gmr.GmrError as e:  # 160 KiB ranges where a workgroup may not hold that much LDS
                assert kw.get("lds_values") == 40960 and "fit the LDS" in str(e), str(e)
                continue
            torch.cuda.synchronize()
            ran += rep == 0
            check(0, Y.cpu().numpy(), y_star, den, f"graph {kw} replay {rep}")
        assert eng.debug()["graphs"] == ran, "one captured graph per distinct option set"


def test_level_chunk_options(hap, geno):
    """EV_LDS windows, XCD-partitioned gather chunks, bit-mode ranges and chunks."""
    for name, (eng, seq, V), rows, cols in (("hap", hap, 4096, 2048), ("geno", geno, 6000, 10000)):
        for sr in (0, 2):
            x = x_for(sr, cols, 400 + sr)
            y_star, den = O.brute_force(sr, rows, cols, seq, V, x)
            for win in (1024, 2048, 8192, 16384):
                for em in (gmr.EMIT_SPOT, gmr.EMIT_GATHER):
                    _say(name, "EV_LDS window", sr, win, em)
                    y = run(eng, sr, x, emit=em, eval_kind=gmr.EVAL_LDS, lds_values=win)
                    check(sr, y, y_star, den, f"{name} window={win} emit={em}")
            for ch in (1, 2, 4):
                for passes in (1, 2, 4):
                    _say(name, "xcd chunks", sr, ch, passes)
                    y = run(eng, sr, x, emit=gmr.EMIT_GATHER, chunks=ch, xcd_passes=passes)
                    check(sr, y, y_star, den, f"{name} xcd chunks={ch} passes={passes}")
        xb = x_for(1, cols, 500)
        yb_star, _ = O.brute_force(1, rows, cols, seq, V, xb)
        for rng in (131072, 262144, 524288, 1048576):
            for ch in (2, 4, 8):
                _say(name, "bits", rng, ch)
                y = run(eng, 1, xb, emit=gmr.EMIT_GATHER, flags=gmr.RUN_BITS, lds_values=rng, chunks=ch)
                check(1, y, yb_star, None, f"{name} bits range={rng} chunks={ch}")
        # the bit mode's levels two to a launch (GMR_RUN_PAIRED_LEVELS), levels and graph
        for var in (gmr.VARIANT_LEVELS, gmr.VARIANT_GRAPH):
            _say(name, "bits paired", var)
            y = run(eng, 1, xb, variant=var, emit=gmr.EMIT_GATHER, flags=gmr.RUN_BITS | gmr.RUN_PAIRED_LEVELS)
            check(1, y, yb_star, None, f"{name} bits paired var={var}")


def test_option_errors(hap):
    eng, seq, V = hap
    x = torch.zeros(2048, device="cuda")
    bad = [dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES, lds_values=3000),          # not a multiple of 1024
           dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES, lds_values=1 << 20),       # too large
           dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES, chunks=3),
           dict(emit=gmr.EMIT_SPOT, flags=gmr.RUN_LDS_RANGES),                              # needs gather
           dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES, variant=gmr.VARIANT_PERSISTENT),
           dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES | gmr.RUN_BITS),
           dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_ROW_ORDER),                             # needs LDS ranges
           dict(emit=gmr.EMIT_GATHER, chunks=8),                                            # XCD gather: 1, 2, 4
           dict(eval_kind=gmr.EVAL_LDS, lds_values=65536)]                                  # window too large
    for kw in bad:
        with pytest.raises(gmr.GmrError) as e:
            eng.run_torch(0, x, **kw)
        assert e.value.code == gmr.ERR_INVALID, kw
    X4 = torch.zeros(2048, 4, device="cuda")
    with pytest.raises(gmr.GmrError):
        eng.run_torch(0, X4, emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES)  # one vector only
    torch.cuda.synchronize()


def test_tropical_large_finite_values_cap_at_infinity():
    """Finite leaf sums above 1e9 (large weights, large distances): the
    reference's y starts at 1e9 and is min-combined, so such a row reads 1e9.
    Every emission form stores that, the whole-row stores included."""
    rng = np.random.default_rng(20261207)
    rows, cols = 600, 300
    M = np.where(rng.random((rows, cols)) < 0.05, rng.choice([5e8, 6e8, 7e8, 3.0], (rows, cols)), 0.0)
    M[:40] = 0.0
    M[:40, :3] = 6e8  # rows whose every leaf sum exceeds 1e9 below
    m = gmr.Matrix.from_dense(M)
    g = m.repair()
    seq, V = m.arrays()
    for B in (1, 5):
        X = rng.uniform(3e8, 9e8, (cols, B)).astype(np.float32)
        X[:3] = 6e8
        X[rng.random((cols, B)) < 0.05] = 1e9
        Xs = X[:, 0] if B == 1 else X
        y_star = O.brute_force_batched(2, rows, cols, seq, V, X)[0]
        y_star = y_star[:, 0] if B == 1 else y_star
        assert (y_star[:40] == 1e9).all()
        for gg in (g, g.inline_top(2)):
            eng = gg.schedule(rows, cols, "resident").engine()
            for var in ("levels", "graph", "persistent"):
                for em in ("spot", "gather"):
                    flags = (0, gmr.RUN_ROW_SEGMENTS, gmr.RUN_SINGLE_LIST, gmr.RUN_LDS_RANGES) \
                        if (em == "gather" and B == 1 and var != "persistent") else (0,)
                    for fl in flags:
                        _say("tropical cap", B, var, em, fl)
                        y = run(eng, 2, Xs, variant=gmr.VARIANTS[var], emit=gmr.EMITS[em], flags=fl)
                        check(2, y, y_star, None, f"tropical cap B={B} {var}/{em}/flags={fl}")
    # the left product's fold stores whole columns the same way
    eng = g.schedule(rows, cols, "resident").engine()
    xr = rng.uniform(3e8, 9e8, rows).astype(np.float32)
    yl_star = O.brute_force_left(2, rows, cols, seq, V, xr[:, None])[0][:, 0]
    yl = eng.run_left_torch(2, torch.from_numpy(xr).cuda()).cpu().numpy()
    check(2, yl, yl_star, None, "tropical cap, left product")


def _fail_then_recover(run_once, check_rc, what):
    """Run with a 1 us barrier limit until a failure is reported (a wait of
    some workgroup outlives 1 us essentially always)."""
    for attempt in range(3):
        run_once()
        torch.cuda.synchronize()  # the failed launch has ended; its host word is set
        try:
            check_rc()
        except gmr.GmrError as e:
            assert e.code == gmr.ERR_BARRIER, f"{what}: code {e.code}"
            return
    pytest.fail(f"{what}: no barrier timeout was reported with a 1 us limit")


def test_persistent_barrier_timeout_is_reported(hap):
    """A grid-barrier wait that gives up (forced from the host with a 1 us
    limit, so every kernel still ends normally) is returned as GMR_ERR_BARRIER
    by the next run, which does not run; later runs are correct again."""
    eng, seq, V = hap
    x = x_for(0, 2048, 600)
    y_star, den = O.brute_force(0, 4096, 2048, seq, V, x)
    xt = torch.from_numpy(x).cuda()
    kw = dict(variant=gmr.VARIANT_PERSISTENT, emit=gmr.EMIT_GATHER)
    eng.set_barrier_timeout(1)
    _fail_then_recover(lambda: eng.run_torch(0, xt, **kw), lambda: eng.run_torch(0, xt, **kw), "next run")
    # a timed run reports its own failure
    _fail_then_recover(lambda: None, lambda: eng.run_torch(0, xt, flags=gmr.RUN_TIMED, **kw), "timed run")
    # gmr_engine_check reports it too
    _fail_then_recover(lambda: eng.run_torch(0, xt, **kw), eng.check, "check")
    eng.set_barrier_timeout(0)
    eng.check()
    check(0, run(eng, 0, x, **kw), y_star, den, "after recovery")
    for _ in range(3):  # and stays correct
        check(0, run(eng, 0, x, **kw), y_star, den, "after recovery, again")


def test_batch_barrier_timeout_is_reported(hap, geno):
    (e1, s1, V1), (e2, s2, V2) = hap, geno
    b = gmr.Batch([e1, e2])
    X1 = torch.from_numpy(x_for(0, 2048, 700, B=3)).cuda()
    X2 = torch.from_numpy(x_for(0, 10000, 701, B=2)).cuda()
    b.set_barrier_timeout(1)
    go = lambda: b.run_torch(0, [X1, X2], variant=gmr.VARIANT_PERSISTENT)  # noqa: E731
    _fail_then_recover(go, go, "batch next run")
    b.set_barrier_timeout(0)
    b.check()
    Ys = b.run_torch(0, [X1, X2], variant=gmr.VARIANT_PERSISTENT)
    torch.cuda.synchronize()
    b.check()
    for Y, X, seq, V, rows, cols in ((Ys[0], X1, s1, V1, 4096, 2048), (Ys[1], X2, s2, V2, 6000, 10000)):
        Y_star, den = O.brute_force_batched(0, rows, cols, seq, V, X.cpu().numpy())
        check(0, Y.cpu().numpy(), Y_star, den, "batch after recovery")


def test_persistent_block_without_headroom_is_refused_or_correct(hap):
    """block = 1024: runs when at least two workgroups fit per CU (one stays
    free as headroom), otherwise is refused with GMR_ERR_INVALID."""
    eng, seq, V = hap
    x = x_for(0, 2048, 800)
    y_star, den = O.brute_force(0, 4096, 2048, seq, V, x)
    try:
        y = run(eng, 0, x, variant=gmr.VARIANT_PERSISTENT, emit=gmr.EMIT_GATHER, block=1024)
    except gmr.GmrError as e:
        assert e.code == gmr.ERR_INVALID and "headroom" in str(e)
        return
    check(0, y, y_star, den, "persistent block=1024")
    assert eng.debug()["per_cu"] >= 1
