"""GPU parity: the HIP sweep against the oracle (FP64 brute force from the
uncompressed matrix) through the C ABI, for every semiring, layout, launch
variant, emission strategy and batch width. Bar: Boolean and Tropical
bit-exact; (+,x) max_i |y_i - y*_i| / sum_j |M_ij x_j| <= 1e-5."""
import ctypes as C
import sys

import numpy as np
import pytest

import gmr
from gmr import _lib
from helpers import fixture, grammar_dict, rel_err, x_for
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


def run_engine(eng, sr, X, **kw):
    Xt = torch.from_numpy(np.ascontiguousarray(X, np.float32)).cuda()
    Y = eng.run_torch(sr, Xt, **kw)
    torch.cuda.synchronize()
    if kw.get("variant") == gmr.VARIANT_PERSISTENT:
        eng.check()  # no grid-barrier wait may have timed out
    return Y.cpu().numpy()


def check(sr, y, y_star, den, what):
    assert np.shape(y) == np.shape(y_star), f"{what}: shape {np.shape(y)} != {np.shape(y_star)}"
    if sr == 0:
        e = rel_err(y, y_star, den)
        _say(f"  {what}: err={e:.3g}")
        assert e <= TOL, f"{what}: FP32 error {e} > {TOL}"
    else:
        ok = np.array_equal(np.asarray(y, np.int64), np.asarray(y_star, np.int64))
        assert ok, f"{what}: not bit-exact ({int((np.asarray(y, np.int64) != y_star).sum())} rows differ)"


CASES = [("levels", "spot"), ("graph", "spot"), ("persistent", "spot")]
RES_CASES = CASES + [("levels", "gather"), ("graph", "gather"), ("persistent", "gather")]


@pytest.fixture(scope="module")
def hap():
    m = gmr.Matrix.haplotype(20261204, 4096, 2048)
    g = m.repair()
    seq, V = m.arrays()
    return m, g, seq, V


def test_unit_grammar_on_gpu():
    f = fixture("unit_grammar")
    g = gmr.Grammar.from_arrays(f["alpha"], f["left"], f["right"], f["C"], f["V"], kind=f["kind"], rep=f["rep"])
    x = np.array(f["x"], np.float32)
    # as given, and with its top rule inlined (the run-length rule it uses stays)
    for gg in (g, g.inline_top(2)):
        for lay in ("refcompat", "wave64", "resident"):
            eng = gg.schedule(2, 3, lay).engine()
            for var, em in (RES_CASES if lay == "resident" else CASES):
                _say("unit", lay, var, em)
                y = run_engine(eng, 0, x, variant=gmr.VARIANTS[var], emit=gmr.EMITS[em])
                np.testing.assert_allclose(y, f["expect_y"], atol=1e-6)
            if lay == "resident":
                # levels two to a launch (the run-length rule's coefficient applied to a
                # recomputed child) and the top levels in one workgroup
                for fl in (gmr.RUN_PAIRED_LEVELS, gmr.RUN_FUSED_TAIL, gmr.RUN_PAIRED_LEVELS | gmr.RUN_FUSED_TAIL):
                    for em in ("spot", "gather"):
                        y = run_engine(eng, 0, x, emit=gmr.EMITS[em], flags=fl)
                        np.testing.assert_allclose(y, f["expect_y"], atol=1e-6)


def test_readme_example_on_gpu():
    f = fixture("readme_8x7")
    m = gmr.Matrix.from_dense(np.array(f["M"]))
    g = m.repair()
    for lay in ("refcompat", "resident"):
        eng = g.schedule(8, 7, lay).engine()
        y = run_engine(eng, 0, np.array(f["x"], np.float32))
        np.testing.assert_allclose(y, f["expect_y"], rtol=1e-6)


def test_integration_matrix_on_gpu():
    """The reference's CI matrix (tests/run_integration_tests.py:110-128: 50 x 50,
    rows drawn from 5 patterns over {0, 1, 2}), compressed by this package's
    RePair, through every layout, launch variant and emission, the single-vector
    gathers (row segments, LDS ranges, bit mode) and batches, for the three
    semirings, against the oracle."""
    f = fixture("integration_50x50")
    rows, cols = f["rows"], f["cols"]
    m = gmr.Matrix.from_dense(np.array(f["M"]))
    g = m.repair()
    seq, V = m.arrays()
    for sr in (0, 1, 2):
        x = x_for(sr, cols, 4200 + sr)
        y_star, den = O.brute_force(sr, rows, cols, seq, V, x)
        X = x_for(sr, cols, 4210 + sr, B=5)
        Y_star, Den = O.brute_force_batched(sr, rows, cols, seq, V, X)
        for lay in ("refcompat", "wave64", "resident"):
            eng = g.schedule(rows, cols, lay).engine()
            for var, em in (RES_CASES if lay == "resident" else CASES):
                _say("ci50", sr, lay, var, em)
                kw = dict(variant=gmr.VARIANTS[var], emit=gmr.EMITS[em])
                check(sr, run_engine(eng, sr, x, **kw), y_star, den, f"ci50 {lay}/{var}/{em}")
                if lay != "refcompat":
                    check(sr, run_engine(eng, sr, X, **kw), Y_star, Den, f"ci50 B=5 {lay}/{var}/{em}")
        eng = g.schedule(rows, cols, "resident").engine()
        for kw in (dict(flags=gmr.RUN_ROW_SEGMENTS), dict(flags=gmr.RUN_LDS_RANGES, lds_values=4096)):
            check(sr, run_engine(eng, sr, x, emit=gmr.EMIT_GATHER, **kw), y_star, den, f"ci50 {kw}")
        if sr == 1:
            check(sr, run_engine(eng, sr, x, emit=gmr.EMIT_GATHER, flags=gmr.RUN_BITS), y_star, den, "ci50 bits")


@pytest.mark.parametrize("sr", [0, 1, 2])
@pytest.mark.parametrize("lay", ["refcompat", "wave64", "resident"])
def test_haplotype_single_vector(hap, sr, lay):
    m, g, seq, V = hap
    eng = g.schedule(4096, 2048, lay).engine()
    x = x_for(sr, 2048, 100 + sr)
    y_star, den = O.brute_force(sr, 4096, 2048, seq, V, x)
    for var, em in (RES_CASES if lay == "resident" else CASES):
        _say("hap B=1", sr, lay, var, em)
        y = run_engine(eng, sr, x, variant=gmr.VARIANTS[var], emit=gmr.EMITS[em])
        check(sr, y, y_star, den, f"{lay}/{var}/{em}")


@pytest.mark.parametrize("B", [4, 16, 64, 256, 3, 7, 100, 125, 9, 27, 50, 91])
def test_haplotype_batched_distinct_columns(hap, B):
    """B vectors with distinct columns, every layout and plan; 9, 27, 50 and 91
    vectors run in value rows padded to 16, 32, 64 and 96 floats (fewer
    128-byte lines per row than the next multiple of 4)."""
    m, g, seq, V = hap
    X = x_for(0, 2048, 7 + B, B=B)
    Y_star, den = O.brute_force_batched(0, 4096, 2048, seq, V, X)
    for lay in ("wave64", "resident"):
        eng = g.schedule(4096, 2048, lay).engine()
        for var, em in (RES_CASES if lay == "resident" else CASES):
            _say("hap B", B, lay, var, em)
            Y = run_engine(eng, 0, X, variant=gmr.VARIANTS[var], emit=gmr.EMITS[em])
            check(0, Y, Y_star, den, f"B={B} {lay}/{var}/{em}")


@pytest.mark.parametrize("sr", [1, 2])
def test_batched_integer_semirings(hap, sr):
    m, g, seq, V = hap
    X = x_for(sr, 2048, 55, B=16)
    Y_star, _ = O.brute_force_batched(sr, 4096, 2048, seq, V, X)
    eng = g.schedule(4096, 2048, "resident").engine()
    for var, em in RES_CASES:
        _say("int batched", sr, var, em)
        Y = run_engine(eng, sr, X, variant=gmr.VARIANTS[var], emit=gmr.EMITS[em])
        check(sr, Y, Y_star, None, f"sr={sr} {var}/{em}")


def test_b_tiles_match_untiled(hap):
    """Vector tiles: one launch sequence per tile (levels, graph) or every tile
    in one persistent launch, including a narrower last tile and widths that
    are not multiples of 4."""
    m, g, seq, V = hap
    eng = g.schedule(4096, 2048, "resident").engine()
    for B, tiles in ((64, (8, 16, 24)), (70, (16, 33))):
        X = x_for(0, 2048, 77 + B, B=B)
        Y_star, den = O.brute_force_batched(0, 4096, 2048, seq, V, X)
        for bt in tiles:
            for var in ("levels", "graph", "persistent"):
                for em in ("gather", "spot"):
                    _say("b_tile", B, bt, var, em)
                    Y = run_engine(eng, 0, X, variant=gmr.VARIANTS[var], emit=gmr.EMITS[em], b_tile=bt)
                    check(0, Y, Y_star, den, f"B={B} b_tile={bt} {var}/{em}")
    for sr in (1, 2):
        X = x_for(sr, 2048, 5 + sr, B=20)
        Y_star, _ = O.brute_force_batched(sr, 4096, 2048, seq, V, X)
        for em in ("gather", "spot"):
            _say("b_tile int", sr, em)
            Y = run_engine(eng, sr, X, variant=gmr.VARIANT_PERSISTENT, emit=gmr.EMITS[em], b_tile=8)
            check(sr, Y, Y_star, None, f"sr={sr} b_tile=8 persistent/{em}")


def test_graph_cache_keys_on_the_gather_form(hap):
    """A captured graph is reused only for the same schedule: another gather
    form or pass count on the same buffers captures its own graph."""
    m, g, seq, V = hap
    eng = g.schedule(4096, 2048, "resident").engine()
    x = x_for(0, 2048, 3)
    y_star, den = O.brute_force(0, 4096, 2048, seq, V, x)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty(4096, device="cuda")
    seen = []
    for kw in ({}, {"xcd_passes": 1}, {"xcd_passes": 2}, {"flags": gmr.RUN_SINGLE_LIST},
               {"flags": gmr.RUN_ROW_SEGMENTS}, {"xcd_passes": 1}):
        eng.run_torch(0, xt, yt, variant=gmr.VARIANT_GRAPH, emit=gmr.EMIT_GATHER, **kw)
        torch.cuda.synchronize()
        check(0, yt.cpu().numpy(), y_star, den, f"graph {kw}")
        seen.append(eng.debug()["graphs"])
    assert seen == [1, 2, 3, 4, 5, 5], seen


def test_graphs_replay_consistently(hap):
    """A captured graph replayed several times, with eager work between the
    replays, gives the eager result every time."""
    m, g, seq, V = hap
    eng = g.schedule(4096, 2048, "resident").engine()
    X = x_for(0, 2048, 9, B=4)
    eager = run_engine(eng, 0, X)
    for i in range(4):
        y = run_engine(eng, 0, X, variant=gmr.VARIANT_GRAPH, emit=gmr.EMIT_SPOT if i % 2 else gmr.EMIT_GATHER)
        _ = run_engine(eng, 0, X)  # eager work in between
        Y_star, den = O.brute_force_batched(0, 4096, 2048, seq, V, X)
        check(0, y, Y_star, den, f"replay {i}")
        assert rel_err(y, eager, den) <= TOL


def test_rmat_graph_semirings():
    for nw, sr in ((1, 1), (100, 2)):
        m = gmr.Matrix.rmat(2026 + nw, 14, 200000, nweights=nw)
        g = m.repair()
        seq, V = m.arrays()
        n = m.rows
        if sr == 1:
            x = (np.random.default_rng(1).random(n) < 0.01).astype(np.float32)
        else:
            x = np.full(n, np.float32(2 ** 31 - 1) // 2, np.float32)
            x[np.random.default_rng(2).choice(n, 64, replace=False)] = 0
        y_star, _ = O.brute_force(sr, n, n, seq, V, x)
        for lay in ("wave64", "resident"):
            eng = g.schedule(n, n, lay).engine()
            for var, em in (RES_CASES if lay == "resident" else CASES):
                _say("rmat", sr, lay, var, em)
                y = run_engine(eng, sr, x, variant=gmr.VARIANTS[var], emit=gmr.EMITS[em])
                check(sr, y, y_star, None, f"rmat sr={sr} {lay}/{var}/{em}")


def test_empty_rows_and_terminal_only_grammar():
    g = gmr.Grammar.from_arrays(4, [], [], [1, 2, 0, 0, 3, 0], [1.0])
    x = np.array([1.0, 2.0, 4.0], np.float32)
    for lay in ("refcompat", "resident"):
        eng = g.schedule(3, 3, lay).engine()
        for var, em in (RES_CASES if lay == "resident" else CASES):
            y = run_engine(eng, 0, x, variant=gmr.VARIANTS[var], emit=gmr.EMITS[em])
            np.testing.assert_array_equal(y, [3.0, 0.0, 4.0])
            yt = run_engine(eng, 2, x, variant=gmr.VARIANTS[var], emit=gmr.EMITS[em])
            np.testing.assert_array_equal(yt, [2, 1000000000, 5])


def test_run_errors_are_reported(hap):
    m, g, seq, V = hap
    eng = g.schedule(4096, 2048, "wave64").engine()
    X = torch.zeros(2048, dtype=torch.float32, device="cuda")
    with pytest.raises(gmr.GmrError, match="Resident"):
        eng.run_torch(0, X, emit=gmr.EMIT_GATHER)
    with pytest.raises(gmr.GmrError, match="semiring"):
        eng.run(7, X.data_ptr(), 1, X.data_ptr())
    with pytest.raises(gmr.GmrError, match="B must"):
        eng.run(0, X.data_ptr(), 0, X.data_ptr())


def _dropin(batched, sched_arrays, rows, cols, V, X, B, semiring_env, monkeypatch):
    monkeypatch.setenv("SEMIRING", semiring_env)
    monkeypatch.setenv("GMR_ENERGY_SECONDS", "0")  # no 1.5 s energy loop per call
    L = _lib.dev()
    s = sched_arrays
    to = s["term_orig"]
    val_idx = np.ascontiguousarray((to[1:] - 1) // cols, np.int32)
    col_idx = np.ascontiguousarray((to[1:] - 1) % cols, np.int32)
    n = len(s["left"])
    rules = (_lib.LayeredRule * max(n, 1))()
    for i in range(n):
        rules[i] = _lib.LayeredRule(int(s["left"][i]), int(s["right"][i]), float(s["coeff"][i]))
    keep = []

    def ip(a):
        a = np.ascontiguousarray(a, np.int32)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_int))
    gs = _lib.GPUSchedule(len(s["level_offsets"]) - 1, int(s["info"]["max_width"]), ip(s["level_offsets"]),
                          ip(s["level_widths"]), ip(s["emit_offsets"]), ip(s["emit_pos"]), ip(s["emit_row"]),
                          len(s["emit_pos"]), ip(s["term_sym"]), ip(s["term_row"]), len(s["term_sym"]))
    Vf = np.ascontiguousarray(V, np.float32)
    Xf = np.ascontiguousarray(X, np.float32)
    Y = np.zeros(rows * B, np.float32)
    fp = C.POINTER(C.c_float)
    if batched:
        L.run_gpu_right_multiply_batched(rows, cols, len(to), n, rules, ip(val_idx), ip(col_idx),
                                         Vf.ctypes.data_as(fp), len(Vf), Xf.ctypes.data_as(fp), C.byref(gs),
                                         Y.ctypes.data_as(fp), B, 2)
    else:
        L.run_gpu_right_multiply_layered(rows, cols, len(to), n, rules, ip(val_idx), ip(col_idx),
                                         Vf.ctypes.data_as(fp), len(Vf), Xf.ctypes.data_as(fp), C.byref(gs),
                                         Y.ctypes.data_as(fp), 2)
    return Y if semiring_env == "plustimes" else Y.view(np.int32)


def test_dropin_entry_points(hap, monkeypatch):
    m, g, seq, V = hap
    s = g.schedule(4096, 2048, "refcompat").arrays()
    for sr, env in ((0, "plustimes"), (1, "boolean"), (2, "tropical")):
        x = x_for(sr, 2048, 300 + sr)
        y_star, den = O.brute_force(sr, 4096, 2048, seq, V, x)
        _say("dropin layered", env)
        y = _dropin(False, s, 4096, 2048, V, x, 1, env, monkeypatch)
        check(sr, y, y_star, den, f"dropin {env}")
    # the drop-in writes top rules out in C at B = 1 (up to 8 uses), 2-4 and >= 64 (up to 2 uses);
    # B = 8 runs the grammar as scheduled
    for sr, env, B in ((0, "plustimes", 4), (0, "plustimes", 8), (0, "plustimes", 64), (1, "boolean", 4),
                       (2, "tropical", 64)):
        X = x_for(sr, 2048, 400 + B + sr, B=B)
        Y_star, den = O.brute_force_batched(sr, 4096, 2048, seq, V, X)
        _say("dropin batched", env, B)
        Y = _dropin(True, s, 4096, 2048, V, X.reshape(-1), B, env, monkeypatch).reshape(4096, B)
        check(sr, Y, Y_star, den, f"dropin batched {env} B={B}")


@pytest.mark.parametrize("B", [1, 3, 4, 64])
def test_eval_kernels_per_level(hap, B):
    """Every node-update kernel, uniform and mixed per level, in every launch
    variant, gives the oracle's product."""
    m, g, seq, V = hap
    X = x_for(0, 2048, 900 + B, B=B) if B > 1 else x_for(0, 2048, 900)
    if B > 1:
        Y_star, den = O.brute_force_batched(0, 4096, 2048, seq, V, X)
    else:
        Y_star, den = O.brute_force(0, 4096, 2048, seq, V, X)
    rng = np.random.default_rng(B)
    for lay in ("wave64", "resident"):
        eng = g.schedule(4096, 2048, lay).engine()
        L = eng.info()["max_depth"]
        mixed = rng.integers(0, 4, L + 1).astype(np.int32)
        for var in ("levels", "graph", "persistent"):
            for kind in (0, 1, 2, 3, None):
                kw = dict(variant=gmr.VARIANTS[var], emit=gmr.EMIT_SPOT)
                if kind is None:
                    kw["level_kind"] = mixed
                else:
                    kw["eval_kind"] = kind
                _say("eval kinds", B, lay, var, kind)
                Y = run_engine(eng, 0, X, **kw)
                check(0, Y, Y_star, den, f"B={B} {lay}/{var}/kind={kind}")


def test_eval_kernels_integer_semirings(hap):
    m, g, seq, V = hap
    eng = g.schedule(4096, 2048, "wave64").engine()
    for sr in (1, 2):
        x = x_for(sr, 2048, 70 + sr)
        y_star, _ = O.brute_force(sr, 4096, 2048, seq, V, x)
        for kind in range(4):
            for var in ("levels", "persistent"):
                _say("int eval kinds", sr, kind, var)
                y = run_engine(eng, sr, x, variant=gmr.VARIANTS[var], eval_kind=kind)
                check(sr, y, y_star, None, f"sr={sr} kind={kind} {var}")


def test_time_levels_reports_every_level(hap):
    m, g, seq, V = hap
    eng = g.schedule(4096, 2048, "resident").engine()
    X = torch.from_numpy(x_for(0, 2048, 5, B=16)).cuda()
    Y = torch.empty(4096, 16, device="cuda")
    for kind in range(4):
        t = eng.time_levels(0, X, Y, reps=2, eval_kind=kind)
        assert len(t) == eng.info()["max_depth"] + 2
        assert np.all(t >= 0) and t[1:-1].sum() > 0


def test_batch_of_uneven_grammars():
    """Grammars of different sizes, depths, layouts and B swept together in one
    launch (persistent), one graph (branches) or one after another: each Y_g
    equals its oracle product."""
    rng = np.random.default_rng(11)
    specs = [("hap", 1, 700, 300, 5), ("hap", 2, 2000, 512, 64), ("geno", 3, 3000, 400, 17),
             ("rmat", 4, 10, 6000, 1), ("hap", 5, 64, 40, 128), ("geno", 6, 900, 256, 8)]
    mats, engs, Xs, stars = [], [], [], []
    for kind, seed, a, b, B in specs:
        if kind == "hap":
            m = gmr.Matrix.haplotype(seed, a, b, founders=12, seg_mean=24.0)
        elif kind == "geno":
            m = gmr.Matrix.genotype(seed, a, b, founders=32, classes=4, seg_mean=32.0)
        else:
            m = gmr.Matrix.rmat(seed, a, b)
        g = m.repair()
        seq, V = m.arrays()
        lay = ["resident", "wave64", "refcompat"][seed % 3]
        eng = g.schedule(m.rows, m.cols, lay).engine()
        X = x_for(0, m.cols, seed, B=B)
        Ys, den = O.brute_force_batched(0, m.rows, m.cols, seq, V, X)
        mats.append((m, g, lay))
        engs.append(eng)
        Xs.append(torch.from_numpy(X).cuda())
        stars.append((Ys, den))
    batch = gmr.Batch(engs)
    for var in ("levels", "graph", "persistent"):
        for emit_gather in (False, True):
            opts = []
            for (m, g, lay), eng in zip(mats, engs):
                L = eng.info()["max_depth"]
                o = {"emit": gmr.EMIT_GATHER if (emit_gather and lay == "resident") else gmr.EMIT_SPOT,
                     "level_kind": rng.integers(0, 4, L + 1).astype(np.int32)}
                opts.append(o)
            _say("batch", var, "gather" if emit_gather else "spot")
            Ys = batch.run_torch(0, Xs, variant=gmr.VARIANTS[var], opts=opts)
            torch.cuda.synchronize()
            if var == "persistent":
                batch.check()
            for i, (Y, (Ystar, den)) in enumerate(zip(Ys, stars)):
                check(0, Y.cpu().numpy(), Ystar, den, f"batch {var} grammar {i}")


def test_batch_integer_semirings():
    ms = [gmr.Matrix.rmat(40 + i, 9 + i % 3, 3000 * (i + 1), nweights=1 + 99 * (i % 2)) for i in range(4)]
    engs, Xs, stars = [], [], []
    for i, m in enumerate(ms):
        g = m.repair()
        seq, V = m.arrays()
        engs.append(g.schedule(m.rows, m.cols, "resident").engine())
        X = x_for(2, m.cols, i, B=4 + i)
        Xs.append(torch.from_numpy(X).cuda())
        stars.append(O.brute_force_batched(2, m.rows, m.cols, seq, V, X)[0])
    batch = gmr.Batch(engs)
    for var in ("graph", "persistent"):
        Ys = batch.run_torch(2, Xs, variant=gmr.VARIANTS[var], opts=[{"emit": gmr.EMIT_GATHER}] * 4)
        torch.cuda.synchronize()
        if var == "persistent":
            batch.check()
        for Y, Ys_ in zip(Ys, stars):
            check(2, Y.cpu().numpy(), Ys_, None, f"batch tropical {var}")


def test_gpu_test_driver_crosscheck_files(tmp_path):
    """The reference-compatible driver: XVEC in, CROSSCHECK out (float32 x,
    float64 y), for every semiring, and its batched self-check."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "g-mm-repair_amd", "bin",
                       "gpu_test")
    m = gmr.Matrix.haplotype(77, 3000, 1024, founders=24, seg_mean=48.0)
    g = m.repair()
    seq, V = m.arrays()
    base = str(tmp_path / "g")
    g.write(base)
    for sr, name in ((0, "plustimes"), (1, "boolean"), (2, "tropical")):
        x = x_for(sr, 1024, 500 + sr)
        xf = tmp_path / f"x{sr}.bin"
        x.tofile(xf)
        prefix = str(tmp_path / "out")
        env = dict(os.environ, SEMIRING=name, XVEC=str(xf), CROSSCHECK=prefix,
                   GMR_ENERGY_SECONDS="0.5" if sr == 0 else "0")
        _say("gpu_test", name)
        r = subprocess.run([exe, base, "3000", "1024", "2", "repair", "4"], capture_output=True, text=True, env=env,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "GPU kernel:" in r.stdout and "Batched verification" in r.stdout and "SUCCESS" in r.stdout
        # the in-run CPU check of the reference driver (gpu_engine_test.cu:143-176, :251-274)
        assert "CPU OpenMP vs sequential: max_rel_diff=" in r.stdout
        assert "SUCCESS: GPU and CPU reference results match" in r.stdout, r.stdout
        # energy per vector between two counter updates >= 0.5 s apart: a
        # positive reading, or -1 where the counter cannot be read
        mj = float(r.stdout.split("ENERGY mJ/vector:")[1].split()[0])
        assert (mj > 0 or mj == -1) if sr == 0 else mj == -1
        xd = np.fromfile(prefix + f".{name}.x", np.float32)
        yd = np.fromfile(prefix + f".{name}.y", np.float64)
        np.testing.assert_array_equal(xd, x)
        y_star, den = O.brute_force(sr, 3000, 1024, seq, V, x)
        check(sr, yd, y_star, den, f"gpu_test {name}")


def test_gather_orders_agree(hap):
    m, g, seq, V = hap
    eng = g.schedule(4096, 2048, "resident").engine()
    X = x_for(0, 2048, 31, B=16)
    Y_star, den = O.brute_force_batched(0, 4096, 2048, seq, V, X)
    for order in (1, 0, 1):
        eng.set_gather_order(order)
        for var in ("levels", "graph", "persistent"):
            Y = run_engine(eng, 0, X, variant=gmr.VARIANTS[var], emit=gmr.EMIT_GATHER)
            check(0, Y, Y_star, den, f"gather order {order} {var}")


@pytest.mark.parametrize("sr", [0, 1, 2])
def test_single_vector_gathers_agree(sr):
    """One vector, row gather: the XCD-partitioned and single-list chunked
    kernels (a wave per 64 top-sequence symbols, segmented scan) and the
    row-segment kernel, on a matrix whose rows are empty, 1 symbol, exactly 64,
    and long enough to cross several chunks (and split into several segments),
    with a symbol count that is not a multiple of 64."""
    rng = np.random.default_rng(77 + sr)
    rows, cols = 40, 300
    lens = [0, 1, 64, 63, 65, 0, 0, 200, 1, 129, 2, 0, 300, 5] + list(rng.integers(0, 90, rows - 14))
    M = np.zeros((rows, cols), np.float64)
    for r, n in enumerate(lens):
        M[r, rng.choice(cols, int(n), replace=False)] = rng.integers(1, 4, int(n)) if sr == 2 else 1.0
    m = gmr.Matrix.from_dense(M)
    g = m.repair()
    seq, V = m.arrays()
    x = x_for(sr, cols, 5 + sr)
    y_star, den = O.brute_force(sr, rows, cols, seq, V, x)
    eng = g.schedule(rows, cols, "resident").engine()
    for flags in (0, gmr.RUN_ROW_SEGMENTS, gmr.RUN_SINGLE_LIST, 0):
        for var in ("levels", "graph", "persistent"):
            _say("one-vector gather", sr, flags, var)
            y = run_engine(eng, sr, x, variant=gmr.VARIANTS[var], emit=gmr.EMIT_GATHER, flags=flags)
            check(sr, y, y_star, den, f"gather flags={flags} {var}")
    for passes in (1, 2, 3, 8, 0):
        for var in ("levels", "graph", "persistent"):
            _say("xcd-partitioned gather", sr, passes, "passes", var)
            y = run_engine(eng, sr, x, variant=gmr.VARIANTS[var], emit=gmr.EMIT_GATHER, xcd_passes=passes)
            check(sr, y, y_star, den, f"xcd passes={passes} {var}")
    with pytest.raises(gmr.GmrError):
        run_engine(eng, sr, x, emit=gmr.EMIT_GATHER, xcd_passes=9)
    # grid shapes the partitioned gather must handle: fewer than 8 workgroups
    # (each takes several lists) and counts that are not multiples of 8
    for cap in (1, 3, 12):
        _say("xcd-partitioned gather, persistent grid cap", cap)
        y = run_engine(eng, sr, x, variant=gmr.VARIANT_PERSISTENT, emit=gmr.EMIT_GATHER, grid_cap=cap, block=64)
        check(sr, y, y_star, den, f"xcd parts grid<={cap}")
    # in a fused batch the partitioned gather falls back to the chunked one
    xt = torch.from_numpy(x).cuda()[:, None]
    eng2 = g.schedule(rows, cols, "resident").engine()
    batch = gmr.Batch([eng, eng2])
    for var in ("persistent", "graph", "levels"):
        _say("batch with xcd-partitioned gather", sr, var)
        Ys = batch.run_torch(sr, [xt, xt], variant=gmr.VARIANTS[var],
                             opts=[{"emit": gmr.EMIT_GATHER}] * 2)
        torch.cuda.synchronize()
        if var == "persistent":
            batch.check()
        for Yb in Ys:
            check(sr, Yb.cpu().numpy()[:, 0], y_star, den, f"batch {var} xcd parts")


def test_wide_batches_loop_over_vectors():
    """B wider than one lane group covers (Bs > 256: every lane takes several
    4-vector chunks) in the node update, both gathers and the left product."""
    m = gmr.Matrix.haplotype(41, 600, 300, founders=12, seg_mean=24.0)
    g = m.repair()
    seq, V = m.arrays()
    eng = g.schedule(600, 300, "resident").engine()
    for B in (520, 1030):
        X = x_for(0, 300, B, B=B)
        Y_star, den = O.brute_force_batched(0, 600, 300, seq, V, X)
        for var in ("levels", "graph", "persistent"):
            for em in ("gather", "spot"):
                _say("wide B", B, var, em)
                Y = run_engine(eng, 0, X, variant=gmr.VARIANTS[var], emit=gmr.EMITS[em])
                check(0, Y, Y_star, den, f"B={B} {var}/{em}")
        XL = x_for(0, 600, B + 1, B=B)
        YL_star, denL = O.brute_force_left(0, 600, 300, seq, V, XL)
        YL = eng.run_left_torch(0, torch.from_numpy(XL).cuda()).cpu().numpy()
        check(0, YL, YL_star, denL, f"left B={B}")


@pytest.mark.parametrize("sr", [0, 1, 2])
def test_inlined_top_rules(hap, sr):
    """The grammar with its once- and twice-used top rules written out in C
    (Grammar.inline_top) gives the same products in every variant, single and
    batched, and the same left product."""
    m, g, seq, V = hap
    g2 = g.inline_top(2)
    assert g2.n_inlined > 0
    eng = g2.schedule(4096, 2048, "resident").engine()
    for B in (1, 64):
        X = x_for(sr, 2048, 70 + B, B=None if B == 1 else B)
        Y_star, den = (O.brute_force if B == 1 else O.brute_force_batched)(sr, 4096, 2048, seq, V, X)
        for var, em in RES_CASES:
            _say("inlined", sr, B, var, em)
            Y = run_engine(eng, sr, X, variant=gmr.VARIANTS[var], emit=gmr.EMITS[em])
            check(sr, Y, Y_star, den if sr == 0 else None, f"inlined sr={sr} B={B} {var}/{em}")
    XL = x_for(sr, 4096, 90)
    YL_star, denL = O.brute_force_left(sr, 4096, 2048, seq, V, XL)
    YL = eng.run_left_torch(sr, torch.from_numpy(XL).cuda()).cpu().numpy()
    check(sr, YL, YL_star, denL if sr == 0 else None, f"inlined left sr={sr}")


@pytest.mark.parametrize("sr", [0, 1, 2])
def test_narrow_tiles_on_the_partitioned_gather(hap, sr):
    """Tiles of 2-4 vectors through the XCD-partitioned gather (their default)
    and the row segments: several pass counts, every variant, and a 10-vector
    product cut into tiles of 4, 4 and 2."""
    m, g, seq, V = hap
    eng = g.schedule(4096, 2048, "resident").engine()
    for B, tile in ((2, 0), (3, 0), (4, 0), (10, 4)):
        X = x_for(sr, 2048, 80 + B, B=B)
        Y_star, den = O.brute_force_batched(sr, 4096, 2048, seq, V, X)
        for var in ("levels", "graph", "persistent"):
            for flags, passes in ((0, 0), (0, 1), (0, 3), (0, 8), (gmr.RUN_ROW_SEGMENTS, 0)):
                _say("narrow", sr, B, tile, var, flags, passes)
                Y = run_engine(eng, sr, X, variant=gmr.VARIANTS[var], emit=gmr.EMIT_GATHER, b_tile=tile,
                               flags=flags, xcd_passes=passes)
                check(sr, Y, Y_star, den if sr == 0 else None,
                      f"narrow sr={sr} B={B} tile={tile} {var} flags={flags} p={passes}")


def test_boolean_bit_mode(hap):
    """Boolean, one vector, values as bits (GMR_RUN_BITS; LDS-range lists,
    and the flat list with GMR_RUN_SINGLE_LIST): bit-exact against the oracle
    and the int32 path on the haplotype grammar (as built and with its top
    rules inlined), ragged rows (empty, exactly 64, crossing several 64-symbol
    chunks), an R-MAT graph, a weighted R-MAT graph whose bitset spans three
    LDS ranges, and a grammar without rules; frontiers of 1%, 30% and all
    ones; levels and graph variants, the graph replayed into the same buffers
    with a new x in between; y not 16-byte aligned, with paired levels."""
    m, g, seq, V = hap
    rng = np.random.default_rng(4242)
    rows, cols = 40, 300
    lens = [0, 1, 64, 63, 65, 0, 0, 200, 1, 129, 2, 0, 300, 5] + list(rng.integers(0, 90, rows - 14))
    M = np.zeros((rows, cols), np.float64)
    for r, n in enumerate(lens):
        M[r, rng.choice(cols, int(n), replace=False)] = 1.0
    ragged = gmr.Matrix.from_dense(M)
    rmat = gmr.Matrix.rmat(2027, 13, 60000)
    # weighted edges: ~1.5M distinct terminals, so the bitset spans three LDS ranges
    wide = gmr.Matrix.rmat(99, 20, 1_500_000, nweights=100)
    cases = [("haplotype", m, g), ("haplotype inlined", m, g.inline_top()), ("ragged", ragged, ragged.repair()),
             ("rmat", rmat, rmat.repair()), ("weighted rmat, three ranges", wide, wide.repair())]
    for name, mm, gg in cases:
        sq, VV = mm.arrays()
        eng = gg.schedule(mm.rows, mm.cols, "resident").engine()
        for dens in (0.01, 0.3, 1.0):
            x = (rng.random(mm.cols) < dens).astype(np.float32)
            y_star, _ = O.brute_force(1, mm.rows, mm.cols, sq, VV, x)
            for var in ("levels", "graph"):
                for fl in (gmr.RUN_BITS, gmr.RUN_BITS | gmr.RUN_SINGLE_LIST):
                    _say("bits", name, dens, var, fl)
                    y = run_engine(eng, 1, x, variant=gmr.VARIANTS[var], emit=gmr.EMIT_GATHER, flags=fl)
                    check(1, y, y_star, None, f"bits {name} {dens} {var} flags={fl}")
            y_int = run_engine(eng, 1, x, variant=gmr.VARIANT_GRAPH, emit=gmr.EMIT_GATHER)
            check(1, y_int, y_star, None, f"int32 path {name} {dens}")
        # graph replays into the same buffers, a new x written in between
        xs = [(rng.random(mm.cols) < 0.05).astype(np.float32) for _ in range(3)]
        Xt = torch.zeros(mm.cols, dtype=torch.float32, device="cuda")
        Yt = torch.full((mm.rows,), 7, dtype=torch.int32, device="cuda")
        for rep, x in enumerate(xs + xs[:1]):
            _say("bits graph replay", name, rep)
            Xt.copy_(torch.from_numpy(x))
            eng.run_torch(1, Xt, Yt, variant=gmr.VARIANT_GRAPH, emit=gmr.EMIT_GATHER, flags=gmr.RUN_BITS)
            torch.cuda.synchronize()
            y_star, _ = O.brute_force(1, mm.rows, mm.cols, sq, VV, x)
            check(1, Yt.cpu().numpy(), y_star, None, f"bits replay {name} {rep}")
        # y at 4, 8 and 12 bytes past a 16-byte boundary (the init zeroes y four
        # rows per store between a head and a tail of single rows), paired levels
        x = (rng.random(mm.cols) < 0.05).astype(np.float32)
        y_star, _ = O.brute_force(1, mm.rows, mm.cols, sq, VV, x)
        Xt = torch.from_numpy(x).cuda()
        base = torch.full((mm.rows + 8,), 7, dtype=torch.int32, device="cuda")
        for shift in (1, 2, 3):
            _say("bits unaligned y", name, shift)
            base.fill_(7)
            Yt = base[shift:shift + mm.rows]
            eng.run_torch(1, Xt, Yt, emit=gmr.EMIT_GATHER, flags=gmr.RUN_BITS | gmr.RUN_PAIRED_LEVELS)
            torch.cuda.synchronize()
            check(1, Yt.cpu().numpy(), y_star, None, f"bits unaligned y {name} shift={shift}")
            got = base.cpu().numpy()
            assert (got[:shift] == 7).all() and (got[shift + mm.rows:] == 7).all(), "wrote outside y"
    # a grammar without rules: terminals only, empty rows
    t = gmr.Grammar.from_arrays(4, [], [], [1, 2, 0, 0, 3, 0], [1.0])
    eng = t.schedule(3, 3, "resident").engine()
    for x, want in (([1.0, 0.0, 0.0], [1, 0, 0]), ([0.0, 0.0, 5.0], [0, 0, 1]), ([0.0, 2.0, 0.0], [1, 0, 0])):
        y = run_engine(eng, 1, np.array(x, np.float32), emit=gmr.EMIT_GATHER, flags=gmr.RUN_BITS)
        np.testing.assert_array_equal(y, want)
    # per-level timing runs the same launches (the bench's phase breakdown)
    eng = g.schedule(4096, 2048, "resident").engine()
    x = (rng.random(2048) < 0.05).astype(np.float32)
    Xt = torch.from_numpy(x).cuda()
    Yt = torch.zeros(4096, dtype=torch.int32, device="cuda")
    t = eng.time_levels(1, Xt, Yt, reps=2, emit=gmr.EMIT_GATHER, flags=gmr.RUN_BITS)
    assert len(t) == eng.info()["max_depth"] + 2 and t[0] > 0 and t[-1] > 0
    torch.cuda.synchronize()
    y_star, _ = O.brute_force(1, 4096, 2048, seq, V, x)
    check(1, Yt.cpu().numpy(), y_star, None, "bits time_levels")
    # what the mode does not run is refused before any launch
    X1 = torch.zeros(2048, dtype=torch.float32, device="cuda")
    X4 = torch.zeros((2048, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(gmr.GmrError, match="Boolean"):
        eng.run_torch(0, X1, emit=gmr.EMIT_GATHER, flags=gmr.RUN_BITS)
    with pytest.raises(gmr.GmrError, match="one vector"):
        eng.run_torch(1, X4, emit=gmr.EMIT_GATHER, flags=gmr.RUN_BITS)
    with pytest.raises(gmr.GmrError, match="one vector"):
        eng.run_torch(1, X1, emit=gmr.EMIT_SPOT, flags=gmr.RUN_BITS)
    with pytest.raises(gmr.GmrError, match="one vector"):
        eng.run_torch(1, X1, variant=gmr.VARIANT_PERSISTENT, emit=gmr.EMIT_GATHER, flags=gmr.RUN_BITS)
    batch = gmr.Batch([eng, g.schedule(4096, 2048, "resident").engine()])
    with pytest.raises(gmr.GmrError, match="levels variant"):
        batch.run_torch(1, [X1[:, None], X1[:, None]], variant=gmr.VARIANT_GRAPH,
                        opts=[{"emit": gmr.EMIT_GATHER, "flags": gmr.RUN_BITS}] * 2)
    Ys = batch.run_torch(1, [Xt[:, None], Xt[:, None]], variant=gmr.VARIANT_LEVELS,
                         opts=[{"emit": gmr.EMIT_GATHER, "flags": gmr.RUN_BITS}] * 2)
    torch.cuda.synchronize()
    for Yb in Ys:
        check(1, Yb.cpu().numpy()[:, 0], y_star, None, "bits in a levels batch")
