"""Host side on the CPU: seeded generators, the RePair builder, the grammar
files and the three schedule layouts, checked against the oracle and, where it
was built from the reference's sources, the reference's own scheduler and
cpu_selftest."""
import os
import subprocess

import numpy as np
import pytest

import gmr
from helpers import fixture, grammar_dict, x_for
from oracle import oracle as O

REF_SELFTEST = os.path.join(os.path.dirname(O.__file__), "_ref", "cpu_selftest")
ARRAY_KEYS = ("left", "right", "coeff", "level_offsets", "level_widths", "emit_offsets", "emit_pos",
              "emit_row", "term_sym", "term_row", "term_orig")


def _small_matrices():
    yield "haplotype", gmr.Matrix.haplotype(7, 512, 256, founders=16, seg_mean=32.0)
    yield "genotype", gmr.Matrix.genotype(8, 800, 300, founders=32, classes=4, seg_mean=40.0)
    yield "rmat", gmr.Matrix.rmat(9, 9, 4000, nweights=1)
    yield "rmat_w", gmr.Matrix.rmat(10, 9, 4000, nweights=100)
    f = fixture("integration_50x50")
    yield "ci50", gmr.Matrix.from_dense(np.array(f["M"]))
    f = fixture("readme_8x7")
    yield "readme", gmr.Matrix.from_dense(np.array(f["M"]))


def test_generators_are_seeded_and_thread_independent():
    a = gmr.Matrix.haplotype(1, 300, 128, founders=8, seg_mean=16.0).arrays()[0]
    b = gmr.Matrix.haplotype(1, 300, 128, founders=8, seg_mean=16.0).arrays()[0]
    c = gmr.Matrix.haplotype(2, 300, 128, founders=8, seg_mean=16.0).arrays()[0]
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    old = os.environ.get("OMP_NUM_THREADS")
    r = subprocess.run(
        ["python", "-c",
         "import sys;sys.path.insert(0,'g-mm-repair_amd');import gmr,hashlib;"
         "s=gmr.Matrix.genotype(5,2000,500,founders=32,classes=4).arrays()[0];"
         "print(hashlib.sha1(s.tobytes()).hexdigest())"],
        capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS="1"),
        cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import hashlib
    s = gmr.Matrix.genotype(5, 2000, 500, founders=32, classes=4).arrays()[0]
    assert r.stdout.strip() == hashlib.sha1(s.tobytes()).hexdigest()
    del old


def test_haplotype_shape_and_density():
    m = gmr.Matrix.haplotype(20261204, 4096, 2048)
    seq, V = m.arrays()
    assert m.rows == 4096 and list(V) == [1.0]
    assert (seq == 0).sum() == 4096
    dens = m.nnz() / (4096 * 2048)
    # 0.08 * 0.99 + 0.92 * 0.01 = 0.0884 (founder density with 1% flips)
    assert abs(dens - 0.0884) < 0.01
    # each row strictly increasing columns
    rows = np.split(seq, np.nonzero(seq == 0)[0] + 1)[:-1]
    assert all(np.all(np.diff(r[:-1]) > 0) for r in rows)


def test_genotype_row_counts_follow_one_over_k():
    m = gmr.Matrix.genotype(3, 20000, 1000, founders=64, classes=8, seg_mean=64.0)
    seq, _ = m.arrays()
    k = np.diff(np.nonzero(seq == 0)[0], prepend=-1) - 1
    assert k.min() >= 1 and k.max() <= 999
    # P(k) ~ 1/k: P(k=1) = 1/H_999 ~ 0.1333
    assert abs((k == 1).mean() - 1 / np.sum(1.0 / np.arange(1, 1000))) < 0.01


def test_rmat_dedup_sorted_and_weighted():
    m = gmr.Matrix.rmat(11, 10, 20000, nweights=100)
    seq, V = m.arrays()
    assert m.rows == 1024 and len(V) == 100
    assert m.nnz() <= 20000
    rows = np.split(seq, np.nonzero(seq == 0)[0] + 1)[:-1]
    for r in rows:
        col = (r[:-1] - 1) % 1024
        assert np.all(np.diff(col) > 0)  # sorted, no duplicates


@pytest.mark.parametrize("name,m", list(_small_matrices()))
def test_repair_round_trip_and_layouts(name, m):
    seq, V = m.arrays()
    g = m.repair()
    ga = g.arrays()
    assert np.array_equal(g.expand(), seq)
    assert np.array_equal(O.expand(ga), seq)
    # topological numbering, no delimiter inside rules
    idx = np.arange(len(ga["left"])) + ga["alpha"]
    assert np.all(ga["left"] < idx) and np.all(ga["right"] < idx)
    assert np.all(ga["left"] > 0) and np.all(ga["right"] > 0)
    rows, cols = m.rows, m.cols
    mine = g.schedule(rows, cols, "refcompat").arrays()
    orc = O.layer_refcompat(ga)
    for k in ARRAY_KEYS:
        np.testing.assert_array_equal(mine[k], orc[k], err_msg=k)
    ref = O.reference_schedule(ga, rows, cols)
    if ref is not None:
        for k in ARRAY_KEYS:
            np.testing.assert_array_equal(mine[k], ref[k], err_msg=k)
        for k in ("max_depth", "max_width", "num_layered", "num_passthrough", "emit_total", "num_term_emit", "U"):
            assert mine["info"][k] == ref["info"][k], k
    # every layout computes the same products (oracle sweep for double-buffered ones)
    for sr in (0, 1, 2):
        x = x_for(sr, cols, 5)
        y_bf, den = O.brute_force(sr, rows, cols, seq, V, x)
        for lay in ("refcompat", "wave64"):
            s = g.schedule(rows, cols, lay).arrays()
            y = O.sweep(sr, s, rows, cols, V, x)
            if sr == 0:
                np.testing.assert_allclose(y, y_bf, rtol=1e-9, atol=1e-9 * (1 + den.max()))
            else:
                np.testing.assert_array_equal(y, y_bf)


def test_layout_invariants():
    m = gmr.Matrix.haplotype(21, 1024, 512, founders=16, seg_mean=64.0)
    g = m.repair()
    ga = g.arrays()
    R = len(ga["left"])
    info = {lay: g.schedule(1024, 512, lay).info() for lay in ("refcompat", "wave64", "resident")}
    assert info["resident"]["num_layered"] == R and info["resident"]["num_passthrough"] == 0
    assert info["wave64"]["num_layered"] <= info["refcompat"]["num_layered"]
    for lay in ("refcompat", "wave64", "resident"):
        s = g.schedule(1024, 512, lay).arrays()
        lo = s["level_offsets"]
        assert lo[0] == 0 and lo[-1] == s["info"]["num_layered"]
        assert np.all(np.diff(s["emit_offsets"]) >= 0)
        # emission lists are row-sorted inside each level
        for k in range(1, len(lo)):
            r = s["emit_row"][s["emit_offsets"][k - 1]:s["emit_offsets"][k]]
            assert np.all(np.diff(r) >= 0)
        for k in range(1, len(lo)):
            L_ = s["left"][lo[k - 1]:lo[k]]
            R_ = s["right"][lo[k - 1]:lo[k]]
            lim = lo[k - 1] if lay == "resident" else (lo[k - 1] - lo[k - 2] if k > 1 else 0)
            for c in (L_, R_):
                assert np.all((c < 0) | (c < lim))
                assert np.all(-c[c < 0] - 1 < s["info"]["U"])
        # the emission count equals the occurrences in C
        Cs = ga["C"]
        assert s["info"]["emit_total"] == int((Cs >= ga["alpha"]).sum())
        assert s["info"]["num_term_emit"] == int(((Cs > 0) & (Cs < ga["alpha"])).sum())
    # Wave64 keeps only readable copies: refcompat minus one copy per carried chain
    lvl = np.zeros(R, np.int64)
    for i in range(R):
        a = lvl[ga["left"][i] - ga["alpha"]] if ga["left"][i] >= ga["alpha"] else 0
        b = lvl[ga["right"][i] - ga["alpha"]] if ga["right"][i] >= ga["alpha"] else 0
        lvl[i] = 1 + max(a, b)
    mref = lvl.copy()
    for i in range(R):
        for c in (ga["left"][i], ga["right"][i]):
            if c >= ga["alpha"]:
                mref[c - ga["alpha"]] = max(mref[c - ga["alpha"]], lvl[i])
    assert info["refcompat"]["num_passthrough"] == int((mref - lvl).sum())
    assert info["wave64"]["num_passthrough"] == int(np.maximum(mref - lvl - 1, 0).sum())


def test_unit_grammar_through_the_engine_scheduler():
    f = fixture("unit_grammar")
    g = gmr.Grammar.from_arrays(f["alpha"], f["left"], f["right"], f["C"], f["V"], kind=f["kind"], rep=f["rep"])
    s = g.schedule(f["rows"], f["cols"], "refcompat")
    i = s.info()
    assert i["max_depth"] == 4 and i["num_passthrough"] == 5
    y = O.sweep(0, s.arrays(), f["rows"], f["cols"], np.array(f["V"]), np.array(f["x"], np.float32))
    np.testing.assert_allclose(y, f["expect_y"], atol=1e-12)
    assert g.schedule(2, 3, "wave64").info()["num_passthrough"] == 2


def test_files_round_trip_and_reference_selftest(tmp_path):
    m = gmr.Matrix.haplotype(31, 700, 300, founders=12, seg_mean=30.0)
    g = m.repair()
    base = str(tmp_path / "hap")
    g.write(base)
    h = gmr.Grammar.read(base)
    a, b = g.arrays(), h.arrays()
    for k in ("left", "right", "C", "V"):
        np.testing.assert_array_equal(a[k], b[k])
    assert a["alpha"] == b["alpha"]
    if not os.path.exists(REF_SELFTEST):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    # The reference's own CUDA-free selftest reads our files, builds its own
    # schedule and checks its sweep against its brute force for 3 semirings.
    r = subprocess.run([REF_SELFTEST, base, "700", "300"], capture_output=True, text=True)
    assert r.returncode == 0 and "REPAIR FILE-TEST PASSED" in r.stdout, r.stdout + r.stderr


def test_bad_inputs_are_rejected(tmp_path):
    with pytest.raises(gmr.GmrError):
        gmr.Grammar.read(str(tmp_path / "missing"))
    m = gmr.Matrix.haplotype(1, 50, 40, founders=4, seg_mean=8.0)
    g = m.repair()
    base = str(tmp_path / "t")
    g.write(base)
    with open(base + ".vc.R", "ab") as fh:
        fh.write(b"\x01\x00\x00\x00")   # half a rule
    with pytest.raises(gmr.GmrError, match="truncated"):
        gmr.Grammar.read(base)
    # a rule referencing itself (not topological)
    bad = gmr.Grammar.from_arrays(3, [1, 4], [2, 2], [3, 0], [1.0])
    with pytest.raises(gmr.GmrError, match="out of range"):
        bad.schedule(1, 2)
    # more rows in C than declared
    g2 = gmr.Grammar.from_arrays(3, [1], [2], [3, 0, 1, 0], [1.0])
    with pytest.raises(gmr.GmrError, match="more rows"):
        g2.schedule(1, 2)
    with pytest.raises(gmr.GmrError):
        gmr.Matrix.from_csrv(2, 3, [1, 0], [1.0])   # 1 delimiter for 2 rows


def test_empty_and_degenerate_grammars():
    # no rules at all: every C symbol is a terminal
    g = gmr.Grammar.from_arrays(4, [], [], [1, 2, 0, 0, 3, 0], [1.0])
    s = g.schedule(3, 3, "refcompat")
    i = s.info()
    assert i["max_depth"] == 0 and i["num_layered"] == 0 and i["num_term_emit"] == 3
    x = np.array([1.0, 2.0, 4.0], np.float32)
    y = O.sweep(0, s.arrays(), 3, 3, np.array([1.0]), x)
    np.testing.assert_array_equal(y, [3.0, 0.0, 4.0])
    # empty matrix
    m = gmr.Matrix.from_dense(np.zeros((4, 5)))
    g = m.repair()
    assert g.info()["nrules"] == 0 and g.schedule(4, 5).info()["num_term_emit"] == 0


def _refs_and_emits(ga):
    alpha, n = ga["alpha"], len(ga["left"])
    refs, emits = np.zeros(n, np.int64), np.zeros(n, np.int64)
    for side in ("left", "right"):
        s = ga[side] if side == "left" else np.where(ga["kind"] == 0, ga["right"], 0)
        np.add.at(refs, s[s >= alpha] - alpha, 1)
    C = ga["C"]
    np.add.at(emits, C[C >= alpha] - alpha, 1)
    return refs, emits


def test_inline_top_known_answers():
    # 4 -> 1 2 is referenced by 5 -> 4 3, which C uses once: only 5 goes.
    g = gmr.Grammar.from_arrays(4, [1, 4], [2, 3], [5, 0, 4, 3, 0], [1.0])
    g2 = g.inline_top(2)
    a = g2.arrays()
    assert g2.n_inlined == 1
    assert list(a["left"]) == [1] and list(a["right"]) == [2] and list(a["C"]) == [4, 3, 0, 4, 3, 0]
    # max_emits = 0 copies; a rule used three times stays at max_emits = 2
    assert list(g.inline_top(0).arrays()["C"]) == [5, 0, 4, 3, 0]
    g = gmr.Grammar.from_arrays(3, [1], [2], [3, 3, 0, 3, 0], [1.0])
    assert g.inline_top(2).n_inlined == 0 and g.inline_top(3).arrays()["C"].tolist() == [1, 2, 1, 2, 0, 1, 2, 0]
    # malformed input is rejected, not read out of bounds
    with pytest.raises(gmr.GmrError):
        gmr.Grammar.from_arrays(4, [1, 9], [2, 3], [5, 0], [1.0]).inline_top(2)
    with pytest.raises(gmr.GmrError):
        gmr.Grammar.from_arrays(4, [1], [2], [7, 0], [1.0]).inline_top(2)
    # the reference's unit grammar: the top rule 8 -> 7 5 goes, the run-length
    # rule 7 = 6^3 (referenced by 8) stays with its coefficient
    f = fixture("unit_grammar")
    g = gmr.Grammar.from_arrays(f["alpha"], f["left"], f["right"], f["C"], f["V"], kind=f["kind"], rep=f["rep"])
    g2 = g.inline_top(2)
    a = g2.arrays()
    assert g2.n_inlined == 1 and list(a["kind"]) == [0, 0, 1] and list(a["rep"]) == [1, 1, 3]
    assert list(a["C"][:2]) == [7, 5]
    assert np.array_equal(g2.expand(), g.expand())
    y = O.sweep(0, g2.schedule(f["rows"], f["cols"], "refcompat").arrays(), f["rows"], f["cols"],
                np.array(f["V"]), np.array(f["x"], np.float32))
    np.testing.assert_allclose(y, f["expect_y"], atol=1e-12)


@pytest.mark.parametrize("name,m", list(_small_matrices()), ids=[n for n, _ in _small_matrices()])
def test_inline_top_keeps_the_matrix(name, m):
    seq, V = m.arrays()
    g = m.repair()
    ga = g.arrays()
    refs, emits = _refs_and_emits(ga)
    for e in (1, 2, 3):
        g2 = g.inline_top(e)
        a = g2.arrays()
        want = int(((refs == 0) & (emits >= 1) & (emits <= e) & (ga["kind"] == 0)).sum())
        assert g2.n_inlined == want and len(a["left"]) == len(ga["left"]) - want
        assert np.array_equal(g2.expand(), seq)
        idx = np.arange(len(a["left"])) + a["alpha"]
        assert np.all(a["left"] < idx) and np.all(a["right"] < idx)
        size = lambda x: 2 * len(x["left"]) + len(x["C"])
        if e <= 2:
            assert size(a) <= size(ga)
        for sr in (0, 1, 2):
            x = x_for(sr, m.cols, 11 + e)
            y_bf, den = O.brute_force(sr, m.rows, m.cols, seq, V, x)
            y = O.sweep(sr, g2.schedule(m.rows, m.cols, "refcompat").arrays(), m.rows, m.cols, V, x)
            if sr == 0:
                np.testing.assert_allclose(y, y_bf, rtol=1e-9, atol=1e-9 * (1 + den.max()))
            else:
                np.testing.assert_array_equal(y, y_bf)


def test_csr_arrays_from_the_csrv_string():
    """The vendor baseline's CSR conversion reproduces the dense matrix,
    including empty rows at the end and in the middle."""
    from gmr import vendor
    rng = np.random.default_rng(5)
    M = rng.integers(1, 4, (40, 17)) * (rng.random((40, 17)) < 0.3)
    M[7] = 0
    M[-3:] = 0
    m = gmr.Matrix.from_dense(M.astype(np.float64))
    seq, V = m.arrays()
    rp, col, vals = vendor.csr_arrays(seq, V, 40, 17)
    assert rp.dtype == np.int32 and len(rp) == 41 and rp[0] == 0 and rp[-1] == len(col)
    D = np.zeros((40, 17))
    for r in range(40):
        for i in range(rp[r], rp[r + 1]):
            D[r, col[i]] += vals[i]
    np.testing.assert_array_equal(D, M)


def test_transpose_of_the_csrv_string():
    """Matrix.transpose() is the CSRV string of M^T (the same string
    from_dense builds from M.T, values and empty rows and columns included), and
    transposing twice gives M back."""
    rng = np.random.default_rng(41)
    for shape in ((37, 23), (5, 60), (64, 1)):
        M = (rng.random(shape) < 0.2) * rng.integers(1, 6, shape).astype(np.float64)
        M[0] = 0.0
        M[:, -1] = 0.0
        m = gmr.Matrix.from_dense(M)
        t = m.transpose()
        s_t, V_t = t.arrays()
        s_d, V_d = gmr.Matrix.from_dense(M.T.copy()).arrays()
        assert (t.rows, t.cols) == (shape[1], shape[0])
        np.testing.assert_array_equal(V_t, V_d)
        np.testing.assert_array_equal(s_t, s_d)
        s_tt, V_tt = t.transpose().arrays()
        s_m, V_m = m.arrays()
        np.testing.assert_array_equal(s_tt, s_m)
        np.testing.assert_array_equal(V_tt, V_m)


def test_row_blocks_partition_the_csrv_string():
    rng = np.random.default_rng(12)
    M = rng.integers(0, 3, (23, 9)) * (rng.random((23, 9)) < 0.4)
    M[5] = 0
    M[-1] = 0
    M[0] = 0
    m = gmr.Matrix.from_dense(M.astype(np.float64))
    seq = m.arrays()[0]
    for n in (1, 2, 3, 5, 23):
        blocks = m.row_blocks(n)
        assert [b[0] for b in blocks][0] == 0 and blocks[-1][1] == 23
        assert all(b.rows == r1 - r0 and b.cols == 9 for r0, r1, b in blocks)
        assert np.array_equal(np.concatenate([b.arrays()[0] for _, _, b in blocks]), seq), n
    sub = m.row_block(4, 9)
    np.testing.assert_array_equal(gmr.Matrix.from_dense(M[4:9].astype(np.float64)).arrays()[0], sub.arrays()[0])
    assert m.row_block(7, 7).rows == 0 and len(m.row_block(7, 7).arrays()[0]) == 0
    for r0, r1 in ((-1, 3), (5, 4), (0, 24)):
        with pytest.raises(gmr.GmrError):
            m.row_block(r0, r1)


@pytest.mark.parametrize("sr", [0, 1, 2])
def test_host_cpu_sweep_matches_the_oracle(sr):
    """The host library's CPU layered sweep (what bin/gpu_test checks the GPU
    against in the same run) against the oracle's brute force, on every layout,
    with 1 and 4 threads: Boolean / min-plus exact, (+,x) within FP32."""
    from helpers import rel_err, x_for
    for m, cols in ((gmr.Matrix.haplotype(20261204, 2048, 1024), 1024),
                    (gmr.Matrix.rmat(77, 12, 40000, nweights=100 if sr == 2 else 1), 4096)):
        g = m.repair()
        seq, V = m.arrays()
        rows = m.rows
        x = x_for(sr, cols, 900 + sr)
        y_star, den = O.brute_force(sr, rows, cols, seq, V, x)
        for lay in ("refcompat", "wave64", "resident"):
            s = g.schedule(rows, cols, lay)
            for threads in (1, 4):
                y = s.cpu_sweep(sr, x, threads=threads)
                if sr == 0:
                    assert rel_err(y, y_star, den) <= 1e-5, (lay, threads)
                else:
                    assert np.array_equal(y.astype(np.int64), y_star.astype(np.int64)), (lay, threads)
