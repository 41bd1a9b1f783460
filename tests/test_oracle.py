"""The oracle against the reference's known answers, the reference's own
scheduler (compiled from its sources into oracle/_ref when present), dense
numpy products and brute force on small inputs."""
import numpy as np
import pytest

from helpers import fixture, grammar_dict, rel_err, x_for
from oracle import oracle as O


def test_unit_grammar_known_answer():
    f = fixture("unit_grammar")
    g = grammar_dict(f)
    s = O.layer_refcompat(g)
    assert s["info"]["max_depth"] == f["expect_max_depth"]
    assert s["info"]["num_passthrough"] == f["expect_num_passthrough"]
    csrv = O.expand(g)
    V = np.array(f["V"])
    x = np.array(f["x"], np.float32)
    y_bf, _ = O.brute_force(0, f["rows"], f["cols"], csrv, V, x)
    y_sw = O.sweep(0, s, f["rows"], f["cols"], V, x)
    np.testing.assert_allclose(y_bf, f["expect_y"], atol=1e-12)
    np.testing.assert_allclose(y_sw, f["expect_y"], atol=1e-12)
    y32 = O.cpu_layered_f32(0, s, f["rows"], f["cols"], V, x)
    np.testing.assert_allclose(y32, f["expect_y"], atol=1e-6)


def test_unit_grammar_matches_reference_scheduler():
    f = fixture("unit_grammar")
    g = grammar_dict(f)
    ref = O.reference_schedule(g, f["rows"], f["cols"])
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    mine = O.layer_refcompat(g)
    for k in ("left", "right", "coeff", "level_offsets", "level_widths", "emit_offsets", "emit_pos",
              "emit_row", "term_sym", "term_row", "term_orig"):
        np.testing.assert_array_equal(mine[k], ref[k], err_msg=k)


def test_readme_example_brute_force():
    f = fixture("readme_8x7")
    M = np.array(f["M"])
    vals = np.unique(M[M != 0])
    rows, cols = M.shape
    seq = []
    for r in range(rows):
        for c in range(cols):
            if M[r, c] != 0:
                seq.append(1 + int(np.searchsorted(vals, M[r, c])) * cols + c)
        seq.append(0)
    y, den = O.brute_force(0, rows, cols, np.array(seq), vals, np.array(f["x"], np.float32))
    np.testing.assert_allclose(y, f["expect_y"], atol=1e-9)
    np.testing.assert_allclose(den, np.abs(M) @ np.abs(np.array(f["x"], float)), atol=1e-9)


def _random_csrv(rng, rows, cols, nvals, density):
    M = np.where(rng.random((rows, cols)) < density, rng.integers(1, nvals + 1, (rows, cols)), 0).astype(float)
    vals = np.unique(M[M != 0])
    seq = []
    for r in range(rows):
        for c in np.nonzero(M[r])[0]:
            seq.append(1 + int(np.searchsorted(vals, M[r, c])) * cols + c)
        seq.append(0)
    return M, vals, np.array(seq, np.int32)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_brute_force_against_dense_products(seed):
    rng = np.random.default_rng(seed)
    M, vals, seq = _random_csrv(rng, 37, 23, 5, 0.3)
    rows, cols = M.shape
    x = x_for(0, cols, seed)
    y, den = O.brute_force(0, rows, cols, seq, vals, x)
    np.testing.assert_allclose(y, M @ x.astype(float), rtol=1e-12, atol=1e-12)
    # Boolean: OR/AND; Tropical: min over stored entries of (int)V + (int)x.
    xb = x_for(1, cols, seed)
    yb, _ = O.brute_force(1, rows, cols, seq, vals, xb)
    np.testing.assert_array_equal(yb, ((M != 0) & (xb[None, :] != 0)).any(1).astype(float))
    xt = x_for(2, cols, seed)
    yt, _ = O.brute_force(2, rows, cols, seq, vals, xt)
    cand = np.where(M != 0, np.where(xt[None, :] >= 1e9, 1e9, M + xt[None, :]), 1e9)
    np.testing.assert_array_equal(yt, cand.min(1))
    # batched form: columns are independent products
    X = x_for(0, cols, seed + 10, B=5)
    Y, _ = O.brute_force_batched(0, rows, cols, seq, vals, X)
    np.testing.assert_allclose(Y, M @ X.astype(float), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_left_brute_force_against_dense_products(seed):
    """x^T M from the CSRV string equals numpy's dense product, and equals the
    right product of the transposed matrix (a dropped or swapped index fails
    one of them)."""
    rng = np.random.default_rng(100 + seed)
    M, vals, seq = _random_csrv(rng, 29, 41, 4, 0.25)
    rows, cols = M.shape
    X = x_for(0, rows, seed, B=3)
    Y, den = O.brute_force_left(0, rows, cols, seq, vals, X)
    np.testing.assert_allclose(Y, M.T @ X.astype(float), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(den, np.abs(M.T) @ np.abs(X.astype(float)), rtol=1e-12)
    # the same through the right product of M^T
    MT, valsT, seqT = M.T.copy(), vals, []
    for r in range(cols):
        for c in np.nonzero(MT[r])[0]:
            seqT.append(1 + int(np.searchsorted(valsT, MT[r, c])) * rows + c)
        seqT.append(0)
    YT, _ = O.brute_force_batched(0, cols, rows, np.array(seqT, np.int32), valsT, X)
    np.testing.assert_allclose(Y, YT, rtol=1e-12, atol=1e-12)
    xb = x_for(1, rows, seed)
    yb, _ = O.brute_force_left(1, rows, cols, seq, vals, xb)
    np.testing.assert_array_equal(yb, ((M != 0) & (xb[:, None] != 0)).any(0).astype(float))
    xt = x_for(2, rows, seed)
    yt, _ = O.brute_force_left(2, rows, cols, seq, vals, xt)
    cand = np.where(M != 0, np.where(xt[:, None] >= 1e9, 1e9, M + xt[:, None]), 1e9)
    np.testing.assert_array_equal(yt, cand.min(0))


def test_rel_err_flags_a_planted_mistake():
    rng = np.random.default_rng(3)
    M, vals, seq = _random_csrv(rng, 20, 10, 3, 0.5)
    x = x_for(0, 10, 3)
    y, den = O.brute_force(0, 20, 10, seq, vals, x)
    assert rel_err(y, y, den) == 0.0
    bad = y.copy()
    bad[7] += 1e-3 * max(den[7], 1.0)
    assert rel_err(bad, y, den) > 1e-5
    # A row with zero denominator must match exactly.
    den0 = den.copy()
    den0[3] = 0.0
    bad2 = y.copy()
    bad2[3] = y[3] + 1e-6
    assert rel_err(bad2, y, den0) == np.inf


def test_sweep_detects_a_planted_schedule_error():
    """Swap one rule's children for a wrong slot: the sweep must disagree with
    brute force (guards against a sweep that ignores its schedule)."""
    f = fixture("unit_grammar")
    g = grammar_dict(f)
    s = O.layer_refcompat(g)
    V = np.array(f["V"])
    x = np.array(f["x"], np.float32)
    good = O.sweep(0, s, 2, 3, V, x)
    s2 = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in s.items()}
    s2["coeff"][np.nonzero(s2["coeff"] == 3.0)[0][0]] = 2.0   # run length 3 -> 2
    bad = O.sweep(0, s2, 2, 3, V, x)
    assert not np.allclose(good, bad)
