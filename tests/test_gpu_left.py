"""GPU parity of the left product Y^T = X^T M (top-down pull sweep over the
Resident layout) against the oracle's brute force from the uncompressed matrix.
Bar: Boolean and Tropical bit-exact; (+,x) max_c |y_c - y*_c| / sum_r |M_rc x_r|
<= 1e-5."""
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
    _lib.dev()


def run_left(eng, sr, X, **kw):
    Xt = torch.from_numpy(np.ascontiguousarray(X, np.float32)).cuda()
    Y = eng.run_left_torch(sr, Xt, **kw)
    torch.cuda.synchronize()
    return Y.cpu().numpy()


def check(sr, y, y_star, den, what):
    assert np.shape(y) == np.shape(y_star), what
    if sr == 0:
        e = rel_err(y, y_star, den)
        _say(f"  {what}: err={e:.3g}")
        assert e <= TOL, f"{what}: FP32 error {e} > {TOL}"
    else:
        diff = np.asarray(y, np.int64) != np.asarray(y_star, np.int64)
        assert not diff.any(), f"{what}: not bit-exact ({int(diff.sum())} entries differ)"


def test_unit_grammar_with_a_run_length_rule():
    """The reference's unit-test grammar (one run-length rule, coefficient 3)."""
    f = fixture("unit_grammar")
    g = gmr.Grammar.from_arrays(f["alpha"], f["left"], f["right"], f["C"], f["V"], kind=f["kind"], rep=f["rep"])
    seq = O.expand(grammar_dict(f))
    rows, cols = 2, 3
    eng = g.schedule(rows, cols, "resident").engine()
    X = np.array([[0.5, -1.0], [2.0, 3.0]], np.float32)
    for XX in (X, X[:, 0].copy()):
        Y_star, den = O.brute_force_left(0, rows, cols, seq, np.array(f["V"], np.float64), XX)
        for var in ("levels", "graph"):
            Y = run_left(eng, 0, XX, variant=gmr.VARIANTS[var])
            check(0, Y, Y_star, den, f"unit {var} B={XX.shape[1] if XX.ndim == 2 else 1}")


@pytest.fixture(scope="module")
def hap():
    m = gmr.Matrix.haplotype(20261204, 4096, 2048)
    g = m.repair()
    seq, V = m.arrays()
    return m, g, seq, V, g.schedule(4096, 2048, "resident").engine()


@pytest.mark.parametrize("sr", [0, 1, 2])
@pytest.mark.parametrize("B", [1, 5, 27, 64])
def test_haplotype_left_product(hap, sr, B):
    m, g, seq, V, eng = hap
    X = x_for(sr, 4096, 40 + B, B=None if B == 1 else B)
    Y_star, den = O.brute_force_left(sr, 4096, 2048, seq, V, X)
    for var in ("levels", "graph", "graph"):
        for tile in (0, 16):
            for flags in ((0, gmr.RUN_SINGLE_LIST) if B == 1 else (0,)):
                _say("left hap", sr, B, var, tile, flags)
                Y = run_left(eng, sr, X, variant=gmr.VARIANTS[var], b_tile=tile, flags=flags)
                check(sr, Y, Y_star, den, f"left sr={sr} B={B} {var} tile={tile} flags={flags}")


def test_rmat_left_semirings():
    """Boolean reachability and min/plus relaxation along edge direction."""
    for nw, sr in ((1, 1), (100, 2)):
        m = gmr.Matrix.rmat(77 + nw, 13, 90000, nweights=nw)
        g = m.repair()
        seq, V = m.arrays()
        n = m.rows
        if sr == 1:
            x = (np.random.default_rng(1).random(n) < 0.02).astype(np.float32)
        else:
            x = np.full(n, np.float32(2 ** 31 - 1) // 2, np.float32)
            x[np.random.default_rng(2).choice(n, 32, replace=False)] = 0
        y_star, _ = O.brute_force_left(sr, n, n, seq, V, x)
        eng = g.schedule(n, n, "resident").engine()
        for var in ("levels", "graph"):
            for flags in (0, gmr.RUN_SINGLE_LIST):
                _say("left rmat", sr, var, flags)
                y = run_left(eng, sr, x, variant=gmr.VARIANTS[var], flags=flags)
                check(sr, y, y_star, None, f"rmat left sr={sr} {var} flags={flags}")


def test_left_equals_right_product_of_the_transpose():
    rng = np.random.default_rng(9)
    M = (rng.random((300, 170)) < 0.1) * rng.integers(1, 4, (300, 170))
    X = x_for(0, 300, 3, B=4)
    gl = gmr.Matrix.from_dense(M).repair()
    gr = gmr.Matrix.from_dense(M.T.copy()).repair()
    yl = run_left(gl.schedule(300, 170, "resident").engine(), 0, X)
    Xt = torch.from_numpy(X).cuda()
    yr = gr.schedule(170, 300, "resident").engine().run_torch(0, Xt, emit=gmr.EMIT_GATHER).cpu().numpy()
    np.testing.assert_allclose(yl, yr, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("sr", [0, 1, 2])
def test_left_fold_with_columns_wider_than_a_chunk(sr):
    """One vector, columns holding up to 150 distinct values (150 terminals):
    the chunked column fold sees columns that start, end and run through
    64-position chunks."""
    rng = np.random.default_rng(17 + sr)
    M = rng.integers(1, 151, (300, 5)) * (rng.random((300, 5)) < 0.9)
    M[:, 3] = 0  # an empty column
    m = gmr.Matrix.from_dense(M.astype(np.float64))
    g = m.repair()
    seq, V = m.arrays()
    eng = g.schedule(300, 5, "resident").engine()
    x = x_for(sr, 300, 23)
    y_star, den = O.brute_force_left(sr, 300, 5, seq, V, x)
    for var in ("levels", "graph"):
        for flags in (0, gmr.RUN_SINGLE_LIST):
            y = run_left(eng, sr, x, variant=gmr.VARIANTS[var], flags=flags)
            check(sr, y, y_star, den if sr == 0 else None, f"wide columns sr={sr} {var} flags={flags}")


def test_left_product_errors():
    m = gmr.Matrix.haplotype(5, 64, 32, founders=4, seg_mean=8.0)
    g = m.repair()
    X = torch.zeros(64, device="cuda")
    with pytest.raises(gmr.GmrError):
        g.schedule(64, 32, "wave64").engine().run_left_torch(0, X)
    eng = g.schedule(64, 32, "resident").engine()
    with pytest.raises(gmr.GmrError):
        eng.run_left_torch(0, X, variant=gmr.VARIANT_PERSISTENT)
    with pytest.raises(ValueError):
        eng.run_left_torch(0, torch.zeros(32, device="cuda"))


def test_left_partitioned_and_single_lists_on_a_larger_grammar():
    """A grammar whose large node groups take the XCD-partitioned lists (and
    the small ones the single list) for one vector, against the oracle."""
    m = gmr.Matrix.genotype(515, 6000, 10000)
    g = m.repair()
    seq, V = m.arrays()
    eng = g.schedule(6000, 10000, "resident").engine()
    for sr in (0, 1, 2):
        x = x_for(sr, 6000, 60 + sr)
        y_star, den = O.brute_force_left(sr, 6000, 10000, seq, V, x)
        for var in ("levels", "graph"):
            for flags in (0, gmr.RUN_SINGLE_LIST):
                _say("left geno", sr, var, flags)
                y = run_left(eng, sr, x, variant=gmr.VARIANTS[var], flags=flags)
                check(sr, y, y_star, den if sr == 0 else None, f"left geno sr={sr} {var} flags={flags}")
