"""GPU fixpoint traversal x <- x (+) x^T M (Boolean reachability, Tropical
shortest distances) against scipy's graph routines and against the oracle's
left product iterated on the CPU."""
import sys

import numpy as np
import pytest

import gmr
from gmr import _lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
csgraph = pytest.importorskip("scipy.sparse.csgraph")
sparse = pytest.importorskip("scipy.sparse")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.dev()


def _csr(m):
    seq, V = m.arrays()
    n = m.rows
    rows, cols, w = [], [], []
    r = 0
    for t in seq:
        if t == 0:
            r += 1
            continue
        rows.append(r)
        cols.append((t - 1) % n)
        w.append(V[(t - 1) // n])
    return sparse.csr_matrix((w, (rows, cols)), shape=(n, n)), seq, V


def _fixpoint_oracle(sr, n, seq, V, x):
    x = x.astype(np.float32).copy()
    for it in range(1, 10 * n):
        y, _ = O.brute_force_left(sr, n, n, seq, V, x)
        new = np.maximum(x, y) if sr == 1 else np.minimum(x, y)
        if np.array_equal(new, x):
            return x, it
        x = new.astype(np.float32)
    raise AssertionError("no fixpoint")


@pytest.mark.parametrize("variant", ["levels", "graph"])
def test_reachability_and_shortest_paths(variant):
    for nw, sr in ((1, 1), (100, 2)):
        m = gmr.Matrix.rmat(300 + nw, 12, 30000, nweights=nw)
        A, seq, V = _csr(m)
        n = m.rows
        g = m.repair()
        eng = g.schedule(n, n, "resident").engine()
        src = np.random.default_rng(nw).choice(n, 3, replace=False)
        if sr == 1:
            x0 = np.zeros(n, np.float32)
            x0[src] = 1
            reach = np.zeros(n, bool)
            for s in src:
                reach[csgraph.breadth_first_order(A, s, directed=True, return_predecessors=False)] = True
            expect = reach.astype(np.float32)
        else:
            x0 = np.full(n, 1e9, np.float32)
            x0[src] = 0
            d = csgraph.dijkstra(A, directed=True, indices=src, min_only=True)
            expect = np.where(np.isinf(d), 1e9, d).astype(np.float32)
        x = torch.from_numpy(x0).cuda()
        iters = eng.traverse_torch(sr, x, variant=gmr.VARIANTS[variant])
        got = x.cpu().numpy()
        print(f"sr={sr} {variant}: {iters} products, {int((got != x0).sum())} entries moved", flush=True)
        np.testing.assert_array_equal(got, expect)
        want, want_it = _fixpoint_oracle(sr, n, seq, V, x0)
        np.testing.assert_array_equal(got, want)
        assert iters == want_it


@pytest.mark.parametrize("variant", ["levels", "graph"])
def test_traversal_by_right_products_of_the_transpose(variant):
    """GMR_RUN_TRANSPOSED: an engine of M^T traverses with right products (the
    bit mode and the LDS-range gather among them) to the same reachable set and
    distances as the left-product traversal of M, scipy and the oracle, in the
    same number of steps (each step after the first multiplies only the
    vertices the step before moved)."""
    for nw, sr in ((1, 1), (100, 2)):
        m = gmr.Matrix.rmat(310 + nw, 12, 30000, nweights=nw)
        A, seq, V = _csr(m)
        n = m.rows
        eng_t = m.transpose().repair().schedule(n, n, "resident").engine()
        src = np.random.default_rng(nw + 7).choice(n, 3, replace=False)
        if sr == 1:
            x0 = np.zeros(n, np.float32)
            x0[src] = 1
            opts = [dict(emit=gmr.EMIT_GATHER), dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_BITS),
                    dict(emit=gmr.EMIT_SPOT)]
        else:
            # +infinity given as INT32_MAX/2, above the 1e9 cap (the frontier's
            # zero is 1e9 itself)
            x0 = np.full(n, np.float32(2 ** 31 - 1) // 2, np.float32)
            x0[src] = 0
            opts = [dict(emit=gmr.EMIT_GATHER),
                    dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES | gmr.RUN_FUSED_TAIL, lds_values=4096),
                    dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES | gmr.RUN_ROW_ORDER, lds_values=16384,
                         chunks=2),
                    dict(emit=gmr.EMIT_SPOT)]
        # every value at or above the 1e9 cap is +infinity: the engine leaves an
        # unreached vertex's INT32_MAX/2 as given, the oracle's min writes 1e9
        want, want_it = _fixpoint_oracle(sr, n, seq, V, np.minimum(x0, np.float32(1e9)))
        for kw in opts:
            x = torch.from_numpy(x0).cuda()
            kw = dict(kw, flags=kw.get("flags", 0) | gmr.RUN_TRANSPOSED)
            iters = eng_t.traverse_torch(sr, x, variant=gmr.VARIANTS[variant], **kw)
            got = np.minimum(x.cpu().numpy(), np.float32(1e9))
            print(f"transposed sr={sr} {variant} {kw}: {iters} products", flush=True)
            np.testing.assert_array_equal(got, want)
            assert iters == want_it


def test_transposed_flag_only_for_traversal():
    m = gmr.Matrix.rmat(9, 6, 300)
    eng = m.repair().schedule(m.rows, m.rows, "resident").engine()
    x = torch.zeros(m.rows, device="cuda")
    with pytest.raises(gmr.GmrError) as e:
        eng.run_torch(1, x, flags=gmr.RUN_TRANSPOSED)
    assert e.value.code == gmr.ERR_INVALID
    with pytest.raises(gmr.GmrError) as e:
        eng.run_left_torch(1, x, flags=gmr.RUN_TRANSPOSED)
    assert e.value.code == gmr.ERR_INVALID
    torch.cuda.synchronize()


def test_traversal_errors():
    m = gmr.Matrix.haplotype(5, 64, 32, founders=4, seg_mean=8.0)
    eng = m.repair().schedule(64, 32, "resident").engine()
    with pytest.raises(gmr.GmrError):
        eng.traverse_torch(1, torch.zeros(64, device="cuda"))  # not square
    sq = gmr.Matrix.rmat(9, 6, 300)
    e2 = sq.repair().schedule(sq.rows, sq.rows, "resident").engine()
    with pytest.raises(gmr.GmrError):
        e2.traverse_torch(0, torch.zeros(sq.rows, device="cuda"))  # (+,x) has no fixpoint here
