"""GPU parity at the configurations' stated sizes, against the oracle (FP64
brute force from the uncompressed matrix, computed on the host cores):

  * 500,000 x 10,000 genotype matrix (nnz ~5.1e8), as built and with its top
    rules inlined (thresholds 2 and 8), at B = 1, 4, 16, 64, 256 (the batched
    products checked on four sampled columns);
  * R-MAT graph with 2^23 vertices from 2^27 edge draws: Boolean one-step
    reachability from a 1% frontier (int32 values and bit mode) and min/plus
    relaxation with weights 1..100 from 1,024 sources (INT32_MAX/2 elsewhere);
  * 32 grammars (genotype matrices and R-MAT graphs, nnz log-uniform in
    [1e5, 1e8], B ~ U{1..128} each) in one fused launch with the shipped
    proposer's choices, persistent and graph variants.

Bar (reference gpu_engine_test.cu:251-274, tightened): Boolean and min/plus
bit-exact; (+,x) max_i |y_i - y*_i| / sum_j |M_ij x_j| <= 1e-5. This file is
named so that it is collected after the smaller GPU tests."""
import sys
import time

import numpy as np
import pytest

import gmr
from gmr import _lib
from helpers import rel_err
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


def _oracle_cols(sr, rows, cols, seq, V, X, pick):
    """The oracle's Y on columns `pick` of X: (Y*, den)."""
    return O.brute_force_batched(sr, rows, cols, seq, V, np.ascontiguousarray(X[:, pick], np.float32))


def _check_cols(sr, Y, ref, pick, what):
    """Y (rows, B) on the device against the oracle's columns `pick`."""
    Ys, den = ref
    Yg = Y[:, pick].cpu().numpy()
    if sr == 0:
        e = rel_err(Yg, Ys, den)
        _say(f"  {what}: err {e:.3g}")
        assert e <= TOL, f"{what}: FP32 error {e} > {TOL}"
    else:
        diff = int((Yg.astype(np.int64) != Ys.astype(np.int64)).sum())
        assert diff == 0, f"{what}: {diff} values differ"


@pytest.fixture(scope="module")
def cfg2():
    t = time.time()
    m = gmr.Matrix.genotype(20261205, 500000, 10000)
    g = m.repair()
    seq, V = m.arrays()
    forms = [("as built", g.schedule(500000, 10000, "resident").engine())]
    forms.append(("inlined", g.inline_top(2).schedule(500000, 10000, "resident").engine()))
    forms.append(("inlined <= 8", g.inline_top(8).schedule(500000, 10000, "resident").engine()))
    _say(f"config 2: nnz {m.nnz()}, {g.info()['nrules']} rules, built in {time.time() - t:.1f}s")
    return seq, V, forms


def test_config2_single_vector(cfg2):
    seq, V, forms = cfg2
    x = np.random.default_rng(20261208).uniform(-1, 1, 10000).astype(np.float32)
    y_star, den = O.brute_force(0, 500000, 10000, seq, V, x)
    xt = torch.from_numpy(x).cuda()
    plans = [dict(emit=gmr.EMIT_GATHER),                                            # XCD-partitioned
             dict(emit=gmr.EMIT_GATHER, xcd_passes=4, chunks=4),
             dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES),                  # LDS ranges
             dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES, lds_values=16384, chunks=2),
             dict(variant=gmr.VARIANT_GRAPH, emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES, lds_values=40960,
                  chunks=4),
             dict(variant=gmr.VARIANT_GRAPH, emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES | gmr.RUN_ROW_ORDER,
                  lds_values=40960, chunks=4),
             dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES | gmr.RUN_ROW_ORDER | gmr.RUN_FUSED_TAIL,
                  lds_values=40960, chunks=8),
             dict(emit=gmr.EMIT_GATHER, lds_values=40960, chunks=4,
                  flags=gmr.RUN_LDS_RANGES | gmr.RUN_ROW_ORDER | gmr.RUN_FUSED_TAIL | gmr.RUN_PAIRED_LEVELS),
             dict(emit=gmr.EMIT_SPOT, flags=gmr.RUN_FUSED_TAIL),
             dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_SINGLE_LIST),
             dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_ROW_SEGMENTS),
             dict(emit=gmr.EMIT_SPOT, eval_kind=gmr.EVAL_LDS),
             dict(variant=gmr.VARIANT_GRAPH, emit=gmr.EMIT_GATHER),
             dict(variant=gmr.VARIANT_PERSISTENT, emit=gmr.EMIT_GATHER)]
    for label, eng in forms:
        for kw in plans:
            y = eng.run_torch(0, xt, **kw)
            torch.cuda.synchronize()
            if kw.get("variant") == gmr.VARIANT_PERSISTENT:
                eng.check()
            e = rel_err(y.cpu().numpy(), y_star, den)
            _say(f"  config 2 B=1 {label} {kw}: err {e:.3g}")
            assert e <= TOL, f"{label} {kw}: FP32 error {e}"


@pytest.mark.parametrize("B", [4, 16, 64, 256])
def test_config2_batched(cfg2, B):
    seq, V, forms = cfg2
    X = np.random.default_rng(20261209 + B).standard_normal((10000, B)).astype(np.float32)
    Xt = torch.from_numpy(X).cuda()
    pick = sorted({0, B // 3, (2 * B) // 3, B - 1})
    ref = _oracle_cols(0, 500000, 10000, seq, V, X, pick)
    plans = [dict(emit=gmr.EMIT_GATHER), dict(variant=gmr.VARIANT_GRAPH, emit=gmr.EMIT_GATHER, b_tile=min(B, 32))]
    if B <= 64:
        plans.append(dict(variant=gmr.VARIANT_PERSISTENT, emit=gmr.EMIT_GATHER))
    else:
        plans.append(dict(emit=gmr.EMIT_GATHER, b_tile=128))
    if B == 4:
        plans.append(dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_ROW_SEGMENTS))
    if B == 16:
        plans.append(dict(emit=gmr.EMIT_SPOT))
    for label, eng in forms:
        for kw in plans:
            Y = eng.run_torch(0, Xt, **kw)
            torch.cuda.synchronize()
            if kw.get("variant") == gmr.VARIANT_PERSISTENT:
                eng.check()
            _check_cols(0, Y, ref, pick, f"config 2 B={B} {label} {kw}")
            del Y


def _rmat(nweights):
    t = time.time()
    m = gmr.Matrix.rmat(20261206 + nweights, 23, 1 << 27, nweights=nweights)
    g = m.repair()
    seq, V = m.arrays()
    n = m.rows
    forms = [("as built", g.schedule(n, n, "resident").engine()),
             ("inlined", g.inline_top(2).schedule(n, n, "resident").engine())]
    _say(f"R-MAT 2^23 (weights {nweights}): nnz {m.nnz()}, {g.info()['nrules']} rules, {time.time() - t:.1f}s")
    return n, seq, V, forms


def test_config4a_boolean_reachability():
    n, seq, V, forms = _rmat(1)
    assert n == 1 << 23
    x = np.zeros(n, np.float32)
    x[np.random.default_rng(2027).choice(n, n // 100, replace=False)] = 1.0
    y_star, _ = O.brute_force(1, n, n, seq, V, x)
    xt = torch.from_numpy(x).cuda()
    for label, eng in forms:
        for kw in (dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_BITS),
                   dict(variant=gmr.VARIANT_GRAPH, emit=gmr.EMIT_GATHER, flags=gmr.RUN_BITS, lds_values=262144),
                   dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_BITS | gmr.RUN_PAIRED_LEVELS, lds_values=262144, chunks=8),
                   dict(emit=gmr.EMIT_GATHER),
                   dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES),
                   dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES | gmr.RUN_ROW_ORDER),
                   dict(emit=gmr.EMIT_SPOT),
                   dict(variant=gmr.VARIANT_PERSISTENT, emit=gmr.EMIT_GATHER)):
            y = eng.run_torch(1, xt, **kw)
            torch.cuda.synchronize()
            if kw.get("variant") == gmr.VARIANT_PERSISTENT:
                eng.check()
            diff = int((y.cpu().numpy().astype(np.int64) != y_star.astype(np.int64)).sum())
            _say(f"  config 4a {label} {kw}: {diff} differ, {int(y_star.sum())} reached")
            assert diff == 0


def test_config4b_minplus_relaxation():
    n, seq, V, forms = _rmat(100)
    x = np.full(n, np.float32(2 ** 31 - 1) // 2, np.float32)  # INT32_MAX/2: +infinity under the 1e9 cap
    x[np.random.default_rng(2126).choice(n, 1024, replace=False)] = 0.0
    y_star, _ = O.brute_force(2, n, n, seq, V, x)
    xt = torch.from_numpy(x).cuda()
    for label, eng in forms:
        for kw in (dict(emit=gmr.EMIT_GATHER),
                   dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES),
                   dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES | gmr.RUN_ROW_ORDER | gmr.RUN_FUSED_TAIL),
                   dict(emit=gmr.EMIT_GATHER, lds_values=32768,
                        flags=gmr.RUN_LDS_RANGES | gmr.RUN_FUSED_TAIL | gmr.RUN_PAIRED_LEVELS),
                   dict(emit=gmr.EMIT_GATHER, lds_values=16384, chunks=2, flags=gmr.RUN_LDS_RANGES),
                   dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_SINGLE_LIST),
                   dict(emit=gmr.EMIT_SPOT),
                   dict(variant=gmr.VARIANT_PERSISTENT, emit=gmr.EMIT_SPOT)):
            y = eng.run_torch(2, xt, **kw)
            torch.cuda.synchronize()
            if kw.get("variant") == gmr.VARIANT_PERSISTENT:
                eng.check()
            diff = int((y.cpu().numpy().astype(np.int64) != y_star.astype(np.int64)).sum())
            _say(f"  config 4b {label} {kw}: {diff} differ, {int((y_star < 1e9).sum())} finite")
            assert diff == 0


def test_config5_heterogeneous_batch_one_fused_launch():
    from gmr import proposer as P
    from gmr.train_proposer import heldout_set
    t = time.time()
    prop = P.Proposer.default()
    mats, engines, Xs, opts = [], [], [], []
    for i, (name, m, B) in enumerate(heldout_set(32, 20261203, 1e5, 1e8)):
        g = m.repair()
        s = g.schedule(m.rows, m.cols, "resident")
        arr = s.arrays()
        choice = prop.choose(arr, B, tiles=P.tile_candidates(B))
        engines.append(s.engine())
        seq, V = m.arrays()
        mats.append((name, m.rows, m.cols, seq, V, B))
        Xs.append(np.random.default_rng(9000 + i).uniform(-1, 1, (m.cols, B)).astype(np.float32))
        opts.append({"emit": P.PLANS[choice[1]][1], "level_kind": np.array([0] + list(choice[2]), np.int32)})
    _say(f"config 5: 32 grammars, {sum(r[1] for r in mats)} rows, built in {time.time() - t:.1f}s")
    batch = gmr.Batch(engines)
    Xt = [torch.from_numpy(X).cuda() for X in Xs]
    picks = [sorted({0, B // 2, B - 1}) for (_, _, _, _, _, B) in mats]
    refs = [_oracle_cols(0, rows, cols, seq, V, X, pick)
            for (_, rows, cols, seq, V, _), X, pick in zip(mats, Xs, picks)]
    for v in (gmr.VARIANT_PERSISTENT, gmr.VARIANT_GRAPH):
        Ys = batch.run_torch(0, Xt, variant=v, opts=opts)
        torch.cuda.synchronize()
        if v == gmr.VARIANT_PERSISTENT:
            batch.check()
        for (name, rows, cols, seq, V, B), Y, ref, pick in zip(mats, Ys, refs, picks):
            _check_cols(0, Y, ref, pick, f"config 5 {name} B={B} variant={v}")
