"""gmr_engine_run_noalloc (the entry point that allocates nothing; its
workspace is the library's static device scratch array) against
gmr_engine_run on the same inputs, through the C ABI: Boolean and Tropical
bit-exact, (+,x) max_i |y_i - y*_i| / sum_j |M_ij x_j| <= 1e-5 against
gmr_engine_run's output. Two calls in a row, the second with a smaller input,
so stale scratch contents from the first cannot leak into the second."""
import sys

import numpy as np
import pytest

import gmr
from gmr import _lib
from helpers import rel_err, x_for
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5
SCRATCH = 64 * 1024 * 1024  # GMR_NOALLOC_SCRATCH_BYTES

torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.dev()


def same(sr, y, y_ref, den, what):
    if sr == 0:
        e = rel_err(y, y_ref, den)
        assert e <= TOL, f"{what}: FP32 difference {e} > {TOL}"
    else:
        diff = int((np.asarray(y, np.int64) != np.asarray(y_ref, np.int64)).sum())
        assert diff == 0, f"{what}: {diff} entries differ"


def pair(eng, sr, X, **kw):
    """(run, run_noalloc) outputs for the same X."""
    Xt = torch.from_numpy(np.ascontiguousarray(X, np.float32)).cuda()
    dt = torch.float32 if sr == 0 else torch.int32
    rows = eng.info()["rows"]
    shape = (rows,) if Xt.dim() == 1 else (rows, Xt.shape[1])
    Y_ref = eng.run_torch(sr, Xt, **kw)
    Y = torch.full(shape, -7, dtype=dt, device="cuda")
    eng.run_noalloc_torch(sr, Xt, Y, **kw)
    torch.cuda.synchronize()
    if kw.get("variant") == gmr.VARIANT_PERSISTENT:
        eng.check()
    return Y.cpu().numpy(), Y_ref.cpu().numpy()


@pytest.fixture(scope="module")
def hap():
    m = gmr.Matrix.haplotype(20261204, 4096, 2048)
    g = m.repair()
    seq, V = m.arrays()
    return g, seq, V


@pytest.mark.parametrize("sr", [0, 1, 2])
def test_noalloc_matches_run_two_calls_second_smaller(hap, sr):
    g, seq, V = hap
    for lay in ("resident", "wave64"):
        eng = g.schedule(4096, 2048, lay).engine()
        emits = ("spot", "gather") if lay == "resident" else ("spot",)
        for var in ("levels", "persistent"):
            for em in emits:
                kw = dict(variant=gmr.VARIANTS[var], emit=gmr.EMITS[em])
                # first call: 64 vectors; second: 3 vectors, then one vector
                for B, seed in ((64, 11), (3, 12), (1, 13)):
                    X = x_for(sr, 2048, seed + 10 * sr, B=None if B == 1 else B)
                    X2 = X if X.ndim == 2 else X[:, None]
                    _, den = O.brute_force_batched(sr, 4096, 2048, seq, V, X2)
                    den = den if X.ndim == 2 else den[:, 0]
                    y, y_ref = pair(eng, sr, X, **kw)
                    _say("noalloc", sr, lay, var, em, B)
                    same(sr, y, y_ref, den, f"{lay}/{var}/{em} B={B}")


def test_noalloc_one_vector_gathers_after_setup(hap):
    """The one-vector gathers' lists are built by gmr_engine_run; until then a
    no-allocation run that needs them is refused, afterwards it matches."""
    g, seq, V = hap
    eng = g.schedule(4096, 2048, "resident").engine()
    x = x_for(0, 2048, 21)
    y_star, den = O.brute_force(0, 4096, 2048, seq, V, x)
    xt = torch.from_numpy(x).cuda()
    Y = torch.empty(4096, device="cuda")
    for kw in (dict(emit=gmr.EMIT_GATHER),  # XCD-partitioned lists
               dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES, lds_values=16384),
               dict(emit=gmr.EMIT_GATHER, flags=gmr.RUN_LDS_RANGES | gmr.RUN_ROW_ORDER, lds_values=16384)):
        with pytest.raises(gmr.GmrError) as e:
            eng.run_noalloc_torch(0, xt, Y, **kw)
        assert e.value.code == gmr.ERR_INVALID and "allocated" in str(e.value)
        y, y_ref = pair(eng, 0, x, **kw)  # run builds them; run_noalloc then reuses them
        same(0, y, y_ref, den, f"one vector {kw}")
        same(0, y, y_star, den, f"one vector {kw} vs oracle")
    # lists built at upload: usable at once
    for fl in (gmr.RUN_SINGLE_LIST, gmr.RUN_ROW_SEGMENTS, gmr.RUN_SINGLE_LIST | gmr.RUN_FUSED_TAIL):
        y, y_ref = pair(eng, 0, x, emit=gmr.EMIT_GATHER, flags=fl)
        same(0, y, y_star, den, f"one vector flags={fl}")


def test_noalloc_allocates_nothing(hap):
    g, seq, V = hap
    eng = g.schedule(4096, 2048, "resident").engine()
    X = torch.from_numpy(x_for(0, 2048, 31, B=16)).cuda()
    Y = torch.empty(4096, 16, device="cuda")
    eng.run_noalloc_torch(0, X, Y, emit=gmr.EMIT_GATHER)  # the scratch array's module is loaded by now
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(5):
        eng.run_noalloc_torch(0, X, Y, emit=gmr.EMIT_GATHER)
        eng.run_noalloc_torch(0, X, Y, variant=gmr.VARIANT_PERSISTENT, emit=gmr.EMIT_GATHER)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0


def test_noalloc_refusals(hap):
    g, seq, V = hap
    eng = g.schedule(4096, 2048, "resident").engine()
    info = eng.info()
    X = torch.zeros(2048, 64, device="cuda")
    Y = torch.empty(4096, 64, device="cuda")
    for kw in (dict(variant=gmr.VARIANT_GRAPH),
               dict(level_kind=[0] * (info["max_depth"] + 1))):
        with pytest.raises(gmr.GmrError) as e:
            eng.run_noalloc_torch(0, X, Y, **kw)
        assert e.value.code == gmr.ERR_INVALID
    # a workspace beyond the scratch array: refused with the sizes named
    big = gmr.Matrix.genotype(20261206, 6000, 10000).repair().schedule(6000, 10000, "resident").engine()
    bi = big.info()
    B = 256
    need = (bi["U"] + bi["num_layered"]) * B * 4
    assert need > SCRATCH
    Xb = torch.from_numpy(x_for(0, 10000, 41, B=B)).cuda()
    Yb = torch.empty(6000, B, device="cuda")
    with pytest.raises(gmr.GmrError) as e:
        big.run_noalloc_torch(0, Xb, Yb, emit=gmr.EMIT_GATHER)
    assert e.value.code == gmr.ERR_INVALID and "scratch" in str(e.value)
    # the same product in narrower tiles fits
    Yr = big.run_torch(0, Xb, emit=gmr.EMIT_GATHER, b_tile=16)
    big.run_noalloc_torch(0, Xb, Yb, emit=gmr.EMIT_GATHER, b_tile=16)
    torch.cuda.synchronize()
    scale = float(Yr.abs().max())
    assert float((Yr - Yb).abs().max()) <= 1e-5 * max(scale, 1.0)
