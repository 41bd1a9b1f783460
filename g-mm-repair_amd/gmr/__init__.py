"""Python mirror of the engine's host interface, over its C ABI.

The reference's interface for this path is C++ (gpu-engine/gpu_engine.h,
grammar.h, layered_types.h, build_schedule.h); that C++ API is provided as-is by
the headers in include/ and the two shared libraries. This module is the thin
Python harness over the same C ABI that the tests and the bench drive:

  Matrix    a CSRV matrix (seeded generators, dense input)          gmr_matrix_*
  Grammar   a RePair grammar (builder, .val/.vc.R/.vc.C files)      gmr_grammar_*
  Schedule  the level schedule (RefCompat / Wave64 / Resident)      gmr_schedule_*
  Engine    the device-resident grammar and its HIP sweep           gmr_engine_*

Nothing here computes y on the CPU: Engine.run is the HIP path, and it raises
if the device library is missing.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import GPUSchedule, LayeredRule, RunOpts

PLUS_TIMES, BOOLEAN, TROPICAL = 0, 1, 2
SEMIRINGS = {"plustimes": PLUS_TIMES, "boolean": BOOLEAN, "bool": BOOLEAN,
             "tropical": TROPICAL, "minplus": TROPICAL}
REFCOMPAT, WAVE64, RESIDENT = 0, 1, 2
LAYOUTS = {"refcompat": REFCOMPAT, "wave64": WAVE64, "resident": RESIDENT}

# Kernel variants (gmr.h gmr_variant) and emission strategies (gmr_emit).
VARIANT_LEVELS, VARIANT_GRAPH, VARIANT_PERSISTENT = 0, 1, 2
EMIT_SPOT, EMIT_GATHER = 0, 1
VARIANTS = {"levels": VARIANT_LEVELS, "graph": VARIANT_GRAPH, "persistent": VARIANT_PERSISTENT}
EMITS = {"spot": EMIT_SPOT, "gather": EMIT_GATHER}
# Per-level node-update kernels (gmr.h gmr_eval_kind).
EVAL_GROUP, EVAL_THREAD, EVAL_WAVE, EVAL_LDS = 0, 1, 2, 3
RUN_TIMED, RUN_ROW_SEGMENTS, RUN_SINGLE_LIST, RUN_BITS, RUN_LDS_RANGES, RUN_ROW_ORDER, RUN_FUSED_TAIL = \
    1, 2, 4, 8, 16, 32, 64
RUN_PAIRED_LEVELS, RUN_TRANSPOSED = 128, 256
EVAL_KINDS = {"group": EVAL_GROUP, "thread": EVAL_THREAD, "wave": EVAL_WAVE, "lds": EVAL_LDS}

TROPICAL_INF = 1000000000


# gmr_status codes (gmr_host.h)
ERR_INVALID, ERR_IO, ERR_NOMEM, ERR_DEVICE, ERR_INTERNAL, ERR_BARRIER = 1, 2, 3, 4, 5, 6


class GmrError(RuntimeError):
    """A C-ABI call failed; .code is its gmr_status."""

    def __init__(self, msg, code=0):
        super().__init__(msg)
        self.code = code


def _check(rc, lib=None):
    if rc != 0:
        msg = _lib.host().gmr_last_error()
        raise GmrError(f"gmr error {rc}: {msg.decode() if msg else ''}", rc)


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Matrix:
    """A matrix as its CSRV string (row terminals, then 0)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    @classmethod
    def haplotype(cls, seed, rows, cols, founders=64, density=0.08, seg_mean=128.0, flip=0.01):
        h = C.c_void_p()
        _check(_lib.host().gmr_matrix_haplotype(seed, rows, cols, founders, density, seg_mean, flip, C.byref(h)))
        return cls(h)

    @classmethod
    def genotype(cls, seed, rows, cols, founders=256, classes=16, seg_mean=512.0):
        h = C.c_void_p()
        _check(_lib.host().gmr_matrix_genotype(seed, rows, cols, founders, classes, seg_mean, C.byref(h)))
        return cls(h)

    @classmethod
    def rmat(cls, seed, scale, edges, a=0.57, b=0.19, c=0.19, nweights=1):
        h = C.c_void_p()
        _check(_lib.host().gmr_matrix_rmat(seed, scale, edges, a, b, c, nweights, C.byref(h)))
        return cls(h)

    @classmethod
    def from_dense(cls, M):
        M = np.ascontiguousarray(M, dtype=np.float64)
        h = C.c_void_p()
        _check(_lib.host().gmr_matrix_from_dense(_ptr(M, C.c_double), M.shape[0], M.shape[1], C.byref(h)))
        return cls(h)

    @classmethod
    def from_csrv(cls, rows, cols, seq, V):
        seq = np.ascontiguousarray(seq, dtype=np.int32)
        V = np.ascontiguousarray(V, dtype=np.float64)
        h = C.c_void_p()
        _check(_lib.host().gmr_matrix_from_csrv(rows, cols, _ptr(seq, C.c_int), len(seq), _ptr(V, C.c_double), len(V), C.byref(h)))
        return cls(h)

    def info(self):
        r, c, nv, al = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        ln = C.c_longlong()
        _check(_lib.host().gmr_matrix_info(self._h, C.byref(r), C.byref(c), C.byref(nv), C.byref(al), C.byref(ln)))
        return dict(rows=r.value, cols=c.value, nvals=nv.value, alpha=al.value, len=ln.value)

    @property
    def rows(self):
        return self.info()["rows"]

    @property
    def cols(self):
        return self.info()["cols"]

    def arrays(self):
        """(seq int32, V float64)."""
        i = self.info()
        seq = np.empty(i["len"], dtype=np.int32)
        V = np.empty(i["nvals"], dtype=np.float64)
        _check(_lib.host().gmr_matrix_copy(self._h, _ptr(seq, C.c_int), _ptr(V, C.c_double)))
        return seq, V

    def nnz(self):
        i = self.info()
        return i["len"] - i["rows"]

    def transpose(self):
        """M^T as a Matrix (cols x rows, the same values), built on the host from
        the CSRV string: terminal 1 + v*cols + c of row r becomes 1 + v*rows + r
        of row c, rows' terminals in column order."""
        i = self.info()
        rows, cols = i["rows"], i["cols"]
        seq, V = self.arrays()
        ends = np.flatnonzero(seq == 0)
        lens = np.diff(np.r_[-1, ends]) - 1
        r = np.repeat(np.arange(rows, dtype=np.int64), lens)
        t = seq[seq != 0].astype(np.int64) - 1
        v, c = t // cols, t % cols
        order = np.lexsort((r, c))
        nt = v[order] * rows + r[order] + 1
        if len(nt) and nt.max() >= 2 ** 31:
            raise ValueError("transpose: terminal ids exceed int32")
        counts = np.bincount(c, minlength=cols)
        out = np.empty(len(nt) + cols, np.int32)
        sep = np.cumsum(counts + 1) - 1
        keep = np.ones(len(out), bool)
        keep[sep] = False
        out[keep] = nt
        out[sep] = 0
        return Matrix.from_csrv(cols, rows, out, V)

    def row_block(self, r0, r1):
        """Rows [r0, r1) as a Matrix of their own (same columns and values)."""
        h = C.c_void_p()
        _check(_lib.host().gmr_matrix_row_block(self._h, int(r0), int(r1), C.byref(h)))
        return Matrix(h)

    def row_blocks(self, n):
        """n contiguous row blocks of near-equal height: [(r0, r1, Matrix)]."""
        rows = self.rows
        cuts = [rows * k // n for k in range(n + 1)]
        return [(cuts[k], cuts[k + 1], self.row_block(cuts[k], cuts[k + 1])) for k in range(n)]

    def repair(self, max_rounds=0, threads=0):
        h = C.c_void_p()
        rounds, secs = C.c_int(), C.c_double()
        _check(_lib.host().gmr_grammar_repair(self._h, max_rounds, threads, C.byref(h), C.byref(rounds), C.byref(secs)))
        g = Grammar(h)
        g.build_rounds, g.build_seconds = rounds.value, secs.value
        return g

    def __del__(self):
        lib = getattr(_lib, "_host", None) if _lib is not None else None
        if getattr(self, "_h", None) and lib is not None:
            lib.gmr_matrix_free(self._h)
            self._h = None


class Grammar:
    """A grammar: alpha, rules (kind, left, right, rep), top sequence C, values V."""

    build_rounds = 0
    build_seconds = 0.0

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    @classmethod
    def from_arrays(cls, alpha, left, right, C_seq, V, kind=None, rep=None):
        left = np.ascontiguousarray(left, dtype=np.int32)
        right = np.ascontiguousarray(right, dtype=np.int32)
        n = len(left)
        kind = np.zeros(n, np.int32) if kind is None else np.ascontiguousarray(kind, dtype=np.int32)
        rep = np.ones(n, np.int32) if rep is None else np.ascontiguousarray(rep, dtype=np.int32)
        C_seq = np.ascontiguousarray(C_seq, dtype=np.int32)
        V = np.ascontiguousarray(V, dtype=np.float64)
        h = C.c_void_p()
        _check(_lib.host().gmr_grammar_from_arrays(
            alpha, n, _ptr(kind, C.c_int), _ptr(left, C.c_int), _ptr(right, C.c_int), _ptr(rep, C.c_int),
            _ptr(C_seq, C.c_int), len(C_seq), _ptr(V, C.c_double), len(V), C.byref(h)))
        return cls(h)

    @classmethod
    def read(cls, base):
        h = C.c_void_p()
        _check(_lib.host().gmr_grammar_read(str(base).encode(), C.byref(h)))
        return cls(h)

    def write(self, base):
        _check(_lib.host().gmr_grammar_write(self._h, str(base).encode()))

    def info(self):
        a, n, nv = C.c_int(), C.c_int(), C.c_int()
        cl = C.c_longlong()
        _check(_lib.host().gmr_grammar_info(self._h, C.byref(a), C.byref(n), C.byref(cl), C.byref(nv)))
        return dict(alpha=a.value, nrules=n.value, clen=cl.value, nvals=nv.value)

    def arrays(self):
        i = self.info()
        n = i["nrules"]
        kind, left, right, rep = (np.empty(n, np.int32) for _ in range(4))
        Cs = np.empty(i["clen"], np.int32)
        V = np.empty(i["nvals"], np.float64)
        _check(_lib.host().gmr_grammar_copy(self._h, _ptr(kind, C.c_int), _ptr(left, C.c_int), _ptr(right, C.c_int),
                                            _ptr(rep, C.c_int), _ptr(Cs, C.c_int), _ptr(V, C.c_double)))
        return dict(alpha=i["alpha"], kind=kind, left=left, right=right, rep=rep, C=Cs, V=V)

    def expand(self):
        n = C.c_longlong()
        _check(_lib.host().gmr_grammar_expand(self._h, None, 0, C.byref(n)))
        out = np.empty(n.value, np.int32)
        _check(_lib.host().gmr_grammar_expand(self._h, _ptr(out, C.c_int), n.value, C.byref(n)))
        return out

    def inline_top(self, max_emits=2):
        """The same matrix with every binary rule that no rule references and
        that C uses at most max_emits times written out in C as its two
        children (fewer node updates for the same row-gather reads: faster
        sweeps over many vectors). .n_inlined = rules removed."""
        h = C.c_void_p()
        n = C.c_longlong()
        _check(_lib.host().gmr_grammar_inline_top(self._h, int(max_emits), C.byref(h), C.byref(n)))
        g = Grammar(h)
        g.n_inlined = n.value
        return g

    def schedule(self, rows, cols, layout=REFCOMPAT):
        if isinstance(layout, str):
            layout = LAYOUTS[layout]
        h = C.c_void_p()
        _check(_lib.host().gmr_schedule_build(self._h, rows, cols, layout, C.byref(h)))
        s = Schedule(h)
        s.V = self.arrays()["V"]
        return s

    def __del__(self):
        lib = getattr(_lib, "_host", None) if _lib is not None else None
        if getattr(self, "_h", None) and lib is not None:
            lib.gmr_grammar_free(self._h)
            self._h = None


class Schedule:
    """A built level schedule (owns the BuiltSchedule)."""

    V = None

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    def info(self):
        a = (C.c_longlong * 11)()
        _check(_lib.host().gmr_schedule_info(self._h, a))
        keys = ["max_depth", "max_width", "num_raw_nt", "num_layered", "num_passthrough",
                "emit_total", "num_term_emit", "U", "rows", "cols", "layout"]
        return dict(zip(keys, list(a)))

    def arrays(self):
        i = self.info()
        L = i["max_depth"]
        rules = (LayeredRule * max(1, i["num_layered"]))()
        lo, lw, eo = (np.empty(L + 1, np.int32) for _ in range(3))
        ep, er = np.empty(i["emit_total"], np.int32), np.empty(i["emit_total"], np.int32)
        ts, tr = np.empty(i["num_term_emit"], np.int32), np.empty(i["num_term_emit"], np.int32)
        to = np.empty(i["U"], np.int32)
        _check(_lib.host().gmr_schedule_copy(self._h, rules, _ptr(lo, C.c_int), _ptr(lw, C.c_int), _ptr(eo, C.c_int),
                                             _ptr(ep, C.c_int), _ptr(er, C.c_int), _ptr(ts, C.c_int), _ptr(tr, C.c_int),
                                             _ptr(to, C.c_int)))
        rl = np.frombuffer(rules, dtype=np.dtype([("left", "<i4"), ("right", "<i4"), ("coeff", "<f4")]),
                           count=i["num_layered"]).copy()
        return dict(left=rl["left"].copy(), right=rl["right"].copy(), coeff=rl["coeff"].copy(),
                    level_offsets=lo, level_widths=lw, emit_offsets=eo, emit_pos=ep, emit_row=er,
                    term_sym=ts, term_row=tr, term_orig=to, info=i)

    def cpu_sweep(self, semiring, x, V=None, threads=0):
        """y = M x on the host CPU over this schedule (the host library's own
        OpenMP level sweep; FP32 for (+,x), int32 otherwise). Not the GPU path:
        it is what the driver checks the device result against."""
        if isinstance(semiring, str):
            semiring = SEMIRINGS[semiring]
        V = self.V if V is None else V
        Vf = np.ascontiguousarray(V, dtype=np.float32)
        x = np.ascontiguousarray(x, dtype=np.float32)
        rows = self.info()["rows"]
        y = np.empty(rows, np.float32 if semiring == PLUS_TIMES else np.int32)
        _check(_lib.host().gmr_schedule_cpu_sweep(self._h, semiring, _ptr(Vf, C.c_float), len(Vf),
                                                  _ptr(x, C.c_float), C.c_void_p(y.ctypes.data), threads))
        return y

    def engine(self, V=None):
        """Upload to the GPU (device-resident engine)."""
        V = self.V if V is None else V
        Vf = np.ascontiguousarray(V, dtype=np.float32)
        h = C.c_void_p()
        _check(_lib.dev().gmr_engine_from_schedule(self._h, _ptr(Vf, C.c_float), len(Vf), C.byref(h)))
        return Engine(h)

    def __del__(self):
        lib = getattr(_lib, "_host", None) if _lib is not None else None
        if getattr(self, "_h", None) and lib is not None:
            lib.gmr_schedule_free(self._h)
            self._h = None


class Engine:
    """A device-resident grammar with its HIP sweep."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    def info(self):
        a = (C.c_longlong * 16)()
        _check(_lib.dev().gmr_engine_info(self._h, a))
        keys = ["rows", "cols", "U", "num_layered", "max_depth", "max_width", "emit_total",
                "num_term_emit", "layout", "has_runlen", "device_bytes", "nvals",
                "n_gseg", "r0", "r1", "r2"]
        return dict(zip(keys, list(a)))

    @staticmethod
    def _opts(variant=VARIANT_LEVELS, emit=EMIT_SPOT, b_tile=0, block=0, grid_cap=0, eval_kind=0,
              level_kind=None, flags=0, xcd_passes=0, lds_values=0, chunks=0):
        o = RunOpts(variant, emit, b_tile, block, grid_cap, eval_kind, flags, xcd_passes, None, lds_values, chunks)
        keep = None
        if level_kind is not None:
            keep = np.ascontiguousarray(level_kind, dtype=np.int32)
            o.level_kind = keep.ctypes.data_as(C.POINTER(C.c_int))
        return o, keep

    def run(self, semiring, X_ptr, B, Y_ptr, stream=0, **kw):
        """Y = M X on device pointers: X is cols x B (X[c*B+b]), Y is rows x B.

        semiring: 0 (+,x) float, 1 Boolean int32, 2 Tropical int32 (Y holds int
        bits; X is float as in the reference API). Keywords: variant, emit,
        b_tile, block, grid_cap, eval_kind, level_kind (max_depth+1 ints)."""
        if isinstance(semiring, str):
            semiring = SEMIRINGS[semiring]
        o, keep = self._opts(**kw)
        _check(_lib.dev().gmr_engine_run(self._h, semiring, C.c_void_p(X_ptr), B, C.c_void_p(Y_ptr),
                                         C.byref(o), C.c_void_p(stream)))
        del keep

    def run_noalloc_torch(self, semiring, X, Y, **kw):
        """gmr_engine_run_noalloc on torch device tensors: X float32 (cols, B)
        or (cols,), Y preallocated (rows, B) or (rows,), float32 for (+,x) and
        int32 otherwise. The workspace is the library's static scratch array."""
        import torch
        if isinstance(semiring, str):
            semiring = SEMIRINGS[semiring]
        assert X.is_cuda and X.dtype == torch.float32 and X.is_contiguous() and Y.is_contiguous()
        B = 1 if X.dim() == 1 else X.shape[1]
        o, keep = self._opts(**kw)
        stream = torch.cuda.current_stream(X.device).cuda_stream
        _check(_lib.dev().gmr_engine_run_noalloc(self._h, semiring, C.c_void_p(X.data_ptr()), B,
                                                 C.c_void_p(Y.data_ptr()), C.byref(o), C.c_void_p(stream)))
        del keep
        return Y

    def time_levels(self, semiring, X, Y, reps=3, **kw):
        """Device times (ms): [0] init and terminal emits, [k] level k's node
        update, [L+1] the last emits or the row gather. X, Y torch tensors."""
        import torch
        if isinstance(semiring, str):
            semiring = SEMIRINGS[semiring]
        B = 1 if X.dim() == 1 else X.shape[1]
        L = self.info()["max_depth"]
        out = np.zeros(L + 2, np.float32)
        o, keep = self._opts(**kw)
        stream = torch.cuda.current_stream(X.device).cuda_stream
        _check(_lib.dev().gmr_engine_time_levels(self._h, semiring, C.c_void_p(X.data_ptr()), B,
                                                 C.c_void_p(Y.data_ptr()), C.byref(o), reps,
                                                 out.ctypes.data_as(C.POINTER(C.c_float)), C.c_void_p(stream)))
        del keep
        return out

    def check(self):
        """Raise if a persistent launch's grid barrier timed out since the last
        check (its results are invalid). Synchronises the device."""
        err = C.c_int()
        _check(_lib.dev().gmr_engine_check(self._h, C.byref(err)))
        if err.value:
            raise GmrError("persistent sweep: a grid-barrier wait timed out; results are invalid", ERR_BARRIER)

    def set_barrier_timeout(self, microseconds):
        """Limit of one grid-barrier wait of the persistent variant (0: 2 s)."""
        _check(_lib.dev().gmr_engine_set_barrier_timeout(self._h, int(microseconds)))

    def set_gather_order(self, order):
        """0: row order; 1: segments sharing their highest rule together."""
        _check(_lib.dev().gmr_engine_set_gather_order(self._h, order))

    def debug(self):
        a = (C.c_longlong * 8)()
        _check(_lib.dev().gmr_engine_debug(self._h, a))
        return dict(zip(["arrived", "error", "expected", "wall_khz", "cus", "grid", "per_cu", "graphs"], list(a)))

    def run_torch(self, semiring, X, Y=None, **kw):
        """Convenience for torch device tensors: X float32 (cols, B) or (cols,)."""
        import torch
        if isinstance(semiring, str):
            semiring = SEMIRINGS[semiring]
        assert X.is_cuda and X.dtype == torch.float32 and X.is_contiguous()
        B = 1 if X.dim() == 1 else X.shape[1]
        i = self.info()
        if X.shape[0] != i["cols"]:
            raise ValueError(f"X has {X.shape[0]} rows, the matrix has {i['cols']} columns")
        dt = torch.float32 if semiring == PLUS_TIMES else torch.int32
        if Y is None:
            Y = torch.empty((i["rows"], B) if X.dim() == 2 else (i["rows"],), dtype=dt, device=X.device)
        stream = torch.cuda.current_stream(X.device).cuda_stream
        self.run(semiring, X.data_ptr(), B, Y.data_ptr(), stream=stream, **kw)
        return Y

    def run_left(self, semiring, X_ptr, B, Y_ptr, stream=0, **kw):
        """Y^T = X^T M on device pointers: X rows x B, Y cols x B (Resident
        layout; variant levels or graph)."""
        if isinstance(semiring, str):
            semiring = SEMIRINGS[semiring]
        o, keep = self._opts(**kw)
        _check(_lib.dev().gmr_engine_run_left(self._h, semiring, C.c_void_p(X_ptr), B, C.c_void_p(Y_ptr),
                                              C.byref(o), C.c_void_p(stream)))
        del keep

    def traverse_torch(self, semiring, x, max_iters=1000, **kw):
        """Fixpoint x <- x (+) x^T M in place on a float32 CUDA vector of
        length rows (square matrix); returns the number of products computed.
        With flags including RUN_TRANSPOSED the engine holds M^T and every
        step is its right product (the other keywords are that product's)."""
        import torch
        if isinstance(semiring, str):
            semiring = SEMIRINGS[semiring]
        assert x.is_cuda and x.dtype == torch.float32 and x.is_contiguous() and x.dim() == 1
        o, keep = self._opts(**kw)
        n = C.c_int()
        stream = torch.cuda.current_stream(x.device).cuda_stream
        _check(_lib.dev().gmr_engine_traverse(self._h, semiring, C.c_void_p(x.data_ptr()), max_iters, C.byref(n),
                                              C.byref(o), C.c_void_p(stream)))
        del keep
        return n.value

    def run_left_torch(self, semiring, X, Y=None, **kw):
        """Left product for torch device tensors: X float32 (rows, B) or (rows,)."""
        import torch
        if isinstance(semiring, str):
            semiring = SEMIRINGS[semiring]
        assert X.is_cuda and X.dtype == torch.float32 and X.is_contiguous()
        B = 1 if X.dim() == 1 else X.shape[1]
        i = self.info()
        if X.shape[0] != i["rows"]:
            raise ValueError(f"X has {X.shape[0]} rows, the matrix has {i['rows']} rows")
        dt = torch.float32 if semiring == PLUS_TIMES else torch.int32
        if Y is None:
            Y = torch.empty((i["cols"], B) if X.dim() == 2 else (i["cols"],), dtype=dt, device=X.device)
        stream = torch.cuda.current_stream(X.device).cuda_stream
        self.run_left(semiring, X.data_ptr(), B, Y.data_ptr(), stream=stream, **kw)
        return Y

    def __del__(self):
        lib = getattr(_lib, "_dev", None) if _lib is not None else None
        if getattr(self, "_h", None) and lib is not None:
            lib.gmr_engine_destroy(self._h)
            self._h = None


class Batch:
    """Several engines swept together: Y_g = M_g X_g with B_g vectors each
    (gmr_batch_*). The engines must outlive the batch."""

    def __init__(self, engines):
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e._h.value for e in self.engines])
        h = C.c_void_p()
        _check(_lib.dev().gmr_batch_create(arr, len(self.engines), C.byref(h)))
        self._h = h

    def run_torch(self, semiring, Xs, Ys=None, variant=VARIANT_PERSISTENT, opts=None, block=0, grid_cap=0):
        """Xs: list of float32 CUDA tensors (cols_g, B_g); returns the Ys.
        opts: optional list of per-grammar dicts (emit, eval_kind, level_kind)."""
        import torch
        if isinstance(semiring, str):
            semiring = SEMIRINGS[semiring]
        n = len(self.engines)
        assert len(Xs) == n
        dt = torch.float32 if semiring == PLUS_TIMES else torch.int32
        if Ys is None:
            Ys = [torch.empty((e.info()["rows"], X.shape[1]), dtype=dt, device=X.device)
                  for e, X in zip(self.engines, Xs)]
        Xp = (C.c_void_p * n)(*[X.data_ptr() for X in Xs])
        Yp = (C.c_void_p * n)(*[Y.data_ptr() for Y in Ys])
        Bs = np.ascontiguousarray([X.shape[1] for X in Xs], np.int32)
        keep = []
        os_ = (RunOpts * n)()
        for g in range(n):
            kw = dict(opts[g]) if opts else {}
            if g == 0:
                kw.setdefault("block", block)
                kw.setdefault("grid_cap", grid_cap)
            o, k = Engine._opts(**kw)
            os_[g] = o
            keep.append(k)
        stream = torch.cuda.current_stream(Xs[0].device).cuda_stream
        _check(_lib.dev().gmr_batch_run(self._h, semiring, Xp, Bs.ctypes.data_as(C.POINTER(C.c_int)), Yp, variant,
                                        os_, C.c_void_p(stream)))
        del keep
        return Ys

    def check(self):
        err = C.c_int()
        _check(_lib.dev().gmr_batch_check(self._h, C.byref(err)))
        if err.value:
            raise GmrError("batched persistent sweep: a grid-barrier wait timed out; results are invalid",
                           ERR_BARRIER)

    def set_barrier_timeout(self, microseconds):
        _check(_lib.dev().gmr_batch_set_barrier_timeout(self._h, int(microseconds)))

    def __del__(self):
        lib = getattr(_lib, "_dev", None) if _lib is not None else None
        if getattr(self, "_h", None) and lib is not None:
            lib.gmr_batch_destroy(self._h)
            self._h = None
