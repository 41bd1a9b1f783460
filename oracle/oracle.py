"""ctypes loader of the oracle (oracle/liboracle.so) and of the reference's own
build_schedule (oracle/_ref/libref_schedule.so). TEST INFRASTRUCTURE ONLY:
imported by tests/, the smoke check and the bench's CPU-baseline leg.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_orc = None
_ref = None


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def lib():
    global _orc
    if _orc is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} is not built (make -C {HERE})")
        _orc = C.CDLL(path)
    return _orc


def ref_lib():
    """The reference's build_schedule, or None where it was not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libref_schedule.so")
        if not os.path.exists(path):
            return None
        _ref = C.CDLL(path)
        _ref.ref_build.restype = C.c_void_p
        _ref.ref_info.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
        _ref.ref_copy.argtypes = [C.c_void_p] + [C.c_void_p] * 11
        _ref.ref_free.argtypes = [C.c_void_p]
    return _ref


def brute_force(semiring, rows, cols, seq, V, x):
    """(y float64[rows], den float64[rows]) straight from the CSRV string."""
    seq = np.ascontiguousarray(seq, np.int32)
    V = np.ascontiguousarray(V, np.float64)
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty(rows, np.float64)
    den = np.empty(rows, np.float64)
    rc = lib().orc_brute_force(C.c_int(semiring), C.c_int(rows), C.c_int(cols), _p(seq, C.c_int),
                               C.c_longlong(len(seq)), _p(V, C.c_double), _p(x, C.c_float),
                               _p(y, C.c_double), _p(den, C.c_double))
    if rc:
        raise ValueError("CSRV has more rows than `rows`")
    return y, den


def brute_force_batched(semiring, rows, cols, seq, V, X):
    """X float32 (cols, B) -> (Y float64 (rows, B), den float64 (rows, B))."""
    seq = np.ascontiguousarray(seq, np.int32)
    V = np.ascontiguousarray(V, np.float64)
    X = np.ascontiguousarray(X, np.float32)
    B = X.shape[1]
    Y = np.empty((rows, B), np.float64)
    den = np.empty((rows, B), np.float64)
    rc = lib().orc_brute_force_batched(C.c_int(semiring), C.c_int(rows), C.c_int(cols), _p(seq, C.c_int),
                                       C.c_longlong(len(seq)), _p(V, C.c_double), _p(X, C.c_float), C.c_int(B),
                                       _p(Y, C.c_double), _p(den, C.c_double))
    if rc:
        raise ValueError("CSRV has more rows than `rows`")
    return Y, den


def brute_force_left(semiring, rows, cols, seq, V, X):
    """X float32 (rows,) or (rows, B) -> (Y float64 (cols[, B]), den): Y^T = X^T M."""
    seq = np.ascontiguousarray(seq, np.int32)
    V = np.ascontiguousarray(V, np.float64)
    X = np.ascontiguousarray(X, np.float32)
    B = 1 if X.ndim == 1 else X.shape[1]
    shape = (cols,) if X.ndim == 1 else (cols, B)
    Y = np.empty(shape, np.float64)
    den = np.empty(shape, np.float64)
    rc = lib().orc_brute_force_left(C.c_int(semiring), C.c_int(rows), C.c_int(cols), _p(seq, C.c_int),
                                    C.c_longlong(len(seq)), _p(V, C.c_double), _p(X, C.c_float), C.c_int(B),
                                    _p(Y, C.c_double), _p(den, C.c_double))
    if rc:
        raise ValueError("CSRV has more rows than `rows`")
    return Y, den


def expand(g):
    """g: dict with alpha, kind, left, right, rep, C -> CSRV int32."""
    f = lib().orc_expand
    f.restype = C.c_longlong
    cap = 1 << 16
    while True:
        out = np.empty(cap, np.int32)
        n = f(C.c_int(g["alpha"]), C.c_int(len(g["left"])), _p(g["kind"], C.c_int), _p(g["left"], C.c_int),
              _p(g["right"], C.c_int), _p(g["rep"], C.c_int), _p(g["C"], C.c_int), C.c_longlong(len(g["C"])),
              _p(out, C.c_int), C.c_longlong(cap))
        if n >= 0:
            return out[:n].copy()
        cap *= 4


def layer_refcompat(g):
    """The reference layout restated (dict like gmr.Schedule.arrays())."""
    kind = np.ascontiguousarray(g["kind"], np.int32)
    left = np.ascontiguousarray(g["left"], np.int32)
    right = np.ascontiguousarray(g["right"], np.int32)
    rep = np.ascontiguousarray(g["rep"], np.int32)
    Cs = np.ascontiguousarray(g["C"], np.int32)
    counts = (C.c_longlong * 6)()
    args = [C.c_int(g["alpha"]), C.c_int(len(left)), _p(kind, C.c_int), _p(left, C.c_int), _p(right, C.c_int),
            _p(rep, C.c_int), _p(Cs, C.c_int), C.c_longlong(len(Cs)), counts]
    rc = lib().orc_layer_refcompat(*args, *([None] * 12))
    if rc:
        raise ValueError("malformed grammar")
    L, W, NL, NE, NT, U = list(counts)
    ol, orr = np.empty(NL, np.int32), np.empty(NL, np.int32)
    oc = np.empty(NL, np.float32)
    lo, lw, eo = (np.empty(L + 1, np.int32) for _ in range(3))
    ep, er = np.empty(NE, np.int32), np.empty(NE, np.int32)
    ts, tr = np.empty(NT, np.int32), np.empty(NT, np.int32)
    to = np.empty(U, np.int32)
    nr = C.c_int()
    lib().orc_layer_refcompat(*args, _p(ol, C.c_int), _p(orr, C.c_int), _p(oc, C.c_float), _p(lo, C.c_int),
                              _p(lw, C.c_int), _p(eo, C.c_int), _p(ep, C.c_int), _p(er, C.c_int), _p(ts, C.c_int),
                              _p(tr, C.c_int), _p(to, C.c_int), C.byref(nr))
    return dict(left=ol, right=orr, coeff=oc, level_offsets=lo, level_widths=lw, emit_offsets=eo,
                emit_pos=ep, emit_row=er, term_sym=ts, term_row=tr, term_orig=to,
                info=dict(max_depth=L, max_width=W, num_layered=NL, emit_total=NE, num_term_emit=NT, U=U,
                          num_raw_nt=len(left), num_passthrough=NL - len(left)))


def reference_schedule(g, rows, cols):
    """The reference's own build_schedule output, or None if not built."""
    R = ref_lib()
    if R is None:
        return None
    kind = np.ascontiguousarray(g["kind"], np.int32)
    left = np.ascontiguousarray(g["left"], np.int32)
    right = np.ascontiguousarray(g["right"], np.int32)
    rep = np.ascontiguousarray(g["rep"], np.int32)
    Cs = np.ascontiguousarray(g["C"], np.int32)
    h = R.ref_build(C.c_int(g["alpha"]), C.c_int(len(left)), _p(kind, C.c_int), _p(left, C.c_int),
                    _p(right, C.c_int), _p(rep, C.c_int), _p(Cs, C.c_int), C.c_longlong(len(Cs)),
                    C.c_int(rows), C.c_int(cols))
    info = (C.c_longlong * 8)()
    R.ref_info(h, info)
    L, W, NR, NL, NP, NE, NT, U = list(info)
    ol, orr = np.empty(NL, np.int32), np.empty(NL, np.int32)
    oc = np.empty(NL, np.float32)
    lo, lw, eo = (np.empty(L + 1, np.int32) for _ in range(3))
    ep, er = np.empty(NE, np.int32), np.empty(NE, np.int32)
    ts, tr = np.empty(NT, np.int32), np.empty(NT, np.int32)
    to = np.empty(U, np.int32)
    R.ref_copy(h, *(a.ctypes.data for a in (ol, orr, oc, lo, lw, eo, ep, er, ts, tr, to)))
    R.ref_free(h)
    return dict(left=ol, right=orr, coeff=oc, level_offsets=lo, level_widths=lw, emit_offsets=eo,
                emit_pos=ep, emit_row=er, term_sym=ts, term_row=tr, term_orig=to,
                info=dict(max_depth=L, max_width=W, num_raw_nt=NR, num_layered=NL, num_passthrough=NP,
                          emit_total=NE, num_term_emit=NT, U=U))


def sweep(semiring, s, rows, cols, V, x):
    """FP64 / exact layered sweep over schedule arrays s (double-buffered layouts)."""
    V = np.ascontiguousarray(V, np.float64)
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty(rows, np.float64)
    L = len(s["level_offsets"]) - 1
    lib().orc_sweep(C.c_int(semiring), C.c_int(rows), C.c_int(cols), C.c_int(len(s["term_orig"])),
                    _p(s["term_orig"], C.c_int), _p(V, C.c_double), _p(x, C.c_float), C.c_int(L),
                    _p(s["left"], C.c_int), _p(s["right"], C.c_int), _p(s["coeff"], C.c_float),
                    _p(s["level_offsets"], C.c_int), _p(s["emit_offsets"], C.c_int), _p(s["emit_pos"], C.c_int),
                    _p(s["emit_row"], C.c_int), C.c_longlong(len(s["term_sym"])), _p(s["term_sym"], C.c_int),
                    _p(s["term_row"], C.c_int), _p(y, C.c_double))
    return y


def set_threads(n):
    """OpenMP threads of cpu_layered_f32."""
    lib().orc_set_threads(C.c_int(n))


def cpu_layered_f32(semiring, s, rows, cols, V, x):
    """The reference's OpenMP FP32/int32 CPU sweep (the CPU baseline)."""
    to = s["term_orig"]
    val_idx = np.ascontiguousarray((to[1:] - 1) // cols, np.int32)
    col_idx = np.ascontiguousarray((to[1:] - 1) % cols, np.int32)
    Vf = np.ascontiguousarray(V, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty(rows, np.float32 if semiring == 0 else np.int32)
    L = len(s["level_offsets"]) - 1
    lib().orc_cpu_layered_f32(C.c_int(semiring), C.c_int(rows), C.c_int(len(to)), _p(val_idx, C.c_int),
                              _p(col_idx, C.c_int), _p(Vf, C.c_float), _p(x, C.c_float), C.c_int(L),
                              _p(s["left"], C.c_int), _p(s["right"], C.c_int), _p(s["coeff"], C.c_float),
                              _p(s["level_offsets"], C.c_int), _p(s["emit_offsets"], C.c_int),
                              _p(s["emit_pos"], C.c_int), _p(s["emit_row"], C.c_int),
                              C.c_longlong(len(s["term_sym"])), _p(s["term_sym"], C.c_int),
                              _p(s["term_row"], C.c_int), y.ctypes.data_as(C.c_void_p))
    return y


def mlp_loss_grad(nin, nh, nout, params, X, T, M, count=None):
    """FP64 loss, gradient and outputs of the proposer network (see oracle.c)."""
    p = np.ascontiguousarray(params, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    T = np.ascontiguousarray(T, np.float64)
    M = np.ascontiguousarray(M, np.float64)
    N = X.shape[0]
    count = float(M.sum()) if count is None else float(count)
    g = np.zeros_like(p)
    out = np.zeros((N, nout), np.float64)
    f = lib().orc_mlp_loss_grad
    f.restype = C.c_double
    loss = f(C.c_int(nin), C.c_int(nh), C.c_int(nout), _p(p, C.c_double), _p(X, C.c_double), _p(T, C.c_double),
             _p(M, C.c_double), C.c_int(N), C.c_double(count), _p(g, C.c_double), _p(out, C.c_double))
    return loss, g, out
