"""Vendor baselines on MI355X (lib/libgmr_vendor.so, include/gmr_vendor.h):
the uncompressed matrix in CSR form multiplied by rocSPARSE -- the
counterpart of the reference's cuSPARSE baseline (gpu-engine/cusparse_test.cu)
-- and the GPU energy counter read through amdsmi (the NVML counterpart)."""
import ctypes as C

import numpy as np

from . import GmrError, _check, _lib

ALGS = {"default": 0, "row": 1, "row_split": 2, "nnz_split": 3}


def csr_arrays(seq, V, rows, cols):
    """(row_ptr int32[rows+1], col int32[nnz], vals float32[nnz]) from a CSRV
    string (row terminals then 0; terminal t = 1 + v*cols + c)."""
    seq = np.asarray(seq)
    ends = np.flatnonzero(seq == 0)
    if len(ends) > rows:
        raise ValueError(f"the string holds {len(ends)} rows, more than {rows}")
    counts = np.diff(np.concatenate([[-1], ends])) - 1
    row_ptr = np.zeros(rows + 1, np.int64)
    row_ptr[1:len(counts) + 1] = np.cumsum(counts)
    row_ptr[len(counts) + 1:] = row_ptr[len(counts)]
    t = seq[seq != 0].astype(np.int64) - 1
    col = (t % cols).astype(np.int32)
    vals = np.asarray(V, np.float64)[t // cols].astype(np.float32)
    if row_ptr[-1] > 0x7FFFFFFF:
        raise ValueError("more than 2^31 non-zeros")
    return row_ptr.astype(np.int32), col, vals


class Csr:
    """A CSR matrix resident on the GPU for rocSPARSE SpMV / SpMM."""

    def __init__(self, rows, cols, row_ptr, col, vals):
        L = _lib.vendor()
        self.rows, self.cols, self.nnz = rows, cols, len(col)
        rp = np.ascontiguousarray(row_ptr, np.int32)
        ci = np.ascontiguousarray(col, np.int32)
        va = np.ascontiguousarray(vals, np.float32)
        h = C.c_void_p()
        _check(L.gmr_csr_create(rows, cols, self.nnz, rp.ctypes.data_as(C.POINTER(C.c_int)),
                                ci.ctypes.data_as(C.POINTER(C.c_int)), va.ctypes.data_as(C.POINTER(C.c_float)),
                                C.byref(h)))
        self._h = h

    @classmethod
    def from_matrix(cls, m):
        """From a gmr.Matrix (its CSRV string)."""
        seq, V = m.arrays()
        return cls(m.rows, m.cols, *csr_arrays(seq, V, m.rows, m.cols))

    def run_torch(self, X, Y=None, alg="default"):
        """Y = M X (X float32 CUDA, (cols,) or (cols, B) row-major) on torch's
        current stream."""
        import torch
        assert X.is_cuda and X.dtype == torch.float32 and X.is_contiguous()
        B = 1 if X.dim() == 1 else X.shape[1]
        if X.shape[0] != self.cols:
            raise ValueError(f"X has {X.shape[0]} rows, the matrix has {self.cols} columns")
        if Y is None:
            Y = torch.empty((self.rows, B) if X.dim() == 2 else (self.rows,), dtype=torch.float32, device=X.device)
        stream = torch.cuda.current_stream(X.device).cuda_stream
        _check(_lib.vendor().gmr_csr_run(self._h, C.c_void_p(X.data_ptr()), B, C.c_void_p(Y.data_ptr()),
                                         ALGS[alg] if isinstance(alg, str) else int(alg), C.c_void_p(stream)))
        return Y

    def energy(self, B=1, alg="default", seconds=1.5):
        """(millijoules per product, mean power in watts) over a loop of at
        least `seconds`; raises GmrError when the counter cannot be read."""
        mj, mw = C.c_double(), C.c_double()
        _check(_lib.vendor().gmr_csr_energy(self._h, B, ALGS[alg] if isinstance(alg, str) else int(alg),
                                            float(seconds), C.byref(mj), C.byref(mw)))
        return mj.value, mw.value * 1e-3

    def device_bytes(self):
        return int(_lib.vendor().gmr_csr_device_bytes(self._h))

    def __del__(self):
        lib = getattr(_lib, "_vendor", None) if _lib is not None else None
        if getattr(self, "_h", None) and lib is not None:
            lib.gmr_csr_free(self._h)
            self._h = None


def energy_read(device=0):
    """(accumulated energy in microjoules, time of the counter's last update in
    ns) from amdsmi, or None when the counter cannot be read."""
    v, ts = C.c_double(), C.c_ulonglong()
    try:
        _check(_lib.vendor().gmr_energy_uj(int(device), C.byref(v), C.byref(ts)))
    except GmrError:
        return None
    return v.value, ts.value


def energy_uj(device=0):
    """The GPU's accumulated energy in microjoules, or None."""
    r = energy_read(device)
    return None if r is None else r[0]


def energy_between_updates(batch, seconds=1.5, device=0):
    """Energy of work between two updates of the counter at least `seconds`
    apart (it advances about once a second): `batch()` runs some products,
    synchronises and returns how many it ran. Returns (millijoules per
    product, mean watts) or None when the counter cannot be read, does not
    advance within 2 x seconds + 5 s, or jumps backwards in three attempts
    (a backwards jump of energy and timestamp has been seen between reads)."""
    for _ in range(3):
        r = _energy_interval(batch, seconds, device)
        if r is None:
            return None
        uj, secs, n = r
        if uj >= 0 and secs > 0:
            return uj * 1e-3 / n, uj * 1e-6 / secs
    return None


def _energy_interval(batch, seconds, device):
    import time
    r = energy_read(device)
    if r is None:
        return None
    t0 = time.time()
    ts_prev = r[1]
    while True:
        batch()
        r0 = energy_read(device)
        if r0 is None or time.time() - t0 > 2 * seconds + 5:
            return None
        if r0[1] != ts_prev:
            break
    n = 0
    while True:
        n += batch()
        r1 = energy_read(device)
        if r1 is None or time.time() - t0 > 2 * seconds + 5:
            return None
        if r1[1] != r0[1] and (r1[1] < r0[1] or (r1[1] - r0[1]) * 1e-9 >= seconds):
            break
    return r1[0] - r0[0], (r1[1] - r0[1]) * 1e-9, n
