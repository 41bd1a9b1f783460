"""ctypes bindings of the engine's C-ABI libraries (lib/libgmr_host.so and
lib/libgmr.so, declared in include/gmr_host.h and include/gmr.h), and of the
vendor baselines (lib/libgmr_vendor.so, include/gmr_vendor.h).

Loading never builds anything: the libraries are built in-tree by the package
Makefile. A missing library raises ImportError naming the file.
"""
import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_DIR, "lib")
HOST_LIB = os.path.join(LIB_DIR, "libgmr_host.so")
DEV_LIB = os.path.join(LIB_DIR, "libgmr.so")
VENDOR_LIB = os.path.join(LIB_DIR, "libgmr_vendor.so")

c_int_p = C.POINTER(C.c_int)
c_float_p = C.POINTER(C.c_float)
c_double_p = C.POINTER(C.c_double)
c_ll_p = C.POINTER(C.c_longlong)


class LayeredRule(C.Structure):
    """layered_types.h LayeredRule: {int left; int right; float coeff;} (12 bytes)."""
    _fields_ = [("left", C.c_int), ("right", C.c_int), ("coeff", C.c_float)]


class GPUSchedule(C.Structure):
    """layered_types.h GPUSchedule (borrowed pointers)."""
    _fields_ = [
        ("max_depth", C.c_int), ("max_width", C.c_int),
        ("level_offsets", c_int_p), ("level_widths", c_int_p),
        ("emit_offsets", c_int_p), ("emit_pos", c_int_p), ("emit_row", c_int_p),
        ("emit_total", C.c_int),
        ("term_emit_sym", c_int_p), ("term_emit_row", c_int_p), ("num_term_emit", C.c_int),
    ]


class RunOpts(C.Structure):
    """gmr.h gmr_run_opts."""
    _fields_ = [
        ("variant", C.c_int), ("emit", C.c_int), ("b_tile", C.c_int), ("block", C.c_int),
        ("grid_cap", C.c_int), ("eval_kind", C.c_int), ("flags", C.c_int), ("xcd_passes", C.c_int),
        ("level_kind", c_int_p), ("lds_values", C.c_int), ("chunks", C.c_int),
    ]


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_host = None
_dev = None
_vendor = None


def host():
    """The host library (no GPU use)."""
    global _host
    if _host is not None:
        return _host
    if not os.path.exists(HOST_LIB):
        raise ImportError(f"{HOST_LIB} is not built (run make in {PKG_DIR})")
    L = C.CDLL(HOST_LIB, mode=C.RTLD_GLOBAL)
    vp, vpp = C.c_void_p, C.POINTER(C.c_void_p)
    _sig(L, "gmr_last_error", C.c_char_p, [])
    _sig(L, "gmr_matrix_haplotype", C.c_int, [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, vpp])
    _sig(L, "gmr_matrix_genotype", C.c_int, [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, vpp])
    _sig(L, "gmr_matrix_rmat", C.c_int, [C.c_uint64, C.c_int, C.c_longlong, C.c_double, C.c_double, C.c_double, C.c_int, vpp])
    _sig(L, "gmr_matrix_from_dense", C.c_int, [c_double_p, C.c_int, C.c_int, vpp])
    _sig(L, "gmr_matrix_from_csrv", C.c_int, [C.c_int, C.c_int, c_int_p, C.c_longlong, c_double_p, C.c_int, vpp])
    _sig(L, "gmr_matrix_info", C.c_int, [vp, c_int_p, c_int_p, c_int_p, c_int_p, c_ll_p])
    _sig(L, "gmr_matrix_copy", C.c_int, [vp, c_int_p, c_double_p])
    _sig(L, "gmr_matrix_row_block", C.c_int, [vp, C.c_int, C.c_int, vpp])
    _sig(L, "gmr_matrix_free", None, [vp])
    _sig(L, "gmr_grammar_repair", C.c_int, [vp, C.c_int, C.c_int, vpp, c_int_p, c_double_p])
    _sig(L, "gmr_grammar_from_arrays", C.c_int, [C.c_int, C.c_int, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p, C.c_longlong, c_double_p, C.c_int, vpp])
    _sig(L, "gmr_grammar_read", C.c_int, [C.c_char_p, vpp])
    _sig(L, "gmr_grammar_write", C.c_int, [vp, C.c_char_p])
    _sig(L, "gmr_grammar_info", C.c_int, [vp, c_int_p, c_int_p, c_ll_p, c_int_p])
    _sig(L, "gmr_grammar_copy", C.c_int, [vp, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p, c_double_p])
    _sig(L, "gmr_grammar_expand", C.c_int, [vp, c_int_p, C.c_longlong, c_ll_p])
    _sig(L, "gmr_grammar_inline_top", C.c_int, [vp, C.c_int, C.POINTER(vp), c_ll_p])
    _sig(L, "gmr_grammar_free", None, [vp])
    _sig(L, "gmr_schedule_build", C.c_int, [vp, C.c_int, C.c_int, C.c_int, vpp])
    _sig(L, "gmr_schedule_info", C.c_int, [vp, c_ll_p])
    _sig(L, "gmr_schedule_copy", C.c_int, [vp, C.POINTER(LayeredRule), c_int_p, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p])
    _sig(L, "gmr_schedule_view", C.c_int, [vp, C.POINTER(GPUSchedule)])
    _sig(L, "gmr_schedule_cpu_sweep", C.c_int, [vp, C.c_int, c_float_p, C.c_int, c_float_p, C.c_void_p, C.c_int])
    _sig(L, "gmr_schedule_free", None, [vp])
    _host = L
    return L


def dev():
    """The device library (HIP). Loading it does not touch the GPU."""
    global _dev
    if _dev is not None:
        return _dev
    host()
    if not os.path.exists(DEV_LIB):
        raise ImportError(f"{DEV_LIB} is not built (run make in {PKG_DIR})")
    L = C.CDLL(DEV_LIB, mode=C.RTLD_GLOBAL)
    vp, vpp = C.c_void_p, C.POINTER(C.c_void_p)
    _sig(L, "gmr_engine_create", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(LayeredRule), c_int_p, c_int_p, c_float_p, C.c_int, C.POINTER(GPUSchedule), C.c_int, vpp])
    _sig(L, "gmr_engine_from_schedule", C.c_int, [vp, c_float_p, C.c_int, vpp])
    _sig(L, "gmr_engine_info", C.c_int, [vp, c_ll_p])
    _sig(L, "gmr_engine_run", C.c_int, [vp, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.POINTER(RunOpts), C.c_void_p])
    _sig(L, "gmr_engine_run_noalloc", C.c_int, [vp, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.POINTER(RunOpts), C.c_void_p])
    _sig(L, "gmr_engine_last_kernel_ms", C.c_int, [vp, c_double_p])
    _sig(L, "gmr_engine_check", C.c_int, [vp, c_int_p])
    _sig(L, "gmr_engine_set_barrier_timeout", C.c_int, [vp, C.c_longlong])
    _sig(L, "gmr_engine_debug", C.c_int, [vp, c_ll_p])
    _sig(L, "gmr_engine_set_gather_order", C.c_int, [vp, C.c_int])
    _sig(L, "gmr_engine_traverse", C.c_int, [vp, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(RunOpts), C.c_void_p])
    _sig(L, "gmr_engine_run_left", C.c_int, [vp, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.POINTER(RunOpts), C.c_void_p])
    _sig(L, "gmr_engine_time_levels", C.c_int, [vp, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.POINTER(RunOpts), C.c_int, c_float_p, C.c_void_p])
    _sig(L, "gmr_engine_destroy", None, [vp])
    _sig(L, "gmr_device_sync", C.c_int, [])
    _sig(L, "gmr_batch_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int, vpp])
    _sig(L, "gmr_batch_run", C.c_int, [vp, C.c_int, C.POINTER(C.c_void_p), c_int_p, C.POINTER(C.c_void_p), C.c_int, C.POINTER(RunOpts), C.c_void_p])
    _sig(L, "gmr_batch_check", C.c_int, [vp, c_int_p])
    _sig(L, "gmr_batch_set_barrier_timeout", C.c_int, [vp, C.c_longlong])
    _sig(L, "gmr_batch_destroy", None, [vp])
    _sig(L, "gmr_mlp_create", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint64, vpp])
    _sig(L, "gmr_mlp_num_params", C.c_longlong, [vp])
    _sig(L, "gmr_mlp_get_params", C.c_int, [vp, c_float_p])
    _sig(L, "gmr_mlp_set_params", C.c_int, [vp, c_float_p])
    _sig(L, "gmr_mlp_forward", C.c_int, [vp, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p])
    _sig(L, "gmr_mlp_loss_grad", C.c_int, [vp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_float, c_float_p, c_float_p])
    _sig(L, "gmr_mlp_train", C.c_int, [vp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.c_int, C.c_float, c_float_p])
    _sig(L, "gmr_mlp_destroy", None, [vp])
    _dev = L
    return L


def vendor():
    """The vendor-baseline library (rocSPARSE CSR products, amdsmi energy).
    Loading it does not touch the GPU."""
    global _vendor
    if _vendor is not None:
        return _vendor
    host()
    if not os.path.exists(VENDOR_LIB):
        raise ImportError(f"{VENDOR_LIB} is not built (run make in {PKG_DIR})")
    L = C.CDLL(VENDOR_LIB, mode=C.RTLD_GLOBAL)
    vp, vpp = C.c_void_p, C.POINTER(C.c_void_p)
    _sig(L, "gmr_csr_create", C.c_int, [C.c_int, C.c_int, C.c_longlong, c_int_p, c_int_p, c_float_p, vpp])
    _sig(L, "gmr_csr_run", C.c_int, [vp, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p])
    _sig(L, "gmr_csr_bench", C.c_int, [vp, c_float_p, C.c_int, c_float_p, C.c_int, C.c_int, c_double_p])
    _sig(L, "gmr_csr_energy", C.c_int, [vp, C.c_int, C.c_int, C.c_double, c_double_p, c_double_p])
    _sig(L, "gmr_csr_device_bytes", C.c_longlong, [vp])
    _sig(L, "gmr_csr_free", None, [vp])
    _sig(L, "gmr_energy_uj", C.c_int, [C.c_int, c_double_p, C.POINTER(C.c_ulonglong)])
    _vendor = L
    return L

