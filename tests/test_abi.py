"""The C-ABI libraries load without a GPU and export every function the
headers in include/ declare. No compute call is made here."""
import ctypes
import os
import re

import pytest

from gmr import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
DECL = re.compile(r"^\s*(?:extern \"C\" )?(?:const )?(?:int|void|char|double|long long)\s*\*?\s*((?:gmr|run_gpu)_\w+)\s*\(",
                  re.M)


def declared(header):
    with open(os.path.join(INC, header)) as f:
        return sorted(set(DECL.findall(f.read())))


def test_headers_declare_the_expected_entry_points():
    assert "run_gpu_right_multiply_layered" in declared("gpu_engine.h")
    assert "run_gpu_right_multiply_batched" in declared("gpu_engine.h")
    assert "gmr_engine_run" in declared("gmr.h")
    assert "gmr_schedule_build" in declared("gmr_host.h")
    assert "gmr_csr_run" in declared("gmr_vendor.h") and "gmr_csr_device_bytes" in declared("gmr_vendor.h")


@pytest.mark.parametrize("header,lib", [("gmr_host.h", "host"), ("gmr.h", "dev"), ("gpu_engine.h", "dev"),
                                        ("gmr_vendor.h", "vendor")])
def test_library_exports_every_declared_symbol(header, lib):
    L = {"host": _lib.host, "dev": _lib.dev, "vendor": _lib.vendor}[lib]()
    names = declared(header)
    assert names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"{header}: not exported by {lib}: {missing}"


def test_struct_layouts_match_the_reference():
    # layered_types.h: LayeredRule {int,int,float} = 12 B; GPUSchedule as in the
    # reference (two ints, two pointers, three pointers + int, two pointers + int).
    assert ctypes.sizeof(_lib.LayeredRule) == 12
    assert ctypes.sizeof(_lib.GPUSchedule) == 8 + 2 * 8 + 3 * 8 + 8 + 2 * 8 + 8


def test_python_constants_match_the_header_enums():
    """Every enumerator of the C headers has the value the Python bindings pass
    (GMR_PLUS_TIMES -> gmr.PLUS_TIMES, GMR_LAYOUT_RESIDENT -> gmr.RESIDENT,
    GMR_RUN_BITS -> gmr.RUN_BITS, GMR_CSR_ALG_ROW -> vendor.ALGS['row'], ...)."""
    import gmr
    from gmr import vendor
    enum = re.compile(r"\b(GMR_[A-Z0-9_]+)\s*=\s*(\d+)")
    values = {}
    for h in ("gmr.h", "gmr_host.h", "gmr_vendor.h"):
        with open(os.path.join(INC, h)) as f:
            values.update({k: int(v) for k, v in enum.findall(f.read())})
    checked = 0
    for name, v in values.items():
        short = name[len("GMR_"):]
        if short.startswith("LAYOUT_"):
            short = short[len("LAYOUT_"):]
        if short.startswith("CSR_ALG_"):
            assert vendor.ALGS[short[len("CSR_ALG_"):].lower()] == v, name
            checked += 1
        elif hasattr(gmr, short):
            assert getattr(gmr, short) == v, name
            checked += 1
    for must in ("GMR_RUN_BITS", "GMR_RUN_SINGLE_LIST", "GMR_VARIANT_PERSISTENT", "GMR_EMIT_GATHER", "GMR_TROPICAL",
                 "GMR_EVAL_LDS", "GMR_LAYOUT_RESIDENT"):
        assert must in values, must
    assert checked >= 20, checked
