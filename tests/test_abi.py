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
