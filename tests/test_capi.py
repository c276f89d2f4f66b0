"""The C ABI library loads and exports every entry point that
include/catn_hip.h declares (no compute calls: the CPU suite has no GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "catn_hip.h")
HIP_LIB = os.path.join(ROOT, "catn.jl_amd", "lib", "libcatn_hip.so")
EMU_LIB = os.path.join(ROOT, "tests", "emul", "libcatn_emul.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(catn_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "catn_contraction" in names and "catn_tsvd" in names and len(names) >= 30


@pytest.mark.parametrize("lib", [HIP_LIB, EMU_LIB])
def test_library_exports_every_declared_symbol(lib):
    assert os.path.exists(lib), f"{lib} not built (run __graft_entry__.build())"
    so = ctypes.CDLL(lib)
    missing = [n for n in declared() if not hasattr(so, n)]
    assert not missing, missing


def test_python_binding_matches_header(catn):
    assert set(catn.EXPORTED_SYMBOLS) <= set(declared())
    assert set(declared()) <= set(catn.EXPORTED_SYMBOLS)


def test_hip_library_has_no_cpu_fallback():
    """The product library contains the gfx950 code object and no emulator."""
    data = open(HIP_LIB, "rb").read()
    assert b"gfx950" in data
    assert b"cpu-emulator" not in data
