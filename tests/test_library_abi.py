"""The C-ABI library loads and exports every entry point the public header declares (no device calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "flosic_hip.h")
LIB = os.path.join(ROOT, "publicrelease_2020_amd", "libflosic_hip.so")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|const char \*)\s*\**\s*(flosic_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "flosic_evaluate" in names and "flosic_fod_forces" in names and len(names) >= 15


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    for name in declared():
        assert hasattr(lib, name), name


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_mirror_binds_every_symbol():
    from publicrelease_2020_amd import flosic as F
    assert sorted(F.exported_symbols()) == declared()
