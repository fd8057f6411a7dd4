"""The C-ABI library loads and exports every symbol include/lenia_hip.h declares.
No compute calls: this runs without a GPU."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lh_[a-z_0-9]+)\s*\(", text)))


def test_header_lists_match_binding():
    from lenia_amd import _lib
    assert sorted(_lib.EXPORTS) == declared("lenia_hip.h")


def test_library_exports_every_declared_symbol():
    from lenia_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("liblenia_hip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared("lenia_hip.h") if not hasattr(lib, s)]
    assert not missing, missing


def test_shape_table_is_host_only():
    """lh_shape_supported consults the compiled shape table only (no HIP call)."""
    from lenia_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("liblenia_hip.so not built")
    L = _lib.load()
    assert L.lh_shape_supported(32, 32, 32) == 1
    assert L.lh_shape_supported(44, 24, 44) == 1
    assert L.lh_shape_supported(30, 30, 30) == 0
    assert L.lh_version().startswith(b"lenia_hip")
