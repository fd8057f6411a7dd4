"""The C-ABI library loads and exports every symbol include/lenia_hip.h and
include/lenia_mlp.h declare.  No compute calls: this runs without a GPU."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(*headers):
    out = set()
    for header in headers:
        text = open(os.path.join(REPO, "include", header)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        out |= set(re.findall(r"\b(l[hm]_[a-z_0-9]+)\s*\(", text))
    return sorted(out)


HEADERS = ("lenia_hip.h", "lenia_mlp.h")


def test_header_lists_match_binding():
    from lenia_amd import _lib
    assert sorted(_lib.EXPORTS) == declared(*HEADERS)


def test_library_exports_every_declared_symbol():
    from lenia_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("liblenia_hip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared(*HEADERS) if not hasattr(lib, s)]
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


def test_shape_modules_found_without_loading(tmp_path, monkeypatch):
    """A shape outside the built-in set is reported as a module (2) when its file is in
    $LENIA_SHAPE_DIR or shapes/ beside the library; nothing is loaded to answer."""
    from lenia_amd import _lib, build as B
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("liblenia_hip.so not built")
    L = _lib.load()
    for shape in B.TEST_SHAPES:
        if os.path.exists(B.shape_module_path(*shape)):
            assert L.lh_shape_supported(*shape) == 2, shape
    assert L.lh_shape_supported(60, 8, 60) == 0
    (tmp_path / "liblenia_shape_60x8x60.so").write_bytes(b"")
    monkeypatch.setenv("LENIA_SHAPE_DIR", str(tmp_path))
    assert L.lh_shape_supported(60, 8, 60) == 2


def test_shape_module_sides_checked():
    from lenia_amd import build as B
    for bad in ((30, 32, 32), (32, 0, 32), (132, 32, 32), (32, 32, 2)):
        with pytest.raises(ValueError):
            B._shape_module_cmd(*bad, "/dev/null")
