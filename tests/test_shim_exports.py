"""liblenia_ref_shim.so exports every entry point lenia_ref_shim.h declares, with the
reference's C++ signatures (lenia.h:5-28, reintegration.h:74-107, transposition.h:5-12)."""
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(REPO, "include", "lenia_ref_shim.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^(?:void|const char \*)\s*(\w+)\s*\(", text, flags=re.M)))


def test_shim_exports_all_declared_functions():
    from lenia_amd import build as B
    if not os.path.exists(B.SHIM):
        pytest.skip("shim not built")
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", B.SHIM], capture_output=True, text=True,
                         check=True).stdout
    names = set(re.findall(r"\b(\w+)\(", out))
    missing = [d for d in declared() if d not in names]
    assert not missing, missing
    # C++ linkage, as the reference declares them (no extern "C")
    assert "fft_step(TD*)" in out and "srt(float*, int, int, int, float, float)" in out
