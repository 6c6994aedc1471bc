"""CPU checks of the C-ABI boundary: the built library loads and exports every declared symbol."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nusht_amd.h")
LIB = os.path.join(ROOT, "cunusht_amd", "libnusht_amd.so")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nusht_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    for n in ("nusht_plan_create", "nusht_nusht2d2", "nusht_nusht2d1", "nusht_gclm2lenmap", "nusht_lenmap2gclm",
              "nusht_ringplan_dlm2pointing", "nusht_doubling_f64", "nusht_adjoint_doubling_f64",
              "nusht_pointing_f64"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (nusht_[a-z0-9_]+)", out))
    missing = [n for n in _declared() if n not in exported]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_ctypes_signatures_cover_header():
    from cunusht_amd import _lib
    lib = _lib.load()
    assert set(_declared()) <= set(_lib.SIGNATURES)
    assert lib.nusht_version() == 1
    # error path without a GPU call: null plan handle
    assert lib.nusht_plan_get_info(None, None) == _lib.NUSHT_OK + 1
    assert b"null" in lib.nusht_last_error()
