"""The C-ABI library loads and exports every symbol include/laminocam.h
declares (no compute calls: runs without a GPU)."""
import os
import re

from laminocam_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "laminocam.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(lc_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_symbols()
    assert names, "no declarations parsed"
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == names


def test_abi_version_and_error_plumbing():
    L = _lib.load()
    assert L.lc_abi_version() == 1
    assert L.lc_last_error() is not None
