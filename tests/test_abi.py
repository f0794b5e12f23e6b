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


def test_struct_layouts_match_the_header(tmp_path):
    # sizes and field offsets of the ctypes mirrors against the C compiler's
    import ctypes as C
    import shutil
    import subprocess
    import pytest
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    mirrors = {"lc_request": _lib.LcRequest, "lc_recon_params": _lib.LcReconParams,
               "lc_report": _lib.LcReport}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "laminocam.h"', 'int main(void) {']
    for cname, py in mirrors.items():
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            cf = "lambda" if f == "lam" else f
            lines.append(f'    printf("{cname} {f} %zu\\n", offsetof({cname}, {cf}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for ln in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        s, f, v = ln.split()
        got[(s, f)] = int(v)
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
