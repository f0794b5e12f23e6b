"""The C-ABI library loads and exports every symbol include/laminocam.h
declares (no compute calls: runs without a GPU)."""
import os
import re

import numpy as np

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


def test_server_request_shapes_are_checked_before_the_library_reads_them():
    # the library copies (or, borrowed, later reads) nslices x the geometry's
    # sizes from each array: every wrong shape must stop at the host check
    import numpy as np
    import pytest
    import laminocam_amd as lc

    g2 = lc.PolarGrid2D(np.linspace(0, np.pi, 30, endpoint=False), 16)
    psf2 = object.__new__(lc.PointSpreadFunction2D)   # fields the check reads; no device
    psf2.grid, psf2.image_size = g2, 16
    g3 = lc.PolarGrid(np.linspace(0, 2 * np.pi, 12, endpoint=False), -0.5, 0.0, 8, 16)
    psf3 = object.__new__(lc.PointSpreadFunction)
    psf3.grid, psf3.recon_dims = g3, (4, 16, 16)

    ns, shape, arr = lc._request_arrays(psf2, np.zeros((3, 30, 16)), None, None)
    assert ns == 3 and shape == (3, 16, 16) and arr["sino"].dtype == np.float32
    ns, shape, arr = lc._request_arrays(psf2, np.zeros((30, 16)), None, np.zeros((16, 16)))
    assert ns == 1 and set(arr) == {"sino", "x0"}
    ns, shape, arr = lc._request_arrays(psf3, np.zeros((12, 8, 16)), None, None)
    assert ns == 4 and shape == (4, 16, 16)
    ns, shape, arr = lc._request_arrays(psf3, None, np.zeros((4, 16, 16)), np.zeros((4, 16, 16)))
    assert set(arr) == {"rhs", "x0"}

    bad = [
        (psf2, np.zeros((3, 29, 16)), None, None),        # one angle short
        (psf2, np.zeros((3, 30, 15)), None, None),        # one bin short
        (psf2, np.zeros((3, 30, 16)), None, np.zeros((2, 16, 16))),   # x0 of 2 slices
        (psf2, None, np.zeros((3, 16, 15)), None),        # rhs too small
        (psf3, np.zeros((12, 8, 15)), None, None),        # projections too small
        (psf3, np.zeros((11, 8, 16)), None, None),        # one projection short
        (psf3, None, np.zeros((4, 16, 16)), np.zeros((4, 16, 15))),   # x0 too small
        (psf3, None, np.zeros((3, 16, 16)), None),        # rhs of 3 slices
        (psf2, None, None, None),                         # no input
        (psf2, np.zeros((3, 30, 16)), np.zeros((3, 16, 16)), None),   # both inputs
    ]
    for case in bad:
        with pytest.raises(ValueError):
            lc._request_arrays(*case)


def test_reference_memory_model_and_fit_rule():
    # scripts/estimate_bregman_memory.py:18-62: Split-Bregman + CG high-water
    # mark 22 N typical, 24 N conservative, 2 N inputs; the fit rule of
    # src/gpu/mbir2d_multi.cu:459-470: 27 N (Split-Bregman) / 11 N (CG)
    # elements below 0.8 of the free memory
    import laminocam_amd as lc
    from laminocam_amd.drivers import ReconParams, Regularizer

    m = lc.estimate_memory((256, 256, 256), np.float32)
    N = 256 ** 3
    assert m["N"] == N
    assert (m["typical_bytes"], m["conservative_bytes"], m["input_bytes"], m["fit_rule_bytes"]) == \
        (22 * 4 * N, 24 * 4 * N, 2 * 4 * N, 27 * 4 * N)
    m = lc.estimate_memory((21, 511, 511), np.float64, Regularizer.UNCONSTRAINED)
    N = 21 * 511 * 511
    assert m["typical_bytes"] == m["fit_rule_bytes"] == 11 * 8 * N
    # 1 GiB volume in float32: SB needs 27 GiB < 0.8 free, CG 11 GiB
    p = ReconParams(regularizer=Regularizer.SPLIT_BREGMAN, recon_dims=(1024, 512, 512))
    need = 27 * 4 * 1024 * 512 * 512
    edge = int(np.ceil(need / 0.8))
    assert lc.volume_fits_one_gpu(p, np.float32, free_bytes=edge + 16)
    assert not lc.volume_fits_one_gpu(p, np.float32, free_bytes=edge - 16)
    assert not lc.volume_fits_one_gpu(p, np.float64, free_bytes=edge + 16)
    p.regularizer = Regularizer.UNCONSTRAINED
    assert lc.volume_fits_one_gpu(p, np.float64, free_bytes=int(np.ceil(11 * 8 * 1024 * 512 * 512 / 0.8)) + 16)
