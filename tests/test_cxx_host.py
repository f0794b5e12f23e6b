"""Host-only parts of the C++ API (include/laminocam_tomocam.hpp): opt::Function
as the std::function alias, cgsolver's host loop over an opaque callable on the
normal equations of tests/test_cg_normal.cpp, and the polar-grid point arrays.
The program (tests/cxx/test_host_api.cpp) makes no device call."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cxx", "_bin", "test_host_api")


def test_cxx_host_api():
    if not os.path.exists(BIN):
        pytest.fail("tests/cxx/_bin/test_host_api is not built (run __graft_entry__.build())")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK: 0 failure(s)" in r.stdout
