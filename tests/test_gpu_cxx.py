"""The reference's C++ API (projection.h, toeplitz*.h, optimize.h, bregman.h)
re-implemented over the C ABI in include/laminocam_tomocam.hpp, driven by a
C++ program built with g++ (tests/cxx/test_shim.cpp)."""
import os
import subprocess

import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cxx", "_bin", "test_shim")


def test_cxx_api_over_c_abi():
    announce("C++ API program")
    if not os.path.exists(BIN):
        pytest.fail("tests/cxx/_bin/test_shim is not built (run __graft_entry__.build())")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK: 0 failure(s)" in r.stdout
