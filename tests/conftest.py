import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def orc64():
    from oracle.oracle import Oracle
    return Oracle(64)


@pytest.fixture(scope="session")
def orc32():
    from oracle.oracle import Oracle
    return Oracle(32)


@pytest.fixture(scope="session")
def gctx():
    import laminocam_amd as lc
    return lc.default_context()


def rel_l2(a, b):
    import numpy as np
    a, b = np.asarray(a), np.asarray(b)
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a, b = a.astype(dt), b.astype(dt)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def announce(msg):
    print(f"[case] {msg}", flush=True)
