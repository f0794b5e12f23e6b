"""lc_cg2d_noalloc, the allocation-free CG solve, against lc_cg2d through the
C ABI: two calls in a row on device buffers, the second on a smaller problem,
each bitwise equal to lc_cg2d's result with the same iteration count, and a
problem beyond the fixed scratch refused with LC_EINVAL."""
import ctypes as C

import numpy as np
import pytest

from conftest import announce
from test_gpu_device_ptrs import Dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lc():
    import laminocam_amd
    return laminocam_amd


def _solve(lib, name, psf, drhs, dx, ns, iters, flags):
    from laminocam_amd import _lib
    rep = _lib.LcReport()
    _lib.check(getattr(lib, name)(psf.ctx.handle, psf.handle, drhs.p, dx.p, ns, iters, 1e-5,
                                  1e-5, flags, C.byref(rep)))
    return dx.get(), rep.as_dict()


def test_noalloc_cg_matches_cg2d(lc):
    from laminocam_amd import _lib
    rng = np.random.default_rng(11)
    dev = _lib.LC_DEVICE_PTRS
    # first a larger problem, then a smaller one: the second call reuses the
    # scratch the first one left behind
    for N, nang, ns, iters, x0 in ((256, 180, 6, 12, False), (96, 60, 2, 9, True)):
        announce(f"noalloc cg N={N} nang={nang} ns={ns} x0={x0}")
        g = lc.PolarGrid2D(np.arange(nang) * np.pi / nang, N)
        psf = lc.PointSpreadFunction2D(g, N)
        rhs = psf.convolve(rng.random((ns, N, N)).astype(np.float32))
        x0a = (0.1 * rng.random((ns, N, N))).astype(np.float32) if x0 else \
            np.zeros((ns, N, N), np.float32)
        flags = dev | (0 if x0 else _lib.LC_X0_ZERO)
        drhs = Dev(rhs)
        xa, ra = _solve(psf.lib, "lc_cg2d", psf, drhs, Dev(x0a), ns, iters, flags)
        xb, rb = _solve(psf.lib, "lc_cg2d_noalloc", psf, drhs, Dev(x0a), ns, iters, flags)
        assert ra["iters"] == rb["iters"] and ra["stop"] == rb["stop"]
        assert ra["res"] == rb["res"]
        assert np.array_equal(xa, xb)
        assert np.all(np.isfinite(xb)) and np.any(xb)


def test_noalloc_cg_refuses_a_problem_beyond_the_scratch(lc):
    from laminocam_amd import _lib
    announce("noalloc cg beyond the scratch")
    N, ns = 1024, 3   # about 90 MB of solver buffers, above the 64 MiB array
    g = lc.PolarGrid2D(np.arange(90) * np.pi / 90, N)
    psf = lc.PointSpreadFunction2D(g, N)
    drhs, dx = Dev(shape=(ns, N, N)), Dev(shape=(ns, N, N))
    rep = _lib.LcReport()
    rc = psf.lib.lc_cg2d_noalloc(psf.ctx.handle, psf.handle, drhs.p, dx.p, ns, 5, 1e-5, 1e-5,
                                 _lib.LC_DEVICE_PTRS | _lib.LC_X0_ZERO, C.byref(rep))
    assert rc == 1   # LC_EINVAL
    assert b"MiB device scratch" in psf.lib.lc_last_error()
    # host pointers are refused too
    h = np.zeros((1, N, N), np.float32)
    rc = psf.lib.lc_cg2d_noalloc(psf.ctx.handle, psf.handle, h.ctypes.data, h.ctypes.data, 1, 5,
                                 1e-5, 1e-5, 0, C.byref(rep))
    assert rc == 1
