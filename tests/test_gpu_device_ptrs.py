"""LC_DEVICE_PTRS: the entry points read and write device buffers in place of
host arrays, with bitwise the same results as the host-pointer calls.  Device
buffers come from the HIP runtime directly (hipMalloc / hipMemcpy)."""
import ctypes as C

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Dev:
    """a device copy of a numpy array (or an uninitialised buffer of a shape)"""

    hip = None

    def __init__(self, a=None, shape=None, dtype=np.float32):
        if Dev.hip is None:
            Dev.hip = C.CDLL("libamdhip64.so")
        self.shape = a.shape if a is not None else tuple(shape)
        self.dtype = a.dtype if a is not None else np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.p = C.c_void_p()
        assert Dev.hip.hipMalloc(C.byref(self.p), C.c_size_t(self.nbytes)) == 0
        if a is not None:
            a = np.ascontiguousarray(a)
            assert Dev.hip.hipMemcpy(self.p, C.c_void_p(a.ctypes.data), C.c_size_t(self.nbytes), H2D) == 0

    def get(self):
        out = np.empty(self.shape, self.dtype)
        assert Dev.hip.hipMemcpy(C.c_void_p(out.ctypes.data), self.p, C.c_size_t(self.nbytes), D2H) == 0
        return out

    def __del__(self):
        if self.p:
            Dev.hip.hipFree(self.p)


@pytest.fixture(scope="module")
def lc():
    import laminocam_amd
    return laminocam_amd


def test_projectors2d_device_ptrs(lc):
    announce("2-D projectors, device pointers")
    from laminocam_amd import _lib
    rng = np.random.default_rng(3)
    N, nang, ns = 64, 90, 3
    g = lc.PolarGrid2D(np.arange(nang) * np.pi / nang, N)
    geo = g._geometry(N)
    vol = rng.random((ns, N, N)).astype(np.float32)
    sino_h = lc.forward2d(vol, g)
    dv, ds = Dev(vol), Dev(shape=(ns, nang, N))
    _lib.check(geo.ctx.lib.lc_forward2d(geo.ctx.handle, geo.handle, dv.p, ns, ds.p, _lib.LC_DEVICE_PTRS))
    assert np.array_equal(ds.get(), sino_h)
    bp_h = lc.backproj2d(sino_h, g, (ns, N, N))
    db = Dev(shape=(ns, N, N))
    _lib.check(geo.ctx.lib.lc_backproj2d(geo.ctx.handle, geo.handle, ds.p, ns, db.p, _lib.LC_DEVICE_PTRS))
    assert np.array_equal(db.get(), bp_h)


def test_cg2d_device_ptrs(lc):
    announce("cg2d, device pointers")
    from laminocam_amd import _lib
    rng = np.random.default_rng(4)
    N, ns = 48, 2
    g = lc.PolarGrid2D(np.arange(60) * np.pi / 60, N)
    psf = lc.PointSpreadFunction2D(g, N)
    rhs = psf.convolve(rng.random((ns, N, N)).astype(np.float32))
    xh, _ = lc.cgsolver(psf, rhs, np.zeros_like(rhs), 7, 1e-5, 1e-5)
    dy, dx = Dev(rhs), Dev(shape=rhs.shape)
    rep = _lib.LcReport()
    _lib.check(psf.lib.lc_cg2d(psf.ctx.handle, psf.handle, dy.p, dx.p, ns, 7, 1e-5, 1e-5,
                               _lib.LC_DEVICE_PTRS | _lib.LC_X0_ZERO, C.byref(rep)))
    assert rep.iters == 7
    assert np.array_equal(dx.get(), xh)


def test_drivers_device_ptrs(lc):
    announce("drivers, device pointers")
    from laminocam_amd import _lib
    rng = np.random.default_rng(5)
    ctx = lc.default_context()
    L = ctx.lib
    # prepare_sinogram over two device datasets
    p0 = (rng.random((7, 4, 32)) * 3).astype(np.float32)
    p1 = (rng.random((5, 4, 32)) * 3).astype(np.float32)
    ref, theta = lc.prepare_sinogram([lc.Dataset(p0, np.arange(7) * 0.2),
                                      lc.Dataset(p1, np.arange(5) * 0.2 + 0.1)])
    d0, d1 = Dev(p0), Dev(p1)
    ds, dm = Dev(shape=ref.shape), Dev(shape=(1,))
    ptrs = (C.c_void_p * 2)(d0.p.value, d1.p.value)
    nang = (C.c_int64 * 2)(7, 5)
    _lib.check(L.lc_prepare_sinogram(ctx.handle, 2, ptrs, nang, 4, 32, ds.p, dm.p, _lib.LC_DEVICE_PTRS))
    assert np.array_equal(ds.get(), ref)
    assert dm.get()[0] == max(p0.max(), p1.max())
    # the 2-D driver on the device sinogram
    params = lc.ReconParams(recon_dims=(4, 32, 32), maxIters=5)
    xh = lc.MBIR2D(ref, theta, params)
    dxr = Dev(shape=(4, 32, 32))
    th = np.ascontiguousarray(theta, np.float64)
    rep = _lib.LcReport()
    _lib.check(L.lc_mbir2d(ctx.handle, ds.p, 4, 12, 32, th.ctypes.data_as(C.POINTER(C.c_double)),
                           C.byref(params.to_c()), dxr.p, C.byref(rep), _lib.LC_DEVICE_PTRS))
    assert np.array_equal(dxr.get(), xh)


def test_mbir3d_device_ptrs_and_shifts(lc):
    announce("3-D driver, device pointers and shifts")
    from laminocam_amd import _lib
    rng = np.random.default_rng(6)
    ctx = lc.default_context()
    nproj, nrows, ncols = 9, 6, 16
    proj = rng.random((nproj, nrows, ncols)).astype(np.float32)
    ang = np.arange(nproj) * 2 * np.pi / nproj
    sh = rng.uniform(-1, 1, (nproj, 2))
    params = lc.ReconParams(recon_dims=(3, 12, 12), maxIters=4)
    xh = lc.MBIR([lc.Dataset(proj, ang, np.deg2rad(-30.0), 0.0, sh)], params)
    dp, dsh, dx = Dev(proj), Dev(np.ascontiguousarray(sh, np.float64)), Dev(shape=(3, 12, 12))
    dpt = C.POINTER(C.c_double)
    th = np.ascontiguousarray(ang, np.float64)
    ga, be = np.array([np.deg2rad(-30.0)]), np.array([0.0])
    _lib.check(ctx.lib.lc_mbir3d(ctx.handle, 1, (C.c_void_p * 1)(dp.p.value), (C.c_int64 * 1)(nproj),
                                 nrows, ncols, th.ctypes.data_as(dpt), ga.ctypes.data_as(dpt),
                                 be.ctypes.data_as(dpt), C.cast(dsh.p, dpt), C.byref(params.to_c()),
                                 dx.p, None, None, _lib.LC_DEVICE_PTRS))
    assert np.array_equal(dx.get(), xh)
