"""laminocam_amd - MI355X (gfx950) LaminoCam reconstruction hot path.

Python mirror of the reference's C++ API over the C ABI of liblaminocam.so
(include/laminocam.h).  Names follow the reference:

  PolarGrid2D            include/polar_grid2d.h:35-66
  PointSpreadFunction2D  include/toeplitz2d.h:40-126
  PolarGrid              include/polar_grid.h:35-94
  PointSpreadFunction    include/toeplitz.h:35-130
  sysmat2d / sysmat      src/gradient.cpp:32-79
  forward2d / backproj2d src/projection.cpp:146-182
  cgsolver               src/conjgrad.cpp:33-76
  split_bregman          src/bregman.cpp:36-94
  grad_u / divergence / laplacian   src/finitediff.cpp:26-90, include/bregman.h:52-55
  prepare_sinogram, MBIR2D, MBIR    include/mbir2d_prep.h:44-65, src/mbir2d.cpp:66-100,
                                    src/mbir.cpp:34-161 (laminocam_amd.drivers)

All compute runs in HIP kernels; nothing here falls back to the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (LC_DEVICE_PTRS, LC_FP64, LC_GRAPH, LC_TV_ANISO, LC_TV_ZERO_PAD,
                   LC_X0_ZERO, LcError, LcReport)

__all__ = ["Context", "PolarGrid2D", "PointSpreadFunction2D", "PolarGrid", "PointSpreadFunction",
           "sysmat", "sysmat2d", "forward", "backproj", "forward2d",
           "backproj2d", "cgsolver",
           "split_bregman", "grad_u", "divergence", "laplacian", "Batch2D", "LcError",
           "default_context", "Server", "Dataset", "Regularizer", "ReconParams", "Sinogram2D",
           "prepare_sinogram", "MBIR2D", "MBIR", "estimate_memory", "volume_fits_one_gpu"]

_default_ctx = None


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _real(a, dtype=None):
    """Contiguous float array for an operator or solver call: float64 input
    (or dtype=float64) selects the FP64 path, like the reference's
    Array<double> instantiations; every other input runs in FP32."""
    a = np.asarray(a)
    if dtype is None:
        dtype = np.float64 if a.dtype == np.float64 else np.float32
    return np.ascontiguousarray(a, dtype=dtype)


def _entry(lib, name, a):
    """the C entry point for the element type of a (name or name_f64)"""
    return getattr(lib, name + "_f64" if a.dtype == np.float64 else name)


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


class Context:
    """One GPU and one HIP stream."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self.lib.lc_ctx_create(device, C.byref(h)))
        self.handle = h

    def sync(self):
        _lib.check(self.lib.lc_ctx_sync(self.handle))

    # ---- Toeplitz kernel shapes per image size ----
    def autotune(self, psf: "PointSpreadFunction2D", nslices: int = 16):
        """Time the candidate kernel shapes for psf's image size on `nslices`
        slices, install the fastest; returns (variant, ms per apply)."""
        v = (C.c_int32 * 3)()
        ms = C.c_double()
        _lib.check(self.lib.lc_autotune2d(self.handle, psf.handle, int(nslices), v, C.byref(ms)))
        return tuple(v), ms.value

    def set_tuning(self, image_n: int, variant):
        v = (C.c_int32 * 3)(*[int(a) for a in variant])
        _lib.check(self.lib.lc_set_tuning2d(self.handle, int(image_n), v))

    def get_tuning(self, image_n: int):
        v = (C.c_int32 * 3)()
        _lib.check(self.lib.lc_get_tuning2d(self.handle, int(image_n), v))
        return tuple(v)

    # ---- 2-D NUFFT spread bin shape per fine-grid size ----
    def autotune2d_nufft(self, psf: "PointSpreadFunction2D", nslices: int = 8):
        """Time the 2-D spread's compiled bin shapes on nslices slices of
        psf's grid, install the fastest for its fine-grid size; returns
        (shape index, ms per back-projection)."""
        sh = C.c_int32()
        ms = C.c_double()
        _lib.check(self.lib.lc_autotune2d_nufft(self.handle, psf.handle, int(nslices), C.byref(sh),
                                                C.byref(ms)))
        return sh.value, ms.value

    def set_tuning2d_nufft(self, psf: "PointSpreadFunction2D", shape: int):
        _lib.check(self.lib.lc_set_tuning2d_nufft(self.handle, psf.handle, int(shape)))

    def get_tuning2d_nufft(self, psf: "PointSpreadFunction2D") -> int:
        sh = C.c_int32()
        _lib.check(self.lib.lc_get_tuning2d_nufft(self.handle, psf.handle, C.byref(sh)))
        return sh.value

    # ---- 3-D NUFFT spread bin shape per grid ----
    def autotune3d(self, psf: "PointSpreadFunction"):
        """Time the 3-D spread's compiled bin shapes on psf's grid (its
        back-projection), install the fastest for that grid's fine-grid
        shape; returns (shape index, ms per back-projection)."""
        sh = C.c_int32()
        ms = C.c_double()
        _lib.check(self.lib.lc_autotune3d(self.handle, psf.handle, C.byref(sh), C.byref(ms)))
        return sh.value, ms.value

    def set_tuning3d(self, psf: "PointSpreadFunction", shape: int):
        _lib.check(self.lib.lc_set_tuning3d(self.handle, psf.handle, int(shape)))

    def get_tuning3d(self, psf: "PointSpreadFunction") -> int:
        sh = C.c_int32()
        _lib.check(self.lib.lc_get_tuning3d(self.handle, psf.handle, C.byref(sh)))
        return sh.value

    def load_tuning(self, path=None):
        """Install a tuning table {"N": [row, col, B], ...} (JSON)."""
        import json
        import os
        path = path or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                    "tuning_gfx950.json")
        if not os.path.exists(path):
            return {}
        with open(path) as f:
            table = json.load(f)
        for n, v in table.get("toeplitz2d", {}).items():
            self.set_tuning(int(n), v)
        return table

    def close(self):
        if self.handle:
            self.lib.lc_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_context() -> Context:
    """Context on device 0 with the shipped gfx950 kernel-shape table installed."""
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
        _default_ctx.load_tuning()
    return _default_ctx


class PolarGrid2D:
    """Radial polar grid: x = r cos(theta), y = r sin(theta),
    r_k = (k + 1/2) 2 pi / ncols - pi (include/polar_grid2d.h:50-65).
    The GPU generates the points on the fly; only theta and ncols are kept."""

    def __init__(self, theta, ncols: int):
        self.theta = np.ascontiguousarray(theta, dtype=np.float64)
        self.ncols = int(ncols)

    def nangles(self):
        return len(self.theta)

    def nradial(self):
        return self.ncols

    def dims(self):
        return (1, len(self.theta), self.ncols)

    def _geometry(self, image_n: int, ctx: "Context | None" = None):
        """device geometry for N x N images on this grid (PSF built on demand)"""
        ctx = ctx or default_context()
        key = (id(ctx), int(image_n))
        cache = self.__dict__.setdefault("_geoms", {})
        if key not in cache:
            h = C.c_void_p()
            _lib.check(ctx.lib.lc_geom2d_create(
                ctx.handle, self.theta.ctypes.data_as(C.POINTER(C.c_double)), len(self.theta),
                self.ncols, int(image_n), _lib.LC_GEOM_LAZY_PSF, C.byref(h)))
            cache[key] = _Geometry(ctx, h)
        return cache[key]


class _Geometry:
    """owning wrapper of an lc_geom2d handle"""

    def __init__(self, ctx, handle):
        self.ctx, self.handle = ctx, handle

    def __del__(self):
        try:
            if self.handle:
                self.ctx.lib.lc_geom2d_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def forward2d(volume, grid: PolarGrid2D, ctx: "Context | None" = None):
    """forward2d(volume, grid) (src/projection.cpp:146-160): sinogram
    {nslices, nangles, nradial} = Re(fftshift1(ifft1(ifftshift1(T2 volume))))/nradial
    for square slices {nslices, N, N}."""
    v = _f32(volume)
    if v.ndim == 2:
        v = v[None]
    if v.ndim != 3 or v.shape[1] != v.shape[2]:
        raise ValueError(f"forward2d: volume must be {{nslices, N, N}}, got {v.shape}")
    ns, N = v.shape[0], v.shape[1]
    g = grid._geometry(N, ctx)
    out = np.empty((ns, grid.nangles(), grid.nradial()), np.float32)
    _lib.check(g.ctx.lib.lc_forward2d(g.ctx.handle, g.handle, _ptr(v), ns, _ptr(out), 0))
    return out


def backproj2d(sinogram, grid: PolarGrid2D, recon_dims, ctx: "Context | None" = None):
    """backproj2d(sinogram, grid, recon_dims) (src/projection.cpp:168-182):
    Re(T1(ifftshift1(fft1(fftshift1(sinogram)))))/nradial on {nslices, N, N}."""
    s = _f32(sinogram)
    if s.ndim == 2:
        s = s[None]
    ns, na, nr = s.shape
    if (na, nr) != (grid.nangles(), grid.nradial()):
        raise ValueError(f"backproj2d: sinogram {s.shape} does not match the grid")
    d = tuple(int(x) for x in recon_dims)
    if len(d) != 3 or d[0] != ns or d[1] != d[2]:
        raise ValueError(f"backproj2d: recon_dims must be ({ns}, N, N), got {d}")
    g = grid._geometry(d[1], ctx)
    out = np.empty(d, np.float32)
    _lib.check(g.ctx.lib.lc_backproj2d(g.ctx.handle, g.handle, _ptr(s), ns, _ptr(out), 0))
    return out


class PointSpreadFunction2D:
    """Toeplitz kernel of A^T A for a 2-D polar grid, shared by every slice
    (include/toeplitz2d.h:81-125).  Built in FP64 on the GPU."""

    def __init__(self, grid: PolarGrid2D, image_size: int, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.grid = grid
        self.image_size = int(image_size)
        h = C.c_void_p()
        th = grid.theta
        _lib.check(self.lib.lc_geom2d_create(
            self.ctx.handle, th.ctypes.data_as(C.POINTER(C.c_double)), len(th), grid.ncols,
            self.image_size, 0, C.byref(h)))
        self.handle = h
        info = (C.c_int64 * 5)()
        _lib.check(self.lib.lc_geom2d_info(h, info))
        self.info = tuple(info)

    def psf_size(self):
        return self.info[3]

    def kernel_hat(self, dtype=np.complex64):
        """reference-layout spectrum {1, M, M/2+1}; dtype=complex128 for the
        PointSpreadFunction2D<double> form"""
        M, H = self.info[3], self.info[4]
        out = np.empty((1, M, H), dtype)
        fn = self.lib.lc_psf2d_kernel_hat_f64 if out.dtype == np.complex128 else \
            self.lib.lc_psf2d_kernel_hat
        _lib.check(fn(self.ctx.handle, self.handle, C.c_void_p(out.ctypes.data)))
        return out

    def convolve(self, x):
        """Toeplitz apply; float64 x runs the FP64 operator"""
        x = _real(x)
        N = self.image_size
        if x.shape[-1] != N or x.shape[-2] != N:
            raise ValueError(f"convolve: slices must be {N}x{N}, got {x.shape}")
        ns = x.size // (N * N)
        out = np.empty_like(x)
        _lib.check(_entry(self.lib, "lc_sysmat2d_psf", x)(self.ctx.handle, self.handle, _ptr(x),
                                                           ns, _ptr(out), 0))
        return out

    def close(self):
        if self.handle:
            self.lib.lc_geom2d_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PolarGrid:
    """3-D laminography grid: points R_i (qX_k, qY_j, 0), R_i =
    RotationTranspose(theta_i, gamma_i, beta_i), qX_k = (k + 1/2) 2 pi / ncols - pi,
    qY_j = (j + 1/2) 2 pi / nrows - pi (include/polar_grid.h:52-94,
    include/rotation.h:28-56).  Points are generated on the GPU; only the angles
    and detector dims are kept."""

    def __init__(self, theta, gamma, beta, nrows: int, ncols: int):
        self.theta = np.ascontiguousarray(theta, dtype=np.float64)
        n = len(self.theta)
        self.gamma = np.ascontiguousarray(np.broadcast_to(gamma, (n,)), dtype=np.float64)
        self.beta = np.ascontiguousarray(np.broadcast_to(beta, (n,)), dtype=np.float64)
        self.nrows, self.ncols = int(nrows), int(ncols)

    def dims(self):
        return (len(self.theta), self.nrows, self.ncols)

    def size(self):
        return int(np.prod(self.dims()))


class PointSpreadFunction:
    """3-D Toeplitz kernel of A^T A for a {n1, n2, n3} volume (n2 == n3)
    (include/toeplitz.h:35-130).  Built in FP64 on the GPU."""

    def __init__(self, grid: PolarGrid, recon_dims, ctx: Context | None = None, lazy: bool = False):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.grid = grid
        self.recon_dims = tuple(int(d) for d in recon_dims)
        rd = (C.c_int64 * 3)(*self.recon_dims)
        h = C.c_void_p()
        f64 = C.POINTER(C.c_double)
        _lib.check(self.lib.lc_geom3d_create(
            self.ctx.handle, grid.theta.ctypes.data_as(f64), grid.gamma.ctypes.data_as(f64),
            grid.beta.ctypes.data_as(f64), len(grid.theta), grid.nrows, grid.ncols, rd,
            _lib.LC_GEOM_LAZY_PSF if lazy else 0, C.byref(h)))
        self.handle = h
        info = (C.c_int64 * 9)()
        _lib.check(self.lib.lc_geom3d_info(h, info))
        self.info = tuple(info)

    def dims(self):
        return self.info[6:9]

    def kernel_hat(self, dtype=np.complex64):
        """reference-layout spectrum {d1, d2, d3/2+1}; dtype=complex128 for the
        PointSpreadFunction<double> form"""
        d1, d2, d3 = self.dims()
        out = np.empty((d1, d2, d3 // 2 + 1), dtype)
        fn = self.lib.lc_psf3d_kernel_hat_f64 if out.dtype == np.complex128 else \
            self.lib.lc_psf3d_kernel_hat
        _lib.check(fn(self.ctx.handle, self.handle, C.c_void_p(out.ctypes.data)))
        return out

    def convolve(self, x):
        """Toeplitz apply; float64 x runs the FP64 operator"""
        x = _real(x)
        if x.shape != self.recon_dims:
            raise ValueError(f"convolve: volume must be {self.recon_dims}, got {x.shape}")
        out = np.empty_like(x)
        _lib.check(_entry(self.lib, "lc_sysmat3d_psf", x)(self.ctx.handle, self.handle, _ptr(x),
                                                           _ptr(out), 0))
        return out

    def close(self):
        if self.handle:
            self.lib.lc_geom3d_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _geom3(grid: PolarGrid, recon_dims, ctx):
    """device geometry of a 3-D grid for recon_dims (PSF built on demand), cached on the grid"""
    ctx = ctx or default_context()
    key = (id(ctx), tuple(int(d) for d in recon_dims))
    cache = grid.__dict__.setdefault("_geoms", {})
    if key not in cache:
        cache[key] = PointSpreadFunction(grid, key[1], ctx=ctx, lazy=True)
    return cache[key]


def forward(volume, grid: PolarGrid, ctx: "Context | None" = None):
    """forward(volume, grid) (src/projection.cpp:42-64): projections
    {nproj, nrows, ncols} = Re(ifftshift2(ifft2(fftshift2(w . T2 volume))))/(nrows ncols)."""
    v = _f32(volume)
    if v.ndim != 3:
        raise ValueError("forward: volume must be {n1, n2, n3}")
    g = _geom3(grid, v.shape, ctx)
    out = np.empty(grid.dims(), np.float32)
    _lib.check(g.lib.lc_forward3d(g.ctx.handle, g.handle, _ptr(v), _ptr(out), 0))
    return out


def backproj(projections, grid: PolarGrid, recon_dims, shifts=None, ctx: "Context | None" = None):
    """backproj(projections, grid, recon_dims[, shifts]) (src/projection.cpp:71-131)."""
    p = _f32(projections)
    if p.shape != grid.dims():
        raise ValueError(f"backproj: projections {p.shape} do not match the grid {grid.dims()}")
    sh = None
    if shifts is not None:
        sh = np.ascontiguousarray(shifts, dtype=np.float64)
        if sh.shape != (p.shape[0], 2):
            # include/projection.h / src/projection.cpp:105-108 throw std::invalid_argument
            raise ValueError("shifts size must match the number of slices in the projection")
    g = _geom3(grid, recon_dims, ctx)
    out = np.empty(tuple(int(d) for d in recon_dims), np.float32)
    _lib.check(g.lib.lc_backproj3d(g.ctx.handle, g.handle, _ptr(p),
                                   None if sh is None else _ptr(sh), _ptr(out), 0))
    return out


def sysmat(x, op, ctx: "Context | None" = None):
    """sysmat(x, psf) = psf.convolve(x) (src/gradient.cpp:49-52);
    sysmat(x, grid) = Re(T1(T2 x))/(nrows ncols) (src/gradient.cpp:32-41)."""
    if isinstance(op, PointSpreadFunction):
        return op.convolve(x)
    if not isinstance(op, PolarGrid):
        raise TypeError("sysmat: second argument must be a PolarGrid or PointSpreadFunction")
    v = _f32(x)
    g = _geom3(op, v.shape, ctx)
    out = np.empty_like(v)
    _lib.check(g.lib.lc_sysmat3d_nufft(g.ctx.handle, g.handle, _ptr(v), _ptr(out), 0))
    return out


def sysmat2d(x, op, ctx: "Context | None" = None):
    """sysmat2d(x, psf) = psf.convolve(x) (src/gradient.cpp:76-79);
    sysmat2d(x, grid) = Re(T1(T2 x))/nradial (src/gradient.cpp:59-69)."""
    if isinstance(op, PointSpreadFunction2D):
        return op.convolve(x)
    if not isinstance(op, PolarGrid2D):
        raise TypeError("sysmat2d: second argument must be a PolarGrid2D or PointSpreadFunction2D")
    v = _f32(x)
    if v.ndim == 2:
        v = v[None]
    if v.ndim != 3 or v.shape[1] != v.shape[2]:
        raise ValueError(f"sysmat2d: x must be {{nslices, N, N}}, got {v.shape}")
    g = op._geometry(v.shape[1], ctx)
    out = np.empty_like(v)
    _lib.check(g.ctx.lib.lc_sysmat2d_nufft(g.ctx.handle, g.handle, _ptr(v), v.shape[0],
                                           _ptr(out), 0))
    return out


def _stack(x, N):
    x = _real(x)
    return x, x.size // (N * N)


def cgsolver(psf, y, x0, max_iter: int, tol: float, xtol: float, graph: bool = False):
    """opt::cgsolver with A = sysmat2d(., psf) on one stack (global alpha/beta
    over every slice, as the reference) or A = sysmat(., psf) on one 3-D
    volume.  float64 y runs cgsolver<double> (FP64 vectors, spectrum and
    FFTs).  Returns (x, report dict)."""
    rep = LcReport()
    flags = LC_GRAPH if graph else 0
    if isinstance(psf, PointSpreadFunction):
        y = _real(y)
        x = _real(x0, y.dtype).reshape(y.shape).copy()
        if not np.any(x):
            flags |= LC_X0_ZERO
        _lib.check(_entry(psf.lib, "lc_cg3d", y)(psf.ctx.handle, psf.handle, _ptr(y), _ptr(x),
                                                 int(max_iter), float(tol), float(xtol), flags,
                                                 C.byref(rep)))
        return x, rep.as_dict()
    y, ns = _stack(y, psf.image_size)
    x = _real(x0, y.dtype).copy()
    if not np.any(x):
        flags |= LC_X0_ZERO
    _lib.check(_entry(psf.lib, "lc_cg2d", y)(psf.ctx.handle, psf.handle, _ptr(y), _ptr(x), ns,
                                             int(max_iter), float(tol), float(xtol), flags,
                                             C.byref(rep)))
    return x, rep.as_dict()


def split_bregman(psf, y, x0, lam: float, mu: float, outer_max: int,
                  inner_max: int, tol: float, xtol: float, zero_pad: bool = False,
                  aniso: bool = False, graph: bool = False):
    """opt::split_bregman with A = sysmat2d(., psf).  zero_pad selects the
    zero-padded finite differences of the reference's GPU path instead of
    the interior-only ones of its CPU path.  float64 y runs
    split_bregman<double>."""
    rep = LcReport()
    flags = (LC_GRAPH if graph else 0) | (LC_TV_ZERO_PAD if zero_pad else 0) | \
        (LC_TV_ANISO if aniso else 0)
    if isinstance(psf, PointSpreadFunction):
        y = _real(y)
        x = _real(x0, y.dtype).reshape(y.shape).copy()
        _lib.check(_entry(psf.lib, "lc_split_bregman3d", y)(
            psf.ctx.handle, psf.handle, _ptr(y), _ptr(x), float(lam), float(mu), int(outer_max),
            int(inner_max), float(tol), float(xtol), flags, C.byref(rep)))
        return x, rep.as_dict()
    y, ns = _stack(y, psf.image_size)
    x = _real(x0, y.dtype).copy()
    _lib.check(_entry(psf.lib, "lc_split_bregman2d", y)(
        psf.ctx.handle, psf.handle, _ptr(y), _ptr(x), ns, float(lam), float(mu), int(outer_max),
        int(inner_max), float(tol), float(xtol), flags, C.byref(rep)))
    return x, rep.as_dict()


def _vol3(u):
    u = _real(u)
    if u.ndim == 2:
        return u, (1,) + u.shape
    if u.ndim != 3:
        raise ValueError("expected a 2-D or 3-D array")
    return u, u.shape


def grad_u(u, zero_pad: bool = False, ctx: Context | None = None):
    ctx = ctx or default_context()
    u, d = _vol3(u)
    g = np.empty((3,) + d, u.dtype)
    _lib.check(_entry(ctx.lib, "lc_grad", u)(ctx.handle, _ptr(u), *d, _ptr(g),
                                             LC_TV_ZERO_PAD if zero_pad else 0))
    return g


def divergence(g, zero_pad: bool = False, ctx: Context | None = None):
    ctx = ctx or default_context()
    g = _real(g)
    d = g.shape[1:]
    if len(d) == 2:
        d = (1,) + d
    out = np.empty(d, g.dtype)
    _lib.check(_entry(ctx.lib, "lc_divergence", g)(ctx.handle, _ptr(g), *d, _ptr(out),
                                                   LC_TV_ZERO_PAD if zero_pad else 0))
    return out


def laplacian(u, zero_pad: bool = False, ctx: Context | None = None):
    ctx = ctx or default_context()
    u, d = _vol3(u)
    out = np.empty(d, u.dtype)
    _lib.check(_entry(ctx.lib, "lc_laplacian", u)(ctx.handle, _ptr(u), *d, _ptr(out),
                                                  LC_TV_ZERO_PAD if zero_pad else 0))
    return out


class Batch2D:
    """Many independent 2-D problems solved together with per-problem CG
    scalars; vectors stay resident on the GPU."""

    X, Y, AX = 0, 1, 2

    def __init__(self, psfs, ns=None, ctx: Context | None = None, fp64: bool = False):
        """psfs: PointSpreadFunction2D (a stack of ns[i] slices) or
        PointSpreadFunction (one 3-D volume) per problem.  fp64: FP64
        vectors, spectra and FFTs (set/get then take float64 arrays)."""
        self.ctx = ctx or psfs[0].ctx
        self.lib = self.ctx.lib
        self.psfs = list(psfs)
        n = len(self.psfs)
        self.ns = [1] * n if ns is None else [int(v) for v in ns]
        is3 = [isinstance(p, PointSpreadFunction) for p in self.psfs]
        for i, p in enumerate(self.psfs):
            if is3[i]:
                self.ns[i] = p.recon_dims[0]
        g2 = (C.c_void_p * n)(*[None if is3[i] else p.handle.value for i, p in enumerate(self.psfs)])
        g3 = (C.c_void_p * n)(*[p.handle.value if is3[i] else None for i, p in enumerate(self.psfs)])
        nss = (C.c_int32 * n)(*self.ns)
        h = C.c_void_p()
        self.fp64 = bool(fp64)
        self.dtype = np.float64 if self.fp64 else np.float32
        _lib.check(self.lib.lc_batch_create_ex(self.ctx.handle, n, g2, g3, nss,
                                               LC_FP64 if self.fp64 else 0, C.byref(h)))
        self.handle = h

    def __len__(self):
        return len(self.psfs)

    def shape(self, i):
        p = self.psfs[i]
        if isinstance(p, PointSpreadFunction):
            return p.recon_dims
        return (self.ns[i], p.image_size, p.image_size)

    def set(self, i, which, arr):
        a = _real(arr, self.dtype)
        if a.size != int(np.prod(self.shape(i))):
            raise ValueError("size mismatch")
        _lib.check(_entry(self.lib, "lc_batch2d_set", a)(self.handle, i, which, _ptr(a), 0))

    def get(self, i, which):
        out = np.empty(self.shape(i), self.dtype)
        _lib.check(_entry(self.lib, "lc_batch2d_get", out)(self.handle, i, which, _ptr(out), 0))
        return out

    def sysmat(self, src=0, dst=2):
        _lib.check(self.lib.lc_batch2d_sysmat(self.handle, src, dst))

    def cg(self, max_iter, tol=1e-5, xtol=1e-5, graph=True, x0_zero=True, host_sync=False):
        """host_sync (with graph=False): the one-problem entry points' mode,
        scalars read back after every iteration"""
        n = len(self)
        mi = [int(max_iter)] * n if np.isscalar(max_iter) else [int(v) for v in max_iter]
        arr = (C.c_int32 * n)(*mi)
        flags = (LC_GRAPH if graph else 0) | (LC_X0_ZERO if x0_zero else 0) | \
            (_lib.LC_HOST_SYNC if host_sync else 0)
        _lib.check(self.lib.lc_batch2d_cg(self.handle, arr, float(tol), float(xtol), flags))

    def split_bregman(self, lam, mu, outer, inner, tol=1e-5, xtol=1e-5, zero_pad=False,
                      aniso=False, graph=True, x0_zero=False, host_sync=False):
        n = len(self)
        la = np.ascontiguousarray(np.broadcast_to(np.float32(lam), (n,)) if np.isscalar(lam) else lam,
                                  dtype=np.float32)
        m = np.ascontiguousarray(np.broadcast_to(np.float32(mu), (n,)) if np.isscalar(mu) else mu,
                                 dtype=np.float32)
        flags = (LC_GRAPH if graph else 0) | (LC_TV_ZERO_PAD if zero_pad else 0) | \
            (LC_TV_ANISO if aniso else 0) | (LC_X0_ZERO if x0_zero else 0) | \
            (_lib.LC_HOST_SYNC if host_sync else 0)
        _lib.check(self.lib.lc_batch2d_split_bregman(
            self.handle, la.ctypes.data_as(C.POINTER(C.c_float)),
            m.ctypes.data_as(C.POINTER(C.c_float)), int(outer), int(inner), float(tol),
            float(xtol), flags))

    def reports(self):
        n = len(self)
        reps = (LcReport * n)()
        _lib.check(self.lib.lc_batch2d_reports(self.handle, reps))
        return [r.as_dict() for r in reps]

    def profile_passes(self, reps=5):
        """HIP-event duration of every Toeplitz pass launch as a CG iteration
        issues it: list of dicts (N, pass, nslices, ms, bytes).  Clobbers the
        solver state and Ax, keeps x and y."""
        cap = 3 * 64 + 1
        rows = (C.c_double * (5 * cap))()
        n = C.c_int32()
        _lib.check(self.lib.lc_batch2d_profile(self.handle, int(reps), cap, rows, C.byref(n)))
        out = []
        for i in range(n.value):
            r = rows[5 * i:5 * i + 5]
            out.append(dict(N=int(r[0]), pass_=int(r[1]), nslices=int(r[2]), ms=r[3], bytes=r[4]))
        return out

    def device_ptr(self, i, which):
        p = C.c_void_p()
        fn = self.lib.lc_batch2d_device_ptr_f64 if self.fp64 else self.lib.lc_batch2d_device_ptr
        _lib.check(fn(self.handle, i, which, C.byref(p)))
        return p.value

    def close(self):
        if self.handle:
            self.lib.lc_batch2d_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _request_arrays(psf, sino, rhs, x0):
    """Shapes of one server request, checked before any pointer reaches the
    library, which copies (or, for a borrowed request, later reads) exactly
    nslices x the geometry's sizes from each array.  Returns (nslices, the
    solution's shape, {field: contiguous float32 array})."""
    if (sino is None) == (rhs is None):
        raise ValueError("give exactly one of sino and rhs")
    arrays = {}
    if isinstance(psf, PointSpreadFunction):
        shape = tuple(psf.recon_dims)
        ns = shape[0]
    else:
        a = np.asarray(sino if sino is not None else rhs)
        ns = int(a.shape[0]) if a.ndim == 3 else 1
        shape = (ns, psf.image_size, psf.image_size)
    if sino is not None:
        s = _f32(sino)
        if isinstance(psf, PointSpreadFunction):
            want = tuple(psf.grid.dims())
        else:
            want = (ns, psf.grid.nangles(), psf.grid.nradial())
        got = s.shape if s.ndim == 3 else (1,) + s.shape
        if tuple(got) != want:
            raise ValueError(f"sinogram shape {s.shape} does not match {want}")
        arrays["sino"] = s
    else:
        y = _f32(rhs)
        if y.size != int(np.prod(shape)):
            raise ValueError(f"rhs size {y.size} does not match {shape}")
        arrays["rhs"] = y
    if x0 is not None:
        x = _f32(x0)
        if x.size != int(np.prod(shape)):
            raise ValueError(f"x0 size {x.size} does not match {shape}")
        arrays["x0"] = x
    return ns, shape, arrays


class Server:
    """Reconstruction server: continuous batching with HBM-budget admission
    over `lanes` device lanes (lc_server_* in laminocam.h).  submit() copies
    the inputs (or, with borrow=True, keeps a reference to them and the lane
    reads them in place: do not modify them until the result is taken) and
    returns a ticket; the lane writes the solution straight into an array
    that result() then returns with (report, (t_submit, t_start, t_done))."""

    def __init__(self, hbm_budget: int = 64 << 30, lanes: int = 2, max_batch: int = 128,
                 ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        h = C.c_void_p()
        _lib.check(self.lib.lc_server_create(self.ctx.handle, int(hbm_budget), int(lanes),
                                             int(max_batch), C.byref(h)))
        self.handle = h
        self._held = {}   # ticket -> (output array, borrowed inputs)

    def submit(self, psf, *, sino=None, rhs=None, x0=None, solver="cg", max_iter=50,
               tol=1e-5, xtol=1e-5, lam=0.1, mu=10.0, outer=1, inner=1, zero_pad=False,
               aniso=False, borrow=False, test_fail=False):
        """psf: PointSpreadFunction2D (a stack) or PointSpreadFunction (a volume).
        sino: sinogram {ns, nang, nrad} / projections {nproj, nrows, ncols}
        (back-projected on the device), or rhs: the right-hand side.
        test_fail (a testing aid) makes the lane raise a host error on a batch
        holding this request, before any device work."""
        ns, shape, arrays = _request_arrays(psf, sino, rhs, x0)
        r = _lib.LcRequest()
        if isinstance(psf, PointSpreadFunction):
            r.g3 = psf.handle.value
        else:
            r.g2 = psf.handle.value
        r.nslices = ns
        r.solver = 0 if solver == "cg" else 1
        r.max_iter, r.outer, r.inner = int(max_iter), int(outer), int(inner)
        r.tol, r.xtol, r.lam, r.mu = float(tol), float(xtol), float(lam), float(mu)
        r.flags = (LC_TV_ZERO_PAD if zero_pad else 0) | (LC_TV_ANISO if aniso else 0) | \
            (_lib.LC_REQ_BORROW if borrow else 0) | (_lib.LC_REQ_TEST_FAIL if test_fail else 0)
        keep = []
        for name, a in arrays.items():
            keep.append(a)
            setattr(r, name, a.ctypes.data)
        out = np.empty(shape, np.float32)
        r.x_out = out.ctypes.data
        t = C.c_uint64()
        _lib.check(self.lib.lc_server_submit(self.handle, C.byref(r), C.byref(t)))
        self._held[t.value] = (out, keep if borrow else None)
        return t.value

    def request_bytes(self, psf, *, sino=None, rhs=None, solver="cg"):
        """device bytes the server charges against its budget for such a
        request (lc_request_bytes)"""
        ns, _, _ = _request_arrays(psf, sino, rhs, None)
        r = _lib.LcRequest()
        if isinstance(psf, PointSpreadFunction):
            r.g3 = psf.handle.value
        else:
            r.g2 = psf.handle.value
        r.nslices = ns
        r.solver = 0 if solver == "cg" else 1
        keep = None
        if sino is not None:
            keep = _f32(sino)
            r.sino = keep.ctypes.data
        out = C.c_uint64()
        _lib.check(self.lib.lc_request_bytes(C.byref(r), C.byref(out)))
        return int(out.value)

    def done(self, ticket) -> bool:
        d = C.c_int32()
        _lib.check(self.lib.lc_server_poll(self.handle, C.c_uint64(ticket), C.byref(d)))
        return bool(d.value)

    def result(self, ticket, timeout: float = -1.0):
        d = C.c_int32()
        _lib.check(self.lib.lc_server_wait(self.handle, C.c_uint64(ticket), float(timeout),
                                           C.byref(d)))
        if not d.value:
            raise TimeoutError(f"request {ticket} not finished after {timeout} s")
        x, _ = self._held.pop(ticket)
        rep = LcReport()
        times = (C.c_double * 3)()
        _lib.check(self.lib.lc_server_result(self.handle, C.c_uint64(ticket), None,
                                             C.byref(rep), times))
        return x, rep.as_dict(), tuple(times)

    def stats(self):
        s = (C.c_double * 10)()
        _lib.check(self.lib.lc_server_stats(self.handle, s))
        return dict(inflight_bytes=s[0], queued=s[1], running=s[2], finished=s[3],
                    batches=s[4], budget=s[5], lane_s_build=s[6], lane_s_input=s[7],
                    lane_s_solve=s[8], lane_s_output=s[9])

    def timeline(self, max_rows: int = 65536):
        """The most recent batches, oldest first: a float64 array of rows
        (lane, requests, host start s, host end s, device solve start ms,
        device solve end ms); device times count from the server's creation."""
        rows = np.zeros((max_rows, 6), np.float64)
        n = C.c_int32()
        _lib.check(self.lib.lc_server_timeline(self.handle, int(max_rows),
                                               rows.ctypes.data_as(C.POINTER(C.c_double)),
                                               C.byref(n)))
        return rows[:n.value].copy()

    def close(self):
        if self.handle:
            self.lib.lc_server_destroy(self.handle)   # finishes every queued request
            self.handle = None
            self._held.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


from .drivers import (MBIR, MBIR2D, Dataset, ReconParams, Regularizer,  # noqa: E402
                      Sinogram2D, estimate_memory, prepare_sinogram, volume_fits_one_gpu)
