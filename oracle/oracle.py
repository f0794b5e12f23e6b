"""ctypes wrapper of the CPU oracle (test infrastructure only).

Only tests/, the smoke check in __graft_entry__.py and the cpu_baseline leg of
bench.py import this module.  The product path (laminocam_amd) never does.

``Oracle(64)`` wraps liboracle64.so (REAL=double, the FP64 oracle);
``Oracle(32)`` wraps liboracle32.so (REAL=float, the FP32 CPU baseline that
restates the reference's float recon_tomo path).  Function names follow the
reference's C++ API (src/projection.cpp, include/toeplitz2d.h, ...).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

NUFFT_TOL64 = 1e-12   # include/finufft_plan.h:48 (double)
NUFFT_TOL32 = 5e-4    # include/finufft_plan.h:71 (float)


def build():
    """Compile the oracle libraries in place (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class Report(C.Structure):
    _fields_ = [("iters", C.c_int64), ("outer", C.c_int64), ("res", C.c_double),
                ("dx", C.c_double), ("stop", C.c_int)]

    def as_dict(self):
        return {"iters": self.iters, "outer": self.outer, "res": self.res,
                "dx": self.dx, "stop": self.stop}


class _PsfOpCtx(C.Structure):
    _fields_ = [("psf", C.c_void_p), ("n1", C.c_int64), ("n2", C.c_int64),
                ("n3", C.c_int64)]


class Oracle:
    def __init__(self, bits: int = 64):
        if bits not in (32, 64):
            raise ValueError("bits must be 32 or 64")
        path = os.path.join(_HERE, f"liboracle{bits}.so")
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        self.bits = bits
        self.dtype = np.float64 if bits == 64 else np.float32
        self.ctype = C.c_double if bits == 64 else C.c_float
        self.cdtype = np.complex128 if bits == 64 else np.complex64
        self.default_tol = NUFFT_TOL64 if bits == 64 else NUFFT_TOL32
        P = C.POINTER(self.ctype)
        L = self.lib
        i64 = C.c_int64
        L.or_next_fast_dim.restype = i64
        L.or_next_fast_dim.argtypes = [i64]
        L.or_fft_lines.argtypes = [P, i64, i64, i64, i64, C.c_int]
        L.or_fft2.argtypes = [P, i64, i64, i64, C.c_int]
        L.or_fft3.argtypes = [P, i64, i64, i64, C.c_int]
        L.or_dft_naive.argtypes = [P, P, i64, C.c_int]
        L.or_shift.argtypes = [P, P, i64, i64, i64, C.c_int, C.c_int, C.c_int]
        L.or_polar_grid2d.argtypes = [P, i64, i64, P, P]
        L.or_polar_grid3d.argtypes = [P, P, P, i64, i64, i64, P, P, P, P]
        L.or_rotation_transpose.argtypes = [C.c_double, C.c_double, C.c_double,
                                            C.POINTER(C.c_double)]
        L.or_nudft2d.argtypes = [C.c_int, P, P, i64, P, P, i64, i64]
        L.or_nudft3d.argtypes = [C.c_int, P, P, P, i64, P, P, i64, i64, i64]
        L.or_nufft2d.argtypes = [C.c_int, P, P, i64, P, P, i64, i64, C.c_double]
        L.or_nufft3d.argtypes = [C.c_int, P, P, P, i64, P, P, i64, i64, i64,
                                 C.c_double]
        L.or_es_params.argtypes = [C.c_double, C.POINTER(C.c_int),
                                   C.POINTER(C.c_double)]
        L.or_forward2d.argtypes = [P, i64, i64, P, i64, i64, P, C.c_double]
        L.or_backproj2d.argtypes = [P, i64, i64, i64, P, i64, P, C.c_double]
        L.or_sysmat2d_nufft.argtypes = [P, i64, i64, P, i64, i64, P, C.c_double]
        L.or_forward3d.argtypes = [P, i64, i64, i64, P, P, P, i64, i64, i64, P,
                                   C.c_double]
        L.or_backproj3d.argtypes = [P, i64, i64, i64, P, P, P, P, i64, i64, i64,
                                    P, C.c_double]
        L.or_sysmat3d_nufft.argtypes = [P, i64, i64, i64, P, P, P, i64, i64, i64,
                                        P, C.c_double]
        L.or_psf2d_create.restype = C.c_void_p
        L.or_psf2d_create.argtypes = [P, i64, i64, i64, C.c_int, C.c_double]
        L.or_psf3d_create.restype = C.c_void_p
        L.or_psf3d_create.argtypes = [P, P, P, i64, i64, i64, i64, i64, i64,
                                      C.c_int, C.c_double]
        L.or_psf_destroy.argtypes = [C.c_void_p]
        L.or_psf_from_real.restype = C.c_void_p
        L.or_psf_from_real.argtypes = [C.c_int, C.POINTER(i64), P]
        L.or_psf_dims.argtypes = [C.c_void_p, C.POINTER(i64)]
        L.or_psf_real.argtypes = [C.c_void_p, P]
        L.or_psf_kernel_hat.argtypes = [C.c_void_p, P]
        L.or_psf_convolve.argtypes = [C.c_void_p, P, i64, i64, i64, P]
        L.or_grad.argtypes = [P, i64, i64, i64, C.c_int, P]
        L.or_div.argtypes = [P, i64, i64, i64, C.c_int, P]
        L.or_laplacian.argtypes = [P, i64, i64, i64, C.c_int, P]
        L.or_cgsolver.argtypes = [C.c_void_p, C.c_void_p, P, P, i64, i64,
                                  C.c_double, C.c_double, P, C.POINTER(Report)]
        L.or_cgsolver_trace.argtypes = [C.c_void_p, C.c_void_p, P, P, i64, i64,
                                        C.c_double, C.c_double, P, C.POINTER(Report),
                                        C.POINTER(i64), i64, P]
        L.or_split_bregman.argtypes = [C.c_void_p, C.c_void_p, P, P, i64, i64,
                                       i64, C.c_double, C.c_double, i64, i64,
                                       C.c_double, C.c_double, C.c_int, C.c_int,
                                       P, C.POINTER(Report)]
        L.or_split_bregman_state.argtypes = [C.c_void_p, C.c_void_p, P, P, i64, i64,
                                             i64, C.c_double, C.c_double, i64, i64,
                                             C.c_double, C.c_double, C.c_int, C.c_int,
                                             P, P, P, C.POINTER(Report)]
        L.or_sb_rhs.argtypes = [P, P, P, i64, i64, i64, C.c_int, C.c_double, P]
        L.or_sb_shrink.argtypes = [P, P, i64, C.c_double, C.c_double, C.c_int, P]
        L.or_sb_bregman_update.argtypes = [P, P, P, i64]
        L.or_num_threads.restype = C.c_int
        L.or_set_num_threads.argtypes = [C.c_int]
        self._opfn_type = C.CFUNCTYPE(None, C.c_void_p, P, P)
        self.psf_apply_op = C.cast(L.or_psf_apply_op, C.c_void_p)

    # ---- helpers ----
    def _r(self, a):
        return np.ascontiguousarray(a, dtype=self.dtype)

    def _p(self, a):
        return a.ctypes.data_as(C.POINTER(self.ctype))

    def _cview(self, a):
        """complex array -> contiguous complex array of the oracle's precision"""
        return np.ascontiguousarray(a, dtype=self.cdtype)

    def num_threads(self):
        return self.lib.or_num_threads()

    def set_num_threads(self, n):
        self.lib.or_set_num_threads(int(n))

    # ---- FFT ----
    def next_fast_dim(self, n):
        return int(self.lib.or_next_fast_dim(int(n)))

    def fft(self, a, sign=-1, axis=-1):
        a = np.moveaxis(self._cview(a), axis, -1).copy()
        n = a.shape[-1]
        how = a.size // n if n else 0
        self.lib.or_fft_lines(self._p(a.view(self.dtype)), n, how, 1, n, sign)
        return np.moveaxis(a, -1, axis)

    def fft2(self, a, sign=-1):
        a = self._cview(a).copy()
        sh = a.shape
        b = a.reshape(-1, sh[-2], sh[-1])
        self.lib.or_fft2(self._p(b.view(self.dtype)), b.shape[0], sh[-2], sh[-1], sign)
        return b.reshape(sh)

    def fft3(self, a, sign=-1):
        a = self._cview(a).copy()
        self.lib.or_fft3(self._p(a.view(self.dtype)), *a.shape, sign)
        return a

    def dft_naive(self, a, sign=-1):
        a = self._cview(a)
        out = np.empty_like(a)
        self.lib.or_dft_naive(self._p(a.view(self.dtype)), self._p(out.view(self.dtype)),
                              a.size, sign)
        return out

    def shift(self, a, axes=1, inverse=False):
        """fftshift/ifftshift over the last one or two axes, reference convention"""
        iscx = np.iscomplexobj(a)
        a = self._cview(a) if iscx else self._r(a)
        sh = a.shape
        a3 = a.reshape((-1,) + sh[-2:]) if a.ndim >= 2 else a.reshape(1, 1, -1)
        out = np.empty_like(a3)
        src = a3.view(self.dtype) if iscx else a3
        dst = out.view(self.dtype) if iscx else out
        self.lib.or_shift(self._p(src), self._p(dst), a3.shape[0], a3.shape[1],
                          a3.shape[2], axes, int(inverse), 2 if iscx else 1)
        return out.reshape(sh)

    # ---- grids ----
    def polar_grid2d(self, theta, nrad):
        th = self._r(theta)
        x = np.empty((len(th), nrad), self.dtype)
        y = np.empty_like(x)
        self.lib.or_polar_grid2d(self._p(th), len(th), nrad, self._p(x), self._p(y))
        return x, y

    def polar_grid3d(self, theta, gamma, beta, nrows, ncols):
        th, ga, be = self._r(theta), self._r(gamma), self._r(beta)
        sh = (len(th), nrows, ncols)
        x, y, z, w = (np.empty(sh, self.dtype) for _ in range(4))
        self.lib.or_polar_grid3d(self._p(th), self._p(ga), self._p(be), len(th),
                                 nrows, ncols, self._p(x), self._p(y), self._p(z),
                                 self._p(w))
        return x, y, z, w

    def rotation_transpose(self, theta, gamma, beta):
        R = np.empty(9, np.float64)
        self.lib.or_rotation_transpose(theta, gamma, beta,
                                       R.ctypes.data_as(C.POINTER(C.c_double)))
        return R.reshape(3, 3)

    # ---- NUDFT / NUFFT ----
    def nudft2d(self, typ, x, y, data, nrows=None, ncols=None):
        x, y = self._r(x).ravel(), self._r(y).ravel()
        if typ == 1:
            c = self._cview(data).ravel()
            f = np.empty((nrows, ncols), self.cdtype)
        else:
            f = self._cview(data)
            nrows, ncols = f.shape
            c = np.empty(x.size, self.cdtype)
        self.lib.or_nudft2d(typ, self._p(x), self._p(y), x.size,
                            self._p(c.view(self.dtype)), self._p(f.view(self.dtype)),
                            nrows, ncols)
        return f if typ == 1 else c

    def nufft2d(self, typ, x, y, data, nrows=None, ncols=None, tol=None):
        tol = self.default_tol if tol is None else tol
        x, y = self._r(x).ravel(), self._r(y).ravel()
        if typ == 1:
            c = self._cview(data).ravel()
            f = np.empty((nrows, ncols), self.cdtype)
        else:
            f = self._cview(data)
            nrows, ncols = f.shape
            c = np.empty(x.size, self.cdtype)
        self.lib.or_nufft2d(typ, self._p(x), self._p(y), x.size,
                            self._p(c.view(self.dtype)), self._p(f.view(self.dtype)),
                            nrows, ncols, tol)
        return f if typ == 1 else c

    def nudft3d(self, typ, x, y, z, data, shape=None):
        x, y, z = (self._r(v).ravel() for v in (x, y, z))
        if typ == 1:
            c = self._cview(data).ravel()
            f = np.empty(shape, self.cdtype)
        else:
            f = self._cview(data)
            c = np.empty(x.size, self.cdtype)
        self.lib.or_nudft3d(typ, self._p(x), self._p(y), self._p(z), x.size,
                            self._p(c.view(self.dtype)), self._p(f.view(self.dtype)),
                            *f.shape)
        return f if typ == 1 else c

    def nufft3d(self, typ, x, y, z, data, shape=None, tol=None):
        tol = self.default_tol if tol is None else tol
        x, y, z = (self._r(v).ravel() for v in (x, y, z))
        if typ == 1:
            c = self._cview(data).ravel()
            f = np.empty(shape, self.cdtype)
        else:
            f = self._cview(data)
            c = np.empty(x.size, self.cdtype)
        self.lib.or_nufft3d(typ, self._p(x), self._p(y), self._p(z), x.size,
                            self._p(c.view(self.dtype)), self._p(f.view(self.dtype)),
                            *f.shape, tol)
        return f if typ == 1 else c

    def es_params(self, tol):
        w = C.c_int()
        b = C.c_double()
        self.lib.or_es_params(tol, C.byref(w), C.byref(b))
        return w.value, b.value

    # ---- 2-D projectors ----
    def forward2d(self, vol, theta, nrad=None, tol=None):
        tol = self.default_tol if tol is None else tol
        v = self._r(vol)
        v3 = v.reshape(-1, v.shape[-2], v.shape[-1])
        ns, N = v3.shape[0], v3.shape[-1]
        nrad = N if nrad is None else nrad
        th = self._r(theta)
        out = np.empty((ns, len(th), nrad), self.dtype)
        self.lib.or_forward2d(self._p(v3), ns, N, self._p(th), len(th), nrad,
                              self._p(out), tol)
        return out

    def backproj2d(self, sino, theta, N, tol=None):
        tol = self.default_tol if tol is None else tol
        s = self._r(sino)
        s3 = s.reshape(-1, s.shape[-2], s.shape[-1])
        th = self._r(theta)
        out = np.empty((s3.shape[0], N, N), self.dtype)
        self.lib.or_backproj2d(self._p(s3), s3.shape[0], s3.shape[1], s3.shape[2],
                               self._p(th), N, self._p(out), tol)
        return out

    def sysmat2d_nufft(self, x, theta, nrad=None, tol=None):
        tol = self.default_tol if tol is None else tol
        v = self._r(x)
        v3 = v.reshape(-1, v.shape[-2], v.shape[-1])
        N = v3.shape[-1]
        nrad = N if nrad is None else nrad
        th = self._r(theta)
        out = np.empty_like(v3)
        self.lib.or_sysmat2d_nufft(self._p(v3), v3.shape[0], N, self._p(th), len(th),
                                   nrad, self._p(out), tol)
        return out

    # ---- 3-D projectors ----
    def forward3d(self, vol, theta, gamma, beta, nrows, ncols, tol=None):
        tol = self.default_tol if tol is None else tol
        v = self._r(vol)
        th, ga, be = self._r(theta), self._r(gamma), self._r(beta)
        out = np.empty((len(th), nrows, ncols), self.dtype)
        self.lib.or_forward3d(self._p(v), *v.shape, self._p(th), self._p(ga),
                              self._p(be), len(th), nrows, ncols, self._p(out), tol)
        return out

    def backproj3d(self, proj, theta, gamma, beta, dims, shifts=None, tol=None):
        tol = self.default_tol if tol is None else tol
        p = self._r(proj)
        th, ga, be = self._r(theta), self._r(gamma), self._r(beta)
        out = np.empty(tuple(dims), self.dtype)
        sh = None
        if shifts is not None:
            sh = self._r(shifts).reshape(-1, 2)
            if sh.shape[0] != p.shape[0]:
                raise ValueError("shifts size must match the number of slices in the projection")
        self.lib.or_backproj3d(self._p(p), *p.shape, self._p(th), self._p(ga),
                               self._p(be), self._p(sh) if sh is not None else None,
                               *dims, self._p(out), tol)
        return out

    def sysmat3d_nufft(self, x, theta, gamma, beta, nrows, ncols, tol=None):
        tol = self.default_tol if tol is None else tol
        v = self._r(x)
        th, ga, be = self._r(theta), self._r(gamma), self._r(beta)
        out = np.empty_like(v)
        self.lib.or_sysmat3d_nufft(self._p(v), *v.shape, self._p(th), self._p(ga),
                                   self._p(be), len(th), nrows, ncols, self._p(out), tol)
        return out

    # ---- PSF ----
    def psf2d(self, theta, nrad, image_n, method=0, tol=None):
        tol = self.default_tol if tol is None else tol
        th = self._r(theta)
        h = self.lib.or_psf2d_create(self._p(th), len(th), nrad, image_n, method, tol)
        return OraclePSF(self, h)

    def psf_from_real(self, P):
        """the Toeplitz operator of given real PSF values (CMCL order): a 2-D
        PSF of shape (M, M) is applied slice by slice, a 3-D one of shape
        (D1, D2, D3) to volumes"""
        P = self._r(P)
        dim = 2 if P.ndim == 2 else 3
        if P.ndim == 2:
            P = P[None]
        d = (C.c_int64 * 3)(*P.shape)
        return OraclePSF(self, self.lib.or_psf_from_real(dim, d, self._p(P)))

    def psf3d(self, theta, gamma, beta, nrows, ncols, recon_dims, method=0, tol=None):
        tol = self.default_tol if tol is None else tol
        th, ga, be = self._r(theta), self._r(gamma), self._r(beta)
        h = self.lib.or_psf3d_create(self._p(th), self._p(ga), self._p(be), len(th),
                                     nrows, ncols, *recon_dims, method, tol)
        return OraclePSF(self, h)

    # ---- finite differences ----
    def grad(self, u, boundary=0):
        u = self._r(u)
        g = np.empty((3,) + u.shape, self.dtype)
        self.lib.or_grad(self._p(u), *u.shape, boundary, self._p(g))
        return g

    def div(self, g, boundary=0):
        g = self._r(g)
        out = np.empty(g.shape[1:], self.dtype)
        self.lib.or_div(self._p(g), *g.shape[1:], boundary, self._p(out))
        return out

    def laplacian(self, u, boundary=0):
        u = self._r(u)
        out = np.empty_like(u)
        self.lib.or_laplacian(self._p(u), *u.shape, boundary, self._p(out))
        return out

    # ---- solvers ----
    def _op(self, A):
        """A: OraclePSF or python callable(ndarray)->ndarray"""
        if isinstance(A, OraclePSF):
            return A, None
        shape_holder = {}

        def cb(_ctx, pin, pout):
            n = shape_holder["n"]
            a = np.ctypeslib.as_array(pin, shape=(n,))
            r = np.ascontiguousarray(A(a.reshape(shape_holder["shape"]).copy()),
                                     dtype=self.dtype).ravel()
            np.ctypeslib.as_array(pout, shape=(n,))[:] = r
        return self._opfn_type(cb), shape_holder

    def cgsolver(self, A, y, x0, max_iter, tol, xtol):
        y = self._r(y)
        x0 = self._r(x0)
        out = np.empty_like(y)
        rep = Report()
        if isinstance(A, OraclePSF):
            ctx = _PsfOpCtx(A.handle, *_as3(y.shape))
            self.lib.or_cgsolver(self.psf_apply_op, C.cast(C.pointer(ctx), C.c_void_p),
                                 self._p(y), self._p(x0), y.size, max_iter, tol, xtol,
                                 self._p(out), C.byref(rep))
        else:
            fn, holder = self._op(A)
            holder["n"] = y.size
            holder["shape"] = y.shape
            self.lib.or_cgsolver(C.cast(fn, C.c_void_p), None, self._p(y), self._p(x0),
                                 y.size, max_iter, tol, xtol, self._p(out), C.byref(rep))
        return out, rep.as_dict()

    def cgsolver_trace(self, A, y, x0, iters, tol, xtol):
        """cgsolver on an OraclePSF up to max(iters) iterations, keeping the
        iterate after each count in iters (ascending).  Returns (array of
        shape (len(iters),) + y.shape, report of the whole run)."""
        y = self._r(y)
        x0 = self._r(x0)
        at = (C.c_int64 * len(iters))(*[int(k) for k in iters])
        xs = np.empty((len(iters),) + y.shape, self.dtype)
        out = np.empty_like(y)
        rep = Report()
        ctx = _PsfOpCtx(A.handle, *_as3(y.shape))
        self.lib.or_cgsolver_trace(self.psf_apply_op, C.cast(C.pointer(ctx), C.c_void_p),
                                   self._p(y), self._p(x0), y.size, int(max(iters)), tol, xtol,
                                   self._p(out), C.byref(rep), at, len(iters), self._p(xs))
        return xs, rep.as_dict()

    # ---- Split-Bregman steps (src/bregman.cpp:58-84) ----
    def sb_rhs(self, y, d, b, mu, boundary=0):
        """yT - mu * div(d - b); d, b of shape (3,) + y.shape"""
        y, d, b = self._r(y), self._r(d), self._r(b)
        out = np.empty_like(y)
        self.lib.or_sb_rhs(self._p(y), self._p(d), self._p(b), *_as3(y.shape), boundary,
                           mu, self._p(out))
        return out

    def sb_shrink(self, g, b, lam, mu, aniso=False):
        """d = shrink(g + b, lam/mu); g, b of shape (3, ...)"""
        g, b = self._r(g), self._r(b)
        d = np.empty_like(g)
        self.lib.or_sb_shrink(self._p(g), self._p(b), g[0].size, lam, mu, int(aniso),
                              self._p(d))
        return d

    def sb_bregman_update(self, b, g, d):
        """returns b + g - d"""
        b = self._r(b).copy()
        g, d = self._r(g), self._r(d)
        self.lib.or_sb_bregman_update(self._p(b), self._p(g), self._p(d), b[0].size)
        return b

    def split_bregman(self, A, y, x0, lam, mu, outer, inner, tol, xtol,
                      boundary=0, aniso=False, state=None):
        """src/bregman.cpp:36-94.  state=(d, b): arrays of shape (3,) + y.shape
        that start the splitting variables and receive their final values
        (the reference starts them at zero)."""
        y = self._r(y)
        x0 = self._r(x0)
        out = np.empty_like(y)
        rep = Report()
        n1, n2, n3 = _as3(y.shape)
        if isinstance(A, OraclePSF):
            ctx = _PsfOpCtx(A.handle, n1, n2, n3)
            fnp, ctxp = self.psf_apply_op, C.cast(C.pointer(ctx), C.c_void_p)
            keep = None
        else:
            keep, holder = self._op(A)
            holder["n"] = y.size
            holder["shape"] = y.shape
            fnp, ctxp = C.cast(keep, C.c_void_p), None
        if state is None:
            self.lib.or_split_bregman(fnp, ctxp, self._p(y), self._p(x0), n1, n2, n3,
                                      lam, mu, outer, inner, tol, xtol, boundary,
                                      int(aniso), self._p(out), C.byref(rep))
        else:
            d, b = state
            for a in (d, b):
                if (not isinstance(a, np.ndarray) or a.dtype != self.dtype
                        or a.shape != (3,) + y.shape or not a.flags.c_contiguous):
                    raise ValueError("state arrays must be C-contiguous, of the oracle's "
                                     "dtype and of shape (3,) + y.shape")
            self.lib.or_split_bregman_state(fnp, ctxp, self._p(y), self._p(x0), n1, n2, n3,
                                            lam, mu, outer, inner, tol, xtol, boundary,
                                            int(aniso), self._p(d), self._p(b),
                                            self._p(out), C.byref(rep))
        return out, rep.as_dict()


def _as3(shape):
    if len(shape) == 3:
        return shape
    if len(shape) == 2:
        return (1,) + tuple(shape)
    return (1, 1, shape[0])


class OraclePSF:
    def __init__(self, orc: Oracle, handle):
        self.o = orc
        self.handle = handle
        d = (C.c_int64 * 3)()
        orc.lib.or_psf_dims(handle, d)
        self.dims = tuple(d)

    def __del__(self):
        try:
            self.o.lib.or_psf_destroy(self.handle)
        except Exception:
            pass

    def real(self):
        out = np.empty(self.dims, self.o.dtype)
        self.o.lib.or_psf_real(self.handle, self.o._p(out))
        return out

    def kernel_hat(self):
        d = self.dims
        out = np.empty((d[0], d[1], d[2] // 2 + 1), self.o.cdtype)
        self.o.lib.or_psf_kernel_hat(self.handle, self.o._p(out.view(self.o.dtype)))
        return out

    def convolve(self, x):
        x = self.o._r(x)
        x3 = x.reshape(_as3(x.shape))
        out = np.empty_like(x3)
        self.o.lib.or_psf_convolve(self.handle, self.o._p(x3), *x3.shape, self.o._p(out))
        return out.reshape(x.shape)
