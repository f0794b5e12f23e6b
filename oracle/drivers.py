"""CPU restatement of the reference's whole-solve drivers (test infrastructure
only: tests/ and nothing in the product imports it).

  pad_size          src/mbir.cpp:41-45 (element type), include/padding.h:19-22 (float)
  pad2d             include/padding.h:16-47   PadType::SYMMETRIC
  crop3d            include/padding.h:100-128 PadType::SYMMETRIC
  prepare_sinogram  include/mbir2d_prep.h:44-65
  mbir2d            src/mbir2d.cpp:66-100 (cpu_mbir2d)
  mbir3d            src/mbir.cpp:34-161 (MBIR)

Pins (tests/test_oracle_drivers.py): pad2d / crop3d against hand-worked
placements and the reference's own pad -> crop round trips
(tests/test_padding.cu:118-131); pad_size against the known padded sizes
(510 -> 721, 63 -> 89, 21 -> 29 at 1.4142) and the even -> odd rule;
prepare_sinogram against an element-by-element loop over the index map.
mbir2d / mbir3d compose these with the oracle's back-projectors, PSFs and
solvers, each pinned on its own in oracle/oracle.py's tests.
"""
from __future__ import annotations

import numpy as np


def pad_size(n: int, factor, dtype=np.float32) -> int:
    """size_t(n * factor) computed in dtype, made odd by -1 when even"""
    t = np.dtype(dtype).type
    m = int(t(n) * t(np.float32(factor)))
    return m - 1 if m % 2 == 0 else m


def pad2d(arr, factor):
    """pad2d(arr, factor, SYMMETRIC): every slice {nrows, ncols} placed at
    ((n2 - nrows) / 2, (n3 - ncols) / 2) in zeros {n2, n3}; a copy when
    factor - 1 < 1e-6"""
    a = np.asarray(arr)
    f = np.float32(factor)
    if float(f - np.float32(1)) < 1e-6:
        return a.copy()
    n2, n3 = pad_size(a.shape[1], f), pad_size(a.shape[2], f)
    out = np.zeros((a.shape[0], n2, n3), a.dtype)
    d1, d2 = (n2 - a.shape[1]) // 2, (n3 - a.shape[2]) // 2
    out[:, d1:d1 + a.shape[1], d2:d2 + a.shape[2]] = a
    return out


def crop3d(arr, new_dims):
    """crop3d(arr, new_dims, SYMMETRIC): the block at (arr.dims - new_dims) / 2"""
    a = np.asarray(arr)
    d = [(a.shape[i] - new_dims[i]) // 2 for i in range(3)]
    return a[d[0]:d[0] + new_dims[0], d[1]:d[1] + new_dims[1], d[2]:d[2] + new_dims[2]].copy()


def normalise(stacked):
    """stacked /= max(stacked) when the maximum is > 0 (element-wise x / m)"""
    m = stacked.max()
    return stacked / m if m > 0 else stacked.copy()


def prepare_sinogram(projs, angles, dtype=np.float32):
    """stack, divide by the global maximum, {n_angles, n_rows, n_cols} ->
    {n_rows, n_angles, n_cols}; returns (sinogram, theta)"""
    stacked = np.concatenate([np.asarray(p, dtype) for p in projs], axis=0)
    theta = np.concatenate([np.ravel(a) for a in angles]).astype(dtype)
    return np.ascontiguousarray(normalise(stacked).transpose(1, 0, 2)), theta


def _solve(orc, A, rhs, p, boundary):
    x0 = np.zeros(rhs.shape)
    if p["regularizer"] == "cg":
        return orc.cgsolver(A, rhs, x0, p["maxIters"], p["tol"], p["xtol"])
    return orc.split_bregman(A, rhs, x0, p["lam"], p["mu"], p["maxIters"], p["innerIters"],
                             p["tol"], p["xtol"], boundary=boundary)


def mbir2d(orc, sino, theta, p, boundary=0):
    """cpu_mbir2d: PolarGrid2D(theta, ncols), PSF2D(n3), rhs = backproj2d of
    the first n1 slices, CG or Split-Bregman from zeros.  p: dict with keys
    regularizer ('cg' | 'sb'), recon_dims, maxIters, innerIters, lam, mu, tol,
    xtol.  Returns (recon, report)."""
    n1, _, N = p["recon_dims"]
    rhs = orc.backproj2d(np.asarray(sino)[:n1], theta, N)
    psf = orc.psf2d(theta, np.asarray(sino).shape[2], N)
    return _solve(orc, psf, rhs, p, boundary)


def mbir3d(orc, projs, angles, gammas, betas, shifts, p, boundary=0, dtype=np.float32):
    """MBIR: stack, normalise (in dtype), pad2d by PAD_FACTOR, grid and PSF on
    the padded detector for the padded volume, rhs = backproj(..., shifts),
    CG or Split-Bregman from zeros, crop3d to recon_dims.  Returns (recon,
    report, padded volume dims)."""
    f = p.get("PAD_FACTOR", 1.4142)
    out_dims = tuple(pad_size(n, f, dtype) for n in p["recon_dims"])
    stacked = np.concatenate([np.asarray(q, dtype) for q in projs], axis=0)
    padded = pad2d(normalise(stacked), f).astype(np.float32)
    theta = np.concatenate([np.ravel(a) for a in angles])
    gamma = np.concatenate([np.full(len(np.ravel(a)), g) for a, g in zip(angles, gammas)])
    beta = np.concatenate([np.full(len(np.ravel(a)), b) for a, b in zip(angles, betas)])
    sh = None if shifts is None else np.concatenate([np.asarray(s, np.float64) for s in shifts])
    nrows, ncols = padded.shape[1:]
    rhs = orc.backproj3d(padded, theta, gamma, beta, out_dims, shifts=sh)
    psf = orc.psf3d(theta, gamma, beta, nrows, ncols, out_dims)
    x, rep = _solve(orc, psf, rhs, p, boundary)
    return crop3d(x, p["recon_dims"]), rep, out_dims
