"""Writes tests/golden/golden2d.npz and golden3d.npz: small input/output
vectors of the reference's 2-D and 3-D projector chains computed by brute
force in numpy (direct non-uniform DFT, numpy FFT, explicit index maps),
independently of the C oracle.

Conventions restated from the reference:
  polar grid 2-D   include/polar_grid2d.h:50-65
  polar grid 3-D   include/polar_grid.h:52-94, include/rotation.h:28-56
  NUFFT            include/nufft.h:37-95 (type 2: exp(-i m.w); type 1: exp(+i m.w);
                   modes m = index - floor(n/2); x <-> columns, y <-> rows, z <-> slices)
  fftshift         out[k] = in[(k + n//2) % n]            include/fftutils.h:34-51
  ifftshift        out[k] = in[(k + n//2 + n%2) % n]      include/fftutils.h:97-113
  forward2d        src/projection.cpp:146-160
  backproj2d       src/projection.cpp:168-182
  forward (3-D)    src/projection.cpp:42-64
  backproj (3-D)   src/projection.cpp:71-131

Run: python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def roll_last(a, d, axis=-1):
    n = a.shape[axis]
    idx = (np.arange(n) + d) % n
    return np.take(a, idx, axis=axis)


def fftshift_ref(a, axis=-1):
    return roll_last(a, a.shape[axis] // 2, axis)


def ifftshift_ref(a, axis=-1):
    n = a.shape[axis]
    return roll_last(a, n // 2 + (n % 2), axis)


def polar2d(theta, nrad):
    r = (np.arange(nrad) + 0.5) * (2 * np.pi / nrad) - np.pi
    return np.cos(theta)[:, None] * r[None, :], np.sin(theta)[:, None] * r[None, :]


def rotT(t, g, b):
    ct, st, cb, sb, cg, sg = np.cos(t), np.sin(t), np.cos(b), np.sin(b), np.cos(g), np.sin(g)
    return np.array([[cb * cg, ct * sg + st * sb * cg, st * sg - ct * sb * cg],
                     [-cb * sg, ct * cg - st * sb * sg, st * cg + ct * sb * sg],
                     [sb, -st * cb, ct * cb]])


def polar3d(theta, gamma, beta, nrows, ncols):
    qX = (np.arange(ncols) + 0.5) * (2 * np.pi / ncols) - np.pi
    qY = (np.arange(nrows) + 0.5) * (2 * np.pi / nrows) - np.pi
    QY, QX = np.meshgrid(qY, qX, indexing="ij")
    xs, ys, zs, ws = [], [], [], []
    for t, g, b in zip(theta, gamma, beta):
        R = rotT(t, g, b)
        x = R[0, 0] * QX + R[0, 1] * QY
        y = R[1, 0] * QX + R[1, 1] * QY
        z = R[2, 0] * QX + R[2, 1] * QY
        w = ~((np.abs(x) > np.pi) | (np.abs(y) > np.pi) | (np.abs(z) > np.pi))
        xs.append(x); ys.append(y); zs.append(z); ws.append(w.astype(float))
    return np.array(xs), np.array(ys), np.array(zs), np.array(ws)


def modes(n):
    return np.arange(n) - n // 2


def forward2d(vol, theta, nrad):
    ns, N, _ = vol.shape
    x, y = polar2d(theta, nrad)
    kx, ky = modes(N), modes(N)
    # c_j = sum_m f_m exp(-i (kx x + ky y))
    ex = np.exp(-1j * x[..., None] * kx)          # nang, nrad, N(cols)
    ey = np.exp(-1j * y[..., None] * ky)          # nang, nrad, N(rows)
    C = np.einsum("sab,ijb,ija->sij", vol.astype(complex), ex, ey)
    C = ifftshift_ref(C)
    C = np.fft.ifft(C, axis=-1) * nrad            # unnormalised backward
    C = fftshift_ref(C)
    return C.real / nrad


def backproj2d(sino, theta, N):
    ns, nang, nrad = sino.shape
    x, y = polar2d(theta, nrad)
    C = fftshift_ref(sino.astype(complex))
    C = np.fft.fft(C, axis=-1)
    C = ifftshift_ref(C)
    kx, ky = modes(N), modes(N)
    ex = np.exp(1j * x[..., None] * kx)
    ey = np.exp(1j * y[..., None] * ky)
    F = np.einsum("sij,ijb,ija->sab", C, ex, ey)
    return F.real / nrad


def forward3d(vol, theta, gamma, beta, nrows, ncols):
    n1, n2, n3 = vol.shape
    x, y, z, w = polar3d(theta, gamma, beta, nrows, ncols)
    kz, ky, kx = modes(n1), modes(n2), modes(n3)
    ex = np.exp(-1j * x[..., None] * kx)
    ey = np.exp(-1j * y[..., None] * ky)
    ez = np.exp(-1j * z[..., None] * kz)
    C = np.einsum("abc,pjkc,pjkb,pjka->pjk", vol.astype(complex), ex, ey, ez) * w
    C = fftshift_ref(fftshift_ref(C, -1), -2)
    C = np.fft.ifft2(C, axes=(-2, -1)) * (nrows * ncols)
    C = ifftshift_ref(ifftshift_ref(C, -1), -2)
    return C.real / (nrows * ncols)


def backproj3d(proj, theta, gamma, beta, dims, shifts=None):
    nproj, nrows, ncols = proj.shape
    x, y, z, w = polar3d(theta, gamma, beta, nrows, ncols)
    C = fftshift_ref(fftshift_ref(proj.astype(complex), -1), -2)
    C = np.fft.fft2(C, axes=(-2, -1))
    C = ifftshift_ref(ifftshift_ref(C, -1), -2)
    if shifts is not None:
        qx = (np.arange(ncols) + 0.5) * (2 * np.pi / ncols) - np.pi
        qy = (np.arange(nrows) + 0.5) * (2 * np.pi / nrows) - np.pi
        ph = qx[None, None, :] * shifts[:, 0][:, None, None] + qy[None, :, None] * shifts[:, 1][:, None, None]
        C = C * np.exp(-1j * ph)
    C = C * w
    n1, n2, n3 = dims
    kz, ky, kx = modes(n1), modes(n2), modes(n3)
    ex = np.exp(1j * x[..., None] * kx)
    ey = np.exp(1j * y[..., None] * ky)
    ez = np.exp(1j * z[..., None] * kz)
    F = np.einsum("pjk,pjkc,pjkb,pjka->abc", C, ex, ey, ez)
    return F.real / (nrows * ncols)


def main():
    rng = np.random.default_rng(20240601)
    out = {}
    # 2-D: even and odd detector widths, ref test angle set
    for tag, (ns, N, nang, nrad) in {"a": (2, 8, 5, 8), "b": (1, 9, 7, 11), "c": (1, 12, 6, 10)}.items():
        th = (np.arange(nang) - nang / 2.0) * np.pi / nang
        vol = rng.random((ns, N, N))
        sino = rng.random((ns, nang, nrad))
        out[f"{tag}_theta"] = th
        out[f"{tag}_vol"] = vol
        out[f"{tag}_sino"] = sino
        out[f"{tag}_fwd"] = forward2d(vol, th, nrad)
        out[f"{tag}_bp"] = backproj2d(sino, th, N)
    np.savez_compressed(os.path.join(HERE, "golden2d.npz"), **out)
    out = {}
    dims = (3, 6, 5)
    nproj, nrows, ncols = 4, 5, 7
    th = np.linspace(-0.6, 0.8, nproj)
    ga = np.full(nproj, -0.5)
    be = np.full(nproj, 0.1)
    vol = rng.random(dims)
    proj = rng.random((nproj, nrows, ncols))
    shifts = rng.uniform(-1.5, 1.5, (nproj, 2))
    out.update(theta=th, gamma=ga, beta=be, vol=vol, proj=proj, shifts=shifts,
               fwd=forward3d(vol, th, ga, be, nrows, ncols),
               bp=backproj3d(proj, th, ga, be, dims),
               bp_shift=backproj3d(proj, th, ga, be, dims, shifts),
               dims=np.array(dims), det=np.array([nrows, ncols]))
    np.savez_compressed(os.path.join(HERE, "golden3d.npz"), **out)


if __name__ == "__main__":
    main()
