"""Pins of the CPU oracle (no GPU).  Each check is chosen so that a plausible
mistake in the restatement (a dropped term, a wrong sign or index, a
transposed operand, a wrong normalisation) fails it:
  * FFT vs numpy's pocketfft and a naive DFT
  * shift index maps of include/fftutils.h
  * projector chains vs brute-force numpy golden vectors (tests/golden)
  * NUFFT vs direct NUDFT, adjointness
  * PSF closed form vs NUFFT of ones vs brute-force sum
  * Toeplitz == NUFFT normal operator (tests/test_sysmat_toeplitz*.cpp)
  * adjoint / symmetry / PSD (tests/test_adjoint*.cpp)
  * CG on normal equations (tests/test_cg_normal.cpp), diagonal system (tests/test_cg.cu)
  * the Split-Bregman steps: tests/test_oracle_bregman.py
"""
import os

import numpy as np
import pytest

from conftest import rel_l2

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def ref_theta(n):
    return (np.arange(n) - n / 2.0) * np.pi / n


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 8, 9, 12, 16, 27, 64, 89, 96, 100, 243, 256, 384,
                               511, 721, 1458, 2048])
def test_fft_vs_numpy(orc64, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(a)
    assert np.abs(orc64.fft(a, -1) - ref).max() <= 1e-13 * max(1, np.abs(ref).max())
    assert np.abs(orc64.fft(a, +1) - np.fft.ifft(a) * n).max() <= 1e-13 * max(1, np.abs(ref).max())


def test_dft_naive_and_fft2(orc64):
    rng = np.random.default_rng(0)
    a = rng.standard_normal(30) + 1j * rng.standard_normal(30)
    assert np.allclose(orc64.dft_naive(a), np.fft.fft(a), atol=1e-12)
    b = rng.standard_normal((3, 12, 18)) + 1j * rng.standard_normal((3, 12, 18))
    assert np.allclose(orc64.fft2(b), np.fft.fft2(b), atol=1e-11)
    c = rng.standard_normal((6, 10, 9)) + 1j * rng.standard_normal((6, 10, 9))
    assert np.allclose(orc64.fft3(c, +1), np.fft.ifftn(c) * c.size, atol=1e-11)


def test_next_fast_dim(orc64):
    # include/toeplitz2d.h:48-57
    assert [orc64.next_fast_dim(v) for v in (255, 511, 1023, 57, 1441, 383, 767, 1535, 2047, 27,
                                              9, 3)] == \
        [256, 512, 1024, 64, 1458, 384, 768, 1536, 2048, 27, 9, 3]


@pytest.mark.parametrize("n", [4, 5, 8, 9])
def test_shift_conventions(orc64, n):
    a = np.arange(n, dtype=float)[None, None, :]
    fs = orc64.shift(a, 1, inverse=False)[0, 0]
    ifs = orc64.shift(a, 1, inverse=True)[0, 0]
    assert np.array_equal(fs, (np.arange(n) + n // 2) % n)
    assert np.array_equal(ifs, (np.arange(n) + n // 2 + n % 2) % n)
    # numpy agrees for even n; for odd n the reference's ifftshift is NOT numpy's
    if n % 2 == 0:
        assert np.array_equal(fs, np.fft.fftshift(np.arange(n)))
    else:
        assert not np.array_equal(ifs, np.fft.ifftshift(np.arange(n)))


# ---------------------------------------------------------------- golden vectors
def test_golden_2d(orc64):
    g = np.load(os.path.join(GOLD, "golden2d.npz"))
    for t in "abc":
        th, vol, sino = g[t + "_theta"], g[t + "_vol"], g[t + "_sino"]
        nrad, N = sino.shape[-1], vol.shape[-1]
        for tol in (0, 1e-12):
            assert rel_l2(orc64.forward2d(vol, th, nrad, tol=tol), g[t + "_fwd"]) < 1e-11
            assert rel_l2(orc64.backproj2d(sino, th, N, tol=tol), g[t + "_bp"]) < 1e-11


def test_golden_3d(orc64):
    g = np.load(os.path.join(GOLD, "golden3d.npz"))
    nr, nc = (int(v) for v in g["det"])
    dims = tuple(int(v) for v in g["dims"])
    for tol in (0, 1e-12):
        assert rel_l2(orc64.forward3d(g["vol"], g["theta"], g["gamma"], g["beta"], nr, nc, tol=tol),
                      g["fwd"]) < 1e-11
        assert rel_l2(orc64.backproj3d(g["proj"], g["theta"], g["gamma"], g["beta"], dims, tol=tol),
                      g["bp"]) < 1e-11
        assert rel_l2(orc64.backproj3d(g["proj"], g["theta"], g["gamma"], g["beta"], dims,
                                       shifts=g["shifts"], tol=tol), g["bp_shift"]) < 1e-11


def test_backproj3d_shift_count_mismatch(orc64):
    g = np.load(os.path.join(GOLD, "golden3d.npz"))
    dims = tuple(int(v) for v in g["dims"])
    with pytest.raises(ValueError):
        orc64.backproj3d(g["proj"], g["theta"], g["gamma"], g["beta"], dims,
                         shifts=g["shifts"][:-1])


# ---------------------------------------------------------------- NUFFT
def test_nudft_tiny_bruteforce(orc64):
    rng = np.random.default_rng(1)
    x, y = rng.uniform(-np.pi, np.pi, 5), rng.uniform(-np.pi, np.pi, 5)
    f = rng.standard_normal((3, 4)) + 1j * rng.standard_normal((3, 4))
    c = orc64.nudft2d(2, x, y, f)
    ky, kx = np.arange(3) - 1, np.arange(4) - 2
    ref = np.array([sum(f[a, b] * np.exp(-1j * (kx[b] * x[j] + ky[a] * y[j]))
                        for a in range(3) for b in range(4)) for j in range(5)])
    assert np.allclose(c, ref, atol=1e-13)


@pytest.mark.parametrize("nr,nc", [(16, 16), (9, 12), (20, 7)])
def test_nufft2d_vs_nudft(orc64, nr, nc):
    rng = np.random.default_rng(nr * nc)
    n = 300
    x, y = rng.uniform(-np.pi, np.pi, n), rng.uniform(-np.pi, np.pi, n)
    f = rng.standard_normal((nr, nc)) + 1j * rng.standard_normal((nr, nc))
    c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel_l2(orc64.nufft2d(2, x, y, f, tol=1e-12), orc64.nudft2d(2, x, y, f)) < 1e-10
    assert rel_l2(orc64.nufft2d(1, x, y, c, nr, nc, tol=1e-12), orc64.nudft2d(1, x, y, c, nr, nc)) < 1e-10
    # adjointness <T2 f, c> = <f, T1 c>
    lhs = np.vdot(c, orc64.nufft2d(2, x, y, f, tol=1e-12))
    rhs = np.vdot(orc64.nufft2d(1, x, y, c, nr, nc, tol=1e-12), f)
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


def test_nufft3d_vs_nudft(orc64):
    rng = np.random.default_rng(5)
    n = 200
    x, y, z = (rng.uniform(-np.pi, np.pi, n) for _ in range(3))
    f = rng.standard_normal((5, 6, 7)) + 1j * rng.standard_normal((5, 6, 7))
    c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel_l2(orc64.nufft3d(2, x, y, z, f, tol=1e-12), orc64.nudft3d(2, x, y, z, f)) < 1e-10
    assert rel_l2(orc64.nufft3d(1, x, y, z, c, (5, 6, 7), tol=1e-12),
                  orc64.nudft3d(1, x, y, z, c, (5, 6, 7))) < 1e-10


def test_nufft_tolerance_tracks_width(orc64):
    # FINUFFT's width rule w = ceil(log10(10/tol)), beta = 2.30 w (upsampling 2)
    # lands within a small factor of tol, including the reference's float
    # tolerance 5e-4 (include/finufft_plan.h:71)
    rng = np.random.default_rng(2)
    x, y = orc64.polar_grid2d(ref_theta(30), 32)
    f = rng.standard_normal((32, 32)) + 0j
    exact = orc64.nudft2d(2, x, y, f)
    for tol in (5e-4, 1e-6, 1e-9):
        assert rel_l2(orc64.nufft2d(2, x, y, f, tol=tol), exact) < 3 * tol
    assert orc64.es_params(5e-4)[0] == 5 and orc64.es_params(1e-12)[0] == 13


# ---------------------------------------------------------------- PSF / Toeplitz
def test_psf_closed_form_vs_bruteforce(orc64):
    th = ref_theta(5)
    nrad, N = 6, 4
    P = orc64.psf2d(th, nrad, N, method=0).real()[0]
    M = P.shape[0]
    x, y = orc64.polar_grid2d(th, nrad)
    ky, kx = np.meshgrid(np.arange(M) - M // 2, np.arange(M) - M // 2, indexing="ij")
    ref = np.exp(1j * (kx[..., None] * x.ravel() + ky[..., None] * y.ravel())).sum(-1)
    assert np.abs(ref.imag).max() < 1e-12
    assert np.allclose(P, ref.real, atol=1e-11)


@pytest.mark.parametrize("N,nang,nrad", [(16, 7, 16), (32, 13, 32), (14, 9, 20), (5, 4, 6)])
def test_psf2d_closed_form_vs_nufft(orc64, N, nang, nrad):
    th = ref_theta(nang)
    a = orc64.psf2d(th, nrad, N, method=0)
    b = orc64.psf2d(th, nrad, N, method=1, tol=1e-12)
    assert rel_l2(a.real(), b.real()) < 1e-10
    assert rel_l2(a.kernel_hat(), b.kernel_hat()) < 1e-10


@pytest.mark.parametrize("vol,nang", [((5, 16, 16), 7), ((8, 32, 32), 13), ((2, 14, 14), 9),
                                      ((1, 27, 27), 11)])
def test_toeplitz2d_equals_sysmat2d(orc64, vol, nang):
    # tests/test_sysmat_toeplitz2d.cpp:61-98 (scale ratio and shape within 1e-4)
    rng = np.random.default_rng(sum(vol))
    th = ref_theta(nang)
    x = rng.random(vol)
    t = orc64.psf2d(th, vol[-1], vol[-1]).convolve(x)
    s = orc64.sysmat2d_nufft(x, th)
    ratio = np.vdot(s, t) / np.vdot(t, t)
    assert abs(ratio - 1) < 1e-10
    assert rel_l2(t, s) < 1e-10


def test_toeplitz2d_is_linear_convolution(orc64):
    rng = np.random.default_rng(4)
    th = ref_theta(6)
    N = 7
    p = orc64.psf2d(th, 9, N)
    P = p.real()[0]
    M = P.shape[0]
    x = rng.random((1, N, N))
    out = p.convolve(x)[0]
    ref = np.zeros((N, N))
    for a in range(N):
        for b in range(N):
            for c in range(N):
                for d in range(N):
                    ref[a, b] += x[0, c, d] * P[a - c + M // 2, b - d + M // 2]
    # the unnormalised FFT pair contributes M^2, cancelled by the 1/(M^2 N) norm
    assert np.allclose(out, ref / N, atol=1e-12)


def test_toeplitz_pin_has_teeth(orc64):
    # a planted mistake (one angle dropped from the PSF) must break the pin
    rng = np.random.default_rng(8)
    th = ref_theta(9)
    x = rng.random((1, 16, 16))
    bad = orc64.psf2d(th[:-1], 16, 16).convolve(x)
    assert rel_l2(bad, orc64.sysmat2d_nufft(x, th)) > 1e-3


def test_psf3d_closed_form_vs_nufft(orc64):
    th = np.linspace(-0.7, 0.9, 5)
    ga = np.full(5, 0.3)
    be = np.full(5, -0.2)
    a = orc64.psf3d(th, ga, be, 7, 9, (3, 5, 4), method=0)
    b = orc64.psf3d(th, ga, be, 7, 9, (3, 5, 4), method=1, tol=1e-12)
    assert rel_l2(a.real(), b.real()) < 1e-10


@pytest.mark.parametrize("gamma", [0.0, 0.3])
def test_toeplitz3d_equals_sysmat3d(orc64, gamma):
    # tests/test_sysmat_toeplitz.cpp:48-127
    rng = np.random.default_rng(3)
    dims = (5, 8, 8)
    n = 7
    th = ref_theta(n)
    ga = np.full(n, gamma)
    be = np.zeros(n)
    x = rng.random(dims)
    t = orc64.psf3d(th, ga, be, 8, 8, dims).convolve(x)
    s = orc64.sysmat3d_nufft(x, th, ga, be, 8, 8)
    # the reference divides Toeplitz by the image n2*n3 and sysmat by the grid's;
    # they agree when the detector matches the image
    assert rel_l2(t, s) < 1e-9


# ---------------------------------------------------------------- adjoint / symmetry / PSD
@pytest.mark.parametrize("vol,nang", [((5, 16, 16), 7), ((8, 32, 32), 13)])
def test_adjoint2d(orc64, vol, nang):
    rng = np.random.default_rng(11)
    th = ref_theta(nang)
    x = rng.random(vol)
    y = rng.random((vol[0], nang, vol[2]))
    lhs = np.vdot(orc64.forward2d(x, th), y)
    rhs = np.vdot(x, orc64.backproj2d(y, th, vol[2]))
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs)) < 1e-5
    u = rng.random(vol)
    Ax, Au = orc64.sysmat2d_nufft(x, th), orc64.sysmat2d_nufft(u, th)
    assert abs(np.vdot(u, Ax) - np.vdot(x, Au)) / abs(np.vdot(u, Ax)) < 1e-5
    assert np.vdot(x, Ax) >= 0


def test_adjoint3d(orc64):
    rng = np.random.default_rng(12)
    dims = (5, 16, 16)
    n = 7
    th = ref_theta(n)
    ga, be = np.full(n, -0.3), np.zeros(n)
    x = rng.random(dims)
    y = rng.random((n, 16, 16))
    lhs = np.vdot(orc64.forward3d(x, th, ga, be, 16, 16), y)
    rhs = np.vdot(x, orc64.backproj3d(y, th, ga, be, dims))
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs)) < 1e-5


# ---------------------------------------------------------------- solvers
def test_cg_normal_equations(orc64, orc32):
    # tests/test_cg_normal.cpp:47-110: sparse 64x32, ~30 % fill
    rng = np.random.default_rng(42)
    m, n = 64, 32
    A = np.where(rng.random((m, n)) < 0.3, rng.uniform(-1, 1, (m, n)), 0.0)
    x_true = rng.uniform(-1, 1, n)
    b = A.T @ (A @ x_true)
    op = lambda v: A.T @ (A @ v)
    for o, tol in ((orc64, 1e-5), (orc32, 1e-3)):
        x, rep = o.cgsolver(op, b, np.zeros(n), 2000, tol, tol / 100)
        assert rel_l2(x, x_true) < tol * 100


def test_cg_diagonal(orc64):
    # tests/test_cg.cu:14-73: diagonal system with condition number 100
    n = 50
    d = np.linspace(1, 100, n)
    xt = np.random.default_rng(0).random(n)
    x, rep = orc64.cgsolver(lambda v: d * v, d * xt, np.zeros(n), 200, 1e-10, 1e-12)
    assert rel_l2(x, xt) < 1e-6


def test_cg_pap_guard(orc64):
    # src/conjgrad.cpp:53-56: |pAp| < 1e-10 stops before updating x
    x, rep = orc64.cgsolver(lambda v: 0.0 * v, np.ones(8), np.zeros(8), 5, 1e-5, 1e-5)
    assert rep["stop"] == 3 and rep["iters"] == 0 and not np.any(x)


def test_finite_differences_closed_forms(orc64):
    n1, n2, n3 = 5, 6, 7
    s, a, b = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
    u = 2.0 * s - 3.0 * a + 0.5 * b
    g = orc64.grad(u, 0)
    inner = (slice(1, -1),) * 3
    assert np.allclose(g[0][inner], 2.0) and np.allclose(g[1][inner], -3.0)
    assert np.allclose(g[2][inner], 0.5)
    assert g[0][0].sum() == 0 and g[2][:, :, -1].sum() == 0
    # single slice: the CPU reference's gradient is identically zero
    assert not np.any(orc64.grad(np.random.default_rng(0).random((1, 8, 8)), 0))
    # zero padded: every point, zero outside
    gz = orc64.grad(u, 1)
    assert np.isclose(gz[2][0, 0, 0], 0.5 * (u[0, 0, 1] - 0))
    # laplacian == div(grad)
    r = np.random.default_rng(1).random((n1, n2, n3))
    for bd in (0, 1):
        assert np.allclose(orc64.laplacian(r, bd), orc64.div(orc64.grad(r, bd), bd))
        # -div is the adjoint of grad: <grad u, g> = -<u, div g>
        gg = np.random.default_rng(2).random((3, n1, n2, n3))
        if bd == 1:
            assert np.isclose(np.vdot(orc64.grad(r, bd), gg), -np.vdot(r, orc64.div(gg, bd)))


def test_split_bregman_single_slice_is_restarted_cg(orc64):
    # interior-only TV vanishes on a single slice: SB(outer, inner=1) == CG restarted
    rng = np.random.default_rng(5)
    th = ref_theta(11)
    N = 16
    P = orc64.psf2d(th, N, N)
    y = P.convolve(rng.random((1, N, N)))
    xs, _ = orc64.split_bregman(P, y, np.zeros_like(y), 0.1, 10.0, 4, 1, 0.0, 0.0, boundary=0)
    x = np.zeros_like(y)
    for _ in range(4):
        x, _ = orc64.cgsolver(P, y, x, 1, 0.0, 0.0)
    assert rel_l2(xs, x) < 1e-12


def test_cg_trace_equals_separate_solves(orc64):
    # the traced iterates are the iterates of separate solves of that length
    # (same arithmetic, bitwise), and a snapshot past an early stop is the
    # final x
    rng = np.random.default_rng(41)
    N, th = 24, np.arange(17) * np.pi / 17
    P = orc64.psf2d(th, N, N, method=0)
    y = P.convolve(rng.random((1, N, N)))
    ks = [1, 3, 4, 9]
    xs, rep = orc64.cgsolver_trace(P, y, np.zeros_like(y), ks, 0.0, 0.0)
    for j, k in enumerate(ks):
        xk, _ = orc64.cgsolver(P, y, np.zeros_like(y), k, 0.0, 0.0)
        assert np.array_equal(xs[j], xk), k
    xe, re = orc64.cgsolver(P, y, np.zeros_like(y), 50, 1e-3, 0.0)
    assert re["iters"] < 40
    xt, _ = orc64.cgsolver_trace(P, y, np.zeros_like(y), [re["iters"], 40, 50], 1e-3, 0.0)
    assert np.array_equal(xt[1], xe) and np.array_equal(xt[2], xe) and np.array_equal(xt[0], xe)


def test_psf_from_real_applies_as_the_computed_psf(orc64):
    # an operator rebuilt from the real PSF values applies exactly as the
    # computed one (2-D and 3-D)
    rng = np.random.default_rng(43)
    th = np.arange(11) * np.pi / 11
    P = orc64.psf2d(th, 20, 20, method=0)
    Q = orc64.psf_from_real(P.real()[0])
    x = rng.random((2, 20, 20))
    assert np.array_equal(Q.convolve(x), P.convolve(x))
    th3 = np.arange(7) * 2 * np.pi / 7
    P3 = orc64.psf3d(th3, np.full(7, -0.5), np.zeros(7), 6, 12, (3, 12, 12), method=0)
    Q3 = orc64.psf_from_real(P3.real())
    x3 = rng.random((3, 12, 12))
    assert np.array_equal(Q3.convolve(x3), P3.convolve(x3))
