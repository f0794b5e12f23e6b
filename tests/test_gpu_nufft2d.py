"""GPU parity of the 2-D NUFFT projectors forward2d / backproj2d
(src/projection.cpp:146-182) against the FP64 oracle and the committed golden
vectors, through the C ABI.  Bar: 1e-4 relative L2 on projections."""
import os

import numpy as np
import pytest

from conftest import announce, rel_l2

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden2d.npz")
PROJ_TOL = 1e-4


@pytest.fixture(scope="module")
def lc():
    import laminocam_amd
    return laminocam_amd


def uniform_theta(n, span=np.pi):
    return np.arange(n) * span / n


@pytest.mark.parametrize("case", ["a", "b", "c"])
def test_projectors_match_golden(lc, case):
    # golden vectors: brute-force NUDFT + numpy FFT restatement of the reference chain
    announce(f"projector golden {case}")
    d = np.load(GOLDEN)
    th, vol, sino = d[f"{case}_theta"], d[f"{case}_vol"], d[f"{case}_sino"]
    g = lc.PolarGrid2D(th, sino.shape[-1])
    fwd = lc.forward2d(vol, g)
    assert fwd.shape == d[f"{case}_fwd"].shape
    assert rel_l2(fwd, d[f"{case}_fwd"]) < PROJ_TOL
    bp = lc.backproj2d(sino, g, vol.shape)
    assert rel_l2(bp, d[f"{case}_bp"]) < PROJ_TOL


# (N, nang, nrad, ns, angular span): even/odd N and nrad, 2^p 3^q and other
# detector widths (direct-DFT path), full-circle angles, stacks
CASES = [(16, 9, 16, 1, np.pi), (24, 30, 24, 2, np.pi), (31, 17, 33, 1, np.pi),
         (64, 90, 64, 3, np.pi), (64, 45, 90, 1, 2 * np.pi), (100, 60, 100, 1, np.pi),
         (128, 180, 128, 1, np.pi), (256, 180, 256, 1, np.pi), (192, 90, 192, 1, np.pi),
         # the image sizes of the batched workload at its largest angle count
         (384, 360, 384, 1, np.pi), (512, 720, 512, 1, np.pi), (768, 720, 768, 1, np.pi),
         (1024, 720, 1024, 1, np.pi)]


@pytest.mark.parametrize("N,nang,nrad,ns,span", CASES)
def test_forward2d_matches_oracle(lc, orc64, N, nang, nrad, ns, span):
    announce(f"forward2d N={N} nang={nang} nrad={nrad} ns={ns}")
    rng = np.random.default_rng(N * 7 + nang)
    th = uniform_theta(nang, span)
    vol = rng.random((ns, N, N))
    y = lc.forward2d(vol, lc.PolarGrid2D(th, nrad))
    yr = orc64.forward2d(vol, th, nrad, tol=1e-12)
    err = rel_l2(y, yr)
    print(f"forward2d rel_l2 = {err:.3e}")
    assert err < PROJ_TOL


@pytest.mark.parametrize("N,nang,nrad,ns,span", CASES)
def test_backproj2d_matches_oracle(lc, orc64, N, nang, nrad, ns, span):
    announce(f"backproj2d N={N} nang={nang} nrad={nrad} ns={ns}")
    rng = np.random.default_rng(N * 11 + nang)
    th = uniform_theta(nang, span)
    sino = rng.standard_normal((ns, nang, nrad))
    x = lc.backproj2d(sino, lc.PolarGrid2D(th, nrad), (ns, N, N))
    xr = orc64.backproj2d(sino, th, N, tol=1e-12)
    err = rel_l2(x, xr)
    print(f"backproj2d rel_l2 = {err:.3e}")
    assert err < PROJ_TOL


@pytest.mark.parametrize("N,nang,nrad", [(64, 50, 64), (96, 120, 128), (512, 360, 512)])
def test_projectors_adjoint(lc, N, nang, nrad):
    # even nrad: fftshift == ifftshift, so backproj2d is the adjoint of forward2d
    # (tests/test_adjoint2d.cpp); holds at any size, no oracle needed
    announce(f"adjoint N={N} nang={nang} nrad={nrad}")
    rng = np.random.default_rng(5)
    g = lc.PolarGrid2D(uniform_theta(nang), nrad)
    x = rng.random((1, N, N))
    y = rng.standard_normal((1, nang, nrad))
    lhs = float(np.vdot(lc.forward2d(x, g).astype(np.float64), y))
    rhs = float(np.vdot(x, lc.backproj2d(y, g, (1, N, N)).astype(np.float64)))
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + abs(rhs))


@pytest.mark.parametrize("N,nang", [(80, 40), (64, 90), (128, 180)])
def test_projector_stack_consistency(lc, N, nang):
    # slice i of a stack == slice i alone (bitwise: per-slice work is independent).
    # N = 80 takes the direct detector DFT; 64 and 128 the LDS FFT, whose
    # blocks hold several sinogram rows (90 and 180 rows are not multiples of
    # the rows per block, so a stack would shift later slices' rows)
    announce(f"projector stack consistency N={N} nang={nang}")
    rng = np.random.default_rng(9)
    g = lc.PolarGrid2D(uniform_theta(nang), N)
    vol = rng.random((4, N, N)).astype(np.float32)
    sino = lc.forward2d(vol, g)
    bp = lc.backproj2d(sino, g, vol.shape)
    for i in range(4):
        assert np.array_equal(lc.forward2d(vol[i:i + 1], g), sino[i:i + 1])
        assert np.array_equal(lc.backproj2d(sino[i:i + 1], g, (1, N, N)), bp[i:i + 1])


@pytest.mark.parametrize("N,nang,ns", [(32, 13, 2), (96, 120, 1), (256, 180, 1)])
def test_sysmat_nufft_matches_oracle_and_toeplitz(lc, orc64, N, nang, ns):
    # tests/test_sysmat_toeplitz2d.cpp:61-98: sysmat2d(x, grid) == sysmat2d(x, psf)
    # (nrad = N, where the two normalisations agree) to NUFFT accuracy
    announce(f"sysmat2d nufft N={N} nang={nang}")
    rng = np.random.default_rng(3 + N)
    th = uniform_theta(nang)
    g = lc.PolarGrid2D(th, N)
    x = rng.random((ns, N, N))
    y1 = lc.sysmat2d(x, g)
    assert rel_l2(y1, orc64.sysmat2d_nufft(x, th, N, tol=1e-12)) < PROJ_TOL
    y2 = lc.sysmat2d(x, lc.PointSpreadFunction2D(g, N))
    assert rel_l2(y1, y2) < PROJ_TOL


def test_backproj2d_bin_shapes_and_autotune(lc, orc64):
    # both compiled bin sizes of the sorted-bin spread (16 and 32 corner
    # cells) give the oracle's back-projection; a slice is the same alone and
    # in a stack for each; the autotuner installs one of them
    N, nang, ns = 128, 180, 3
    announce(f"backproj2d bin shapes N={N}")
    rng = np.random.default_rng(19)
    th = uniform_theta(nang)
    sino = rng.standard_normal((ns, nang, N)).astype(np.float32)
    xr = orc64.backproj2d(sino, th, N, tol=1e-12)
    ctx = lc.Context(0)
    try:
        g = lc.PolarGrid2D(th, N)
        psf = lc.PointSpreadFunction2D(g, N, ctx=ctx)
        xs = []
        for shape in range(2):
            ctx.set_tuning2d_nufft(psf, shape)
            assert ctx.get_tuning2d_nufft(psf) == shape
            x = lc.backproj2d(sino, g, (ns, N, N), ctx=ctx)
            err = rel_l2(x, xr)
            print(f"shape {shape}: rel_l2 = {err:.3e}")
            assert err < PROJ_TOL
            assert np.array_equal(lc.backproj2d(sino[1:2], g, (1, N, N), ctx=ctx), x[1:2])
            xs.append(x)
        assert rel_l2(xs[1], xs[0]) < 1e-5
        best, ms = ctx.autotune2d_nufft(psf, 4)
        print(f"autotune2d_nufft: shape {best}, {ms:.3f} ms")
        assert best in (0, 1) and ms > 0.0
        assert ctx.get_tuning2d_nufft(psf) == best
    finally:
        ctx.close()
