"""GPU parity of the 3-D laminography path against the FP64 oracle and the
committed golden vectors, through the C ABI: PointSpreadFunction and
sysmat(x, psf) (include/toeplitz.h:61-129), CG and Split-Bregman
(src/conjgrad.cpp:33-76, src/bregman.cpp:36-94), forward / backproj / NUFFT
sysmat (src/projection.cpp:42-131, src/gradient.cpp:32-41)."""
import os

import numpy as np
import pytest

from conftest import announce, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lc():
    import laminocam_amd
    return laminocam_amd


def lamino_angles(n, tilt_deg=-30.0, span=2 * np.pi):
    th = np.arange(n) * span / n
    return th, np.full(n, np.deg2rad(tilt_deg)), np.zeros(n)


# (recon dims, nproj, nrows, ncols, tilt): odd/even dims, PSF sizes 2^p 3^q,
# a single slice (M1 = 1), no tilt
SHAPES = [((5, 16, 16), 7, 9, 16, -30.0), ((3, 33, 33), 11, 12, 30, -30.0),
          ((8, 40, 40), 30, 16, 40, 20.0), ((1, 24, 24), 9, 6, 24, -30.0),
          ((6, 32, 32), 13, 10, 32, 0.0)]


@pytest.mark.parametrize("dims,nproj,nrows,ncols,tilt", SHAPES)
def test_kernel_hat3d_matches_oracle(lc, orc64, dims, nproj, nrows, ncols, tilt):
    announce(f"kernel_hat3d dims={dims} nproj={nproj}")
    th, ga, be = lamino_angles(nproj, tilt)
    psf = lc.PointSpreadFunction(lc.PolarGrid(th, ga, be, nrows, ncols), dims)
    ref = orc64.psf3d(th, ga, be, nrows, ncols, dims, method=0)
    assert tuple(psf.dims()) == tuple(ref.dims)
    kh, kr = psf.kernel_hat(), ref.kernel_hat()
    assert kh.shape == kr.shape
    assert rel_l2(kh, kr) < 2e-6


@pytest.mark.parametrize("dims,nproj,nrows,ncols,tilt", SHAPES)
def test_toeplitz3d_matches_oracle(lc, orc64, dims, nproj, nrows, ncols, tilt):
    announce(f"toeplitz3d dims={dims} nproj={nproj}")
    rng = np.random.default_rng(sum(dims) + nproj)
    th, ga, be = lamino_angles(nproj, tilt)
    psf = lc.PointSpreadFunction(lc.PolarGrid(th, ga, be, nrows, ncols), dims)
    x = rng.random(dims).astype(np.float32)   # float32: the FP32 operator
    y = psf.convolve(x)
    yr = orc64.psf3d(th, ga, be, nrows, ncols, dims, method=0).convolve(x)
    err = rel_l2(y, yr)
    print(f"toeplitz3d rel_l2 = {err:.3e}")
    assert err < 1e-5


@pytest.mark.parametrize("dims,nproj,nrows", [((3, 721, 721), 6, 8), ((2, 576, 576), 5, 6)])
def test_toeplitz3d_large_inplane_matches_oracle(lc, orc64, dims, nproj, nrows):
    # in-plane PSF lengths above 1024 (1458 as in the laminography volume,
    # 1152), which take the 512-thread row passes, on a short slice axis
    announce(f"toeplitz3d large in-plane dims={dims}")
    rng = np.random.default_rng(dims[1])
    th, ga, be = lamino_angles(nproj)
    psf = lc.PointSpreadFunction(lc.PolarGrid(th, ga, be, nrows, dims[2]), dims)
    assert psf.dims()[1] > 1024
    x = rng.random(dims).astype(np.float32)
    y = psf.convolve(x)
    yr = orc64.psf3d(th, ga, be, nrows, dims[2], dims, method=0).convolve(x)
    err = rel_l2(y, yr)
    print(f"toeplitz3d {dims} rel_l2 = {err:.3e}")
    assert err < 1e-5


def test_toeplitz3d_lamino_size_symmetric_psd(lc):
    # the padded laminography volume of the 3-D configuration: 29 x 721 x 721
    # on 360 projections of 89 x 721 at 30 degrees tilt.  A^T A is symmetric and
    # positive semi-definite whatever the size (tests/test_adjoint.cpp)
    announce("toeplitz3d 29x721x721 symmetry")
    th, ga, be = lamino_angles(360)
    psf = lc.PointSpreadFunction(lc.PolarGrid(th, ga, be, 89, 721), (29, 721, 721))
    assert tuple(psf.dims()) == (64, 1458, 1458)
    rng = np.random.default_rng(11)
    x = rng.standard_normal((29, 721, 721)).astype(np.float32)
    z = rng.standard_normal((29, 721, 721)).astype(np.float32)
    ax, az = psf.convolve(x), psf.convolve(z)
    lhs = float(np.vdot(ax.astype(np.float64), z))
    rhs = float(np.vdot(x.astype(np.float64), az))
    assert abs(lhs - rhs) <= 1e-4 * np.linalg.norm(ax) * np.linalg.norm(z)
    assert float(np.vdot(ax.astype(np.float64), x)) > 0.0


@pytest.mark.parametrize("dims,nproj,nrows,ncols,tilt", SHAPES[:3])
def test_cg3d_matches_oracle(lc, orc64, dims, nproj, nrows, ncols, tilt):
    announce(f"cg3d dims={dims}")
    rng = np.random.default_rng(5)
    th, ga, be = lamino_angles(nproj, tilt)
    psf = lc.PointSpreadFunction(lc.PolarGrid(th, ga, be, nrows, ncols), dims)
    P = orc64.psf3d(th, ga, be, nrows, ncols, dims, method=0)
    xt = rng.random(dims)
    y = P.convolve(xt).astype(np.float32)
    x, rep = lc.cgsolver(psf, y, np.zeros(dims, np.float32), 5, 1e-12, 0.0)
    xr, rr = orc64.cgsolver(P, y, np.zeros(dims), 5, 1e-12, 0.0)
    assert rep["iters"] == rr["iters"]
    err = rel_l2(x, xr)
    print(f"cg3d rel_l2 = {err:.3e}")
    assert err < 1e-3


@pytest.mark.parametrize("zero_pad", [False, True])
def test_split_bregman3d_matches_oracle(lc, orc64, zero_pad):
    announce(f"split_bregman3d zero_pad={zero_pad}")
    dims = (5, 16, 16)
    rng = np.random.default_rng(8)
    th, ga, be = lamino_angles(9)
    psf = lc.PointSpreadFunction(lc.PolarGrid(th, ga, be, 9, 16), dims)
    P = orc64.psf3d(th, ga, be, 9, 16, dims, method=0)
    y = P.convolve(rng.random(dims)).astype(np.float32)
    x, rep = lc.split_bregman(psf, y, np.zeros(dims, np.float32), 0.1, 10.0, 6, 1, 1e-12, 0.0,
                              zero_pad=zero_pad)
    xr, rr = orc64.split_bregman(P, y, np.zeros(dims), 0.1, 10.0, 6, 1, 1e-12, 0.0,
                                 boundary=1 if zero_pad else 0)
    err = rel_l2(x, xr)
    print(f"sb3d rel_l2 = {err:.3e}")
    assert err < 1e-3


def test_batch_mixed_2d_3d_same_alone_and_in_mix(lc):
    # a 3-D volume batched with 2-D stacks keeps its own CG scalars: its result
    # is bitwise the result of solving it alone
    announce("mixed 2-D / 3-D batch")
    rng = np.random.default_rng(21)
    dims = (6, 32, 32)
    th, ga, be = lamino_angles(13)
    p3 = lc.PointSpreadFunction(lc.PolarGrid(th, ga, be, 10, 32), dims)
    p2a = lc.PointSpreadFunction2D(lc.PolarGrid2D(np.arange(30) * np.pi / 30, 48), 48)
    p2b = lc.PointSpreadFunction2D(lc.PolarGrid2D(np.arange(20) * np.pi / 20, 32), 32)
    f32 = np.float32
    y3 = p3.convolve(rng.random(dims).astype(f32))
    y2a = p2a.convolve(rng.random((2, 48, 48)).astype(f32))
    y2b = p2b.convolve(rng.random((1, 32, 32)).astype(f32))
    mix = lc.Batch2D([p2a, p3, p2b], ns=[2, 0, 1])
    for i, y in enumerate([y2a, y3, y2b]):
        mix.set(i, mix.Y, y)
    mix.cg(8, 0.0, 0.0, graph=True)
    alone = lc.Batch2D([p3])
    alone.set(0, alone.Y, y3)
    alone.cg(8, 0.0, 0.0, graph=True)
    assert np.array_equal(mix.get(1, mix.X), alone.get(0, alone.X))
    single, _ = lc.cgsolver(p3, y3, np.zeros(dims, np.float32), 8, 0.0, 0.0)
    assert np.array_equal(single, alone.get(0, alone.X))


# ---------------------------------------------------------------- projectors
GOLDEN3 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden3d.npz")
PROJ_TOL = 1e-4


def test_projectors3d_match_golden(lc):
    # brute-force NUDFT + numpy FFT restatement of the reference chain (odd,
    # non-2^p3^q detector: Bluestein path; rectangular volume)
    announce("projector3d golden")
    d = np.load(GOLDEN3)
    g = lc.PolarGrid(d["theta"], d["gamma"], d["beta"], int(d["det"][0]), int(d["det"][1]))
    fwd = lc.forward(d["vol"], g)
    assert fwd.shape == d["fwd"].shape
    assert rel_l2(fwd, d["fwd"]) < PROJ_TOL
    assert rel_l2(lc.backproj(d["proj"], g, tuple(d["dims"])), d["bp"]) < PROJ_TOL
    assert rel_l2(lc.backproj(d["proj"], g, tuple(d["dims"]), d["shifts"]), d["bp_shift"]) < PROJ_TOL


# (recon dims, nproj, nrows, ncols, tilt)
PSHAPES = [((6, 24, 24), 20, 16, 24, -30.0), ((5, 21, 19), 13, 11, 23, -30.0),
           ((9, 32, 32), 36, 13, 31, 20.0), ((1, 30, 30), 12, 8, 32, -30.0)]


@pytest.mark.parametrize("dims,nproj,nrows,ncols,tilt", PSHAPES)
def test_forward3d_matches_oracle(lc, orc64, dims, nproj, nrows, ncols, tilt):
    announce(f"forward3d dims={dims} det=({nrows},{ncols})")
    rng = np.random.default_rng(31 + nproj)
    th, ga, be = lamino_angles(nproj, tilt)
    vol = rng.random(dims)
    y = lc.forward(vol, lc.PolarGrid(th, ga, be, nrows, ncols))
    yr = orc64.forward3d(vol, th, ga, be, nrows, ncols, tol=1e-12)
    err = rel_l2(y, yr)
    print(f"forward3d rel_l2 = {err:.3e}")
    assert err < PROJ_TOL


@pytest.mark.parametrize("dims,nproj,nrows,ncols,tilt", PSHAPES)
def test_backproj3d_matches_oracle(lc, orc64, dims, nproj, nrows, ncols, tilt):
    announce(f"backproj3d dims={dims} det=({nrows},{ncols})")
    rng = np.random.default_rng(41 + nproj)
    th, ga, be = lamino_angles(nproj, tilt)
    proj = rng.standard_normal((nproj, nrows, ncols))
    g = lc.PolarGrid(th, ga, be, nrows, ncols)
    x = lc.backproj(proj, g, dims)
    xr = orc64.backproj3d(proj, th, ga, be, dims, tol=1e-12)
    err = rel_l2(x, xr)
    print(f"backproj3d rel_l2 = {err:.3e}")
    assert err < PROJ_TOL
    sh = rng.uniform(-2, 2, (nproj, 2))
    xs = lc.backproj(proj, g, dims, sh)
    xsr = orc64.backproj3d(proj, th, ga, be, dims, shifts=sh, tol=1e-12)
    assert rel_l2(xs, xsr) < PROJ_TOL


@pytest.mark.parametrize("dims,nproj,nrows,ncols,tilt", [((7, 36, 53), 90, 24, 53, -30.0),
                                                          ((16, 96, 96), 90, 32, 96, -30.0)])
def test_backproj3d_mid_size_matches_oracle(lc, orc64, dims, nproj, nrows, ncols, tilt):
    # fine grids (16, 72, 108) and (32, 192, 192): an odd number of spread bins
    # with a partial last bin, and many points per bin
    announce(f"backproj3d mid size dims={dims} det=({nrows},{ncols})")
    rng = np.random.default_rng(51 + nproj)
    th, ga, be = lamino_angles(nproj, tilt)
    proj = rng.standard_normal((nproj, nrows, ncols))
    g = lc.PolarGrid(th, ga, be, nrows, ncols)
    sh = rng.uniform(-2, 2, (nproj, 2))
    x = lc.backproj(proj, g, dims, sh)
    xr = orc64.backproj3d(proj, th, ga, be, dims, shifts=sh, tol=1e-9)
    err = rel_l2(x, xr)
    print(f"backproj3d rel_l2 = {err:.3e}")
    assert err < PROJ_TOL
    # deterministic: the same call gives the same bits
    assert np.array_equal(lc.backproj(proj, g, dims, sh), x)


def test_backproj3d_bin_shapes_and_autotune(lc, orc64):
    # every compiled bin shape of the spread (0: 4x4x8, 1: 8x4x4, 2: 8x8x8
    # corner cells) gives the oracle's back-projection, deterministically;
    # the autotuner installs one of them for the grid's fine-grid shape
    dims, nproj, nrows, ncols = (16, 96, 96), 90, 32, 96
    announce(f"backproj3d bin shapes dims={dims}")
    rng = np.random.default_rng(77)
    th, ga, be = lamino_angles(nproj, -30.0)
    proj = rng.standard_normal((nproj, nrows, ncols))
    g = lc.PolarGrid(th, ga, be, nrows, ncols)
    sh = rng.uniform(-2, 2, (nproj, 2))
    xr = orc64.backproj3d(proj, th, ga, be, dims, shifts=sh, tol=1e-9)
    ctx = lc.Context(0)
    try:
        psf = lc.PointSpreadFunction(g, dims, ctx=ctx)
        xs = []
        for shape in range(3):
            ctx.set_tuning3d(psf, shape)
            assert ctx.get_tuning3d(psf) == shape
            x = lc.backproj(proj, g, dims, sh, ctx=ctx)
            err = rel_l2(x, xr)
            print(f"shape {shape}: rel_l2 = {err:.3e}")
            assert err < PROJ_TOL
            assert np.array_equal(lc.backproj(proj, g, dims, sh, ctx=ctx), x)
            xs.append(x)
        for x in xs[1:]:
            assert rel_l2(x, xs[0]) < 1e-5
        best, ms = ctx.autotune3d(psf)
        print(f"autotune3d: shape {best}, {ms:.3f} ms")
        assert best in (0, 1, 2) and ms > 0.0
        assert ctx.get_tuning3d(psf) == best
        assert np.array_equal(lc.backproj(proj, g, dims, sh, ctx=ctx), xs[best])
    finally:
        ctx.close()


def test_backproj3d_shift_count_mismatch(lc):
    # src/projection.cpp:105-108: std::invalid_argument
    th, ga, be = lamino_angles(4)
    g = lc.PolarGrid(th, ga, be, 4, 8)
    with pytest.raises(ValueError):
        lc.backproj(np.zeros((4, 4, 8)), g, (2, 8, 8), np.zeros((3, 2)))


@pytest.mark.parametrize("dims,nproj,nrows,ncols", [((4, 32, 32), 24, 16, 32), ((8, 48, 48), 60, 32, 48)])
def test_projectors3d_adjoint(lc, dims, nproj, nrows, ncols):
    # even detector dims: fftshift2 == ifftshift2, so backproj is the adjoint of forward
    announce(f"adjoint3d dims={dims}")
    rng = np.random.default_rng(2)
    th, ga, be = lamino_angles(nproj)
    g = lc.PolarGrid(th, ga, be, nrows, ncols)
    x = rng.random(dims)
    y = rng.standard_normal((nproj, nrows, ncols))
    lhs = float(np.vdot(lc.forward(x, g).astype(np.float64), y))
    rhs = float(np.vdot(x, lc.backproj(y, g, dims).astype(np.float64)))
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + abs(rhs))


@pytest.mark.parametrize("tilt", [0.0, -30.0])
def test_sysmat3d_nufft_matches_oracle_and_toeplitz(lc, orc64, tilt):
    # tests/test_sysmat_toeplitz.cpp:48-127: NUFFT sysmat == Toeplitz sysmat when
    # the detector matches the image (nrows * ncols = n2 * n3)
    announce(f"sysmat3d nufft tilt={tilt}")
    rng = np.random.default_rng(4)
    dims = (5, 16, 16)
    th, ga, be = lamino_angles(9, tilt)
    g = lc.PolarGrid(th, ga, be, 16, 16)
    x = rng.random(dims)
    y1 = lc.sysmat(x, g)
    assert rel_l2(y1, orc64.sysmat3d_nufft(x, th, ga, be, 16, 16, tol=1e-12)) < PROJ_TOL
    y2 = lc.sysmat(x, lc.PointSpreadFunction(g, dims))
    assert rel_l2(y1, y2) < PROJ_TOL


def test_projectors3d_lamino_size(lc):
    # the laminography configuration's shapes: volume 29 x 721 x 721,
    # 360 projections of 89 x 721 (Bluestein on both detector axes); adjoint
    # pairing holds up to the odd-length shift convention, so check the
    # forward/backward sizes and finiteness plus linearity
    announce("projectors3d 29x721x721 / 360x89x721")
    th, ga, be = lamino_angles(360)
    g = lc.PolarGrid(th, ga, be, 89, 721)
    rng = np.random.default_rng(6)
    x = rng.random((29, 721, 721)).astype(np.float32)
    p = lc.forward(x, g)
    assert p.shape == (360, 89, 721) and np.all(np.isfinite(p))
    p2 = lc.forward(2.0 * x, g)
    assert rel_l2(p2, 2.0 * p) < 1e-6
    b = lc.backproj(p, g, (29, 721, 721))
    assert b.shape == (29, 721, 721) and np.all(np.isfinite(b))
    assert float(np.vdot(b.astype(np.float64), x)) > 0.0
