"""GPU parity of the whole-solve drivers against the oracle's restatement
(oracle/drivers.py) and against the same steps called one by one through the
library: prepare_sinogram (include/mbir2d_prep.h:44-65), the 2-D slice-stack
solve (src/mbir2d.cpp:66-100) and the 3-D MBIR (src/mbir.cpp:34-161)."""
import numpy as np
import pytest

from conftest import announce, rel_l2
from oracle import drivers as od

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lc():
    import laminocam_amd
    return laminocam_amd


def _params(lc, reg, dims, iters, inner=1, pad=1.4142):
    return lc.ReconParams(regularizer=lc.Regularizer.SPLIT_BREGMAN if reg == "sb" else
                          lc.Regularizer.UNCONSTRAINED, recon_dims=dims, maxIters=iters,
                          innerIters=inner, lam=0.1, mu=10.0, tol=1e-5, xtol=1e-5, PAD_FACTOR=pad)


def _oparams(reg, dims, iters, inner=1, pad=1.4142):
    # the float32 values the reference's ReconParams holds
    f = lambda v: float(np.float32(v))  # noqa: E731
    return {"regularizer": reg, "recon_dims": dims, "maxIters": iters, "innerIters": inner,
            "lam": f(0.1), "mu": f(10.0), "tol": f(1e-5), "xtol": f(1e-5), "PAD_FACTOR": pad}


# ---- prepare_sinogram ----
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("nangs,nrows,ncols,sign", [((7, 5), 4, 33, 1), ((1,), 1, 1, 1),
                                                    ((90, 91, 1), 17, 64, 1), ((3, 2), 5, 8, -1)])
def test_prepare_sinogram_bit_exact(lc, dtype, nangs, nrows, ncols, sign):
    announce(f"prepare_sinogram {dtype.__name__} nangs={nangs} {nrows}x{ncols} sign={sign}")
    rng = np.random.default_rng(sum(nangs) + nrows)
    projs = [sign * (rng.random((n, nrows, ncols)) * 7).astype(dtype) for n in nangs]
    angles = [rng.random(n) for n in nangs]
    ds = [lc.Dataset(p, a) for p, a in zip(projs, angles)]
    sino, theta = lc.prepare_sinogram(ds)
    ref, rth = od.prepare_sinogram(projs, angles, dtype)
    assert sino.dtype == dtype and sino.shape == ref.shape
    assert np.array_equal(sino, ref)          # x / max rounded once, as the host divides
    assert np.array_equal(theta, rth)


# ---- 2-D driver ----
def _sino2d(orc64, nangs, ns, N, seed):
    # projections of random-ellipse phantoms with 1% noise, split into datasets
    from laminocam_amd import synth
    rng = synth.Stream(seed)
    ntot = sum(nangs)
    theta = synth.angles(ntot)
    vol = np.stack([synth.ellipse_phantom(N, 10, rng) for _ in range(ns)])
    sino = synth.add_gaussian_noise(orc64.forward2d(vol, theta, N), 0.01, rng)  # {ns, ntot, N}
    proj = sino.transpose(1, 0, 2).astype(np.float32)     # {ntot, rows = ns, N}
    out, o = [], 0
    for n in nangs:
        out.append(np.ascontiguousarray(proj[o:o + n]))
        o += n
    return out, np.split(theta, np.cumsum(nangs)[:-1])


@pytest.mark.parametrize("reg,iters", [("cg", 6), ("sb", 6)])
def test_mbir2d_matches_oracle(lc, orc64, reg, iters):
    announce(f"mbir2d {reg}")
    N, ns = 32, 3
    projs, angles = _sino2d(orc64, (60, 30), ns, N, 4)
    sino, theta = lc.prepare_sinogram([lc.Dataset(p, a) for p, a in zip(projs, angles)])
    rep = {}
    x = lc.MBIR2D(sino, theta, _params(lc, reg, (ns, N, N), iters), report=rep)
    xr, rr = od.mbir2d(orc64, sino, theta, _oparams(reg, (ns, N, N), iters))
    err = rel_l2(x, xr)
    print(f"mbir2d {reg} rel_l2 = {err:.3e} iters {rep['iters']} / {rr['iters']}")
    assert rep["iters"] == rr["iters"]
    assert err < 1e-3


def test_mbir2d_equals_its_steps(lc, orc64):
    # the driver is exactly grid -> PSF -> backproj2d -> cgsolver from zeros
    announce("mbir2d == steps")
    N, ns, n1 = 48, 4, 3
    projs, angles = _sino2d(orc64, (45,), ns, N, 9)
    sino, theta = lc.prepare_sinogram([lc.Dataset(projs[0], angles[0])])
    x = lc.MBIR2D(sino, theta, _params(lc, "cg", (n1, N, N), 10))
    grid = lc.PolarGrid2D(theta, N)
    rhs = lc.backproj2d(sino[:n1], grid, (n1, N, N))
    xs, _ = lc.cgsolver(lc.PointSpreadFunction2D(grid, N), rhs, np.zeros_like(rhs), 10,
                        np.float32(1e-5), np.float32(1e-5))
    assert np.array_equal(x, xs)


def test_mbir2d_fp64(lc, orc64):
    announce("mbir2d fp64")
    N, ns = 24, 2
    projs, angles = _sino2d(orc64, (40,), ns, N, 5)
    sino, theta = lc.prepare_sinogram([lc.Dataset(projs[0].astype(np.float64), angles[0])])
    assert sino.dtype == np.float64
    x = lc.MBIR2D(sino, theta, _params(lc, "cg", (ns, N, N), 6))
    # the oracle on the FP32-rounded sinogram the FP32 back-projector reads
    xr, _ = od.mbir2d(orc64, sino.astype(np.float32), theta, _oparams("cg", (ns, N, N), 6))
    assert x.dtype == np.float64
    # the right-hand side comes from the FP32 back-projector: the projection bar
    assert rel_l2(x, xr) < 1e-4


def test_mbir2d_rejects_bad_dims(lc):
    announce("mbir2d errors")
    sino = np.ones((2, 10, 16), np.float32)
    th = np.arange(10) * 0.3
    with pytest.raises(lc.LcError):
        lc.MBIR2D(sino, th, _params(lc, "cg", (3, 16, 16), 2))     # more slices than rows
    with pytest.raises(lc.LcError):
        lc.MBIR2D(sino, th, _params(lc, "cg", (2, 16, 17), 2))     # not square in-plane


# ---- 3-D driver ----
def _lamino_datasets(orc64, nangs, nrows, ncols, tilts, seed, shifts=True, dtype=np.float32):
    # per dataset: projections of one random-ellipsoid phantom at its own
    # tilt, 1% noise, and random sub-pixel shifts
    from laminocam_amd import synth
    srng = synth.Stream(seed)
    rng = np.random.default_rng(seed)
    vol = synth.ellipsoid_phantom((max(2, nrows // 2), ncols, ncols), 6, srng)
    out = []
    for k, (n, t) in enumerate(zip(nangs, tilts)):
        a = np.arange(n) * 2 * np.pi / n + 0.1 * k
        p = orc64.forward3d(vol, a, np.full(n, np.deg2rad(t)), np.full(n, 0.05 * k), nrows, ncols)
        p = synth.add_gaussian_noise(p, 0.01, srng).astype(dtype)
        sh = rng.uniform(-1, 1, (n, 2)) if shifts else None
        out.append((p, a, np.deg2rad(t), 0.05 * k, sh))
    return out


@pytest.mark.parametrize("reg,iters,pad,dims", [("cg", 5, 1.4142, (3, 12, 12)),
                                                ("sb", 4, 1.4142, (3, 12, 12)),
                                                ("cg", 5, 1.0, (5, 11, 11)),
                                                ("cg", 4, 2.0, (2, 8, 8))])
def test_mbir3d_matches_oracle(lc, orc64, reg, iters, pad, dims):
    announce(f"mbir3d {reg} pad={pad} dims={dims}")
    ds = _lamino_datasets(orc64, (7, 6), 6, 16, (-30.0, -25.0), 3)
    rep = {}
    x = lc.MBIR([lc.Dataset(*d) for d in ds], _params(lc, reg, dims, iters, pad=pad), report=rep)
    xr, rr, odims = od.mbir3d(orc64, [d[0] for d in ds], [d[1] for d in ds], [d[2] for d in ds],
                              [d[3] for d in ds], [d[4] for d in ds],
                              _oparams(reg, dims, iters, pad=pad))
    assert rep["padded_dims"][:3] == odims
    err = rel_l2(x, xr)
    print(f"mbir3d rel_l2 = {err:.3e} padded {rep['padded_dims']}")
    assert rep["iters"] == rr["iters"] and rep["outer"] == rr["outer"]
    assert x.shape == dims
    assert err < 1e-3


def test_mbir3d_equals_its_steps(lc, orc64):
    # normalise + pad on the host, then backproj (shifts) -> PSF -> cgsolver
    # from zeros -> crop, each through the library
    announce("mbir3d == steps")
    ds = _lamino_datasets(orc64, (9, 4), 7, 18, (-30.0, -20.0), 7)
    dims = (4, 13, 13)
    p = _params(lc, "cg", dims, 6)
    x = lc.MBIR([lc.Dataset(*d) for d in ds], p)
    stacked = np.concatenate([d[0] for d in ds])
    padded = od.pad2d(stacked / stacked.max(), 1.4142)
    theta = np.concatenate([d[1] for d in ds])
    gamma = np.concatenate([np.full(len(d[1]), d[2]) for d in ds])
    beta = np.concatenate([np.full(len(d[1]), d[3]) for d in ds])
    shifts = np.concatenate([d[4] for d in ds])
    odims = tuple(od.pad_size(n, 1.4142) for n in dims)
    grid = lc.PolarGrid(theta, gamma, beta, *padded.shape[1:])
    rhs = lc.backproj(padded, grid, odims, shifts=shifts)
    xs, _ = lc.cgsolver(lc.PointSpreadFunction(grid, odims), rhs, np.zeros(odims, np.float32), 6,
                        np.float32(1e-5), np.float32(1e-5))
    assert np.array_equal(x, od.crop3d(xs, dims))


def test_mbir3d_fp64_and_no_shifts(lc, orc64):
    announce("mbir3d fp64, no shifts")
    ds = _lamino_datasets(orc64, (8,), 5, 12, (-30.0,), 2, shifts=False, dtype=np.float64)
    dims = (3, 9, 9)
    x = lc.MBIR([lc.Dataset(*d) for d in ds], _params(lc, "cg", dims, 5))
    assert x.dtype == np.float64
    xr, _, _ = od.mbir3d(orc64, [d[0] for d in ds], [d[1] for d in ds], [d[2] for d in ds],
                         [d[3] for d in ds], None, _oparams("cg", dims, 5), dtype=np.float64)
    # the right-hand side comes from the FP32 back-projector: the projection bar
    assert rel_l2(x, xr) < 1e-4


def test_mbir3d_rejects_bad_input(lc, orc64):
    announce("mbir3d errors")
    ds = _lamino_datasets(orc64, (4,), 4, 8, (-30.0,), 1)
    with pytest.raises(lc.LcError):       # padded volume 7 < 8 at factor 1
        lc.MBIR([lc.Dataset(*d) for d in ds], _params(lc, "cg", (3, 8, 8), 2, pad=1.0))
    with pytest.raises(ValueError):       # shifts count mismatch (src/projection.cpp:105-108)
        p, a, g, b, sh = ds[0]
        lc.MBIR([lc.Dataset(p, a, g, b, sh[:2])], _params(lc, "cg", (3, 8, 8), 2))


def test_mbir3d_lamino_size_scale_invariant(lc):
    # the laminography problem at full size: 360 projections of 64 x 511 at
    # 30 degrees tilt, recon_dims 21 x 511 x 511 -> padded volume 29 x 721 x 721
    # and detector 89 x 721.  Normalisation by the maximum makes the result
    # independent of a power-of-two scale of the data, bit for bit.
    announce("mbir3d laminography size, 2 CG iterations")
    rng = np.random.default_rng(12)
    proj = rng.random((360, 64, 511), dtype=np.float32)
    ang = np.arange(360) * 2 * np.pi / 360
    p = _params(lc, "cg", (21, 511, 511), 2)
    rep = {}
    x1 = lc.MBIR([lc.Dataset(proj, ang, np.deg2rad(-30.0))], p, report=rep)
    assert rep["padded_dims"] == (29, 721, 721, 89, 721)
    assert rep["iters"] == 2 and np.isfinite(x1).all() and np.abs(x1).max() > 0
    x2 = lc.MBIR([lc.Dataset(proj * np.float32(4), ang, np.deg2rad(-30.0))], p)
    assert np.array_equal(x1, x2)
