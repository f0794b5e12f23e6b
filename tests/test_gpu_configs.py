"""Parity at the sizes of the benchmark workloads, through the C ABI.

- The batched 2-D workload: one slice per image size N in {128, ..., 1024}
  and angle count in {90, 180, 360, 720} (its 28 size classes), solved in one
  batch for 8 CG iterations; each slice against the FP64 oracle and bitwise
  against itself solved alone.
- Split-Bregman on large single slices with zero-padded differences (the
  reference GPU boundary, where the TV shrink is active on one slice).
- Served requests at the sizes of the served stream: a 128^2 slice and an
  8 x 256^2 laminography volume (180 projections, 30 degree tilt), CG and
  Split-Bregman, each bitwise equal to its solo solve.
- Split-Bregman at the padded laminography size (29 x 721^2), FP32 against
  the FP64 path on the same device (the FP64 path matches the oracle to 1e-9
  at the sizes the oracle can run).
"""
import numpy as np
import pytest

from conftest import announce, rel_l2

pytestmark = pytest.mark.gpu

N_SET = (128, 192, 256, 384, 512, 768, 1024)
ANG_SET = (90, 180, 360, 720)


@pytest.fixture(scope="module")
def lc():
    import laminocam_amd
    return laminocam_amd


def _slice_problem(orc64, N, nang, seed):
    """noisy ellipse phantom -> FP64 oracle sinogram -> FP64 back-projection"""
    from laminocam_amd import synth
    rng = synth.Stream(seed)
    th = synth.angles(nang)
    vol = synth.ellipse_phantom(N, 10, rng)[None]
    sino = orc64.forward2d(vol, th, N)
    sino = synth.add_gaussian_noise(sino, 0.01, rng).astype(np.float64)
    return th, sino, orc64.backproj2d(sino, th, N).astype(np.float32)


def test_batch_at_the_workload_size_classes(lc, orc64):
    announce("batched CG, 28 size classes of the batched workload")
    iters = 8
    probs = []
    for i, N in enumerate(N_SET):
        for j, nang in enumerate(ANG_SET):
            th, _, rhs = _slice_problem(orc64, N, nang, 4000 + 10 * i + j)
            probs.append((N, nang, th, rhs))
    psfs = [lc.PointSpreadFunction2D(lc.PolarGrid2D(th, N), N) for N, _, th, _ in probs]
    batch = lc.Batch2D(psfs)
    for k, (_, _, _, rhs) in enumerate(probs):
        batch.set(k, batch.Y, rhs)
    batch.cg([iters] * len(probs), 1e-5, 1e-5, graph=True)
    reps = batch.reports()
    worst = 0.0
    for k, (N, nang, th, rhs) in enumerate(probs):
        xb = batch.get(k, batch.X)
        alone = lc.Batch2D([psfs[k]])
        alone.set(0, alone.Y, rhs)
        alone.cg([iters], 1e-5, 1e-5, graph=False)
        assert np.array_equal(alone.get(0, alone.X), xb), f"N={N} nang={nang}: alone != in batch"
        alone.close()
        # the oracle's PSF from a type-1 NUFFT of ones (method 1, equal to its
        # closed form to 2e-12 here and 8x faster at M = 2048)
        P = orc64.psf2d(th, N, N, method=1)
        xr, rr = orc64.cgsolver(P, rhs, np.zeros_like(rhs), iters, 1e-5, 1e-5)
        del P
        err = rel_l2(xb, xr)
        worst = max(worst, err)
        print(f"N={N:5d} nang={nang:4d}: rel_l2 {err:.3e}, iters {reps[k]['iters']}/{rr['iters']}",
              flush=True)
        assert reps[k]["iters"] == rr["iters"]
        assert err < 1e-3, f"N={N} nang={nang}"
    print(f"worst {worst:.3e}")


@pytest.mark.parametrize("N,nang", [(512, 180), (1024, 360)])
def test_split_bregman_large_slice_zero_pad(lc, orc64, N, nang):
    announce(f"split-bregman zero-padded N={N} nang={nang}")
    th, _, rhs = _slice_problem(orc64, N, nang, 4100 + N)
    psf = lc.PointSpreadFunction2D(lc.PolarGrid2D(th, N), N)
    x, rep = lc.split_bregman(psf, rhs, np.zeros_like(rhs), 0.1, 10.0, 3, 1, 1e-5, 1e-5,
                              zero_pad=True)
    P = orc64.psf2d(th, N, N, method=1)
    xr, rr = orc64.split_bregman(P, rhs, np.zeros_like(rhs), 0.1, 10.0, 3, 1, 1e-5, 1e-5,
                                 boundary=1)
    err = rel_l2(x, xr)
    # the shrink did work: the result differs from plain CG with the same
    # number of operator solves
    xc, _ = orc64.cgsolver(P, rhs, np.zeros_like(rhs), 3, 1e-5, 1e-5)
    print(f"rel_l2 vs oracle {err:.3e}; distance to CG {rel_l2(xr, xc):.3e}")
    assert rep["outer"] == rr["outer"]
    assert err < 1e-3
    assert rel_l2(xr, xc) > 1e-3


def _served_problems(lc, orc64):
    from laminocam_amd import synth
    rng = synth.Stream(4200)
    th2 = synth.angles(180)
    g2 = lc.PolarGrid2D(th2, 128)
    p2 = lc.PointSpreadFunction2D(g2, 128)
    sino2 = lc.forward2d(synth.ellipse_phantom(128, 10, rng)[None], g2)
    dims = (8, 256, 256)
    th3 = np.arange(180) * (2 * np.pi / 180)
    g3 = lc.PolarGrid(th3, np.full(180, -np.pi / 6), np.zeros(180), 64, 256)
    p3 = lc.PointSpreadFunction(g3, dims)
    proj = lc.forward(synth.ellipsoid_phantom(dims, 12, rng), g3)
    proj = synth.add_gaussian_noise(proj, 0.01, rng).astype(np.float32)
    reqs = []
    for solver in ("cg", "sb"):
        kw = dict(solver="cg", max_iter=30) if solver == "cg" else \
            dict(solver="sb", lam=0.1, mu=10.0, outer=5, inner=1)
        reqs.append(dict(psf=p2, grid=g2, sino=sino2, kw=kw))
        reqs.append(dict(psf=p3, grid=g3, sino=proj, kw=kw))
    return reqs


def _solo(lc, r):
    p, kw = r["psf"], r["kw"]
    if isinstance(p, lc.PointSpreadFunction):
        rhs = lc.backproj(r["sino"], r["grid"], p.recon_dims)
    else:
        rhs = lc.backproj2d(r["sino"], r["grid"], (1, p.image_size, p.image_size))
    x0 = np.zeros(rhs.shape, np.float32)
    if kw["solver"] == "cg":
        return lc.cgsolver(p, rhs, x0, kw["max_iter"], 1e-5, 1e-5)[0]
    return lc.split_bregman(p, rhs, x0, kw["lam"], kw["mu"], kw["outer"], kw["inner"], 1e-5,
                            1e-5)[0]


def test_served_requests_at_stream_sizes(lc, orc64):
    announce("served requests: 128^2 slice, 8x256^2 volume, CG and SB")
    reqs = _served_problems(lc, orc64)
    ref = [_solo(lc, r) for r in reqs]
    with lc.Server(hbm_budget=32 << 30, lanes=2) as srv:
        tk = [srv.submit(r["psf"], sino=r["sino"], **r["kw"]) for r in reqs]
        out = [srv.result(t, timeout=120.0)[0] for t in tk]
    for r, x, xr in zip(reqs, out, ref):
        assert x.shape == xr.shape
        assert np.array_equal(x, xr), f"{r['kw']['solver']} {x.shape}: served != solo"


def test_split_bregman_lamino_size_fp32_vs_fp64(lc):
    # 29 x 721^2 on 360 projections of 89 x 721 (the padded laminography
    # volume), 3 outer x 1 inner: the FP32 path against the FP64 one
    announce("split-bregman 29x721^2 FP32 vs FP64")
    from laminocam_amd import synth
    th = np.arange(360) * (2 * np.pi / 360)
    grid = lc.PolarGrid(th, np.full(360, -np.pi / 6), np.zeros(360), 89, 721)
    psf = lc.PointSpreadFunction(grid, (29, 721, 721))
    ph = synth.ellipsoid_phantom((29, 721, 721), 12, synth.Stream(4300), piecewise=True)
    y = psf.convolve(ph.astype(np.float32))
    x32, r32 = lc.split_bregman(psf, y, np.zeros_like(y), 0.1, 10.0, 3, 1, 1e-12, 0.0)
    y64 = y.astype(np.float64)
    x64, r64 = lc.split_bregman(psf, y64, np.zeros_like(y64), 0.1, 10.0, 3, 1, 1e-12, 0.0)
    err = rel_l2(x32, x64)
    print(f"FP32 vs FP64 rel_l2 {err:.3e}; outer {r32['outer']}/{r64['outer']}")
    assert x64.dtype == np.float64 and r32["outer"] == r64["outer"] == 3
    assert err < 1e-4
