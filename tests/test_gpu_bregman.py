"""GPU Split-Bregman where the TV step does work, against the oracle whose
steps tests/test_oracle_bregman.py pins, and the reference's end-to-end
sanity reconstruction (tests/test_mbir2d_sanity.cpp:98-205) on the GPU."""
import numpy as np
import pytest

from conftest import announce, rel_l2
from test_oracle_bregman import sanity_scores, two_disk_phantom

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lc():
    import laminocam_amd
    return laminocam_amd


def _problem(orc, N, nang, ns, seed):
    from laminocam_amd import synth
    th = synth.angles(nang)
    rng = synth.Stream(seed)
    vol = np.stack([synth.ellipse_phantom(N, 8, rng) for _ in range(ns)]).astype(np.float32)
    P = orc.psf2d(th, N, N, method=0)
    return th, P, P.convolve(vol).astype(np.float32)


@pytest.mark.parametrize("zero_pad,aniso", [(False, False), (True, False), (True, True)])
def test_split_bregman_tv_active_matches_oracle(lc, orc64, zero_pad, aniso):
    """5 x 32^2 stack, lambda = 0.5, mu = 10, 4 outer x 3 inner: between 30%
    and 70% of the oracle's d values are shrunk to zero, the rest are not, so
    the threshold, the Bregman update and the rhs all act on the result."""
    announce(f"sb tv active zero_pad={zero_pad} aniso={aniso}")
    N, ns, lam, mu = 32, 5, 0.5, 10.0
    th, P, rhs = _problem(orc64, N, 45, ns, 21)
    bd = 1 if zero_pad else 0
    d = np.zeros((3,) + rhs.shape)
    b = np.zeros_like(d)
    xr, rr = orc64.split_bregman(P, rhs, np.zeros_like(rhs), lam, mu, 4, 3, 1e-5, 1e-5,
                                 boundary=bd, aniso=aniso, state=(d, b))
    act = (np.abs(d) > 0).mean() if aniso else (np.sqrt((d ** 2).sum(0)) > 0).mean()
    assert 0.3 < act < 0.7, act
    assert np.any(b)
    psf = lc.PointSpreadFunction2D(lc.PolarGrid2D(th, N), N)
    for graph in (False, True):
        x, rep = lc.split_bregman(psf, rhs, np.zeros_like(rhs), lam, mu, 4, 3, 1e-5, 1e-5,
                                  zero_pad=zero_pad, aniso=aniso, graph=graph)
        assert rep["outer"] == rr["outer"]
        err = rel_l2(x, xr)
        print(f"graph={graph} rel_l2={err:.3e}", flush=True)
        assert err < 1e-3
    # FP64 path: the same solve to FP64 rounding
    x64, _ = lc.split_bregman(psf, rhs.astype(np.float64), np.zeros(rhs.shape), lam, mu, 4, 3,
                              1e-5, 1e-5, zero_pad=zero_pad, aniso=aniso)
    assert rel_l2(x64, xr) < 1e-9


def test_mbir2d_sanity_split_bregman_gpu(lc, orc64):
    """tests/test_mbir2d_sanity.cpp:98-205 on the GPU path: forward2d,
    backproj2d, Split-Bregman (lambda 0.1, mu 5, 50 x 25, tol = xtol = 1e-5)
    with the 2-D Toeplitz operator; the reference's bar (finite, cos > 0.85,
    rel. err < 0.5), and the same scores as the oracle's run to 0.01."""
    announce("mbir2d sanity 5x128^2")
    N, nz, nt = 128, 5, 180
    ph = two_disk_phantom(nz, N)
    th = np.arange(nt) * np.pi / nt
    grid = lc.PolarGrid2D(th, N)
    sino = lc.forward2d(ph, grid)
    rhs = lc.backproj2d(sino, grid, (nz, N, N))
    psf = lc.PointSpreadFunction2D(grid, N)
    x, rep = lc.split_bregman(psf, rhs, np.zeros_like(rhs), 0.1, 5.0, 50, 25, 1e-5, 1e-5)
    rel, cos = sanity_scores(x, ph)
    assert np.all(np.isfinite(x))
    assert cos > 0.85 and rel < 0.5, (rel, cos, rep)
    so = orc64.forward2d(ph, th, N)
    ro = orc64.backproj2d(so, th, N)
    xo, _ = orc64.split_bregman(orc64.psf2d(th, N, N), ro, np.zeros_like(ro), 0.1, 5.0, 50, 25,
                                1e-5, 1e-5, boundary=0)
    rel_o, cos_o = sanity_scores(xo, ph)
    assert abs(rel - rel_o) < 0.01 and abs(cos - cos_o) < 0.01
