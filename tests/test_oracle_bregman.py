"""Pins of the oracle's Split-Bregman steps (no GPU).

The oracle restates src/bregman.cpp:36-94 as three exposed steps, composed by
or_split_bregman:
  rhs = yT - mu div(d - b)                     (src/bregman.cpp:58-61)
  d   = shrink(grad x + b, lambda / mu)        (src/bregman.cpp:66-80)
  b  += grad x - d                             (src/bregman.cpp:82-84)
Each step is pinned here by what the mathematics fixes, not by retyping it:
  * the shrink is the proximal map of (lambda/mu)|.|_2 per point (|.|_1 per
    component when anisotropic), checked through its optimality condition;
  * rhs is checked piece by piece through linearity against the pinned
    divergence, so a sign on d or b that is wrong fails;
  * the composed solver with A = I and zero-padded differences (where -div is
    the adjoint of grad) must reach the minimiser of the ROF energy
    0.5|x - y|^2 + lambda sum_i |(grad x)_i|_2, computed independently by a
    primal-dual (Chambolle-Pock) iteration; planted mistakes in any step fail;
  * at a fixed point d = grad x and mu b is a subgradient of lambda |d|;
  * the reference's own end-to-end sanity reconstruction
    (tests/test_mbir2d_sanity.cpp:68-205) holds its bar.
"""
import numpy as np
import pytest

from conftest import rel_l2


def _fields(seed, shape=(3, 4, 6, 7), scale=1.0):
    rng = np.random.default_rng(seed)
    return scale * rng.standard_normal(shape), scale * rng.standard_normal(shape)


# ---------------------------------------------------------------- shrink
@pytest.mark.parametrize("bits", [64, 32])
def test_sb_shrink_closed_form(request, bits):
    """tests/test_bregman_gpu_cpu.cu:144-212: compute_sk and shrink on random
    fields, with lambda/mu between the smallest and largest |g + b| so that both
    branches of the max occur."""
    o = request.getfixturevalue(f"orc{bits}")
    g, b = _fields(1)
    lam, mu = 1.1, 0.7
    d = o.sb_shrink(g, b, lam, mu)
    v = (g + b).astype(o.dtype)
    sk = np.sqrt((v.astype(np.float64) ** 2).sum(0))
    want = np.maximum(0.0, sk - lam / mu) * v / (sk + 1e-8)
    tol = 1e-14 if bits == 64 else 2e-6
    assert np.abs(d - want).max() <= tol * np.abs(want).max()
    zero = sk <= lam / mu
    assert 0.2 < zero.mean() < 0.8
    assert not np.any(d[:, zero])


@pytest.mark.parametrize("aniso", [False, True])
def test_sb_shrink_is_proximal_map(orc64, aniso):
    """d = argmin_z t N(z) + |z - v|^2 / 2 with t = lambda/mu and N the point's
    Euclidean norm (isotropic) or the component's absolute value
    (anisotropic).  Optimality: v - d = t d/N(d) where d != 0, N(v) <= t where
    d = 0.  The 1e-8 in the reference's denominator moves d by t*1e-8/N(v)."""
    g, b = _fields(2)
    lam, mu = 0.9, 0.6
    t = lam / mu
    d = orc64.sb_shrink(g, b, lam, mu, aniso=aniso)
    v = g + b
    if aniso:
        nz = d != 0
        assert np.all(np.abs(v[~nz]) <= t)
        assert np.abs((v - d)[nz] - t * np.sign(d[nz])).max() < 1e-12
        assert 0.2 < nz.mean() < 0.8
    else:
        nd = np.sqrt((d ** 2).sum(0))
        nv = np.sqrt((v ** 2).sum(0))
        nz = nd > 0
        assert np.all(nv[~nz] <= t)
        resid = (v - d)[:, nz] - t * d[:, nz] / nd[nz]
        assert np.abs(resid).max() < 1e-7
        assert 0.2 < nz.mean() < 0.8
    # and it is the minimiser: random perturbations never lower the objective
    rng = np.random.default_rng(3)

    def obj(z):
        n = np.abs(z) if aniso else np.sqrt((z ** 2).sum(0))
        return t * n.sum() + 0.5 * ((z - v) ** 2).sum()
    f0 = obj(d)
    for _ in range(20):
        assert obj(d + 1e-3 * rng.standard_normal(d.shape)) >= f0 - 1e-9


def test_sb_bregman_update(orc64):
    g, b = _fields(4)
    d = orc64.sb_shrink(g, b, 0.3, 1.0)
    bn = orc64.sb_bregman_update(b, g, d)
    # b + g - d, and a fixed point d = g leaves b unchanged
    assert np.array_equal(bn, b + (g - d))
    assert np.array_equal(orc64.sb_bregman_update(b, g, g), b)


# ---------------------------------------------------------------- rhs
@pytest.mark.parametrize("boundary", [0, 1])
def test_sb_rhs_by_linearity(orc64, boundary):
    """rhs(y, d, b) = y - mu div(d) + mu div(b): each term on its own against
    the divergence (pinned by closed forms and adjointness elsewhere)."""
    rng = np.random.default_rng(5)
    sh = (4, 6, 7)
    y = rng.standard_normal(sh)
    d, b = _fields(6, (3,) + sh)
    z3 = np.zeros((3,) + sh)
    mu = 2.5
    assert np.array_equal(orc64.sb_rhs(y, z3, z3, mu, boundary), y)
    assert np.allclose(orc64.sb_rhs(0 * y, d, z3, mu, boundary),
                       -mu * orc64.div(d, boundary), rtol=0, atol=1e-13)
    assert np.allclose(orc64.sb_rhs(0 * y, z3, b, mu, boundary),
                       +mu * orc64.div(b, boundary), rtol=0, atol=1e-13)
    full = orc64.sb_rhs(y, d, b, mu, boundary)
    parts = y - mu * orc64.div(d, boundary) + mu * orc64.div(b, boundary)
    assert np.allclose(full, parts, rtol=0, atol=1e-12)
    assert np.any(orc64.div(d, boundary))


# ---------------------------------------------------------------- composition
@pytest.mark.parametrize("boundary,aniso", [(0, False), (1, False), (1, True)])
def test_sb_solver_is_composition_of_steps(orc64, boundary, aniso):
    """or_split_bregman == the three pinned steps and CG on A - mu lap, composed
    here outer iteration by outer iteration (bitwise)."""
    rng = np.random.default_rng(7)
    N, nang = 12, 9
    th = (np.arange(nang) - nang / 2) * np.pi / nang
    P = orc64.psf2d(th, N, N)
    y = P.convolve(rng.random((3, N, N)))
    lam, mu, outer, inner = 0.05, 2.0, 4, 3
    xs, _ = orc64.split_bregman(P, y, np.zeros_like(y), lam, mu, outer, inner, 0.0, 0.0,
                                boundary=boundary, aniso=aniso)
    op = lambda v: P.convolve(v) - orc64.laplacian(v, boundary) * mu
    d = np.zeros((3,) + y.shape)
    b = np.zeros_like(d)
    x = np.zeros_like(y)
    for _ in range(outer):
        rhs = orc64.sb_rhs(y, d, b, mu, boundary)
        x, _ = orc64.cgsolver(op, rhs, x, inner, 0.0, 0.0)
        g = orc64.grad(x, boundary)
        d = orc64.sb_shrink(g, b, lam, mu, aniso=aniso)
        b = orc64.sb_bregman_update(b, g, d)
    assert np.array_equal(x, xs)
    assert np.any(d)   # the TV step did work


# ---------------------------------------------------------------- ROF minimiser
ROF_SHAPE = (4, 10, 10)
ROF_LAM, ROF_MU = 0.2, 3.0


def _rof_data():
    rng = np.random.default_rng(3)
    y = np.zeros(ROF_SHAPE)
    y[:, 2:7, 3:8] = 1.0
    y[1:3, 5:9, 1:5] += 0.6
    return y + 0.1 * rng.standard_normal(ROF_SHAPE)


def _rof_energy(orc, x, y, lam):
    g = orc.grad(x, 1)
    return 0.5 * ((x - y) ** 2).sum() + lam * np.sqrt((g ** 2).sum(0)).sum()


@pytest.fixture(scope="module")
def rof_minimiser(orc64):
    """Chambolle-Pock primal-dual iteration for min 0.5|x-y|^2 + lam |G x|_{2,1}
    with G the zero-padded central differences (|G|^2 <= 3, so tau = sigma =
    0.5 satisfies tau sigma |G|^2 < 1)."""
    y = _rof_data()
    G = lambda u: orc64.grad(u, 1)
    Gt = lambda p: -orc64.div(p, 1)
    x = y.copy()
    xb = x.copy()
    p = np.zeros((3,) + y.shape)
    tau = sig = 0.5
    for _ in range(20000):
        p = p + sig * G(xb)
        p = p / np.maximum(1.0, np.sqrt((p ** 2).sum(0)) / ROF_LAM)
        xn = (x - tau * Gt(p) + tau * y) / (1 + tau)
        xb = 2 * xn - x
        x = xn
    return y, x


def _sb_composed(orc, y, lam, mu, outer, inner, plant=None):
    """Split-Bregman with A = I from the oracle's steps, with one planted mistake"""
    d = np.zeros((3,) + y.shape)
    b = np.zeros_like(d)
    x = np.zeros_like(y)
    op = lambda v: v - orc.laplacian(v, 1) * mu
    for _ in range(outer):
        if plant == "rhs_b_sign":      # y - mu div(d + b)
            rhs = orc.sb_rhs(y, d + b, 0 * b, mu, 1)
        else:
            rhs = orc.sb_rhs(y, d, b, mu, 1)
        x, _ = orc.cgsolver(op, rhs, x, inner, 0.0, 0.0)
        g = orc.grad(x, 1)
        if plant == "threshold_lambda_mu":   # lambda * mu in place of lambda / mu
            d = orc.sb_shrink(g, b, lam * mu * mu, mu)
        else:
            d = orc.sb_shrink(g, b, lam, mu)
        if plant == "b_sign":
            b = b - (g - d)
        else:
            b = orc.sb_bregman_update(b, g, d)
    return x


def test_sb_reaches_rof_minimiser(orc64, rof_minimiser):
    y, xp = rof_minimiser
    d = np.zeros((3,) + y.shape)
    b = np.zeros_like(d)
    x, rep = orc64.split_bregman(lambda v: v, y, np.zeros_like(y), ROF_LAM, ROF_MU, 600, 10,
                                 0.0, 0.0, boundary=1, state=(d, b))
    assert rep["outer"] == 600
    assert rel_l2(x, xp) < 2e-4
    e_sb, e_pd = _rof_energy(orc64, x, y, ROF_LAM), _rof_energy(orc64, xp, y, ROF_LAM)
    assert abs(e_sb - e_pd) < 1e-5 * e_pd
    assert e_sb < _rof_energy(orc64, y, y, ROF_LAM)
    # fixed point: the constraint d = G x holds and mu b is a subgradient of lam |d|
    g = orc64.grad(x, 1)
    assert rel_l2(d, g) < 1e-4
    nb = np.sqrt(((ROF_MU * b) ** 2).sum(0))
    nd = np.sqrt((d ** 2).sum(0))
    act = nd > 1e-6
    assert nb.max() <= ROF_LAM * (1 + 1e-6)
    assert 0.3 < act.mean() < 0.99
    assert np.abs(ROF_MU * b[:, act] - ROF_LAM * d[:, act] / nd[act]).max() < 1e-6
    # the composition of the steps is the solver
    assert rel_l2(_sb_composed(orc64, y, ROF_LAM, ROF_MU, 600, 10), x) < 1e-12


@pytest.mark.parametrize("plant", ["rhs_b_sign", "threshold_lambda_mu", "b_sign"])
def test_sb_planted_mistakes_miss_rof_minimiser(orc64, rof_minimiser, plant):
    y, xp = rof_minimiser
    with np.errstate(all="ignore"):
        xq = _sb_composed(orc64, y, ROF_LAM, ROF_MU, 600, 10, plant)
        err = rel_l2(xq, xp) if np.all(np.isfinite(xq)) else np.inf
    assert err > 100 * 2e-4


@pytest.mark.parametrize("aniso", [False, True])
def test_sb_fixed_point_interior_boundary(orc64, aniso):
    """REF_CPU differences (interior only): -div is not grad's adjoint there, so
    no ROF energy is minimised, but a converged run is a Split-Bregman fixed
    point: d = grad x, and mu b is a subgradient of lam |d|."""
    y = _rof_data()
    d = np.zeros((3,) + y.shape)
    b = np.zeros_like(d)
    x, _ = orc64.split_bregman(lambda v: v, y, np.zeros_like(y), ROF_LAM, ROF_MU, 600, 10,
                               0.0, 0.0, boundary=0, aniso=aniso, state=(d, b))
    g = orc64.grad(x, 0)
    assert rel_l2(d, g) < 1e-5
    mb = ROF_MU * b
    if aniso:
        act = np.abs(d) > 1e-6
        assert np.abs(mb).max() <= ROF_LAM * (1 + 1e-6)
        assert np.abs(mb[act] - ROF_LAM * np.sign(d[act])).max() < 1e-6
    else:
        nd = np.sqrt((d ** 2).sum(0))
        act = nd > 1e-6
        assert np.sqrt((mb ** 2).sum(0)).max() <= ROF_LAM * (1 + 1e-6)
        assert np.abs(mb[:, act] - ROF_LAM * d[:, act] / nd[act]).max() < 1e-6
    assert 0.1 < act.mean() < 0.9


# ---------------------------------------------------------------- end to end
def two_disk_phantom(nslices, n):
    """tests/test_mbir2d_sanity.cpp:68-96: a disk of 1 (radius 0.35 n) with a
    disk of 2 (radius 0.15 n) whose centre drifts with the slice index."""
    ph = np.zeros((nslices, n, n), np.float32)
    c = n / 2.0
    j, k = np.meshgrid(np.arange(n, dtype=np.float32), np.arange(n, dtype=np.float32),
                       indexing="ij")
    for i in range(nslices):
        drift = (i / (nslices - 1) - 0.5) * n * 0.2 if nslices > 1 else 0.0
        v = np.zeros((n, n), np.float32)
        v[(j - c) ** 2 + (k - c) ** 2 < (n * 0.35) ** 2] = 1.0
        v[(j - (c - drift)) ** 2 + (k - (c + drift)) ** 2 < (n * 0.15) ** 2] = 2.0
        ph[i] = v
    return ph


def sanity_scores(recon, phantom):
    """least-squares scale fit, then relative L2 error and cosine similarity
    (tests/test_mbir2d_sanity.cpp:139-148)"""
    r = recon.astype(np.float64)
    p = phantom.astype(np.float64)
    num, den = np.vdot(r, p), np.vdot(r, r)
    scale = num / den if den > 0 else 1.0
    rel = np.linalg.norm(p - scale * r) / np.linalg.norm(p)
    cos = num / (np.linalg.norm(p) * np.linalg.norm(r) + 1e-8)
    return rel, cos


def test_mbir2d_sanity_split_bregman(orc64):
    """tests/test_mbir2d_sanity.cpp:98-205: 5 x 128^2 two-disk phantom, 180
    angles over [0, pi), forward2d, backproj2d, Split-Bregman (lambda 0.1, mu 5,
    50 outer x 25 inner, tol = xtol = 1e-5) with the 2-D Toeplitz operator;
    passes when finite, cos > 0.85 and rel. err < 0.5 after the scale fit."""
    N, nz, nt = 128, 5, 180
    ph = two_disk_phantom(nz, N)
    th = np.arange(nt) * np.pi / nt
    sino = orc64.forward2d(ph, th, N)
    rhs = orc64.backproj2d(sino, th, N)
    P = orc64.psf2d(th, N, N)
    x, rep = orc64.split_bregman(P, rhs, np.zeros_like(rhs), 0.1, 5.0, 50, 25, 1e-5, 1e-5,
                                 boundary=0)
    rel, cos = sanity_scores(x, ph)
    assert np.all(np.isfinite(x))
    assert cos > 0.85 and rel < 0.5, (rel, cos, rep)
