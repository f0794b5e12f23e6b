"""Pins of the oracle's driver restatement (oracle/drivers.py) and the host
side of the driver boundary (ReconParams layout and defaults, padded sizes).
CPU only."""
import ctypes as C

import numpy as np
import pytest

from oracle import drivers as od


def iota(shape, start=1.0):
    return (start + np.arange(int(np.prod(shape)), dtype=np.float32)).reshape(shape)


# ---- pad_size ----
def test_pad_size_known_values():
    # the 3-D laminography sizes: volume 21 x 511 x 511 -> 29 x 721 x 721,
    # detector 64 x 511 -> 89 x 721 (511 * 1.4142 = 722.66 -> 722, even -> 721)
    assert [od.pad_size(n, 1.4142) for n in (21, 511, 64)] == [29, 721, 89]
    assert od.pad_size(510, 1.4142) == 721
    # even products step down to the odd size below; odd ones stay
    assert od.pad_size(3, 2.0) == 5 and od.pad_size(5, 2.0) == 9
    assert od.pad_size(7, 1.0) == 7 and od.pad_size(8, 1.0) == 7
    # float vs double product: both truncate 100 * 1.4142 the same way here
    assert od.pad_size(100, 1.4142, np.float64) == od.pad_size(100, 1.4142, np.float32) == 141


# ---- pad2d / crop3d ----
def test_pad2d_symmetric_placement_by_hand():
    # {2, 3, 5} at factor 2: n2 = 6 -> 5, n3 = 10 -> 9; offsets (1, 2)
    a = iota((2, 3, 5))
    p = od.pad2d(a, 2.0)
    assert p.shape == (2, 5, 9)
    assert p[0, 1, 2] == 1.0 and p[0, 1, 6] == 5.0 and p[0, 3, 2] == 11.0
    assert p[1, 1, 2] == 16.0 and p[1, 3, 6] == 30.0
    mask = np.zeros(p.shape, bool)
    mask[:, 1:4, 2:7] = True
    assert np.all(p[~mask] == 0) and np.array_equal(p[mask].reshape(a.shape), a)


def test_pad2d_factor_one_is_a_copy():
    a = iota((2, 4, 6))
    assert np.array_equal(od.pad2d(a, 1.0), a)
    assert np.array_equal(od.pad2d(a, 1.0000001), a)


def test_crop3d_symmetric_by_hand():
    # {3, 7, 11} -> {2, 3, 5}: crop (1, 4, 6), offsets (0, 2, 3)
    a = iota((3, 7, 11))
    c = od.crop3d(a, (2, 3, 5))
    assert c.shape == (2, 3, 5)
    assert c[0, 0, 0] == a[0, 2, 3] and c[1, 2, 4] == a[1, 4, 7]


@pytest.mark.parametrize("shape,factor", [((2, 3, 5), 2.0), ((3, 8, 9), 1.4142), ((1, 64, 17), 1.4142)])
def test_pad2d_then_crop3d_recovers(shape, factor):
    # tests/test_padding.cu:118-131: pad2d(SYMMETRIC) -> crop3d(SYMMETRIC) is the identity
    a = iota(shape)
    assert np.array_equal(od.crop3d(od.pad2d(a, factor), shape), a)


# ---- prepare_sinogram ----
def test_prepare_sinogram_index_map_by_loops():
    rng = np.random.default_rng(1)
    projs = [rng.random((3, 4, 5), dtype=np.float32), rng.random((2, 4, 5), dtype=np.float32)]
    angles = [np.arange(3) * 0.1, np.arange(2) * 0.2 + 1]
    sino, theta = od.prepare_sinogram(projs, angles)
    m = max(float(p.max()) for p in projs)
    stacked = [p[a] for p in projs for a in range(p.shape[0])]
    assert sino.shape == (4, 5, 5)
    for r in range(4):
        for a in range(5):
            for c in range(5):
                assert sino[r, a, c] == np.float32(stacked[a][r, c]) / np.float32(m)
    assert np.allclose(theta, [0, 0.1, 0.2, 1.0, 1.2])


def test_prepare_sinogram_nonpositive_max_is_not_divided():
    projs = [-np.ones((2, 3, 4), np.float32), np.zeros((1, 3, 4), np.float32)]
    sino, _ = od.prepare_sinogram(projs, [np.zeros(2), np.zeros(1)])
    assert sino.shape == (3, 3, 4)
    assert np.array_equal(sino[:, :2], -np.ones((3, 2, 4))) and not sino[:, 2].any()


# ---- host side of the boundary ----
def test_recon_params_layout_and_defaults():
    import laminocam_amd as lc
    from laminocam_amd import _lib
    assert C.sizeof(_lib.LcReconParams) == 72
    L = _lib.load()
    p = _lib.LcReconParams()
    L.lc_recon_params_default(C.byref(p))     # host-only call: no GPU needed
    q = lc.ReconParams().to_c()
    for f, _ in _lib.LcReconParams._fields_:
        a, b = getattr(p, f), getattr(q, f)
        if f == "recon_dims":
            a, b = list(a), list(b)
        assert a == b, f
    # include/recon_params.h:45-54
    assert (p.regularizer, p.max_iters, p.inner_iters) == (_lib.LC_REG_UNCONSTRAINED, 100, 1)
    assert np.float32(p.lam) == np.float32(0.1) and p.mu == 10.0
    assert np.float32(p.pad_factor) == np.float32(1.4142)


def test_padded_dims_host_arithmetic_matches_oracle():
    import laminocam_amd as lc
    for dims, nr, nc, f in [((21, 511, 511), 64, 511, 1.4142), ((3, 12, 12), 6, 16, 1.4142),
                            ((5, 9, 9), 7, 9, 1.0), ((4, 10, 10), 5, 10, 2.0)]:
        got = lc.drivers.padded_dims(lc.ReconParams(recon_dims=dims, PAD_FACTOR=f), nr, nc)
        vol = tuple(od.pad_size(n, f) for n in dims)
        det = od.pad2d(np.zeros((1, nr, nc), np.float32), f).shape[1:]
        assert got == vol + det
