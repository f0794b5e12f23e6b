"""Seeded synthetic inputs: a project-owned SplitMix64 stream, Box-Muller
normals, random-ellipse (2-D) and ellipsoid (3-D) phantoms.

The generator is counter based (value i of stream s is a pure function of
(seed, s, i)), so the same data come out of any process or machine.
Base seed 0x1A41C0DE; sub-stream = config id * 2**16 + request index.
"""
from __future__ import annotations

import numpy as np

BASE_SEED = 0x1A41C0DE
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n 64-bit outputs of SplitMix64 started at state seed ^ (stream * golden)."""
    with np.errstate(over="ignore"):
        state0 = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) ^ (np.uint64(stream & 0xFFFFFFFFFFFFFFFF) * _GOLDEN)
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = state0 + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


class Stream:
    """Sequential draws from one SplitMix64 sub-stream."""

    def __init__(self, stream: int, seed: int = BASE_SEED):
        self.seed = seed
        self.stream = stream
        self.pos = 0

    def bits(self, n):
        out = splitmix64(self.seed, self.stream, n, self.pos)
        self.pos += n
        return out

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self.bits(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n):
        m = (n + 1) // 2
        u1 = 1.0 - self.uniform(m)          # (0, 1]
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]

    def choice(self, values, n):
        idx = (self.bits(n) % np.uint64(len(values))).astype(np.int64)
        return np.asarray(values)[idx]


def ellipse_phantom(N: int, n_ellipses: int, rng: Stream, piecewise: bool = False) -> np.ndarray:
    """N x N float32 sum of ellipses on [-1, 1]^2 (pixel centres).
    centres U[-0.5, 0.5], semi-axes U[0.05, 0.4], rotation U[0, pi),
    intensity U[0, 1]; piecewise=True paints instead of adding."""
    c = (np.arange(N) + 0.5) * (2.0 / N) - 1.0
    img = np.zeros((N, N), np.float64)
    for _ in range(n_ellipses):
        cx, cy = rng.uniform(2, -0.5, 0.5)
        a, b = rng.uniform(2, 0.05, 0.4)
        phi = rng.uniform(1, 0.0, np.pi)[0]
        val = rng.uniform(1)[0]
        # only the bounding square of the circumscribed circle can be inside
        rmax = max(a, b)
        ix = np.nonzero(np.abs(c - cx) <= rmax)[0]
        iy = np.nonzero(np.abs(c - cy) <= rmax)[0]
        if ix.size == 0 or iy.size == 0:
            continue
        yy, xx = np.meshgrid(c[iy], c[ix], indexing="ij")
        xr = (xx - cx) * np.cos(phi) + (yy - cy) * np.sin(phi)
        yr = -(xx - cx) * np.sin(phi) + (yy - cy) * np.cos(phi)
        inside = (xr / a) ** 2 + (yr / b) ** 2 <= 1.0
        sub = img[iy[0]:iy[-1] + 1, ix[0]:ix[-1] + 1]
        if piecewise:
            sub[inside] = val
        else:
            sub[inside] += val
    return img.astype(np.float32)


def ellipsoid_phantom(dims, n_ellipsoids: int, rng: Stream, piecewise: bool = False) -> np.ndarray:
    """n1 x n2 x n3 float32 sum of ellipsoids on [-1, 1]^3 (same draws per axis)."""
    n1, n2, n3 = dims
    z = (np.arange(n1) + 0.5) * (2.0 / n1) - 1.0
    y = (np.arange(n2) + 0.5) * (2.0 / n2) - 1.0
    x = (np.arange(n3) + 0.5) * (2.0 / n3) - 1.0
    vol = np.zeros(dims, np.float64)
    for _ in range(n_ellipsoids):
        cz, cy, cx = rng.uniform(3, -0.5, 0.5)
        az, ay, ax = rng.uniform(3, 0.05, 0.4)
        phi = rng.uniform(1, 0.0, np.pi)[0]
        val = rng.uniform(1)[0]
        # only the bounding box of the circumscribed cylinder can be inside
        rxy = max(ax, ay)
        iz = np.nonzero(np.abs(z - cz) <= az)[0]
        iy = np.nonzero(np.abs(y - cy) <= rxy)[0]
        ix = np.nonzero(np.abs(x - cx) <= rxy)[0]
        if iz.size == 0 or iy.size == 0 or ix.size == 0:
            continue
        zz, yy, xx = np.meshgrid(z[iz], y[iy], x[ix], indexing="ij")
        xr = (xx - cx) * np.cos(phi) + (yy - cy) * np.sin(phi)
        yr = -(xx - cx) * np.sin(phi) + (yy - cy) * np.cos(phi)
        inside = (xr / ax) ** 2 + (yr / ay) ** 2 + ((zz - cz) / az) ** 2 <= 1.0
        sub = vol[iz[0]:iz[-1] + 1, iy[0]:iy[-1] + 1, ix[0]:ix[-1] + 1]
        if piecewise:
            sub[inside] = val
        else:
            sub[inside] += val
    return vol.astype(np.float32)


def add_gaussian_noise(data: np.ndarray, rel_sigma: float, rng: Stream) -> np.ndarray:
    """data + N(0, (rel_sigma * max|data|)^2)"""
    s = rel_sigma * float(np.max(np.abs(data)))
    return (data + s * rng.normal(data.size).reshape(data.shape)).astype(np.float32)


def angles(n: int, span: float = np.pi) -> np.ndarray:
    """n angles uniform in [0, span)"""
    return np.arange(n, dtype=np.float64) * (span / n)
