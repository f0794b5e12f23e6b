"""Whole-solve drivers over the C ABI (include/laminocam.h, lc_mbir2d / lc_mbir3d /
lc_prepare_sinogram), mirroring the reference's callers of the hot path:

  Dataset           Dataset_t           include/recon_params.h:32-39
  Regularizer       enum Regularizer    include/recon_params.h:41
  ReconParams       ReconParams         include/recon_params.h:43-54
  prepare_sinogram  prepare_sinogram    include/mbir2d_prep.h:44-65
  MBIR2D            cpu_mbir2d / gpu::MBIR2D   src/mbir2d.cpp:66-100, src/gpu/mbir2d.cu:36-85
  MBIR              MBIR                src/mbir.cpp:34-161
  broadcast_shifts, apply_alignment     include/config.h:284-343 (per-dataset shifts
                                        and the alignment-file override)

Normalisation, transposes, padding, back-projection, the solve and the crop
all run on the GPU; the host only stacks pointers and angles.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field
from typing import NamedTuple, Sequence

import numpy as np

from . import _lib, _ptr, default_context
from ._lib import (LC_GRAPH, LC_REG_SPLIT_BREGMAN, LC_REG_UNCONSTRAINED, LC_TV_ANISO,
                   LC_TV_ZERO_PAD, LcReconParams, LcReport)

__all__ = ["Dataset", "Regularizer", "ReconParams", "Sinogram2D", "prepare_sinogram", "MBIR2D",
           "MBIR", "padded_dims", "estimate_memory", "volume_fits_one_gpu", "broadcast_shifts",
           "apply_alignment", "load_alignment"]


class Regularizer(enum.IntEnum):
    """enum class Regularizer { SPLIT_BREGMAN, UNCONSTRAINED }"""
    SPLIT_BREGMAN = LC_REG_SPLIT_BREGMAN
    UNCONSTRAINED = LC_REG_UNCONSTRAINED


@dataclass
class ReconParams:
    """ReconParams with the reference's field names and defaults (`lambda` is
    `lam` here).  Floating-point fields are held as float32 at the boundary,
    as the reference's struct stores them."""
    regularizer: Regularizer = Regularizer.UNCONSTRAINED
    recon_dims: tuple = (0, 0, 0)
    maxIters: int = 100
    innerIters: int = 1
    lam: float = 0.1
    mu: float = 10.0
    tol: float = 1e-5
    xtol: float = 1e-5
    PAD_FACTOR: float = 1.4142
    # the reference's multi-device streaming fields (include/recon_params.h:56-60):
    # read from configuration files; this library runs one device
    numGpus: int = 0
    chunkSlices: int = 0
    streamed: bool = False

    def print(self, file=None) -> None:
        """ReconParams::print (include/recon_params.h:62-81)"""
        import sys
        f = file or sys.stdout
        sb = self.regularizer == Regularizer.SPLIT_BREGMAN
        g = lambda x: f"{float(np.float32(x)):g}"
        lines = ["Reconstruction Parameters:", f"  max_outer_iters: {self.maxIters}",
                 "  recon_dims: [{}, {}, {}]".format(*self.recon_dims), f"  tol: {g(self.tol)}",
                 f"  xtol: {g(self.xtol)}",
                 f"  regularizer: {'Split-Bregman' if sb else 'Unconstrained'}"]
        if sb:
            lines += [f"    inner_iters: {self.innerIters}", f"    lambda: {g(self.lam)}",
                      f"    mu: {g(self.mu)}"]
        if self.numGpus > 0:
            lines.append(f"  num_gpus: {self.numGpus}")
        if self.chunkSlices > 0:
            lines.append(f"  chunk_slices: {self.chunkSlices}")
        if self.streamed:
            lines.append("  streamed: true")
        print("\n".join(lines), file=f)

    def to_c(self) -> LcReconParams:
        p = LcReconParams()
        p.regularizer = int(self.regularizer)
        for i in range(3):
            p.recon_dims[i] = int(self.recon_dims[i])
        p.max_iters = int(self.maxIters)
        p.inner_iters = int(self.innerIters)
        p.lam, p.mu, p.tol, p.xtol = float(self.lam), float(self.mu), float(self.tol), float(self.xtol)
        p.pad_factor = float(self.PAD_FACTOR)
        return p


@dataclass
class Dataset:
    """Dataset_t: projections {nang, nrows, ncols}, one angle per projection
    (radians), the dataset's tilt gamma and beta, and per-projection shifts
    {nang, 2} as [dx, dy] in pixels (None: no shifts)."""
    projs: np.ndarray
    angles: np.ndarray
    gamma: float = 0.0
    beta: float = 0.0
    shifts: "np.ndarray | None" = field(default=None)


class Sinogram2D(NamedTuple):
    """Sinogram2D: data {n_slices, n_angles, n_cols}, theta one angle per projection"""
    data: np.ndarray
    theta: np.ndarray


def _dtype_of(datasets):
    return np.float64 if all(np.asarray(d.projs).dtype == np.float64 for d in datasets) \
        else np.float32


def _stack(datasets: Sequence[Dataset]):
    if len(datasets) == 0:
        raise ValueError("no datasets")
    dt = _dtype_of(datasets)
    projs = [np.ascontiguousarray(d.projs, dtype=dt) for d in datasets]
    shape = projs[0].shape
    for p, d in zip(projs, datasets):
        if p.ndim != 3 or p.shape[1:] != shape[1:]:
            raise ValueError("every dataset needs projections {nang, nrows, ncols} of one detector size")
        if len(np.ravel(d.angles)) != p.shape[0]:
            raise ValueError("one angle per projection")
    nang = (C.c_int64 * len(projs))(*[p.shape[0] for p in projs])
    ptrs = (C.c_void_p * len(projs))(*[p.ctypes.data for p in projs])
    theta = np.ascontiguousarray(np.concatenate([np.ravel(d.angles) for d in datasets]),
                                 dtype=np.float64)
    return dt, projs, nang, ptrs, theta, shape[1], shape[2]


def prepare_sinogram(datasets: Sequence[Dataset], ctx=None) -> Sinogram2D:
    """Stack the datasets, divide by the global maximum (when > 0) and reorder
    {n_angles, n_rows, n_cols} -> {n_rows, n_angles, n_cols}, on the GPU."""
    ctx = ctx or default_context()
    dt, projs, nang, ptrs, theta, nrows, ncols = _stack(datasets)
    out = np.empty((nrows, len(theta), ncols), dt)
    fn = ctx.lib.lc_prepare_sinogram_f64 if dt == np.float64 else ctx.lib.lc_prepare_sinogram
    _lib.check(fn(ctx.handle, len(projs), ptrs, nang, nrows, ncols, _ptr(out), None, 0))
    return Sinogram2D(out, theta.astype(dt))


def _flags(graph, tv_zero_pad, tv_aniso):
    return (LC_GRAPH if graph else 0) | (LC_TV_ZERO_PAD if tv_zero_pad else 0) | \
        (LC_TV_ANISO if tv_aniso else 0)


def MBIR2D(sinogram, theta, params: ReconParams, ctx=None, *, graph=False, tv_zero_pad=False,
           tv_aniso=False, report: "dict | None" = None):
    """The 2-D slice-stack reconstruction of a prepared sinogram {n_slices,
    n_angles, n_cols}: grid, PSF, rhs = backproj2d, CG or Split-Bregman from
    zeros over recon_dims.  float64 input solves in FP64.  tv_zero_pad selects
    the zero-padded differences of the reference's GPU path (default: its CPU
    path's interior-only differences)."""
    ctx = ctx or default_context()
    s = np.asarray(sinogram)
    dt = np.float64 if s.dtype == np.float64 else np.float32
    s = np.ascontiguousarray(s, dtype=dt)
    if s.ndim != 3:
        raise ValueError("sinogram must be {n_slices, n_angles, n_cols}")
    th = np.ascontiguousarray(theta, dtype=np.float64)
    if th.size != s.shape[1]:
        raise ValueError("one angle per sinogram row")
    out = np.empty(tuple(int(d) for d in params.recon_dims), dt)
    rep = LcReport()
    fn = ctx.lib.lc_mbir2d_f64 if dt == np.float64 else ctx.lib.lc_mbir2d
    _lib.check(fn(ctx.handle, _ptr(s), s.shape[0], s.shape[1], s.shape[2],
                  th.ctypes.data_as(C.POINTER(C.c_double)), C.byref(params.to_c()), _ptr(out),
                  C.byref(rep), _flags(graph, tv_zero_pad, tv_aniso)))
    if report is not None:
        report.update(rep.as_dict())
    return out


def MBIR(datasets: Sequence[Dataset], params: ReconParams, ctx=None, *, graph=False,
         tv_zero_pad=False, tv_aniso=False, report: "dict | None" = None):
    """The 3-D laminography reconstruction of src/mbir.cpp:34-161: stack,
    normalise, pad the detector and the volume by PAD_FACTOR to odd sizes,
    back-project with the shifts, solve from zeros, crop to recon_dims.
    float64 projections solve in FP64.  report (if given) also receives the
    padded sizes under "padded_dims"."""
    ctx = ctx or default_context()
    dt, projs, nang, ptrs, theta, nrows, ncols = _stack(datasets)
    nds = len(projs)
    gamma = np.array([float(d.gamma) for d in datasets], np.float64)
    beta = np.array([float(d.beta) for d in datasets], np.float64)
    shifts = None
    if any(d.shifts is not None for d in datasets):
        parts = []
        for p, d in zip(projs, datasets):
            sh = np.zeros((p.shape[0], 2)) if d.shifts is None else np.asarray(d.shifts, np.float64)
            if sh.shape != (p.shape[0], 2):
                # src/projection.cpp:105-108 throws std::invalid_argument
                raise ValueError("shifts size must match the number of slices in the projection")
            parts.append(sh)
        shifts = np.ascontiguousarray(np.concatenate(parts), dtype=np.float64)
    out = np.empty(tuple(int(d) for d in params.recon_dims), dt)
    pd = (C.c_int64 * 5)()
    rep = LcReport()
    dp = C.POINTER(C.c_double)
    fn = ctx.lib.lc_mbir3d_f64 if dt == np.float64 else ctx.lib.lc_mbir3d
    _lib.check(fn(ctx.handle, nds, ptrs, nang, nrows, ncols, theta.ctypes.data_as(dp),
                  gamma.ctypes.data_as(dp), beta.ctypes.data_as(dp),
                  None if shifts is None else shifts.ctypes.data_as(dp), C.byref(params.to_c()),
                  _ptr(out), pd, C.byref(rep), _flags(graph, tv_zero_pad, tv_aniso)))
    if report is not None:
        report.update(rep.as_dict())
        report["padded_dims"] = tuple(pd)
    return out


def padded_dims(params: ReconParams, nrows: int, ncols: int, dtype=np.float32):
    """Host arithmetic of the padded sizes the 3-D driver uses: the volume's
    {n1, n2, n3} (size_t(n * PAD_FACTOR) in the element type, made odd by -1
    when even; src/mbir.cpp:41-45) and the detector's {nrows, ncols} (the
    same in float; include/padding.h:19-22, no padding when PAD_FACTOR - 1 <
    1e-6)."""
    def pad(n, f, t):
        m = int(t(n) * t(f))
        return m - 1 if m % 2 == 0 else m
    f32 = np.float32(params.PAD_FACTOR)
    t = np.float64 if dtype == np.float64 else np.float32
    vol = tuple(pad(int(n), f32, t) for n in params.recon_dims)
    if float(f32 - np.float32(1)) < 1e-6:
        det = (int(nrows), int(ncols))
    else:
        det = (pad(int(nrows), f32, np.float32), pad(int(ncols), f32, np.float32))
    return vol + det


def estimate_memory(shape, dtype=np.float32, regularizer=Regularizer.SPLIT_BREGMAN):
    """The reference's analytic device-memory model of its single-device
    solvers (scripts/estimate_bregman_memory.py:18-62 and the fit rule of
    src/gpu/mbir2d_multi.cu:459-470) for a volume `shape`: bytes of the typical
    and conservative high-water marks (Split-Bregman 22 N / 24 N; CG 11 N),
    of the inputs (2 N) and of the fit rule's working set (27 N / 11 N)."""
    L = _lib.load()
    d = (C.c_int64 * 3)(*[int(v) for v in shape])
    out = (C.c_double * 5)()
    _lib.check(L.lc_reference_memory_model(d, int(np.dtype(dtype).itemsize), int(regularizer), out))
    return {"typical_bytes": out[0], "conservative_bytes": out[1], "input_bytes": out[2],
            "fit_rule_bytes": out[3], "N": int(out[4])}


def volume_fits_one_gpu(params: ReconParams, dtype=np.float32, free_bytes=None, ctx=None) -> bool:
    """gpu::volume_fits_one_gpu (include/gpu/tomocam.h:91-93): the fit rule's
    working set (Split-Bregman 27 N, CG 11 N elements) below 0.8 of the free
    device memory; free_bytes given: against that figure instead (no device
    call)."""
    L = _lib.load()
    p = params.to_c()
    fits = C.c_int32()
    eb = int(np.dtype(dtype).itemsize)
    if free_bytes is not None:
        _lib.check(L.lc_volume_fits(C.byref(p), eb, int(free_bytes), C.byref(fits)))
    else:
        ctx = ctx or default_context()
        _lib.check(L.lc_volume_fits_one_gpu(ctx.handle, C.byref(p), eb, C.byref(fits)))
    return bool(fits.value)


def broadcast_shifts(shifts, nproj: int):
    """Per-projection [dx, dy] shifts of one [[input]] dataset
    (include/config.h:284-296): none gives zeros, a single pair (the legacy
    'cor-offset') is repeated for every projection, otherwise there must be one
    pair per projection (ValueError)."""
    s = np.zeros((0, 2)) if shifts is None else np.asarray(shifts, dtype=np.float64).reshape(-1, 2)
    if len(s) == 0:
        return np.zeros((nproj, 2))
    if len(s) == 1:
        return np.repeat(s, nproj, axis=0)
    if len(s) != nproj:
        raise ValueError(f"shifts file has {len(s)} entries but {nproj} projections were loaded")
    return s.copy()


def apply_alignment(dataset: Dataset, table: dict) -> Dataset:
    """The alignment override of an [[input]] dataset (include/config.h:298-343):
    `table` is the parsed alignment file.  'gamma_deg' and 'beta_deg' (optional)
    replace the tilts, gamma with the parser's sign flip (gamma = -gamma_deg,
    radians); 'corrected_angles_deg' (required) replaces the angles; 'shifts_px'
    (optional, [dx, dy] pairs) replaces the shifts.  The shifts must then
    number one per corrected angle.  The dataset's shifts are first completed
    as broadcast_shifts does.  Returns a new Dataset; projections are shared."""
    import math
    gamma, beta = float(dataset.gamma), float(dataset.beta)
    if "gamma_deg" in table:
        gamma = -float(table["gamma_deg"]) * math.pi / 180.0
    if "beta_deg" in table:
        beta = float(table["beta_deg"]) * math.pi / 180.0
    if "corrected_angles_deg" not in table:
        raise ValueError("Alignment TOML missing 'corrected_angles_deg'")
    angles = np.asarray([float(a) * math.pi / 180.0 for a in table["corrected_angles_deg"]])
    shifts = broadcast_shifts(dataset.shifts, len(np.asarray(dataset.angles)))
    if "shifts_px" in table:
        pairs = []
        for e in table["shifts_px"]:
            if not hasattr(e, "__len__") or len(e) < 2:
                raise ValueError("Alignment TOML 'shifts_px' entries must be [dx, dy] pairs")
            pairs.append((float(e[0]), float(e[1])))
        shifts = np.asarray(pairs, dtype=np.float64).reshape(-1, 2)
    if len(shifts) != len(angles):
        raise ValueError(f"Alignment TOML: shifts_px has {len(shifts)} entries but "
                         f"corrected_angles_deg has {len(angles)}")
    return Dataset(projs=dataset.projs, angles=angles, gamma=gamma, beta=beta, shifts=shifts)


def load_alignment(path: str) -> dict:
    """Parse an alignment TOML file (the table apply_alignment takes)."""
    import tomli
    with open(path, "rb") as f:
        return tomli.load(f)
