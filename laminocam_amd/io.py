"""File formats of the reconstruction callers, over the library's own codecs
(lc_tiff_* / lc_h5_* of include/laminocam.h; host only, no device needed):

  tiff_read, tiff_write             include/tiff.h:31-117 (tomocam::tiff)
  h5_read_images, h5_read_angles    include/hdfread.h:34-96 (tomocam::h5)
  h5_read, h5_shape                 any numeric HDF5 dataset
  mask_infs_nans, mask_support      include/mask.h:43-112

TIFF stacks hold single 32-bit float samples, one page per slice.  HDF5
datasets are read from integer or IEEE float storage and converted.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from . import _lib

__all__ = ["tiff_info", "tiff_read", "tiff_write", "h5_shape", "h5_read", "h5_read_images",
           "h5_read_angles", "mask_infs_nans", "mask_support"]


def _path(p) -> bytes:
    return os.fsencode(os.fspath(p))


def tiff_info(filename):
    """(pages, height, width) of a TIFF stack."""
    L = _lib.load()
    n, h, w = C.c_int64(), C.c_int64(), C.c_int64()
    _lib.check(L.lc_tiff_info(_path(filename), C.byref(n), C.byref(h), C.byref(w)))
    return n.value, h.value, w.value


def tiff_read(filename, odd=True) -> np.ndarray:
    """tomocam::tiff::read: every page as a float32 {pages, rows, cols} array.
    As the reference's reader does, an even height or width is made odd by
    dropping the last row or column (odd=False keeps the full pages)."""
    L = _lib.load()
    n, h, w = tiff_info(filename)
    if odd:
        h -= 1 - h % 2
        w -= 1 - w % 2
    out = np.empty((n, h, w), np.float32)
    _lib.check(L.lc_tiff_read(_path(filename), out.ctypes.data, n, h, w))
    return out


def tiff_write(filename, data) -> None:
    """tomocam::tiff::write: a {pages, rows, cols} array (a 2-D array is one
    page) as float32 pages of one uncompressed strip each."""
    L = _lib.load()
    a = np.ascontiguousarray(data, dtype=np.float32)
    if a.ndim == 2:
        a = a[None]
    if a.ndim != 3:
        raise ValueError("tiff_write: data must be {pages, rows, cols}")
    _lib.check(L.lc_tiff_write(_path(filename), a.ctypes.data, *a.shape))


def h5_shape(filename, dataset) -> tuple:
    L = _lib.load()
    r = C.c_int32()
    dims = (C.c_int64 * 32)()
    _lib.check(L.lc_h5_shape(_path(filename), dataset.encode(), C.byref(r), dims))
    return tuple(dims[i] for i in range(r.value))


def h5_read(filename, dataset, dtype=np.float32) -> np.ndarray:
    """The dataset at absolute path `dataset`, converted to float32 (the
    reference's H5T_NATIVE_FLOAT reads) or float64."""
    L = _lib.load()
    shape = h5_shape(filename, dataset)
    dt = np.float64 if np.dtype(dtype) == np.float64 else np.float32
    out = np.empty(shape, dt)
    _lib.check(L.lc_h5_read(_path(filename), dataset.encode(), out.ctypes.data, out.size,
                            1 if dt == np.float64 else 0))
    return out


def h5_read_images(filename, dataset="/images") -> np.ndarray:
    """tomocam::h5::read_images: a 3-D dataset {nangles, nrows, ncols} as float32."""
    if len(h5_shape(filename, dataset)) != 3:
        raise ValueError(f"expected 3-D dataset: {dataset}")
    return h5_read(filename, dataset)


def h5_read_angles(filename, dataset="/coords/alpha") -> np.ndarray:
    """tomocam::h5::read_angles: a 1-D dataset as float32."""
    if len(h5_shape(filename, dataset)) != 1:
        raise ValueError(f"expected 1-D dataset: {dataset}")
    return h5_read(filename, dataset)


def mask_infs_nans(projs) -> np.ndarray:
    """A copy with every NaN and infinity set to 0 (include/mask.h:43-52)."""
    out = np.array(projs, copy=True)
    out[~np.isfinite(out)] = 0
    return out


def mask_support(projs, dims, theta, gamma) -> np.ndarray:
    """Projections with the detector pixels outside the rectangular support
    of a {n1, n2, n3} object set to 0 (include/mask.h:67-110), in the
    projections' element type: centre (ncols / 2, nrows / 2) in integer
    division, limits n2 / 2 and n3 / 2 likewise; a projection angle whose
    cosine is below 1e-6 in magnitude is a ValueError."""
    p = np.asarray(projs)
    T = p.dtype.type if p.dtype in (np.float32, np.float64) else np.float64
    out = np.array(p, copy=True)
    cg, sg = T(math.cos(T(gamma))), T(math.sin(T(gamma)))
    xcen, ycen = T(p.shape[2] // 2), T(p.shape[1] // 2)
    ylim, xlim = T(int(dims[1]) // 2), T(int(dims[2]) // 2)
    xr = np.arange(p.shape[2], dtype=T) - xcen
    yr = (np.arange(p.shape[1], dtype=T) - ycen)[:, None]
    for i, a in enumerate(np.ravel(theta)):
        ct = T(math.cos(T(a)))
        if abs(ct) < 1e-6:
            raise ValueError("Projection angle too close to 90 degrees, support mask is undefined.")
        xc = (xr * cg * ct + yr * sg) / ct
        yc = (-xr * sg * ct + yr * cg) / ct
        out[i][(np.abs(xc) > xlim) | (np.abs(yc) > ylim)] = 0
    return out
