"""The file-format codecs (lc_tiff_* / lc_h5_* of include/laminocam.h, through
laminocam_amd.io), on the CPU.

TIFF (include/tiff.h:31-117) is checked against Pillow, an independent TIFF
implementation: Pillow reads what the library writes, and the library reads
what Pillow writes uncompressed, LZW, deflate and PackBits, in one strip and
in many.  A big-endian file and a BigTIFF file built here are checked against
Pillow's reading of them first.  The reader's crop of even sizes to odd ones
is the reference's.

HDF5 (include/hdfread.h:34-96) is checked against files written by libhdf5
(tests/golden/h5/, from the PyTables test suite, with the contents its tests
state): big- and little-endian float64, int32 and int64 arrays, float16 /
float32 / float64 arrays and a chunked, extendible big-endian int32 array.
Deflate, shuffle and edge chunks run through files from tests/h5write.py.
"""
import os
import struct

import numpy as np
import pytest

from laminocam_amd import io as lio
from laminocam_amd._lib import LcError

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "h5")


def _pil_pages(path):
    from PIL import Image
    im = Image.open(path)
    out = []
    for i in range(im.n_frames):
        im.seek(i)
        out.append(np.array(im, dtype=np.float32))
    return np.stack(out)


def _pil_write(path, a, **kw):
    from PIL import Image
    ims = [Image.fromarray(np.ascontiguousarray(x), mode="F") for x in a]
    ims[0].save(path, save_all=True, append_images=ims[1:], **kw)


@pytest.fixture
def rng():
    return np.random.default_rng(20261203)


def test_tiff_round_trip_and_odd_crop(tmp_path, rng):
    a = rng.standard_normal((3, 6, 8)).astype(np.float32)
    p = str(tmp_path / "a.tif")
    lio.tiff_write(p, a)
    assert lio.tiff_info(p) == (3, 6, 8)
    assert np.array_equal(lio.tiff_read(p, odd=False), a)
    # the reference's reader drops the last row / column of even sizes
    r = lio.tiff_read(p)
    assert r.shape == (3, 5, 7) and np.array_equal(r, a[:, :5, :7])
    b = rng.standard_normal((2, 5, 7)).astype(np.float32)
    lio.tiff_write(p, b)
    assert np.array_equal(lio.tiff_read(p), b)
    # one 2-D page
    lio.tiff_write(p, b[0])
    assert lio.tiff_info(p) == (1, 5, 7)


def test_tiff_written_here_reads_in_pillow(tmp_path, rng):
    a = rng.standard_normal((4, 33, 17)).astype(np.float32)
    a[1, 2, 3] = np.inf
    a[2, 0, 0] = np.nan
    p = str(tmp_path / "w.tif")
    lio.tiff_write(p, a)
    assert np.array_equal(_pil_pages(p), a, equal_nan=True)


@pytest.mark.parametrize("compression", [None, "tiff_lzw", "tiff_adobe_deflate", "packbits"])
@pytest.mark.parametrize("kind", ["noise", "smooth", "constant"])
def test_tiff_written_by_pillow(tmp_path, rng, compression, kind):
    shape = (2, 257, 301)
    if kind == "noise":
        a = rng.standard_normal(shape).astype(np.float32)
    elif kind == "smooth":
        a = np.sin(np.arange(np.prod(shape)).reshape(shape) / 50.0).astype(np.float32)
    else:
        a = np.full(shape, 3.25, np.float32)
    p = str(tmp_path / "p.tif")
    _pil_write(p, a, **({} if compression is None else {"compression": compression}))
    assert np.array_equal(lio.tiff_read(p, odd=False), a)
    assert np.array_equal(lio.tiff_read(p), a[:, :257, :301])   # odd sizes stay


def test_tiff_many_strips(tmp_path, rng):
    # Pillow writes compressed images in strips of a few rows
    a = rng.standard_normal((1, 1024, 700)).astype(np.float32)
    p = str(tmp_path / "s.tif")
    _pil_write(p, a, compression="tiff_lzw")
    from PIL import Image
    assert len(Image.open(p).tag_v2[273]) > 1
    assert np.array_equal(lio.tiff_read(p, odd=False), a)


def _ifd_entries(rows, cols, data_at, fmt_short, fmt_long, big):
    return [(256, 4, cols), (257, 4, rows), (258, 3, 32), (259, 3, 1), (262, 3, 1),
            (273, 16 if big else 4, data_at), (277, 3, 1), (278, 4, rows),
            (279, 16 if big else 4, rows * cols * 4), (339, 3, 3)]


def _hand_tiff(a, endian, big):
    """one page, one strip; endian '<' or '>'; big: BigTIFF"""
    rows, cols = a.shape
    e = endian
    bo = b"II" if e == "<" else b"MM"
    data = a.astype(e + "f4").tobytes()
    hdr = bo + (struct.pack(e + "HHHQ", 43, 8, 0, 0) if big else struct.pack(e + "HI", 42, 0))
    data_at = len(hdr)
    ifd_at = data_at + len(data)
    ents = _ifd_entries(rows, cols, data_at, None, None, big)
    ifd = struct.pack(e + ("Q" if big else "H"), len(ents))
    for tag, typ, val in ents:
        if big:
            vb = struct.pack(e + {3: "H", 4: "I", 16: "Q"}[typ], val).ljust(8, b"\0")
            ifd += struct.pack(e + "HHQ", tag, typ, 1) + vb
        else:
            vb = struct.pack(e + {3: "H", 4: "I"}[typ], val).ljust(4, b"\0")
            ifd += struct.pack(e + "HHI", tag, typ, 1) + vb
    ifd += struct.pack(e + ("Q" if big else "I"), 0)
    hdr = bo + (struct.pack(e + "HHHQ", 43, 8, 0, ifd_at) if big else struct.pack(e + "HI", 42, ifd_at))
    return hdr + data + ifd


@pytest.mark.parametrize("endian,big", [(">", False), ("<", True), (">", True)])
def test_tiff_big_endian_and_bigtiff(tmp_path, rng, endian, big):
    a = rng.standard_normal((9, 11)).astype(np.float32)
    p = str(tmp_path / "h.tif")
    with open(p, "wb") as f:
        f.write(_hand_tiff(a, endian, big))
    assert np.array_equal(lio.tiff_read(p, odd=False)[0], a)
    if endian == ">" and big:
        return   # Pillow does not read big-endian BigTIFF; the two other forms pin the layout
    assert np.array_equal(_pil_pages(p)[0], a)   # the file is what it claims to be


def test_tiff_errors(tmp_path):
    from PIL import Image
    p = str(tmp_path / "u16.tif")
    Image.fromarray(np.arange(12, dtype=np.uint16).reshape(3, 4)).save(p)
    with pytest.raises(LcError, match="32-bit IEEE floats"):
        lio.tiff_read(p)
    q = str(tmp_path / "not.tif")
    with open(q, "wb") as f:
        f.write(b"plain text, not an image")
    with pytest.raises(LcError, match="not a TIFF"):
        lio.tiff_read(q)
    with pytest.raises(LcError, match="cannot open"):
        lio.tiff_read(str(tmp_path / "missing.tif"))
    r = str(tmp_path / "sizes.tif")
    ims = [Image.fromarray(np.zeros((3, 4), np.float32), mode="F"),
           Image.fromarray(np.zeros((5, 4), np.float32), mode="F")]
    ims[0].save(r, save_all=True, append_images=ims[1:])
    with pytest.raises(LcError, match="pages differ in size"):
        lio.tiff_read(r)
    with pytest.raises(ValueError):
        lio.tiff_write(str(tmp_path / "x.tif"), np.zeros(5, np.float32))


# ---- HDF5 ----

EXP_6x5 = np.add.outer(np.arange(6), np.arange(5))


@pytest.mark.parametrize("name", ["smpl_f64le", "smpl_f64be", "smpl_i32le", "smpl_i32be", "smpl_i64be"])
def test_h5_libhdf5_numeric(name):
    p = os.path.join(GOLD, name + ".h5")
    assert lio.h5_shape(p, "/TestArray") == (6, 5)
    a = lio.h5_read(p, "/TestArray", np.float64)
    assert np.array_equal(a, EXP_6x5)
    assert lio.h5_read(p, "TestArray").dtype == np.float32


def test_h5_libhdf5_float_kinds():
    p = os.path.join(GOLD, "float.h5")
    v = np.add.outer(np.arange(5), np.arange(6))
    for name in ("float16", "float32", "float64"):
        assert np.array_equal(lio.h5_read(p, "/" + name, np.float64), v), name


def test_h5_libhdf5_chunked_extendible():
    # chunked storage (version-1 B-tree index), big-endian int32, a partly written extent
    p = os.path.join(GOLD, "smpl_SDSextendible.h5")
    exp = np.array([[1, 1, 1, 3, 3], [1, 1, 1, 3, 3], [1, 1, 1, 0, 0]] + [[2, 0, 0, 0, 0]] * 7)
    assert np.array_equal(lio.h5_read(p, "/ExtendibleArray", np.float64), exp)


def test_h5_images_and_angles(tmp_path, rng):
    import h5write
    im = rng.standard_normal((7, 5, 9)).astype(np.float32)
    al = np.linspace(-60, 60, 7)
    p = str(tmp_path / "t.h5")
    h5write.write(p, {"images": im, "coords/alpha": al})
    assert np.array_equal(lio.h5_read_images(p), im)
    got = lio.h5_read_angles(p)
    assert got.dtype == np.float32 and np.array_equal(got, al.astype(np.float32))
    assert np.array_equal(lio.h5_read(p, "/coords/alpha", np.float64), al)
    with pytest.raises(ValueError, match="expected 3-D dataset"):
        lio.h5_read_images(p, "/coords/alpha")
    with pytest.raises(ValueError, match="expected 1-D dataset"):
        lio.h5_read_angles(p, "/images")
    with pytest.raises(LcError, match="no object /nothere"):
        lio.h5_read(p, "/nothere")
    with pytest.raises(LcError, match="not a dataset"):
        lio.h5_read(p, "/coords")
    with pytest.raises(LcError, match="not an HDF5 file"):
        lio.h5_read(os.path.join(GOLD, "..", "..", "test_io.py"), "/images")


@pytest.mark.parametrize("shuffle", [False, True])
@pytest.mark.parametrize("deflate", [None, 6])
def test_h5_chunked_filters(tmp_path, rng, shuffle, deflate):
    import h5write
    im = rng.standard_normal((7, 5, 9))
    al = (np.arange(10) - 3).astype("<i2")
    p = str(tmp_path / "c.h5")
    h5write.write(p, {"images": im.astype(">f8"), "coords/alpha": al},
                  chunks={"images": (3, 2, 4), "coords/alpha": (4,)}, deflate=deflate, shuffle=shuffle)
    assert np.array_equal(lio.h5_read(p, "/images", np.float64), im)   # edge chunks clipped
    assert np.array_equal(lio.h5_read(p, "/coords/alpha", np.float64), al)


def test_masks():
    a = np.array([[[1.0, np.nan], [np.inf, -np.inf]]], np.float32)
    m = lio.mask_infs_nans(a)
    assert np.array_equal(m, [[[1.0, 0.0], [0.0, 0.0]]]) and np.isnan(a[0, 0, 1])
    # untilted, theta = 0: the support is the rectangle |x - xc| <= n3/2, |y - yc| <= n2/2
    p = np.ones((1, 21, 31), np.float32)
    s = lio.mask_support(p, (5, 11, 15), [0.0], 0.0)
    xs = np.abs(np.arange(31) - 15) <= 7
    ys = np.abs(np.arange(21) - 10) <= 5
    assert np.array_equal(s[0], np.outer(ys, xs).astype(np.float32))
    # untilted, at theta the rows keep |y - yc| <= (n2/2) cos(theta) (here 2.5)
    s2 = lio.mask_support(p, (5, 11, 15), [np.pi / 3], 0.0)
    ys2 = np.abs(np.arange(21) - 10) <= 2.5
    assert np.array_equal(s2[0], np.outer(ys2, xs).astype(np.float32))
    with pytest.raises(ValueError, match="too close to 90"):
        lio.mask_support(p, (5, 11, 15), [np.pi / 2], 0.0)
