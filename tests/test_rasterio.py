"""Raster files on the host side (map_amd/csrc/rasterio.cpp): .bin and tiled TIFF / BigTIFF.

The TIFF reader and writer are checked against two independent implementations
present in this image: the system libtiff (the library the reference's
file/tiff.cpp uses; called through ctypes, with the calls the reference makes:
TIFFOpen "w8"/"r", the tags of tiff::setMeta / setOthers, TIFFWriteTile /
TIFFReadTile at absolute pixel coordinates of each block) and Pillow's TIFF
decoder. Files written here must read back identically in both; files libtiff
writes (BigTIFF and classic TIFF, little- and big-endian) must read back here.
No device is needed (ma_fileLoad / ma_fileStore / ma_read).
"""
import ctypes as ct
import os

import numpy as np
import pytest

from map_amd import map as M

TIFF_TYPES = [np.uint8, np.uint16, np.uint32, np.int8, np.int16, np.int32, np.float32, np.float64]
BIN_TYPES = TIFF_TYPES + [np.uint64, np.int64]
GEOMS = [((100, 70), (32, 16)), ((64, 64), (64, 64)), ((33, 17), (48, 32)), ((130, 90), (40, 40))]


def _array(dt, w, h, seed=0):
    rng = np.random.default_rng(seed)
    if np.issubdtype(dt, np.floating):
        return (rng.standard_normal((h, w)) * 1000).astype(dt)
    info = np.iinfo(dt)
    return rng.integers(info.min, info.max, size=(h, w), dtype=dt, endpoint=True)


@pytest.mark.parametrize("ext,types", [("bin", BIN_TYPES), ("tif", TIFF_TYPES)])
@pytest.mark.parametrize("ds,bs", GEOMS)
def test_store_load_round_trip(tmp_path, ext, types, ds, bs):
    for i, dt in enumerate(types):
        a = _array(dt, ds[0], ds[1], i)
        p = str(tmp_path / ("r%d.%s" % (i, ext)))
        M.file_store(p, a, bs)
        r = M.read(p)
        assert (tuple(r.datasize()), tuple(r.blocksize())) == (ds, bs)
        assert M._NP[r.datatype()] == np.dtype(dt)
        b = M.file_load(p)
        assert b.dtype == a.dtype and np.array_equal(a, b), (ext, dt)


def test_tiff_rejects_types_it_cannot_hold(tmp_path):
    for dt in (np.uint64, np.int64):
        with pytest.raises(RuntimeError, match="not supported by TIFF"):
            M.file_store(str(tmp_path / "x.tif"), np.zeros((8, 8), dt), (8, 8))
        assert not os.path.exists(str(tmp_path / "x.tif"))


def test_unknown_extension_is_an_error(tmp_path):
    with pytest.raises(RuntimeError, match="unknown raster file format"):
        M.file_store(str(tmp_path / "x.png"), np.zeros((8, 8), np.float32), (8, 8))


# ---- libtiff (system library) --------------------------------------------------------

def _libtiff():
    try:
        lt = ct.CDLL("libtiff.so.5")
    except OSError:
        pytest.skip("libtiff.so.5 not present")
    lt.TIFFOpen.restype = ct.c_void_p
    lt.TIFFOpen.argtypes = [ct.c_char_p, ct.c_char_p]
    lt.TIFFSetWarningHandler.restype = ct.c_void_p
    lt.TIFFSetWarningHandler(None)
    lt.TIFFReadTile.restype = ct.c_int64
    lt.TIFFWriteTile.restype = ct.c_int64
    lt.TIFFIsTiled.argtypes = [ct.c_void_p]
    lt.TIFFClose.argtypes = [ct.c_void_p]
    return lt


# TIFF tag numbers and values (TIFF 6.0; tiffio.h names)
IMAGEWIDTH, IMAGELENGTH, BITSPERSAMPLE, COMPRESSION, PHOTOMETRIC = 256, 257, 258, 259, 262
ORIENTATION, SAMPLESPERPIXEL, PLANARCONFIG, TILEWIDTH, TILELENGTH, SAMPLEFORMAT = 274, 277, 284, 322, 323, 339
SAMPLEFORMAT_OF = {"u": 1, "i": 2, "f": 3}


def _libtiff_write(lt, path, a, bs, mode, compression=1):
    """Tiled TIFF written the way file/tiff.cpp writes one: setMeta + setOthers tags,
    then one TIFFWriteTile per block at absolute pixel coordinates (writeBlock)."""
    t = lt.TIFFOpen(path.encode(), mode)
    assert t
    T = ct.c_void_p(t)
    h, w = a.shape
    sf = SAMPLEFORMAT_OF[a.dtype.kind]
    U16, U32 = ct.c_int, ct.c_uint32  # uint16 arguments promote to int in the variadic call
    for tag, val in [(SAMPLESPERPIXEL, U16(1)), (SAMPLEFORMAT, U16(sf)), (BITSPERSAMPLE, U16(a.dtype.itemsize * 8)),
                     (IMAGEWIDTH, U32(w)), (IMAGELENGTH, U32(h)), (TILEWIDTH, U32(bs[0])), (TILELENGTH, U32(bs[1])),
                     (PHOTOMETRIC, U16(1)), (COMPRESSION, U16(compression)), (ORIENTATION, U16(1)),
                     (PLANARCONFIG, U16(1))]:
        assert lt.TIFFSetField(T, ct.c_uint32(tag), val) == 1, tag
    nb = ((w + bs[0] - 1) // bs[0], (h + bs[1] - 1) // bs[1])
    for by in range(nb[1]):
        for bx in range(nb[0]):
            tile = np.zeros((bs[1], bs[0]), a.dtype)
            part = a[by * bs[1]:(by + 1) * bs[1], bx * bs[0]:(bx + 1) * bs[0]]
            tile[:part.shape[0], :part.shape[1]] = part
            n = lt.TIFFWriteTile(T, tile.ctypes.data_as(ct.c_void_p), ct.c_uint32(bx * bs[0]), ct.c_uint32(by * bs[1]),
                                 ct.c_uint32(0), ct.c_uint16(0))
            assert n == tile.nbytes
    lt.TIFFClose(T)


def _libtiff_read(lt, path, ds, bs, dt):
    t = lt.TIFFOpen(path.encode(), b"r")
    assert t
    T = ct.c_void_p(t)
    assert lt.TIFFIsTiled(T) == 1
    w, h = ds
    out = np.zeros((h, w), dt)
    nb = ((w + bs[0] - 1) // bs[0], (h + bs[1] - 1) // bs[1])
    for by in range(nb[1]):
        for bx in range(nb[0]):
            tile = np.zeros((bs[1], bs[0]), dt)
            n = lt.TIFFReadTile(T, tile.ctypes.data_as(ct.c_void_p), ct.c_uint32(bx * bs[0]), ct.c_uint32(by * bs[1]),
                                ct.c_uint32(0), ct.c_uint16(0))
            assert n == tile.nbytes
            part = out[by * bs[1]:(by + 1) * bs[1], bx * bs[0]:(bx + 1) * bs[0]]
            part[:] = tile[:part.shape[0], :part.shape[1]]
    lt.TIFFClose(T)
    return out


@pytest.mark.parametrize("ds,bs", [((100, 70), (32, 16)), ((64, 64), (64, 64)), ((130, 90), (48, 32))])
def test_written_tiff_reads_back_in_libtiff(tmp_path, ds, bs):
    lt = _libtiff()
    for i, dt in enumerate(TIFF_TYPES):
        a = _array(dt, ds[0], ds[1], 10 + i)
        p = str(tmp_path / ("w%d.tif" % i))
        M.file_store(p, a, bs)
        b = _libtiff_read(lt, p, ds, bs, dt)
        assert np.array_equal(a, b), dt


@pytest.mark.parametrize("mode", [b"w8", b"w", b"w8b", b"wb"])  # BigTIFF / classic, little / big endian
@pytest.mark.parametrize("ds,bs", [((100, 70), (32, 16)), ((48, 48), (48, 48))])
def test_libtiff_written_tiff_reads_here(tmp_path, mode, ds, bs):
    lt = _libtiff()
    for i, dt in enumerate(TIFF_TYPES):
        a = _array(dt, ds[0], ds[1], 20 + i)
        p = str(tmp_path / ("l%d.tif" % i))
        _libtiff_write(lt, p, a, bs, mode)
        r = M.read(p)
        assert (tuple(r.datasize()), tuple(r.blocksize())) == (ds, bs) and M._NP[r.datatype()] == np.dtype(dt)
        assert np.array_equal(M.file_load(p), a), (mode, dt)


def test_compressed_tiff_is_rejected(tmp_path):
    lt = _libtiff()
    p = str(tmp_path / "c.tif")
    _libtiff_write(lt, p, _array(np.float32, 64, 32), (32, 32), b"w8", compression=8)  # deflate
    with pytest.raises(RuntimeError, match="compressed"):
        M.read(p)


# ---- Pillow ---------------------------------------------------------------------------

def test_written_tiff_reads_back_in_pillow(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    for dt in (np.uint8, np.uint16, np.int32, np.float32):
        a = _array(dt, 100, 70, 3)
        p = str(tmp_path / "p.tif")
        M.file_store(p, a, (32, 16))
        with Image.open(p) as im:
            assert im.size == (100, 70)
            assert np.array_equal(np.array(im), a), dt


def test_striped_tiff_is_rejected(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    p = str(tmp_path / "s.tif")
    Image.fromarray(_array(np.float32, 40, 30)).save(p)  # Pillow writes strips
    with pytest.raises(RuntimeError, match="only tiled TIFF"):
        M.read(p)
