"""The .bin ROW+BLK raster format and block tiling, in numpy.

Layout (file/binary.cpp:221-390 of the reference): a 4096-byte header of
little-endian int32 {data_type, num_dim, ds0, ds1, mem_order, bs0, bs1}, zero
padded, then the blocks in proj order (bx + by*NB0), each a full BS0*BS1
row-major block; edge blocks are stored full size (padded). The same
block-tiled layout is used in HBM.
"""
import numpy as np

PAGE = 4096
D2 = 0x06
ROW_BLK = 0x03

DTYPES = {1: np.float32, 2: np.float64, 3: np.uint8, 4: np.uint8, 5: np.uint16, 6: np.uint32, 7: np.uint64,
          8: np.int8, 9: np.int16, 10: np.int32, 11: np.int64}
CODES = {np.dtype(np.float32): 1, np.dtype(np.float64): 2, np.dtype(np.uint8): 4, np.dtype(np.uint16): 5,
         np.dtype(np.uint32): 6, np.dtype(np.uint64): 7, np.dtype(np.int8): 8, np.dtype(np.int16): 9,
         np.dtype(np.int32): 10, np.dtype(np.int64): 11}


def num_blocks(ds, bs):
    return ((ds[0] + bs[0] - 1) // bs[0], (ds[1] + bs[1] - 1) // bs[1])


def to_blocks(a, bs, pad=0):
    """2D array (rows=y, cols=x) -> 1D block-tiled array with padded edge blocks."""
    h, w = a.shape
    nb0, nb1 = num_blocks((w, h), bs)
    full = np.full((nb1 * bs[1], nb0 * bs[0]), pad, dtype=a.dtype)
    full[:h, :w] = a
    t = full.reshape(nb1, bs[1], nb0, bs[0]).transpose(0, 2, 1, 3)
    return np.ascontiguousarray(t).reshape(-1)


def from_blocks(flat, ds, bs):
    """1D block-tiled array -> 2D array of the data size (padding dropped)."""
    nb0, nb1 = num_blocks(ds, bs)
    t = np.asarray(flat).reshape(nb1, nb0, bs[1], bs[0]).transpose(0, 2, 1, 3)
    return np.ascontiguousarray(t.reshape(nb1 * bs[1], nb0 * bs[0])[:ds[1], :ds[0]])


def write_bin(path, a, bs, dtype_code=None):
    a = np.asarray(a)
    code = dtype_code or CODES[a.dtype]
    hdr = np.zeros(PAGE // 4, dtype='<i4')
    hdr[:7] = [code, D2, a.shape[1], a.shape[0], ROW_BLK, bs[0], bs[1]]
    with open(path, 'wb') as f:
        f.write(hdr.tobytes())
        f.write(to_blocks(a, bs).astype(a.dtype.newbyteorder('<'), copy=False).tobytes())


def read_bin(path):
    """Returns (2D array, block size, dtype code)."""
    with open(path, 'rb') as f:
        hdr = np.frombuffer(f.read(PAGE), dtype='<i4')
        code, nd, ds0, ds1, mo, bs0, bs1 = (int(x) for x in hdr[:7])
        if nd != D2 or mo != ROW_BLK:
            raise ValueError("only D2 ROW+BLK rasters are supported")
        dt = np.dtype(DTYPES[code]).newbyteorder('<')
        nb0, nb1 = num_blocks((ds0, ds1), (bs0, bs1))
        flat = np.frombuffer(f.read(nb0 * nb1 * bs0 * bs1 * dt.itemsize), dtype=dt)
    return from_blocks(flat, (ds0, ds1), (bs0, bs1)), (bs0, bs1), code
