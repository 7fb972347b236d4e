"""Pins of the oracle's FocalFunc (fmax / fmin and the other reductions) and of the
zonal PROD / rAND / rOR reductions, by hand-worked cells, brute force over neighbour
rasters, and the C wrap of integer types.

FocalFunc (runtime/skeleton/FocalSkeleton.cpp:338-372) reduces the window of the mask's
size from the reduction's neutral value (ReductionType.cpp:66-78) and ignores the mask's
values (quirk B3); ZonalReduc PROD / rAND / rOR (ZonalSkeleton.cpp:51-288) accumulate in
the raster's type, so U8 products wrap modulo 256.
"""
import numpy as np

from oracle import ref as O


def test_fmax_fmin_hand_cells():
    # one block; the centre cell (1, 1) sees the whole 3x3 raster
    a = np.array([[3, -7, 5], [0, 2, 9], [-1, 4, 8]], np.int32)
    r = O.from_array(a, (3, 3))
    assert O.fmax(r, [[1] * 3] * 3).val[1, 1] == 9
    assert O.fmin(r, [[1] * 3] * 3).val[1, 1] == -7
    # the weights are not used (quirk B3): a zero mask gives the same window maximum
    assert O.fmax(r, [[0] * 3] * 3).val[1, 1] == 9
    # a 1-row mask (3 wide, 1 high) only sees the centre row
    assert O.fmax(r, [[1, 1, 1]]).val[1, 1] == 9 and O.fmin(r, [[1, 1, 1]]).val[1, 1] == 0
    assert O.fmax(r, [[1], [1], [1]]).val[1, 1] == 4 and O.fmin(r, [[1], [1], [1]]).val[1, 1] == -7


def test_fmax_skips_nan_and_keeps_type():
    a = np.array([[1.5, np.nan, -2.0], [np.nan, np.nan, 0.25], [7.0, -3.5, np.nan]], np.float32)
    r = O.from_array(a, (3, 3))
    mx, mn = O.fmax(r, [[1] * 3] * 3), O.fmin(r, [[1] * 3] * 3)
    assert mx.val.dtype == np.float32 and mx.val[1, 1] == np.float32(7.0) and mn.val[1, 1] == np.float32(-3.5)
    u = O.from_array(np.array([[200, 13], [255, 0]], np.uint8), (2, 2))
    assert O.fmax(u, [[1] * 3] * 3).val.dtype == np.uint8
    assert O.fmax(u, [[1] * 3] * 3).val[0, 0] == 255 and O.fmin(u, [[1] * 3] * 3).val[0, 0] == 0


def test_fmax_equals_max_over_neighbour_rasters():
    """At every cell, including block seams and raster edges, the window maximum equals the
    elementwise maximum of the nine neighbour rasters (each read with the border rule)."""
    rng = np.random.default_rng(3)
    a = rng.integers(-1000, 1000, (37, 53)).astype(np.int32)
    r = O.from_array(a, (16, 8))
    want = None
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            n = O.neighbor(r, (dx, dy)).val
            want = n if want is None else np.maximum(want, n)
    assert np.array_equal(O.fmax(r, [[1] * 3] * 3).val, want)
    want5 = None
    for dy in (-1, 0, 1):
        for dx in (-2, -1, 0, 1, 2):
            n = O.neighbor(r, (dx, dy)).val
            want5 = n if want5 is None else np.minimum(want5, n)
    assert np.array_equal(O.fmin(r, [[1] * 5] * 3).val, want5)


def test_focal_sum_and_prod_wrap():
    """FocalFunc SUM over a 3x3 window equals convolution with ones (same order, same type);
    U8 PROD wraps modulo 256."""
    rng = np.random.default_rng(4)
    a = rng.integers(0, 256, (20, 30)).astype(np.uint8)
    r = O.from_array(a, (8, 8))
    s = O.ffunc(r, [[1] * 3] * 3, O.SUM)
    assert s.val.dtype == np.uint8
    assert np.array_equal(s.val, (O.convolve(r, [[1] * 3] * 3).val.astype(np.int64) % 256).astype(np.uint8))
    p = O.ffunc(O.from_array(np.full((3, 3), 3, np.uint8), (3, 3)), [[1] * 3] * 3, O.PROD)
    assert p.val[1, 1] == (3 ** 9) % 256


def test_zonal_prod_and_or_hand_cases():
    def z(vals, dt, red):
        a = np.array(vals, dt).reshape(1, -1)
        return O.value(O._zred(O.from_array(a, (len(vals), 1)), red))
    assert z([3, 5, 7], np.uint8, O.PROD) == 105
    assert z([16, 17], np.uint8, O.PROD) == 16           # 272 mod 256
    assert z([2] * 9, np.uint8, O.PROD) == 0              # 512 mod 256
    assert z([-3, 5, 7], np.int32, O.PROD) == -105
    assert z([65536, 65536], np.int32, O.PROD) == 0       # 2^32 wraps to 0
    assert z([0xF0, 0x3C], np.uint8, O.rAND) == 0x30
    assert z([0xF0, 0x3C], np.uint8, O.rOR) == 0xFC
    assert z([-1, 6], np.int32, O.rAND) == 6
    assert z([-8, 3], np.int32, O.rOR) == -5              # ...11111000 | 011
    # the reference folds from the neutral, which for signed rAND / rOR is the type's
    # maximum / minimum (ReductionType.cpp:80-90): the sign bit is cleared / set
    assert z([-1, -1], np.int32, O.rAND) == 2 ** 31 - 1
    assert z([5, 3], np.int32, O.rOR) == -(2 ** 31) | 7
    assert z([-1], np.int8, O.rAND) == 127 and z([1], np.int16, O.rOR) == -32768 + 1
    assert z([0.5, 4.0, 3.0], np.float32, O.PROD) == 6.0
    assert z([2.0] * 10, np.float32, O.PROD) == 1024.0


def test_reduction_neutrals():
    """The neutral of each reduction (ReductionType.cpp:80-90): a one-cell raster reduces to
    its own value."""
    for dt, v in ((np.uint8, 77), (np.int32, -12345), (np.float32, 2.5)):
        for red in (O.SUM, O.PROD, O.MAX, O.MIN) + ((O.rAND, O.rOR) if dt == np.uint8 else ()):
            a = np.array([[v]], dt)
            assert O.value(O._zred(O.from_array(a, (1, 1)), red)) == v, (dt, red)
    assert O._neutral(O.rAND, O.U8) == 255 and O._neutral(O.rOR, O.S32) == np.iinfo(np.int32).min
    assert O._neutral(O.MAX, O.F32) == -np.inf and O._neutral(O.PROD, O.S16) == 1
