"""Pins the oracle's FocalPercent and Diversity (oracle/ref.py) with hand-worked cells.

FocalPercent divides by the reference's neighbourhood size (2*h0+1)*h1*2+1 (7 for a 3x3
window: nbh_size in runtime/skeleton/util.cpp:321-327 misses parentheses), counts in the
raster's type and divides with C semantics. Diversity keeps distinct values in order of first
appearance; MEAN divides by the number of distinct values, accumulating in the type (U8 wraps).
"""
import numpy as np

from oracle import ref as O


def test_focal_percent_closed_forms():
    O.set_mode("f32")
    const = O.from_array(np.full((20, 30), 7, np.float32), (16, 16))
    age = O.focal_percent(const, [[1, 1, 1], [1, 1, 1], [1, 1, 1]], O.AGE).val
    assert np.all(age == np.float32(9) / np.float32(7))
    assert np.all(O.focal_percent(const, [[1, 1, 1]] * 3, O.ILE).val == 0)
    wide = O.focal_percent(const, [[1] * 5] * 3, O.AGE).val  # 5x3 window: 15 cells / ((5)*1*2+1 = 11)
    assert np.all(wide == np.float32(15) / np.float32(11))
    ramp = O.from_array(np.tile(np.arange(30, dtype=np.int32), (20, 1)), (16, 16))
    ile = O.focal_percent(ramp, [[1, 1, 1]] * 3, O.ILE).val
    assert np.all(ile[1:-1, 1:-1] == 3 // 7) and ile.dtype == np.int32  # 3 cells below the centre
    fr = O.from_array(np.tile(np.arange(30, dtype=np.float32), (20, 1)), (16, 16))
    assert np.all(O.focal_percent(fr, [[1, 1, 1]] * 3, O.ILE).val[1:-1, 1:-1] == np.float32(3) / np.float32(7))


def _div(cells, dt, kind):
    xs = [O.from_array(np.array([[c[k] for c in cells]], dtype=dt), (8, 8)) for k in range(len(cells[0]))]
    return O.diversity(xs, kind).val[0]


def test_diversity_hand_worked_cells():
    cells = [(1, 1, 2, 3), (5, 5, 5, 5), (4, 2, 2, 4), (9, 8, 7, 6)]
    assert list(_div(cells, np.int32, O.VARI)) == [3, 1, 2, 4]
    assert list(_div(cells, np.int32, O.MAJO)) == [1, 5, 4, 9]  # ties: the first in order of appearance
    assert list(_div(cells, np.int32, O.MINO)) == [2, 5, 4, 9]
    assert list(_div(cells, np.int32, O.MEAN)) == [7 // 3, 20, 12 // 2, 30 // 4]
    f = _div(cells, np.float32, O.MEAN)
    assert f[0] == np.float32(7) / np.float32(3) and f[3] == np.float32(7.5)
    assert _div([(-7, -7, 1, 1)], np.int32, O.MEAN)[0] == -6  # (-7*2 + 1*2) / 2
    assert _div([(-7, 0, 0, 0)], np.int32, O.MEAN)[0] == -3  # -7 / 2 truncates toward zero
    assert _div([(200, 200, 100, 50)], np.uint8, O.MEAN)[0] == ((400 + 100 + 50) % 256) // 3  # 38 // 3
    nan = np.float32("nan")
    v = [_div([(nan, nan, 1, 1)], np.float32, k)[0] for k in (O.VARI, O.MAJO, O.MINO, O.MEAN)]
    assert v[0] == 3 and v[1] == 1 and np.isnan(v[2]) and np.isnan(v[3])  # NaN never equals itself
