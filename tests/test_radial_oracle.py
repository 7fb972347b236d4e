"""Pins the oracle's radial scan (oracle/ref.py _radial) before it is trusted as the checker.

* Rays along the axes and diagonals from the start cell never interpolate: their values are
  the running reduction of the finite inputs along the ray, from the neutral element
  (closed form: a sequential accumulation).
* With MAX and the input max(|dx|, |dy|) (the distance d0 in every octant) the scan returns its
  input exactly: every interpolated predecessor is at most d0 - 1.
* A per-cell recursion written independently (memoized, the octant chosen from the cell's own
  offset) gives the same bits as the column-by-column evaluation, for every reduction, both
  floating types and start cells in corners, on edges and inside; the same recursion with the
  two interpolation operands swapped (a planted mistake) does not.
"""
import sys

import numpy as np
import pytest

from oracle import ref as O

sys.setrecursionlimit(100000)
NEUTRAL = {O.SUM: 0.0, O.PROD: 1.0, O.MAX: -np.inf, O.MIN: np.inf}
OPS = {O.SUM: np.add, O.PROD: np.multiply, O.MAX: np.fmax, O.MIN: np.fmin}
SCAN = {O.SUM: O.rsum, O.PROD: O.rprod, O.MAX: O.rmax, O.MIN: O.rmin}


def _input(h, w, dt, seed, specials=True):
    rng = np.random.default_rng(seed)
    a = (rng.standard_normal((h, w)) * 3).astype(dt)
    if specials:
        a[rng.random((h, w)) < 0.03] = np.nan
        a[rng.random((h, w)) < 0.01] = np.inf
    return a


def _recursive(a, red, s, swap=False):
    T = a.dtype.type
    H, W = a.shape
    op = OPS[red]
    memo = {}

    def R(x, y):
        if (x, y) in memo:
            return memo[(x, y)]
        dx, dy = x - s[0], y - s[1]
        if dx == 0 and dy == 0:
            base = T(NEUTRAL[red])
        else:
            if abs(dx) >= abs(dy):
                F, S, d0, d1 = (np.sign(dx), 0), (0, np.sign(dy)), abs(dx), abs(dy)
            else:
                F, S, d0, d1 = (0, np.sign(dy)), (np.sign(dx), 0), abs(dy), abs(dx)
            if d1 == 0:
                base = R(x - F[0], y - F[1])
            elif d1 == d0:
                base = R(x - F[0] - S[0], y - F[1] - S[1])
            else:
                pa, pb = R(x - F[0], y - F[1]), R(x - F[0] - S[0], y - F[1] - S[1])
                if swap:
                    pa, pb = pb, pa
                w = T(d1) / T(d0) * T(d0 - 1) - T(d1 - 1)
                base = T(T(pa * w) + T(pb * (T(1) - w)))
        v = a[y, x]
        with np.errstate(all="ignore"):
            r = T(op(base, v)) if np.isfinite(v) else base
        memo[(x, y)] = r
        return r

    out = np.empty_like(a)
    for y in range(H):
        for x in range(W):
            out[y, x] = R(x, y)
    return out


def _same(a, b):
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("red", [O.SUM, O.MAX, O.MIN, O.PROD])
def test_rays_on_axes_and_diagonals_are_running_reductions(red):
    O.set_mode("f32")
    a = _input(41, 53, np.float32, 1)
    if red == O.PROD:
        a = np.where(np.isfinite(a), 1 + a / 64, a).astype(np.float32)
    s = (20, 17)
    out = SCAN[red](O.from_array(a, (16, 16)), s).val
    for ux, uy in [(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (1, -1), (-1, 1), (-1, -1)]:
        acc = np.float32(NEUTRAL[red])
        k = 0
        while 0 <= s[0] + k * ux < a.shape[1] and 0 <= s[1] + k * uy < a.shape[0]:
            x, y = s[0] + k * ux, s[1] + k * uy
            if np.isfinite(a[y, x]):
                with np.errstate(all="ignore"):
                    acc = np.float32(OPS[red](acc, a[y, x]))
            assert _same(np.array([out[y, x]]), np.array([acc])), (red, ux, uy, k)
            k += 1


def test_max_of_the_octant_distance_is_itself():
    O.set_mode("f32")
    for s in [(0, 0), (37, 5), (12, 29), (63, 47)]:
        yy, xx = np.mgrid[0:48, 0:64]
        d = np.maximum(abs(xx - s[0]), abs(yy - s[1])).astype(np.float32)
        assert np.array_equal(O.rmax(O.from_array(d, (32, 16)), s).val, d), s


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("red", [O.SUM, O.MAX, O.MIN, O.PROD])
def test_columns_match_a_per_cell_recursion(dt, red):
    O.set_mode("f32")
    a = _input(19, 23, dt, 7)
    if red == O.PROD:
        a = np.where(np.isfinite(a), 1 + a / 64, a).astype(dt)
    for s in [(0, 0), (22, 0), (0, 18), (22, 18), (11, 9), (5, 14), (22, 7)]:
        got = SCAN[red](O.from_array(a, (8, 8)), s).val
        assert got.dtype == dt
        assert _same(got, _recursive(a, red, s)), (red, s)


def test_swapped_interpolation_operands_are_caught():
    """A planted mistake in the checker (a and b exchanged) must not go unnoticed."""
    O.set_mode("f32")
    a = _input(19, 23, np.float32, 9, specials=False)
    got = O.rsum(O.from_array(a, (8, 8)), (4, 6)).val
    assert _same(got, _recursive(a, O.SUM, (4, 6)))
    assert not _same(got, _recursive(a, O.SUM, (4, 6), swap=True))
