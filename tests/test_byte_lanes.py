"""Byte-lane form of U8 focal groups (four cells to a dword, one byte lane per cell).

CPU side: which groups the code generator gives a byte-lane form and with which bound on
the staged values (255: any data; 1: tiles of 0/1 values), and the lane formulas of the device header (lane_eq, lane_ne, lane_lt,
lane_ge, lane_sel, lane_pick) checked against the comparisons they stand for, over every
pair of lane values they are used on and with neighbouring lanes at random. The GPU parity
tests (test_gpu_parity.py) run both forms against the oracle.
"""
import numpy as np

from map_amd import workloads as scripts


def _lane_r(M, expr):
    gs = [g for g in M.plan(expr)["groups"] if "kernel" in g]
    assert len(gs) == 1, [g["kinds"] for g in gs]
    return gs[0]["lane_r"], gs[0]["source"]


def _u8(M, w=300, h=200, bs=(128, 64)):
    return M.from_array(np.zeros((h, w), np.uint8), bs)


def lane_scripts(M, u):
    """Scripts of a U8 raster with a byte-lane form (name -> (expression, bound))."""
    nbh = M.convolve(u, scripts.LIFE_MASK)
    cross = u([1, 0]) + u([-1, 0]) + u([0, 1]) + u([0, -1])
    return {
        "life": (scripts.life(M, u, 1), 1),
        "rule with conditional": (M.con(u == 1, (nbh >= 2) & (nbh <= 3), nbh == 3), 1),
        "cross majority": ((cross > 2) | (u != 0), 1),
        "xor of neighbours": (u([1, 0]) ^ u([-1, 0]) ^ u, 255),
        "7x7 count": (M.convolve(u, [[1] * 7] * 7) > 20, 1),
    }


def test_lane_bounds(M):
    u = _u8(M)
    for name, (expr, bound) in lane_scripts(M, u).items():
        r, src = _lane_r(M, expr)
        assert r == bound, (name, r)
        assert "post_lanes" in src and "static constexpr unsigned LANE_R = %du;" % bound in src


def test_groups_without_a_lane_form(M):
    u = _u8(M)
    f = M.from_array(np.zeros((200, 300), np.float32), (128, 64))
    nbh = M.convolve(u, scripts.LIFE_MASK)
    for name, expr in [
        ("weight 2", M.convolve(u, [[0, 1, 0], [1, 2, 1], [0, 1, 0]]) == 6),
        ("sum past a byte", M.convolve(u, [[1] * 17] * 17) > 3),  # 289 taps, past 255 even for 0/1 data
        ("float data", scripts.hillshade(M, f)),
        ("negative difference", nbh - u),
        ("product of two sums", nbh * nbh),
        ("division", nbh / 2),
    ]:
        r, src = _lane_r(M, expr)
        assert r == 0, name
        assert "static constexpr unsigned LANE_R = 0u;" in src and "post_lanes" not in src


def _pack(l):
    l = np.asarray(l, np.uint64)
    return (l[..., 0] | l[..., 1] << 8 | l[..., 2] << 16 | l[..., 3] << 24).astype(np.uint64)


def _lanes(v):
    v = np.asarray(v, np.uint64)
    return np.stack([(v >> (8 * i)) & 0xff for i in range(4)], -1)


M32 = np.uint64(0xffffffff)


def lane_eq(x, y):
    return ((~((x ^ y) + np.uint64(0x7f7f7f7f)) & M32) >> np.uint64(7)) & np.uint64(0x01010101)


def lane_ne(x, y):
    return ((((x ^ y) + np.uint64(0x7f7f7f7f)) & M32) >> np.uint64(7)) & np.uint64(0x01010101)


def lane_ge(x, y):
    return (((((x | np.uint64(0x80808080)) - y) & M32)) >> np.uint64(7)) & np.uint64(0x01010101)


def lane_lt(x, y):
    return ((~((x | np.uint64(0x80808080)) - y) & M32) >> np.uint64(7)) & np.uint64(0x01010101)


def lane_sel(b, x):
    return (((b << np.uint64(8)) - b) & M32) & x


def lane_pick(b, x, y):
    m = ((b << np.uint64(8)) - b) & M32
    return (x & m) | (y & (~m & M32))


def test_lane_formulas_match_the_comparisons():
    rng = np.random.Generator(np.random.PCG64(5))
    a, b = np.meshgrid(np.arange(128), np.arange(128), indexing="ij")
    a, b = a.ravel(), b.ravel()
    for pos in range(4):  # the pair in each lane, the other lanes random in [0, 127]
        xs = rng.integers(0, 128, (a.size, 4))
        ys = rng.integers(0, 128, (a.size, 4))
        xs[:, pos], ys[:, pos] = a, b
        x, y = _pack(xs), _pack(ys)
        for f, g in [(lane_eq, np.equal), (lane_ne, np.not_equal), (lane_ge, np.greater_equal), (lane_lt, np.less)]:
            assert np.array_equal(_lanes(f(x, y)), g(xs, ys).astype(np.uint64)), f.__name__
    bs = rng.integers(0, 2, (100000, 4))
    xs = rng.integers(0, 256, (100000, 4))
    ys = rng.integers(0, 256, (100000, 4))
    assert np.array_equal(_lanes(lane_sel(_pack(bs), _pack(xs))), np.where(bs == 1, xs, 0))
    assert np.array_equal(_lanes(lane_pick(_pack(bs), _pack(xs), _pack(ys))), np.where(bs == 1, xs, ys))


def test_lane_sums_do_not_carry():
    """Eight neighbour lanes of values in [0, 31] add as eight byte sums: every lane
    sum is at most 248, so no lane carries into the next."""
    rng = np.random.Generator(np.random.PCG64(6))
    taps = rng.integers(0, 32, (8, 50000, 4))
    total = sum(_pack(t) for t in taps) & M32
    assert np.array_equal(_lanes(total), taps.sum(0))
