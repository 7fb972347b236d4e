"""Host emulation of the spreading passes of map_amd/csrc/spread.hip (k_tile, k_walk), run
cell by cell in Python on small rasters with small tiles, against the oracle: checks the
closed/open split, the counts assembled from both sides of tile edges (contributions in a
shuffled order, a source listed by the contribution that leaves zero) and the walk hand-off
(every cell computed once, the fixed summation order) without a GPU. Walks are run one after
another in a shuffled order, the schedule the atomics allow."""
import numpy as np
import pytest

from map_amd import workloads as scripts

INFLOW = [[2, 4, 8], [1, 0, 16], [128, 64, 32]]
OUT = {INFLOW[i1 + 1][i0 + 1]: (-i0, -i1) for i1 in (-1, 0, 1) for i0 in (-1, 0, 1) if i0 or i1}
CLOSED, SRC = 0x200, 0x100


def emulate(v, d, ts, seed):
    h, w = d.shape
    T = v.dtype.type
    out = np.zeros_like(v)
    cnt = np.zeros((h, w), np.int64)

    def inr(x, y):
        return 0 <= x < w and 0 <= y < h

    def down(x, y):
        o = OUT.get(int(d[y, x]))
        return None if o is None or not inr(x + o[0], y + o[1]) else (x + o[0], y + o[1])

    contrib = []
    # k_tile
    for ty0 in range(0, h, ts):
        for tx0 in range(0, w, ts):
            cells = [(x, y) for y in range(ty0, min(ty0 + ts, h)) for x in range(tx0, min(tx0 + ts, w))]
            inside = lambda x, y: tx0 <= x < tx0 + ts and ty0 <= y < ty0 + ts  # noqa: E731
            left, opn, done, tot = {}, {}, {}, {}
            for (x, y) in cells:
                kin, ext = 0, False
                for i1 in (-1, 0, 1):
                    for i0 in (-1, 0, 1):
                        u = (x + i0, y + i1)
                        if (i0 or i1) and inr(*u) and d[u[1], u[0]] == INFLOW[i1 + 1][i0 + 1]:
                            if inside(*u):
                                kin += 1
                            else:
                                ext = True
                left[(x, y)], opn[(x, y)], done[(x, y)], tot[(x, y)] = kin, ext, False, T(v[y, x])
            # walks from the tile's sources (fixed before any count falls), in a shuffled order;
            # the walk that takes a downstream count to zero goes on
            tsrcs = [c for c in cells if left[c] == 0]
            np.random.default_rng(seed + tx0 + 7 * ty0).shuffle(tsrcs)
            for c in tsrcs:
                x, y = c
                while True:
                    assert not done[(x, y)]
                    if not opn[(x, y)]:
                        acc = tot[(x, y)]
                        for i1 in (-1, 0, 1):
                            for i0 in (-1, 0, 1):
                                u = (x + i0, y + i1)
                                if (i0 or i1) and inr(*u) and d[u[1], u[0]] == INFLOW[i1 + 1][i0 + 1]:
                                    assert inside(*u) and done[u] and not opn[u]
                                    acc = T(acc + tot[u])
                        tot[(x, y)] = acc
                    done[(x, y)] = True
                    dn = down(x, y)
                    if dn is None or not inside(*dn):
                        break
                    opn[dn] = opn[dn] or opn[(x, y)]
                    left[dn] -= 1
                    if left[dn] != 0:
                        break
                    x, y = dn
            for c in cells:
                closed = done[c] and not opn[c]
                out[c[1], c[0]] = tot[c] if closed and down(*c) is not None else T(0)
                if closed:
                    dn = down(*c)
                    if dn is not None and not inside(*dn):
                        contrib.append((dn, -1))
                else:
                    k = 0
                    for i1 in (-1, 0, 1):
                        for i0 in (-1, 0, 1):
                            u = (c[0] + i0, c[1] + i1)
                            if (i0 or i1) and inr(*u) and d[u[1], u[0]] == INFLOW[i1 + 1][i0 + 1]:
                                k += 1 if (not inside(*u) or not done[u] or opn[u]) else 0
                    contrib.append((c, k))
    # the atomics of all tiles in any order: the one that leaves zero lists a source
    np.random.default_rng(seed + 1).shuffle(contrib)
    srcs = []
    for (x, y), k in contrib:
        cnt[y, x] += k
        if cnt[y, x] == 0:
            srcs.append((x, y))
    # k_walk, sources in a shuffled order
    np.random.default_rng(seed).shuffle(srcs)
    computed = set()
    for (x, y) in srcs:
        while True:
            assert (x, y) not in computed
            computed.add((x, y))
            acc = T(v[y, x])
            for i1 in (-1, 0, 1):
                for i0 in (-1, 0, 1):
                    u = (x + i0, y + i1)
                    if (i0 or i1) and inr(*u) and d[u[1], u[0]] == INFLOW[i1 + 1][i0 + 1]:
                        acc = T(acc + out[u[1], u[0]])
            dn = down(x, y)
            if dn is None:
                break
            out[y, x] = acc
            cnt[dn[1], dn[0]] -= 1
            if cnt[dn[1], dn[0]] != 0:
                break
            x, y = dn
    return out


@pytest.mark.parametrize("ts,seed", [(4, 0), (8, 1), (16, 2), (5, 3)])
def test_three_passes_match_oracle(O, ts, seed):
    rng = np.random.default_rng(seed)
    h, w = 37, 45
    codes = np.array([0, 1, 2, 4, 8, 16, 32, 64, 128, 3], np.uint8)
    for d in (scripts.d8_directions(rng.normal(size=(h, w)).cumsum(0).cumsum(1)), rng.choice(codes, (h, w))):
        for v in (rng.integers(0, 50, (h, w)).astype(np.int32), rng.uniform(0, 1, (h, w)).astype(np.float32)):
            got = emulate(v, d, ts, seed)
            want = O.spread_scan(O.from_array(v, (16, 16)), O.from_array(d, (16, 16))).val
            assert np.array_equal(got, want), (ts, v.dtype)


def _hood(BS0, BS1, NB0, DS0, DS1, x, y):
    """spread.hip hood(): neighbour element indices from one division pair per axis and carries."""
    by, bx = y // BS1, x // BS0
    ly, lx = y - by * BS1, x - bx * BS0
    blk, row = BS1 * BS0, NB0 * BS1 * BS0
    r1 = by * row + ly * BS0
    r = [r1 - BS0 if ly > 0 else r1 - row + (BS1 - 1) * BS0, r1,
         r1 + BS0 if ly + 1 < BS1 else r1 + row - (BS1 - 1) * BS0]
    c1 = bx * blk + lx
    c = [c1 - 1 if lx > 0 else c1 - blk + (BS0 - 1), c1, c1 + 1 if lx + 1 < BS0 else c1 + blk - (BS0 - 1)]
    out = []
    for k in range(9):
        u0, u1 = x + k % 3 - 1, y + k // 3 - 1
        inr = 0 <= u0 < DS0 and 0 <= u1 < DS1
        out.append((inr, r[k // 3] + c[k % 3] if inr else r1 + c1))
    return out


@pytest.mark.parametrize("ds,bs", [((70, 50), (16, 8)), ((33, 47), (32, 16)), ((64, 64), (64, 64)), ((9, 5), (4, 2))])
def test_hood_indices_are_block_tiled_indices(ds, bs):
    """Every neighbour index hood() forms is the block-tiled index of that cell, and cell_of()
    inverts the index of every cell (the padding of edge blocks is never reached)."""
    DS0, DS1 = ds
    BS0, BS1 = bs
    NB0, NB1 = -(-DS0 // BS0), -(-DS1 // BS1)

    def at(x, y):
        bx, by = x // BS0, y // BS1
        return ((by * NB0 + bx) * BS1 + (y - by * BS1)) * BS0 + (x - bx * BS0)

    n = NB0 * NB1 * BS0 * BS1
    for y in range(DS1):
        for x in range(DS0):
            e = at(x, y)
            b, r = e // (BS0 * BS1), e % (BS0 * BS1)
            assert ((b % NB0) * BS0 + r % BS0, (b // NB0) * BS1 + r // BS0) == (x, y)
            for k, (inr, idx) in enumerate(_hood(BS0, BS1, NB0, DS0, DS1, x, y)):
                u0, u1 = x + k % 3 - 1, y + k // 3 - 1
                assert 0 <= idx < n
                assert idx == (at(u0, u1) if inr else e)


@pytest.mark.parametrize("TS", [32, 64, 16])
def test_tile_cell_mapping_spreads_rows_and_columns(TS):
    """spread.hip tile_cell(): the TS*TS cells of a tile go to 256 threads in equal shares,
    and every row and every column of the tile meets TS different threads."""
    NT = 256
    owner = {}
    for t in range(NT):
        for j in range(TS * TS // NT):
            ly = t // TS + (NT // TS) * j
            i = ly * TS + ((t % TS - ly) & (TS - 1))
            assert i not in owner
            owner[i] = t
    assert len(owner) == TS * TS
    for r in range(TS):
        assert len({owner[r * TS + c] for c in range(TS)}) == TS
        assert len({owner[c * TS + r] for c in range(TS)}) == TS
