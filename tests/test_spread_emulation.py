"""Host emulation of the three spreading passes of map_amd/csrc/spread.hip (k_tile,
k_count, k_walk), run cell by cell in Python on small rasters with small tiles, against
the oracle: checks the closed/open split, the counts and the walk hand-off (every cell
computed once, the fixed summation order) without a GPU. Walks are run one after
another in a shuffled order, the schedule the atomics allow."""
import numpy as np
import pytest

from tests import scripts

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
            while True:
                ready = [c for c in cells if not done[c] and left[c] == 0]
                if not ready:
                    break
                for (x, y) in ready:
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
                    if dn is not None and inside(*dn):
                        opn[dn] = opn[dn] or opn[(x, y)]
                        left[dn] -= 1
            for c in cells:
                closed = done[c] and not opn[c]
                out[c[1], c[0]] = tot[c] if closed and down(*c) is not None else T(0)
                cnt[c[1], c[0]] = CLOSED if closed else 0
    # k_count
    new = cnt.copy()
    for y in range(h):
        for x in range(w):
            if cnt[y, x] == CLOSED:
                continue
            c = sum(1 for i1 in (-1, 0, 1) for i0 in (-1, 0, 1)
                    if (i0 or i1) and inr(x + i0, y + i1) and d[y + i1, x + i0] == INFLOW[i1 + 1][i0 + 1]
                    and cnt[y + i1, x + i0] != CLOSED)
            new[y, x] = SRC if c == 0 else c
    cnt = new
    # k_walk, sources in a shuffled order
    srcs = [(x, y) for y in range(h) for x in range(w) if cnt[y, x] == SRC]
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
