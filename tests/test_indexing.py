"""Host-side check of the focal template's global indexing.

Emulates, for small raster geometries, every global access the focal kernel
body makes (device/mapx_dev.hiph: tiles of each block, body vectors through
row_src from clamped coordinates, vectors outside the central block's
columns through ldv's border rule inside [-H, E+H), per-thread
cell vectors and output stores)
and checks that each lands inside the raster allocation and inside the block
it claims. Geometries include tiles wider than blocks and partial edge
blocks (the cases that would read out of bounds without the halo bound).
"""
import itertools

import pytest


def invert(i, d):
    return -i - 1 if i < 0 else (d - (i - d + 1) if i >= d else i)


def modulo(i, d):
    return i + d if i < 0 else (i - d if i >= d else i)


def ext(ds, bs, b, i):
    return min(bs[i], ds[i] - b * bs[i])


def ldf_index(ds, bs, nb, bx, by, x, y):
    """mx::ldf: the rule applied from each block reached until the lookup is inside one
    (64 steps at most, then clamped)."""
    e0, e1 = ext(ds, bs, bx, 0), ext(ds, bs, by, 1)
    for _ in range(64):
        i0 = 0 if x < 0 else (2 if x >= e0 else 1)
        i1 = 0 if y < 0 else (2 if y >= e1 else 1)
        if i0 == 1 and i1 == 1:
            break
        nbx, nby = bx + i0 - 1, by + i1 - 1
        if nbx < 0 or nby < 0 or nbx >= nb[0] or nby >= nb[1]:
            x, y = invert(x, e0), invert(y, e1)
        else:
            x, y = modulo(x, bs[0]), modulo(y, bs[1])
            bx, by = nbx, nby
            e0, e1 = ext(ds, bs, bx, 0), ext(ds, bs, by, 1)
    else:
        raise AssertionError("border rule did not end")
    return by * nb[0] + bx, x, y


def row_src(ds, bs, nb, bx, by, e1, y):
    """(element offset, ok) of halo row y as mx::row_src computes them: k.off + row * BS0 with
    row = s*y + d (clamped to 0 for a mirrored row past a narrow last block), and whether the
    row lies inside its block's valid rows."""
    has_up, has_dn = by > 0, by + 1 < nb[1]
    du = bs[1] * (1 - nb[0]) if has_up else -1
    dd = bs[1] * (nb[0] - 1) if has_dn else 2 * e1 - 1
    up, dn = y < 0, y >= e1
    flip = (up and not has_up) or (dn and not has_dn)
    d = du if up else (dd if dn else 0)
    row = (-y if flip else y) + d
    if dn and not has_dn and row < 0:
        row = 0
    koff = (by * nb[0] + bx) * bs[0] * bs[1]
    if up:
        ok = has_up or -y - 1 < e1
    elif dn:
        ok = (y - bs[1] < ext(ds, bs, by + 1, 1)) if has_dn else (2 * e1 - 1 - y >= 0)
    else:
        ok = True
    return koff + row * bs[0], ok


def tile_for(bs, tx, ty, vx):
    """Mirror of Runtime::tile_for's clamping."""
    while bs[0] % vx:
        vx //= 2
        tx //= 2
    while tx > bs[0] and tx // vx > 16:
        tx //= 2
        ty *= 2
    while tx > bs[0] and vx > 1:
        tx //= 2
        vx //= 2
    while ty > bs[1] and ty > 256 // (tx // vx):
        ty //= 2
    return tx, ty, vx


def ldv_reads(ds, bs, nb, bx, by, x, y, vx, h0):
    """Global (block, x, y) reads of mx::ldv for the vector at (x..x+vx-1, y)."""
    e0 = min(bs[0], ds[0] - bx * bs[0])
    e1 = min(bs[1], ds[1] - by * bs[1])
    xl = x + vx - 1
    i0a = 0 if x < 0 else (2 if x >= e0 else 1)
    i0b = 0 if xl < 0 else (2 if xl >= e0 else 1)
    i1 = 0 if y < 0 else (2 if y >= e1 else 1)
    if i0a == i0b:
        nbx, nby = bx + i0a - 1, by + i1 - 1
        if (0 <= nbx < nb[0] and 0 <= nby < nb[1] and modulo(x, bs[0]) + vx <= ext(ds, bs, nbx, 0)
                and modulo(y, bs[1]) < ext(ds, bs, nby, 1)):
            xx, yy = modulo(x, bs[0]), modulo(y, bs[1])
            assert xx % vx == 0, "misaligned vector"
            return [(nby * nb[0] + nbx, xx + v, yy, True) for v in range(vx)]
    out = []
    for v in range(vx):
        xv = x + v
        if -h0 <= xv < e0 + h0:
            blk, sx, sy = ldf_index(ds, bs, nb, bx, by, xv, y)
            out.append((blk, sx, sy, False))
    return out


GEOMS = [((512, 384), (128, 128), 1, 1, (256, 16, 4)),
         ((200, 150), (64, 32), 1, 1, (256, 16, 4)),
         ((300, 220), (128, 64), 1, 1, (256, 16, 4)),
         ((70, 41), (32, 16), 2, 2, (256, 16, 4)),
         ((33, 17), (16, 16), 1, 1, (256, 16, 4)),
         ((40, 40), (40, 40), 3, 3, (256, 16, 4)),
         ((101, 57), (50, 30), 1, 1, (256, 16, 4)),
         ((98, 60), (49, 20), 2, 1, (256, 16, 4)),
         ((2048, 1040), (1024, 1024), 1, 1, (1024, 16, 16)),
         ((300, 220), (128, 64), 1, 1, (512, 16, 16)),
         ((200, 150), (64, 32), 1, 1, (256, 16, 16)),
         # edge blocks narrower than the halo (the rule applied again from the narrow block)
         ((65, 33), (32, 16), 2, 2, (256, 16, 4)),
         ((129, 66), (64, 32), 3, 3, (256, 16, 4)),
         ((33, 17), (16, 16), 3, 2, (256, 16, 4)),
         ((3, 2), (4, 4), 3, 3, (256, 16, 4)),
         ((2, 1), (8, 8), 1, 1, (256, 16, 4)),
         ((1025, 1026), (1024, 1024), 3, 3, (1024, 16, 8)),
         ((258, 130), (128, 64), 2, 2, (512, 16, 16))]


@pytest.mark.parametrize("ds,bs,h0,h1,tile", GEOMS)
def test_focal_reads_in_bounds(ds, bs, h0, h1, tile):
    tx, ty, vx = tile_for(bs, *tile)
    twx = tx // vx
    assert 256 % twx == 0 and ty % (256 // twx) == 0 and bs[0] % vx == 0
    pad = 0 if h0 == 0 else vx * ((h0 + vx - 1) // vx)
    sw, hh = tx + 2 * pad, ty + 2 * h1
    nv = sw // vx
    nb = ((ds[0] + bs[0] - 1) // bs[0], (ds[1] + bs[1] - 1) // bs[1])
    total = nb[0] * nb[1] * bs[0] * bs[1]
    tb = ((bs[0] + tx - 1) // tx, (bs[1] + ty - 1) // ty)
    for b in range(nb[0] * nb[1]):
        bx, by = b % nb[0], b // nb[0]
        e0 = min(bs[0], ds[0] - bx * bs[0])
        e1 = min(bs[1], ds[1] - by * bs[1])
        for t in range(tb[0] * tb[1]):
            x0, y0 = (t % tb[0]) * tx, (t // tb[0]) * ty
            if x0 >= e0 or y0 >= e1:
                continue
            pv = pad // vx
            nbv = tx // vx
            # fetch: every slot loads one vector through row_src from clamped
            # coordinates (rows to e1+h1-1, columns to [0, bs0-vx]), also slots
            # past the tile's vector count (they repeat its last vector)
            fetched = {}
            for kind, count, nper in (("body", hh * nbv, nbv), ("pad", hh * 2 * pv, 2 * pv)):
                if count == 0:
                    continue
                for e in range(0, ((count + 255) // 256) * 256):
                    if (e & ~63) >= count:
                        continue  # the whole wavefront is past the count: no load
                    ec = min(e, count - 1)
                    r, j = ec // nper, ec % nper
                    if kind == "body":
                        x = x0 + j * vx
                    else:
                        x = x0 - pad + j * vx if j < pv else x0 + tx + (j - pv) * vx
                    y = min(y0 - h1 + r, e1 + h1 - 1)
                    xc = max(0, min(x, bs[0] - vx))
                    off = row_src(ds, bs, nb, bx, by, e1, y)[0] + xc
                    assert off % vx == 0 and 0 <= off and off + vx <= total, (b, t, kind, e)
                    blk = off // (bs[0] * bs[1])
                    assert blk == (off + vx - 1) // (bs[0] * bs[1]), "vector crosses a block"
                    fetched[(kind, r, j)] = off
            vecs = [("body", r, j, x0 + j * vx) for r in range(hh) for j in range(nbv)]
            vecs += [("pad", r, j, x0 - pad + j * vx if j < pv else x0 + tx + (j - pv) * vx)
                     for r in range(hh) for j in range(2 * pv)]
            for kind, r, j, x in vecs:
                y = y0 - h1 + r
                if y >= e1 + h1 or x + vx <= -h0 or x >= e0 + h0:
                    continue  # zero
                if 0 <= x and x + vx <= e0 and row_src(ds, bs, nb, bx, by, e1, y)[1]:
                    # the fetched word: the row's block (central, above/below) or mirrored row
                    if 0 <= y < e1:
                        blk, yy = b, y
                    elif 0 <= by + (-1 if y < 0 else 1) < nb[1]:
                        blk, yy = (by + (-1 if y < 0 else 1)) * nb[0] + bx, modulo(y, bs[1])
                    else:
                        blk, yy = b, invert(y, e1)
                    be1 = min(bs[1], ds[1] - (blk // nb[0]) * bs[1])
                    assert 0 <= yy < be1
                    assert fetched[(kind, r, j)] == (blk * bs[1] + yy) * bs[0] + x, (b, t, kind, r, j)
                    for v in range(vx):  # the cells match the per-cell border rule
                        assert ldf_index(ds, bs, nb, bx, by, x + v, y) == (blk, x + v, yy)
                    reads = [(blk, x + v, yy, False) for v in range(vx)]
                else:
                    reads = ldv_reads(ds, bs, nb, bx, by, x, y, vx, h0)
                for rb, rx, ry, _vec in reads:  # every cell read is a valid cell of its block
                    assert 0 <= rx < ext(ds, bs, rb % nb[0], 0) and 0 <= ry < ext(ds, bs, rb // nb[0], 1), \
                        (b, t, kind, r, j, rb, rx, ry)
            # LDS window reads of every thread stay inside the halo tile
            for tid in range(256):
                ly, lx = tid // twx, (tid % twx) * vx
                for ry in range(ty // (256 // twx)):
                    base = (ly + ry * (256 // twx)) * sw + lx
                    assert base + 2 * h1 * sw + vx + 2 * pad - 1 < hh * sw
            # per-thread cell vectors (loads at the clamped row, stores of the valid cells)
            for xv, ly in itertools.product(range(x0, x0 + tx, vx), range(ty)):
                y = y0 + ly
                yc = min(y, e1 - 1)
                xl = xv if xv < e0 else ((e0 - 1) // vx) * vx
                assert xl + vx <= bs[0] and 0 <= (b * bs[1] + yc) * bs[0] + xl + vx <= total
                nvalid = (vx if e0 - xv >= vx else max(0, e0 - xv)) if y < e1 else 0
                if nvalid:
                    assert (b * bs[1] + y) * bs[0] + xv + nvalid <= total and y < bs[1]


def test_border_rule_matches_oracle():
    """The emulated rule and the oracle's C restatement agree on every lookup, also past edge
    blocks narrower than the halo, and every lookup ends inside a block's valid extent."""
    from oracle import ref as O
    for ds, bs, h in [((70, 41), (32, 16), 1), ((64, 64), (16, 16), 1), ((9, 7), (4, 4), 1), ((65, 33), (32, 16), 3),
                      ((9, 7), (4, 4), 4), ((3, 2), (4, 4), 4), ((1, 1), (8, 8), 3), ((17, 18), (8, 8), 7)]:
        nb = ((ds[0] + bs[0] - 1) // bs[0], (ds[1] + bs[1] - 1) // bs[1])
        for b in range(nb[0] * nb[1]):
            bx, by = b % nb[0], b // nb[0]
            e0 = min(bs[0], ds[0] - bx * bs[0])
            e1 = min(bs[1], ds[1] - by * bs[1])
            for x in range(-h, e0 + h):
                for y in range(-h, e1 + h):
                    blk, sx, sy = ldf_index(ds, bs, nb, bx, by, x, y)
                    assert 0 <= sx < ext(ds, bs, blk % nb[0], 0) and 0 <= sy < ext(ds, bs, blk // nb[0], 1)
                    gx = (blk % nb[0]) * bs[0] + sx
                    gy = (blk // nb[0]) * bs[1] + sy
                    assert (gx, gy) == O.src_cell(ds, bs, bx, by, x, y)


# ---- chunk split, partial slots and task lookup (device/mapx_dev.hiph run_chunks) ----

def make_walk(nwork, G, g):
    if G % 8 == 0 and G >= 8:
        x, r, g8 = g & 7, g >> 3, G >> 3
        c = (nwork + 7) >> 3
        b, e = x * c, min(x * c + c, nwork)
        return range(b + r, max(e, b + r), g8)
    return range(g, nwork, G)


def chunk_tiles(L, cl):
    """Tiles (block, tile) of chunk cl of a launch, and its partial slot (run_chunks)."""
    b = L["blk0"] + cl // L["CPB"]
    band = cl % L["CPB"]
    tiles = L["TB0"] * L["TB1"]
    t0, t1 = band * L["CH"], min(band * L["CH"] + L["CH"], tiles)
    return [(b, t) for t in range(t0, t1)], L["part0"] + cl


def task_of(starts, c, t):
    if starts[t + 1] > c:
        return t
    lo, hi = t + 1, len(starts) - 2
    while lo < hi:
        m = (lo + hi + 1) >> 1
        if starts[m] <= c:
            lo = m
        else:
            hi = m - 1
    return lo


@pytest.mark.parametrize("nb,tb,ch", [((4, 3), (4, 64), 16), ((1, 1), (1, 1), 1), ((5, 2), (3, 7), 4),
                                      ((2, 2), (4, 16), 64), ((3, 3), (2, 5), 1)])
def test_chunks_cover_tiles_once_and_slots_match_forms(nb, tb, ch):
    tiles = tb[0] * tb[1]
    ch = min(ch, tiles)
    cpb = (tiles + ch - 1) // ch
    nblk = nb[0] * nb[1]
    L = dict(TB0=tb[0], TB1=tb[1], CH=ch, CPB=cpb, blk0=0, part0=0)
    batched = {}
    for G in (8, 16, 24, 1, 3):  # static split: every chunk exactly once for any grid
        seen = []
        for g in range(G):
            seen += list(make_walk(nblk * cpb, G, g))
        assert sorted(seen) == list(range(nblk * cpb)), G
    for cl in range(nblk * cpb):
        ts, slot = chunk_tiles(L, cl)
        for bt in ts:
            assert bt not in batched
            batched[bt] = slot
    assert len(batched) == nblk * tiles
    # plain form: one launch per block, chunks land in the same slots
    for b in range(nblk):
        Lp = dict(L, blk0=b, part0=b * cpb)
        for cl in range(cpb):
            ts, slot = chunk_tiles(Lp, cl)
            for bt in ts:
                assert batched[bt] == slot


def test_task_lookup_in_batched_launch():
    rng = __import__("numpy").random.default_rng(0)
    sizes = [int(s) for s in rng.integers(0, 9, 300)]
    sizes = [s for s in sizes if s > 0]
    starts = [0]
    for s in sizes:
        starts.append(starts[-1] + s)
    total = starts[-1]
    want = [i for i, s in enumerate(sizes) for _ in range(s)]
    for G in (8, 64, 3):
        for g in range(G):
            t = 0
            for c in make_walk(total, G, g):
                t = task_of(starts, c, t)
                assert t == want[c]
    t = 0
    for c in sorted(rng.choice(total, 200, replace=False)):  # queue order: increasing per work-group
        t = task_of(starts, int(c), t)
        assert t == want[c]


def span_pad(h0, size):
    unit = 1 if size >= 4 else 4 // size
    return unit * ((h0 + unit - 1) // unit)


@pytest.mark.parametrize("ds,bs,h0,h1,tile,size", [((512, 384), (128, 128), 1, 1, (512, 16, 8), 1),
                                                   ((200, 150), (64, 32), 1, 1, (512, 16, 8), 1),
                                                   ((300, 220), (128, 64), 1, 1, (128, 32, 4), 4),
                                                   ((70, 41), (32, 16), 2, 2, (128, 32, 4), 4),
                                                   ((101, 57), (52, 27), 2, 2, (256, 16, 4), 4),
                                                   ((98, 60), (48, 20), 1, 1, (256, 32, 8), 1),
                                                   ((65, 33), (32, 16), 2, 2, (128, 32, 4), 4),
                                                   ((130, 67), (64, 32), 3, 3, (512, 16, 8), 1),
                                                   ((3, 2), (8, 8), 3, 3, (128, 32, 4), 4),
                                                   ((1025, 1026), (1024, 1024), 3, 3, (1024, 16, 8), 1)])
def test_stream_reads_in_bounds(ds, bs, h0, h1, tile, size):
    """Emulates the global reads of stream_tile: every row a thread fetches
    (central rows directly, others through the row source or zeros) and its
    span (one dword run inside the central block's columns, else own cells +
    border-rule halo cells) stay inside the raster and on valid cells."""
    tx_, ty_, vx = tile_for(bs, *tile)
    a = span_pad(h0, size)
    if (vx * size) % 4:
        return  # the clamped tile is not eligible for the streaming form (codegen stages it in LDS)
    assert (a * size) % 4 == 0
    twx = tx_ // vx
    twy = 256 // twx
    ry_n = ty_ // twy
    nb = ((ds[0] + bs[0] - 1) // bs[0], (ds[1] + bs[1] - 1) // bs[1])
    total = nb[0] * nb[1] * bs[0] * bs[1]
    tb = ((bs[0] + tx_ - 1) // tx_, (bs[1] + ty_ - 1) // ty_)
    for b in range(nb[0] * nb[1]):
        bx, by = b % nb[0], b // nb[0]
        e0 = min(bs[0], ds[0] - bx * bs[0])
        e1 = min(bs[1], ds[1] - by * bs[1])
        for t in range(tb[0] * tb[1]):
            x0, y0 = (t % tb[0]) * tx_, (t // tb[0]) * ty_
            if x0 >= e0 or y0 >= e1:
                continue
            for tid in range(256):
                x = x0 + (tid % twx) * vx
                ys = y0 + (tid // twx) * ry_n
                if x >= e0:
                    continue  # fetches nothing (zeros), still runs the loop
                rows = list(range(ys - h1, ys + ry_n + h1))
                for y in rows:
                    if y >= e1 + h1:
                        continue  # zeros
                    off, ok = (((by * nb[0] + bx) * bs[1] + y) * bs[0], True) if 0 <= y < e1 else \
                        row_src(ds, bs, nb, bx, by, e1, y)
                    assert 0 <= off and off + bs[0] <= total  # the row source stays in the allocation
                    blk, yy = off // (bs[0] * bs[1]), (off // bs[0]) % bs[1]
                    if ok:
                        be1 = min(bs[1], ds[1] - (blk // nb[0]) * bs[1])
                        assert 0 <= yy < be1
                    if x - a >= 0 and x + vx + a <= e0 and ok:
                        cells = [(blk, c, yy) for c in range(x - a, x + vx + a)]
                        for c in range(x - a, x + vx + a):
                            assert ldf_index(ds, bs, nb, bx, by, c, y) == (blk, c, yy)
                    else:
                        cells = []
                        for xc in range(x - a, x + vx + a):
                            if 0 <= xc < e0 and ok:
                                cells.append((blk, xc, yy))
                            elif -h0 <= xc < e0 + h0:
                                cells.append(ldf_index(ds, bs, nb, bx, by, xc, y))
                    for cb, cx, cy in cells:
                        be0 = min(bs[0], ds[0] - (cb % nb[0]) * bs[0])
                        cbe1 = min(bs[1], ds[1] - (cb // nb[0]) * bs[1])
                        assert 0 <= cx < be0 and 0 <= cy < cbe1, (b, x, y, cx, cy)
                        assert 0 <= (cb * bs[1] + cy) * bs[0] + cx < total


# ---- radial scan (map_amd/csrc/radial.hip) ---------------------------------------------

RADIAL_DIRS = [(0, -1, 1, 0), (1, 0, 0, -1), (1, 0, 0, 1), (0, 1, 1, 0),
               (0, 1, -1, 0), (-1, 0, 0, 1), (-1, 0, 0, -1), (0, -1, -1, 0)]


def _radial_launches(ds, s, K, NT=256):
    """The host loop of radial::run: per band, the octants' work-group counts."""
    HS = NT - K
    octs = []
    for fx, fy, sx, sy in RADIAL_DIRS:
        L0 = (ds[0] - 1 - s[0] if fx > 0 else s[0]) if fx else (ds[1] - 1 - s[1] if fy > 0 else s[1])
        L1 = (ds[0] - 1 - s[0] if sx > 0 else s[0]) if sx else (ds[1] - 1 - s[1] if sy > 0 else s[1])
        octs.append((fx, fy, sx, sy, L0, L1))
    maxL0 = max(o[4] for o in octs)
    for c0 in range(0, maxL0 + 1, K):
        wgs = []
        for o, (fx, fy, sx, sy, L0, L1) in enumerate(octs):
            if c0 > L0:
                continue
            rmax = min(c0 + K - 1, L1)
            for j in range((rmax + 1 + HS - 1) // HS):
                wgs.append((o, j))
        if wgs:
            yield c0, octs, wgs


@pytest.mark.parametrize("ds,bs,s,K", [((300, 220), (128, 64), (150, 73), 64), ((700, 90), (256, 32), (0, 0), 64),
                                      ((257, 129), (64, 64), (256, 128), 32), ((33, 17), (16, 16), (16, 0), 64),
                                      ((520, 390), (128, 128), (200, 150), 64), ((5, 300), (4, 64), (2, 299), 32)])
def test_radial_reads_and_writes_in_bounds(ds, bs, s, K):
    """Emulates every global access of k_radial: loads of the band's inputs, the read of the
    previous column, the owners' stores; all land in the allocation at the intended cell, the
    previous column was stored by an earlier band, and the stores cover every cell."""
    NT = 256
    HS = NT - K
    nb = ((ds[0] + bs[0] - 1) // bs[0], (ds[1] + bs[1] - 1) // bs[1])
    total = nb[0] * nb[1] * bs[0] * bs[1]

    def off(x, y):
        assert 0 <= x < ds[0] and 0 <= y < ds[1], (x, y)
        bx, by = x // bs[0], y // bs[1]
        o = ((by * nb[0] + bx) * bs[1] + y % bs[1]) * bs[0] + x % bs[0]
        assert 0 <= o < total
        return o

    stored = set()
    for c0, octs, wgs in _radial_launches(ds, s, K, NT):
        band = set()
        for o, j in wgs:
            fx, fy, sx, sy, L0, L1 = octs[o]
            cend = min(c0 + K, L0 + 1)
            for t in range(NT):
                r = j * HS - K + t
                if not (0 <= r <= L1):
                    continue
                if c0 > 0 and r <= c0 - 1:  # previous column, stored by an earlier band
                    assert off(s[0] + (c0 - 1) * fx + r * sx, s[1] + (c0 - 1) * fy + r * sy) in stored
                for c in range(c0, cend):
                    if r <= c:
                        a = off(s[0] + c * fx + r * sx, s[1] + c * fy + r * sy)
                        if t >= K:
                            band.add(a)
        stored |= band
    cells = {off(x, y) for y in range(ds[1]) for x in range(ds[0])}
    assert stored == cells
