"""Map-algebra scripts of the benchmark workloads, written once against a
module M with the map.py API: the product front end (map_amd.map) or any
evaluator with the same API. They follow the reference's example scripts
(python/hill.py, python/life.py, python/wsum.py)."""

PI = 3.141593  # python/hill.py:14


def hori(M, dem, dist=1):
    return M.convolve(dem, [[-1, 0, 1], [-2, 0, 2], [-1, 0, 1]]) / (8 * dist)


def vert(M, dem, dist=1):
    return M.convolve(dem, [[-1, -2, -1], [0, 0, 0], [1, 2, 1]]) / (8 * dist)


def slope(M, dem, zf=1, dist=1):
    x = hori(M, dem, dist)
    y = vert(M, dem, dist)
    return M.atan(zf * M.sqrt(x * x + y * y))


def aspect(M, dem, dist=1):
    x = hori(M, dem, dist)
    y = vert(M, dem, dist)
    z1 = (x != 0) * M.atan2(y, -x)
    z1 = z1 + (z1 < 0) * (PI * 2)
    z0 = (x == 0) * ((y > 0) * (PI / 2) + (y < 0) * (PI * 2 - PI / 2))
    return z1 + z0


def hillshade_terms(M, dem, zenith=45, azimuth=315):
    """The two terms of the last combination step (for the cancellation bound)."""
    zr = (90 - float(zenith)) / 180 * PI
    ar = 360 - float(azimuth) + 90
    ar = ar - 360 if (ar > 360) else ar
    ar = ar / 180 * PI
    t1 = M.cos(zr) * M.cos(slope(M, dem))
    t2 = M.sin(zr) * M.sin(slope(M, dem)) * M.cos(ar - aspect(M, dem))
    return t1, t2


def hillshade(M, dem, zenith=45, azimuth=315):
    t1, t2 = hillshade_terms(M, dem, zenith, azimuth)
    return t1 + t2


LIFE_MASK = [[1, 1, 1], [1, 0, 1], [1, 1, 1]]


def life(M, state, n):
    for _ in range(n):
        nbh = M.convolve(state, LIFE_MASK)
        state = (nbh == 3) + (nbh == 2) * state
    return state


def wsum(M, r1, r2, r3, r4):
    return 0 + 0.1 * r1 + r2 * 0.2 + 0.3 * r3 + r4 * 0.4 + 0  # python/wsum.py:11


def batch_tile(M, dem, zones, weights, nz=16, thr=0.5):
    """Per-tile pipeline of the batched workload: a 5x5 weighted focal sum,
    the slope of it, a threshold mask, and per-zone statistics of the mask
    and of the slope."""
    s = M.convolve(dem, weights)
    sl = slope(M, s)
    mask = sl > thr
    return mask, M.zstats(mask, zones, nz), M.zstats(sl, zones, nz)


def focal_mean7(M, dem):
    return M.convolve(dem, [[1] * 7 for _ in range(7)]) / 49


def serve_request(M, kind, dem, zones=None):
    """Targets of one request of the served mix."""
    if kind == "hillshade":
        return (hillshade(M, dem),)
    if kind == "slope_aspect":
        return (slope(M, dem), aspect(M, dem))
    if kind == "mean7":
        return (focal_mean7(M, dem),)
    if kind == "zstats":
        return (M.zstats(dem, zones, 64),)
    raise ValueError(kind)


def batch_tile_sums(M, dem, zones, weights, nz=16, thr=0.5):
    """The batched workload's outputs: per-zone statistics of the threshold
    mask of the slope of a 5x5 weighted focal sum (the mask is not kept)."""
    s = M.convolve(dem, weights)
    return M.zstats(slope(M, s) > thr, zones, nz)


def hillshade_zonal(M, dem, zones, nz=256):
    """Per-zone statistics (sum, min, max, count, mean) of the hillshade."""
    return M.zstats(hillshade(M, dem), zones, nz)


# ---- seeded synthetic inputs --------------------------------------------------------

def dem_gaussians(rng, w, h, n=64, ref_side=4096.0):
    """(cx, cy, sigma, amplitude) rows: centres uniform in the raster, sigma
    U[50, 800] px scaled by max(w, h) / ref_side, amplitude U[0, 3000] m."""
    import numpy as np
    s = max(w, h) / ref_side
    return np.stack([rng.uniform(0, w, n), rng.uniform(0, h, n), rng.uniform(50, 800, n) * s,
                     rng.uniform(0, 3000, n)], 1)


def life_input(side, seed, p=0.3):
    """side x side U8 cells, Bernoulli(p) from rng.random() < p (numpy PCG64, rows in order)."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.empty((side, side), np.uint8)
    rows = max(1, (1 << 24) // side)
    for r in range(0, side, rows):
        n = min(rows, side - r)
        out[r:r + n] = rng.random((n, side)) < p
    return out


# D8 codes of the eight neighbours at offset (dx, dy), y growing south (ESRI order:
# 1 E, 2 SE, 4 S, 8 SW, 16 W, 32 NW, 64 N, 128 NE), the codes the reference's spreading
# kernel and flowAccuGather read (runtime/skeleton/util.cpp:628, front/Raster.cpp:586-589).
D8_CODES = [((1, 0), 1), ((1, 1), 2), ((0, 1), 4), ((-1, 1), 8), ((-1, 0), 16), ((-1, -1), 32),
            ((0, -1), 64), ((1, -1), 128)]


def d8_directions(dem):
    """Steepest-descent D8 flow directions of a DEM (rows = y) as a U8 raster: the code of the
    neighbour with the largest drop per unit distance (the first in D8_CODES order on ties),
    0 where no neighbour inside the raster is lower (pits and flats)."""
    import numpy as np
    z = np.asarray(dem, np.float64)
    h, w = z.shape
    best = np.zeros((h, w))
    code = np.zeros((h, w), np.uint8)
    pad = np.pad(z, 1, constant_values=np.inf)
    for (dx, dy), c in D8_CODES:
        nb = pad[1 + dy:1 + dy + h, 1 + dx:1 + dx + w]
        drop = (z - nb) / (np.sqrt(2.0) if dx and dy else 1.0)
        take = drop > best
        best = np.where(take, drop, best)
        code = np.where(take, np.uint8(c), code)
    return code


def flow_gather(M, water, dirs):
    """One gathering step written with neighbour and conditional ops (the loop body of the
    reference's flowAccuGather, front/Raster.cpp:595-602): the sum of the neighbours' water
    whose direction points at each cell."""
    inflow = [[2, 4, 8], [1, 0, 16], [128, 64, 32]]
    new = M.zeros_like(water)
    for y in (-1, 0, 1):
        for x in (-1, 0, 1):
            if x == 0 and y == 0:
                continue
            new = M.con(dirs([x, y]) == inflow[y + 1][x + 1], new + water([x, y]), new)
    return new


def d8_script(M, dem):
    """Steepest-descent D8 directions as a map-algebra script (one fused focal group): the code
    of the neighbour with the largest drop per unit distance, the first in D8_CODES order on
    ties, 0 where no neighbour is lower; neighbours outside the raster follow the focal border
    rule (d8_directions treats them as higher instead)."""
    best = M.zeros_like(dem)
    code = M.zeros_like(dem, M.S32)
    for (dx, dy), c in D8_CODES:
        drop = (dem - dem([dx, dy])) / (1.4142135 if dx and dy else 1.0)
        take = drop > best
        best = M.con(take, drop, best)
        code = M.con(take, c, code)
    return M.astype(code, M.U8)
