"""Map-algebra scripts written once against a module M: the product front end
(map_amd.map) or the oracle evaluator (oracle.ref). They follow the
reference's example scripts (python/hill.py, python/life.py)."""

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
