"""Front end, DAG typing, value numbering, fusion grouping and kernel
generation (compiled with hiprtc for gfx950; no GPU needed)."""
import numpy as np
import pytest

from map_amd import workloads as scripts


def dem(M, w=64, h=48, bs=(32, 16)):
    return M.from_array(np.zeros((h, w), np.float32), bs)


def test_constant_typing(M):
    r = dem(M)
    assert (r * 2).datatype() == M.F32
    assert (r * 2.5).datatype() == M.F32
    u = M.from_array(np.zeros((4, 4), np.uint8), (4, 4))
    assert (u + 1).datatype() == M.S32          # promote(U8, S32)
    assert (u == 1).datatype() == M.U8           # relational -> U8
    assert (u * u).datatype() == M.U8
    assert M.convolve(u, scripts.LIFE_MASK).datatype() == M.S32
    assert M.convolve(r, scripts.LIFE_MASK).datatype() == M.F32
    assert M.zsum(u).numdim() == M.D0 and M.zsum(u).datatype() == M.U8
    assert M.sqrt(r).datatype() == M.F32
    assert u.no().datatype() == M.U8


def test_value_numbering(M):
    r = dem(M)
    a = M.convolve(r, [[1, 2, 1], [0, 0, 0], [-1, -2, -1]])
    b = M.convolve(r, [[1, 2, 1], [0, 0, 0], [-1, -2, -1]])
    assert a.nodeid() == b.nodeid()
    assert (a + 1).nodeid() == (a + 1).nodeid()
    assert (a + 1).nodeid() != (a + 1.0).nodeid()


def test_front_end_constant_and_mask_reuse_keeps_distinct_values_apart(M):
    """Scalar operands and masks are reused by the front end per value; values that compare
    equal but differ in type or sign (1 / 1.0 / True, 0.0 / -0.0) stay distinct nodes."""
    r = dem(M)
    ids = lambda *xs: [x._node.value for x in xs]
    a, b, c = r * 1, r * 1.0, r * True
    assert len(set(ids(a, b, c))) == 3
    assert (r * 1.0)._node.value == b._node.value
    assert (r * 0.0)._node.value != (r * -0.0)._node.value
    assert (r * -0.0).datatype() == M.F32
    m1, m2 = M.convolve(r, [[1, 0, 1]]), M.convolve(r, [[1.0, 0, 1]])
    assert m1._node.value != m2._node.value
    assert M.convolve(r, [[1, 0, 1]])._node.value == m1._node.value
    assert M.convolve(r, [[1], [0], [1]])._node.value != m1._node.value  # same values, other shape


def test_errors_raise(M):
    a = dem(M, 64, 48)
    b = dem(M, 32, 48)
    with pytest.raises(M.MapError):
        a + b
    with pytest.raises(M.MapError):
        M.value(M.zsum(a))  # not evaluated
    with pytest.raises(M.MapError):
        M.convolve(a, [[1, 1], [1, 1]])  # even mask


def test_hillshade_is_one_fused_group(M):
    p = M.plan(scripts.hillshade(M, dem(M)))
    assert len(p["groups"]) == 1
    g = p["groups"][0]
    assert g["pattern"] & 0x04 and g["pattern"] & 0x02  # FOCAL and LOCAL
    assert g["kinds"].count("Convolution") == 2
    assert g["halo"] == [1, 1]
    # D0 constants enter as kernel arguments (Fusioner.cpp:494-508)
    assert len(g["ins"]) == 10


def test_life_generations_share_one_kernel(M):
    s = M.from_array(np.zeros((64, 64), np.uint8), (32, 32))
    p = M.plan(scripts.life(M, s, 6))
    assert len(p["groups"]) == 6
    assert len({g["kernel"] for g in p["groups"]}) == 1


def test_scalar_chain_is_a_host_group(M):
    r = dem(M)
    mean = M.zsum(r) / 3072
    p = M.plan(mean)
    kinds = [g["kinds"] for g in p["groups"]]
    assert ["ZonalReduc"] in kinds and ["Binary"] in kinds
    assert any(g["numdim"] == M.D0 for g in p["groups"])


def test_focal_zonal_fusion_is_opt_in(M):
    r = dem(M)
    z = M.zsum(scripts.hillshade(M, r))
    assert len(M.plan(z)["groups"]) == 2  # FOCAL -> ZONAL not fused (Pattern.cpp:87)
    M.set_option("focal_zonal", 1)
    try:
        assert len(M.plan(z)["groups"]) == 1
    finally:
        M.set_option("focal_zonal", 0)


def test_focal_zonal_keeps_focal_on_focal_apart(M):
    """With the focal->zonal extension a focal producer read through a halo
    (a focal op on a focal result) still stays in its own group."""
    d = M.from_array(np.zeros((115, 52), np.float32), (64, 64))
    z = M.from_array(np.zeros((115, 52), np.int32), (64, 64))
    w = (np.ones((5, 5)) / 64).tolist()
    for fz, ngroups in [(0, 4), (1, 2)]:
        M.set_option("focal_zonal", fz)
        try:
            p = M.plan(*scripts.batch_tile(M, d, z, w))
        finally:
            M.set_option("focal_zonal", 0)
        assert len(p["groups"]) == ngroups, [g["kinds"] for g in p["groups"]]
        assert all("error" not in g for g in p["groups"]), [g.get("error") for g in p["groups"]]
        assert all(g["kinds"].count("Convolution") <= 2 and g["halo"] in ([2, 2], [1, 1], [0, 0])
                   for g in p["groups"] if "halo" in g)


def test_kernels_compile_for_gfx950(M):
    r = dem(M)
    u = M.from_array(np.zeros((64, 64), np.uint8), (32, 32))
    zones = M.from_array(np.zeros((48, 64), np.int32), (32, 16))
    exprs = [scripts.hillshade(M, r), scripts.life(M, u, 2), M.zsum(r), M.zmax(u), M.zmin(r * 2),
             M.zstats(r, zones, 16), M.fmax(r, [[1, 1, 1], [1, 1, 1], [1, 1, 1]]), r(  [1, -1]) + r([-1, 0]),
             M.con(r > 0, r, -r), M.astype(r, M.S32) % 7]
    exprs += list(scripts.batch_tile(M, r, zones, (np.ones((5, 5)) / 64).tolist()))
    for e in exprs:
        assert M.compile_plan(e) >= 1


def test_tile_fits_the_block(M):
    import re
    for side, bs in [(96, 48), (1024, 1024), (64, 16)]:
        u = M.from_array(np.zeros((side, side), np.uint8), (bs, bs))
        src = M.plan(scripts.life(M, u, 1))["groups"][0]["source"]
        tx, ty, vx = map(int, re.search(r"run_chunks<P, (\d+), (\d+), (\d+), mx::(?:STAGED|STREAM), (?:true|false), ", src).groups())
        assert tx <= bs and ty <= bs and 256 % (tx // vx) == 0, (bs, tx, ty, vx)


def test_iterated_groups_share_tune_key_and_kernel(M):
    """Every generation of an iterated script (kept or not) tunes and compiles as one computation."""
    u = M.from_array(np.zeros((128, 128), np.uint8), (64, 64))
    gs = M.plan(scripts.life(M, u, 4))["groups"]
    assert len(gs) == 4
    assert len({g["tune_key"] for g in gs}) == 1 and len({g["kernel"] for g in gs}) == 1


def test_float_mask_weights_do_not_change_the_kernel(M):
    """Tiles with different floating 5x5 weights compile to one kernel (the
    weights are kernel arguments), so they can share one batched launch."""
    rng = np.random.default_rng(3)
    d = M.from_array(np.zeros((100, 90), np.float32), (64, 64))
    z = M.from_array(np.zeros((100, 90), np.int32), (64, 64))
    names = set()
    for _ in range(3):
        w = (rng.integers(-64, 65, (5, 5)) / 64.0).tolist()
        gs = M.plan(scripts.batch_tile_sums(M, d, z, w))["groups"]
        names.add(tuple((g["kernel"], g["tune_key"]) for g in gs if "kernel" in g))
    assert len(names) == 1


def test_d2_raster_has_no_truth_value():
    """Only D0 rasters have a truth value (the loop condition of a `while` over rasters)."""
    from map_amd import map as M
    r = M.from_array(np.zeros((8, 8), np.float32), (8, 8))
    with pytest.raises(M.MapError, match="ambiguous"):
        bool(r)


def test_radial_scan_front_end_checks():
    """Radial scans take floating rasters and a start cell inside the raster; the plan shows
    a radial group by itself (its locals are not fused into it)."""
    from map_amd import map as M
    f = M.from_array(np.zeros((40, 60), np.float32), (32, 32))
    with pytest.raises(M.MapError, match="floating"):
        M.rmax(M.from_array(np.zeros((40, 60), np.int32), (32, 32)), [1, 1])
    with pytest.raises(M.MapError, match="inside"):
        M.rmax(f, [60, 3])
    v = (M.rmax(f * 2.0, [10, 20]) + 1.0) * f
    groups = M.plan(v)["groups"]
    radial = [g for g in groups if "radial" in g]
    assert len(radial) == 1 and radial[0]["kinds"] == ["RadialScan"] and radial[0]["radial"]["start"] == [10, 20]


def test_spreading_front_end_checks_and_plans(M):
    """Spreading scans need U8 directions and SUM, and run as a group of their own; a
    spreading step, Stats, Barrier and LhsAccess compile as generated kernels for gfx950."""
    f = M.from_array(np.zeros((40, 60), np.float32), (32, 32))
    d = M.from_array(np.zeros((40, 60), np.uint8), (32, 32))
    with pytest.raises(M.MapError, match="U8"):
        M.spread_scan(f, M.from_array(np.zeros((40, 60), np.int32), (32, 32)))
    with pytest.raises(M.MapError, match="SUM"):
        M.spread_scan(f, d, M.MAX)
    with pytest.raises(M.MapError, match="shape"):
        M.spread_scan(f, M.from_array(np.zeros((40, 61), np.uint8), (32, 32)))
    v = M.spread_scan(f * 2.0, d) + 1.0
    spread = [g for g in M.plan(v)["groups"] if "spread" in g]
    assert len(spread) == 1 and spread[0]["kinds"] == ["SpreadScan"]
    c = f * 1.0
    c[[3, 4]] = 2.5
    for e in (f(d) * 2.0, M.spread_neighbor(f, d, M.MAX), M.stats(f + 1.0), M.barrier(f) * 2.0, c):
        assert M.compile_plan(e) >= 1
    # a Barrier keeps its producer and consumer in two kernels
    assert len([g for g in M.plan(M.barrier(f + 1.0) * 2.0)["groups"] if "kernel" in g]) == 2


def test_library_exports_every_declared_entry_point():
    """libmapx.so exports every function include/mapx.h declares (loaded without a GPU)."""
    import ctypes
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    names = re.findall(r"\b(ma_\w+)\s*\(", open(os.path.join(root, "include", "mapx.h")).read())
    lib = ctypes.CDLL(os.path.join(root, "map_amd", "libmapx.so"))
    missing = [n for n in sorted(set(names)) if not hasattr(lib, n)]
    assert len(set(names)) > 50 and not missing, missing


def test_kernel_shape_key_determines_the_source(M):
    """Groups with equal shape keys (the key under which a compiled kernel is reused for a new
    raster) have identical generated sources: checked over the workload scripts applied to
    rasters of different sizes, types and block sizes, and scripts that differ only in a
    constant, a mask or an output."""
    import numpy as np
    from map_amd import workloads as W
    rasters = [M.from_array(np.zeros((h, w), dt), bs) for (w, h, dt, bs) in
               [(300, 200, np.float32, (128, 128)), (517, 389, np.float32, (256, 128)), (96, 70, np.float32, (32, 32)),
                (300, 200, np.int32, (128, 128)), (64, 64, np.uint8, (64, 64)), (200, 100, np.uint8, (64, 32))]]
    scripts = [lambda r: (W.hillshade(M, r),), lambda r: (W.slope(M, r), W.aspect(M, r)),
               lambda r: (W.focal_mean7(M, r),), lambda r: (r * 2 + 1,), lambda r: (r * 3 + 1,),
               lambda r: (M.convolve(r, [[1, 2, 1]] * 3),), lambda r: (M.convolve(r, [[1, 2, 2]] * 3),),
               lambda r: (W.life(M, r, 1),), lambda r: (M.zsum(r * 2),)]
    seen = {}
    for r in rasters:
        for sc in scripts:
            for g in M.plan(*sc(r))["groups"]:
                if "shape_key" not in g:
                    continue
                if g["shape_key"] in seen:
                    assert seen[g["shape_key"]] == g["source"], g["shape_key"]
                seen[g["shape_key"]] = g["source"]
    assert len(seen) >= 12
    # the same script over two rasters of one type and block size shares its key
    a = M.plan(W.hillshade(M, rasters[0]))["groups"]
    b = M.plan(W.hillshade(M, M.from_array(np.zeros((389, 517), np.float32), (128, 128))))["groups"]
    assert [g.get("shape_key") for g in a if "kernel" in g][0] == [g.get("shape_key") for g in b if "kernel" in g][0]
