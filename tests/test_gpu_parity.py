"""GPU parity of the HIP path against the CPU oracle (same seeded inputs).

Integer and boolean results are compared bit for bit; FP32 results against
the FP64 oracle with the ULP / cancellation bound of tests/parity.py, and
against the FP32-semantics oracle for reference.
"""
import numpy as np
import pytest

from map_amd import workloads as scripts
from tests.parity import assert_exact, assert_fp32, fp32_report

pytestmark = pytest.mark.gpu


def dem_array(w, h, seed):
    from oracle import ref as O
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 16
    g = np.stack([rng.uniform(0, w, n), rng.uniform(0, h, n), rng.uniform(8, 80, n), rng.uniform(0, 3000, n)], 1)
    return O.dem(0, 0, w, h, g, seed)


def _ulps(got, ref):
    got = np.asarray(got, np.float32).astype(np.float64)
    ref = np.asarray(ref, np.float32).astype(np.float64)
    sp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(got - ref) / sp


@pytest.mark.parametrize("w,h,bs", [(256, 192, (64, 64)), (300, 220, (128, 64)), (1024, 1024, (256, 256))])
def test_hillshade(G, O, w, h, bs):
    """Hillshade through the HIP path vs the reference-semantics FP32 oracle.

    Exact FP32 stages (Sobel sums in reference order, /8, x*x+y*y) are
    bit-exact; sqrt within 1 ULP and every transcendental node (atan, atan2,
    sin, cos) within 4 ULP of the correctly rounded FP32 result on the same
    FP32 input; the
    composed output's error against the FP64 oracle is reported for both the
    GPU and the FP32-semantics oracle.
    """
    print("case hillshade", w, h, bs, flush=True)
    dem = dem_array(w, h, 7)
    D = G.from_array(dem, bs)
    od = O.from_array(dem, bs)
    # exact stages: bit-exact; sqrt (v_sqrt_f32 on its normal range): within 1 ulp
    m_g = G.to_array(scripts.hori(G, D) * scripts.hori(G, D) + scripts.vert(G, D) * scripts.vert(G, D))
    x_o, y_o = scripts.hori(O, od), scripts.vert(O, od)
    m_o = x_o * x_o + y_o * y_o
    assert_exact(m_g, m_o.val, "sobel, squares and sum")
    r_o = O.sqrt(m_o)
    u = _ulps(G.to_array(G.sqrt(G.from_array(m_o.val, bs))), r_o.val)
    print("  sqrt: max %.2f ulp" % u.max(), flush=True)
    assert u.max() <= 1.0
    # transcendental nodes on identical FP32 inputs: <= 4 ULP of correctly rounded
    R_in = G.from_array(r_o.val, bs)
    X_in, Y_in = G.from_array(x_o.val, bs), G.from_array(y_o.val, bs)
    s_o = O.atan(r_o)
    S_in = G.from_array(s_o.val, bs)
    asp_o = scripts.aspect(O, od)
    arg_o = O.const(float(np.float32((360 - 315.0 + 90) / 180 * scripts.PI))) - asp_o
    A_in = G.from_array(arg_o.val, bs)
    for name, g, o in [("atan", G.atan(R_in), s_o), ("atan2", G.atan2(Y_in, -X_in), O.atan2(y_o, -x_o)),
                       ("sin", G.sin(S_in), O.sin(s_o)), ("cos", G.cos(S_in), O.cos(s_o)),
                       ("cos(ar-aspect)", G.cos(A_in), O.cos(arg_o))]:
        u = _ulps(G.to_array(g), o.val)
        print("  %s: max %.2f ulp, mean %.4f" % (name, u.max(), u.mean()), flush=True)
        assert u.max() <= 4.0, "%s: %.2f ulp" % (name, u.max())
    # composed output: FP64 error reported for the GPU and for the FP32-semantics oracle
    got = G.to_array(scripts.hillshade(G, D))
    O.set_mode("f64")
    t1, t2 = scripts.hillshade_terms(O, O.from_array(dem, bs))
    ref64 = (t1 + t2).val
    scale = np.abs(t1.val) + np.abs(t2.val)
    O.set_mode("f32")
    ref32 = scripts.hillshade(O, od).val
    print("  gpu vs fp64:", fp32_report(got, ref64, scale), flush=True)
    print("  fp32 oracle vs fp64:", fp32_report(ref32, ref64, scale), flush=True)
    print("  gpu vs fp32 oracle: max %.1f ulp, max abs %.3g" % (_ulps(got, ref32).max(),
                                                                np.abs(got.astype(np.float64) - ref32).max()), flush=True)
    assert np.all(np.isfinite(got))
    assert np.abs(got.astype(np.float64) - ref64).max() <= 2 * np.abs(ref32.astype(np.float64) - ref64).max() + 1e-6


@pytest.mark.parametrize("w,h,bs,gens", [(200, 150, (64, 32), 12), (130, 90, (40, 40), 9), (512, 512, (128, 128), 40),
                                         (1024, 768, (256, 256), 25)])
def test_life(G, O, w, h, bs, gens):
    print("case life", w, h, bs, gens, flush=True)
    rng = np.random.Generator(np.random.PCG64(1))
    s0 = (rng.random((h, w)) < 0.3).astype(np.uint8)
    got = G.to_array(scripts.life(G, G.from_array(s0, bs), gens))
    ref = O.life(s0, bs, gens)
    assert_exact(got, ref, "life")


def test_life_forms_identical(G, O):
    """Plain (one launch per block), batched and graph-replayed forms agree bit for bit."""
    rng = np.random.Generator(np.random.PCG64(2))
    s0 = (rng.random((384, 512)) < 0.3).astype(np.uint8)
    bs = (128, 128)
    outs = []
    try:
        for form, graph in [(0, 0), (1, 0), (1, 1), (1, 1)]:
            G.set_option("form", form)
            G.set_option("graph", graph)
            outs.append(G.to_array(scripts.life(G, G.from_array(s0, bs), 10)))
    finally:
        G.set_option("form", 1)
        G.set_option("graph", 0)
    ref = O.life(s0, bs, 10)
    for o in outs:
        assert_exact(o, ref, "life forms")


@pytest.mark.parametrize("case", ["01", "mixed", "u8", "narrow"])
def test_byte_lane_groups_match_oracle(G, O, case):
    """Groups with a byte-lane form (tests/test_byte_lanes.py) against the oracle on 0/1
    rasters (every tile in byte lanes), on 0/1 rasters with scattered larger values (the
    tiles holding one, through their halo too, take the per-cell form), on full-range U8
    data (the per-cell form, except the group exact for any byte) and on a raster whose
    edge blocks are narrower than the 7x7 halo. Right-edge vectors are partial (300 % 8)."""
    from tests.test_byte_lanes import lane_scripts
    print("case byte lanes", case, flush=True)
    rng = np.random.Generator(np.random.PCG64(11))
    w, h, bs = (258, 129, (128, 128)) if case == "narrow" else (300, 200, (128, 64))
    a = (rng.random((h, w)) < 0.35).astype(np.uint8)
    if case == "mixed":
        idx = rng.integers(0, a.size, 40)
        a.flat[idx] = rng.integers(2, 256, 40)
    if case == "u8":
        a = rng.integers(0, 256, (h, w)).astype(np.uint8)
    gs, os_ = lane_scripts(G, G.from_array(a, bs)), lane_scripts(O, O.from_array(a, bs))
    for name in gs:
        assert_exact(G.to_array(gs[name][0]), os_[name][0].val, name)


def test_life_byte_lanes_with_larger_values_and_constants(G, O):
    """Life over 12 generations from 0/1 cells with a few larger values (they persist where
    a cell has two neighbours, so tiles move between the two forms from one generation to
    the next), and a rule whose constants lie outside the lanes' [0, 127] (the kernel's
    argument check sends every tile to the per-cell form)."""
    rng = np.random.Generator(np.random.PCG64(12))
    a = (rng.random((384, 520)) < 0.3).astype(np.uint8)
    a[100:108, 200:260] = 5
    a[300, 17] = 200
    bs = (128, 128)
    got = G.to_array(scripts.life(G, G.from_array(a, bs), 12))
    ref = scripts.life(O, O.from_array(a, bs), 12).val
    assert_exact(got, ref, "life with larger values")

    def rule(M, u):
        nbh = M.convolve(u, scripts.LIFE_MASK)
        return (nbh == 300) + (nbh == -1) * u + (nbh == 3)
    b = (rng.random((200, 300)) < 0.3).astype(np.uint8)
    assert_exact(G.to_array(rule(G, G.from_array(b, bs))), rule(O, O.from_array(b, bs)).val, "constants out of range")


def test_local_integer_ops(G, O):
    rng = np.random.Generator(np.random.PCG64(3))
    a = rng.integers(0, 256, (100, 130)).astype(np.uint8)
    b = rng.integers(-1000, 1000, (100, 130)).astype(np.int32)
    bs = (32, 64)

    def script(M, A, B):
        c = A * 3 + B
        d = (c > 100) * (A - 7) + (c <= 100) * B / 3
        e = M.con(A < 128, A * 2, A + 1)
        f = (B % 7) + (-A) + (A == 5)
        return d + e * 2 + f
    got = G.to_array(script(G, G.from_array(a, bs), G.from_array(b, bs)))
    ref = script(O, O.from_array(a, bs), O.from_array(b, bs))
    assert_exact(got, ref.val, "integer local ops")


def test_zonal_global(G, O):
    rng = np.random.Generator(np.random.PCG64(4))
    a = rng.integers(0, 256, (300, 200)).astype(np.uint8)
    b = rng.integers(-5000, 5000, (300, 200)).astype(np.int32)
    f = rng.normal(0, 100, (300, 200)).astype(np.float32)
    bs = (64, 64)
    A, B, F = G.from_array(a, bs), G.from_array(b, bs), G.from_array(f, bs)
    rs = [G.zsum(A), G.zsum(B), G.zmax(B), G.zmin(B), G.zmax(F), G.zmin(F), G.zsum(F), G.zsum(G.astype(A, G.S32))]
    G.eval(*rs)
    vals = [G.value(r) for r in rs]
    oa, ob, of = O.from_array(a, bs), O.from_array(b, bs), O.from_array(f, bs)
    want = [O.value(O.zsum(oa)), O.value(O.zsum(ob)), O.value(O.zmax(ob)), O.value(O.zmin(ob)),
            O.value(O.zmax(of)), O.value(O.zmin(of))]
    assert vals[:6] == want, (vals, want)
    s64 = float(np.sum(f.astype(np.float64)))
    bound = 4 * np.finfo(np.float32).eps * float(np.sum(np.abs(f.astype(np.float64))))
    assert abs(vals[6] - s64) <= bound, (vals[6], s64)
    assert vals[7] == int(a.astype(np.int64).sum())


def test_zstats(G, O):
    rng = np.random.Generator(np.random.PCG64(5))
    w, h, nz = 520, 390, 37
    px, py = rng.integers(0, w, nz), rng.integers(0, h, nz)
    zones = O.voronoi(0, 0, w, h, px, py)
    f = rng.normal(10, 50, (h, w)).astype(np.float32)
    bs = (128, 128)
    zs = G.zstats(G.from_array(f, bs), G.from_array(zones, bs), nz)
    G.eval(zs)
    t = zs.table()
    ref = O.zstats(O.from_array(f, bs), O.from_array(zones, bs), nz)
    assert np.array_equal(t["count"], ref["count"])
    assert np.array_equal(t["min"], ref["min"])
    assert np.array_equal(t["max"], ref["max"])
    tol = 1e-12 * np.maximum(1.0, np.abs(ref["sum"])) * 64
    assert np.all(np.abs(t["sum"] - ref["sum"]) <= tol), np.max(np.abs(t["sum"] - ref["sum"]))


def _batch_cases(n, seed, bs=64):
    """Tiles of uneven sizes (not block multiples; edge remainders 0 or >= the
    5x5 halo), each with its DEM, 16 Voronoi zones and 5x5 weights in 1/64."""
    from oracle import ref as O
    rng = np.random.Generator(np.random.PCG64(seed))
    cases = []
    while len(cases) < n:
        w, h = (int(round(np.exp(rng.uniform(np.log(40), np.log(330))))) for _ in range(2))
        if w % bs == 1 or h % bs == 1:
            continue
        g = np.stack([rng.uniform(0, w, 6), rng.uniform(0, h, 6), rng.uniform(5, 60, 6), rng.uniform(0, 300, 6)], 1)
        dem = O.dem(0, 0, w, h, g, int(rng.integers(1 << 30)))
        zones = O.voronoi(0, 0, w, h, rng.integers(0, w, 16), rng.integers(0, h, 16))
        wts = (rng.integers(-64, 65, (5, 5)) / 64.0).tolist()
        cases.append((dem, zones, wts))
    return cases


def _run_batch(G, cases, bs, together):
    built = [scripts.batch_tile(G, G.from_array(d, bs), G.from_array(z, bs), w) for d, z, w in cases]
    if together:
        G.eval(*[x for b in built for x in b])
    else:
        for b in built:
            G.eval(*b)
    return [(G.to_array(m), zm.table(), zs.table()) for m, zm, zs in built]


def test_batched_tasks_match_alone_plain_and_oracle(G, O):
    """Uneven tiles evaluated together (one launch per kernel over a device
    task table and chunk queue), one by one, and in the plain per-block form
    give identical masks and identical per-zone tables (also the FP64 sums:
    partial records are per chunk, so results do not depend on the launch);
    masks and tables agree with the oracle."""
    bs = (64, 64)
    cases = _batch_cases(10, 11)
    print("case batched tasks", [c[0].shape for c in cases], flush=True)
    runs = {}
    try:
        for name, form, together, fz in [("batched", 1, True, 0), ("alone", 1, False, 0), ("plain", 0, True, 0),
                                         ("fused", 1, True, 1)]:
            G.set_option("form", form)
            G.set_option("focal_zonal", fz)
            runs[name] = _run_batch(G, cases, bs, together)
            print("  ran", name, G.stats()["launches"], flush=True)
    finally:
        G.set_option("form", 1)
        G.set_option("focal_zonal", 0)
    base = runs["batched"]
    for name, r in runs.items():
        for i, ((m, zm, zs), (m0, zm0, zs0)) in enumerate(zip(r, base)):
            assert_exact(m, m0, "%s tile %d mask" % (name, i))
            for k in ("sum", "min", "max", "count"):
                assert np.array_equal(zm[k], zm0[k]), (name, i, "mask", k)
                assert np.array_equal(zs[k], zs0[k], equal_nan=True), (name, i, "slope", k)
    for i, (d, z, w) in enumerate(cases):
        od = O.from_array(d, bs)
        sl = scripts.slope(O, O.convolve(od, w))
        m_ref = (sl.val > np.float32(0.5)).astype(np.uint8)
        # cells whose slope lies within 4 ulp of the threshold may round either way
        amb = np.abs(sl.val.astype(np.float64) - 0.5) <= 4 * np.spacing(np.float32(0.5))
        m, zm, zs = base[i]
        assert np.array_equal(m[~amb], m_ref[~amb]), "tile %d mask vs oracle" % i
        ref = O.zstats(O.from_array(m_ref, bs), O.from_array(z, bs), 16)
        assert np.array_equal(zm["count"], ref["count"])
        assert np.all(np.abs(zm["sum"] - ref["sum"]) <= np.bincount(z[amb], minlength=16)[:16]), i
        rs = O.zstats(sl, O.from_array(z, bs), 16)
        assert np.array_equal(zs["count"], rs["count"])
        tol = 4 * 2.0 ** -23 * np.maximum(1.0, np.abs(rs["sum"])) + 1e-9
        # slope values: within 4 ulp of the oracle's, so sums within 4 ulp relative
        assert np.all(np.abs(zs["sum"] - rs["sum"]) <= tol * 2), i


def _serve_results(G, outs):
    res = []
    for o in outs:
        res.append(o.table() if isinstance(o, G.ZoneStats) else G.to_array(o))
    return res


def test_serve_results_same_alone_and_in_mix(G, O):
    """Requests of the served mix (hillshade, slope+aspect, 7x7 mean, per-zone
    statistics) over DEMs of different sizes, submitted in waves while earlier
    ones run (continuous batching, a small HBM budget so that admission waits
    for memory, a repeated request that shares nodes with one in flight): every
    result equals the result of the same request evaluated alone, bit for bit."""
    bs = (256, 256)
    sizes = [(300, 200), (517, 389), (1024, 640), (96, 700), (777, 515)]
    rng = np.random.Generator(np.random.PCG64(21))
    dem_np = [dem_array(w, h, 30 + i) for i, (w, h) in enumerate(sizes)]
    zone_np = [O.voronoi(0, 0, w, h, rng.integers(0, w, 64), rng.integers(0, h, 64)) for (w, h) in sizes]
    kinds = ["hillshade", "slope_aspect", "mean7", "zstats"]
    reqs = [(kinds[int(rng.choice(4, p=[0.4, 0.3, 0.2, 0.1]))], int(rng.integers(len(sizes)))) for _ in range(30)]
    reqs += [reqs[0], reqs[0]]  # identical requests: shared nodes
    print("case serve", len(reqs), "requests", flush=True)
    dems = [G.from_array(d, bs) for d in dem_np]
    zones = [G.from_array(z, bs) for z in zone_np]
    srv = G.Server(hbm_budget=6 << 20, max_inflight=6)
    got = {}
    try:
        ids = []
        for wave in range(0, len(reqs), 5):
            for kind, di in reqs[wave:wave + 5]:
                ids.append(srv.submit(*scripts.serve_request(G, kind, dems[di], zones[di])))
            for rid, *_ in srv.step(0):
                got[rid] = _serve_results(G, srv.result(rid))
        while len(got) < len(ids):
            for rid, *_ in srv.step(2000):
                got[rid] = _serve_results(G, srv.result(rid))
        st = srv.stats()
        print("  serve stats", st, flush=True)
    finally:
        srv.close()
    assert st["completed"] == len(reqs) and st["waits_memory"] > 0
    for i, (kind, di) in enumerate(reqs):
        d = G.from_array(dem_np[di], bs)
        z = G.from_array(zone_np[di], bs)
        outs = scripts.serve_request(G, kind, d, z)
        G.eval(*outs)
        alone = _serve_results(G, outs)
        for a, b in zip(got[ids[i]], alone):
            if isinstance(a, dict):
                for k in ("sum", "min", "max", "count"):
                    assert np.array_equal(a[k], b[k], equal_nan=True), (i, kind, k)
            else:
                assert_exact(a, b, "request %d (%s)" % (i, kind))
    # zonal counts against the oracle
    for i, (kind, di) in enumerate(reqs):
        if kind == "zstats":
            ref = O.zstats(O.from_array(dem_np[di], bs), O.from_array(zone_np[di], bs), 64)
            assert np.array_equal(got[ids[i]][0]["count"], ref["count"])
            assert np.array_equal(got[ids[i]][0]["min"], ref["min"])


def test_serve_request_with_held_results_completes_without_work(G, O):
    """A request whose targets are the nodes of an earlier request that still holds them
    (value numbering gives identical scripts over the same DEM the same nodes) completes at
    admission without planning or launching, with the same results."""
    bs = (256, 256)
    dem = G.from_array(dem_array(700, 500, 51), bs)
    srv = G.Server(hbm_budget=64 << 20, max_inflight=4)
    try:
        first = scripts.serve_request(G, "slope_aspect", dem)
        r1 = srv.submit(*first)
        done = {}
        while r1 not in done:
            done.update({rid: 1 for rid, *_ in srv.step(2000)})
        kept = srv.result(r1)  # the caller still holds the results
        ref = [G.to_array(x) for x in kept]
        l0 = G.stats()["launches"]
        r2 = srv.submit(*scripts.serve_request(G, "slope_aspect", dem))
        while r2 not in done:
            done.update({rid: 1 for rid, *_ in srv.step(2000)})
        again = srv.result(r2)
        st = srv.stats()
    finally:
        srv.close()
    assert st["done_held"] == 1 and G.stats()["launches"] == l0
    for a, b in zip(again, ref):
        assert_exact(G.to_array(a), b, "held request result")
    del kept


def test_sin_cos_pair_within_4ulp(G, O):
    """sin and cos of one value in one group share one sincos call; each
    result stays within 4 ulp of the correctly rounded value."""
    rng = np.random.Generator(np.random.PCG64(8))
    parts = [rng.uniform(0, np.pi / 2, 40000), rng.uniform(-10, 10, 40000), rng.uniform(-1e3, 1e3, 40000),
             rng.uniform(-1e-3, 1e-3, 11072)]
    x = np.concatenate(parts).astype(np.float32).reshape(256, 512)
    X = G.from_array(x, (128, 128))
    s, c = G.sin(X), G.cos(X)
    srcs = [g.get("source", "") for g in G.plan(s, c)["groups"]]
    assert any("mx::sincos_f32(" in t for t in srcs) or any("sincosf(" in t for t in srcs)
    G.eval(s, c)
    ox = O.from_array(x, (128, 128))
    for name, g, o in [("sin", s, O.sin(ox)), ("cos", c, O.cos(ox))]:
        u = _ulps(G.to_array(g), o.val)
        print("  paired %s: max %.2f ulp" % (name, u.max()), flush=True)
        assert u.max() <= 4.0, (name, u.max())


@pytest.mark.parametrize("st", [1])
def test_register_window_form_matches(G, O, st):
    """The register-window focal form (tile_st=1) gives the staged form's results:
    Life bit-exact with the oracle on edge-block geometries, hillshade bit-exact
    with the staged kernels, per-zone sums of a fused focal+zonal group equal."""
    rng = np.random.Generator(np.random.PCG64(12))
    s0 = (rng.random((150, 203)) < 0.3).astype(np.uint8)
    dem = dem_array(300, 220, 9)
    zones = O.voronoi(0, 0, 300, 220, rng.integers(0, 300, 12), rng.integers(0, 220, 12))
    outs = {}
    try:
        for form in (0, st):
            G.set_option("tile_st", form)
            life = G.to_array(scripts.life(G, G.from_array(s0, (64, 32)), 7))
            D = G.from_array(dem, (128, 64))
            hs = G.to_array(scripts.hillshade(G, D))
            G.set_option("focal_zonal", 1)
            zs = G.zstats(scripts.hillshade(G, G.from_array(dem, (128, 64))), G.from_array(zones, (128, 64)), 12)
            G.eval(zs)
            G.set_option("focal_zonal", 0)
            outs[form] = (life, hs, zs.table())
    finally:
        G.set_option("tile_st", 0)
        G.set_option("focal_zonal", 0)
    assert_exact(outs[st][0], O.life(s0, (64, 32), 7), "life (register window)")
    assert_exact(outs[st][1], outs[0][1], "hillshade (register window vs staged)")
    for k in ("count", "min", "max"):
        assert np.array_equal(outs[st][2][k], outs[0][2][k]), k
    assert np.allclose(outs[st][2]["sum"], outs[0][2]["sum"], rtol=1e-12, atol=1e-9)


def test_raster_files_stream_through_the_device(G, O, tmp_path):
    """.bin and tiled BigTIFF inputs larger than one 32 MiB staging chunk (the chunk
    boundary falls inside a tile) stream to HBM, compute exactly as the same raster
    from memory does, and stream back out through ma_write."""
    print("case raster files 3000x2900 f32, blocks 1000x1000", flush=True)
    dem = dem_array(3000, 2900, 11)
    bs = (1000, 1000)
    ref = G.to_array(scripts.hillshade(G, G.from_array(dem, bs)))
    for ext in ("bin", "tif"):
        src = str(tmp_path / ("dem." + ext))
        G.file_store(src, dem, bs)
        D = G.read(src)
        assert G.to_array(D).tobytes() == dem.tobytes(), ext
        hs = scripts.hillshade(G, D)
        got = G.to_array(hs)
        assert_exact(got, ref, "hillshade from ." + ext)
        out = str(tmp_path / ("hs." + ext))
        G.write(hs, out)
        assert_exact(G.file_load(out), got, "written ." + ext)
        del D, hs


def test_while_loop_over_a_d0_condition(G, O):
    """`while <D0 raster>:` runs unmodified: the condition is evaluated on the device each
    iteration; the loop ends exactly when the zonal maximum reaches 0, and the state after
    each iteration matches the oracle's."""
    print("case while loop 130x90 S32, blocks 64x32", flush=True)
    rng = np.random.default_rng(8)
    s0 = rng.integers(0, 7, (90, 130)).astype(np.int32)
    bs = (64, 32)
    s = G.from_array(s0, bs)
    o = O.from_array(s0, bs)
    n = 0
    while G.zmax(s) > 0:
        s = (s > 0) * (s - 1) + 0 * G.convolve(s, [[0, 1, 0], [1, 1, 1], [0, 1, 0]])
        o = (o > 0) * (o - 1) + 0 * O.convolve(o, [[0, 1, 0], [1, 1, 1], [0, 1, 0]])
        n += 1
        assert_exact(G.to_array(s), o.val, "iteration %d" % n)
    assert n == int(s0.max())
    assert not G.to_array(s).any()


@pytest.mark.parametrize("w,h,bs,dt", [(300, 220, (128, 64), np.float32), (257, 129, (64, 64), np.float64),
                                       (700, 90, (256, 32), np.float32)])
def test_radial_scans_match_oracle(G, O, w, h, bs, dt):
    """rsum / rprod / rmax / rmin from start cells in corners, on edges, on block seams and
    inside: bit-exact against the oracle, inputs with NaN and infinities sprinkled in."""
    print("case radial", w, h, bs, np.dtype(dt).name, flush=True)
    rng = np.random.default_rng(w + h)
    a = (rng.standard_normal((h, w)) * 3).astype(dt)
    a[rng.random((h, w)) < 0.02] = np.nan
    a[rng.random((h, w)) < 0.01] = np.inf
    p = np.where(np.isfinite(a), 1 + a / 256, a).astype(dt)  # for the product
    starts = [(0, 0), (w - 1, h - 1), (w // 2, h // 3), (bs[0], bs[1] - 1), (w - 1, 0), (3, h - 2)]
    for red, gf, of, x in [("sum", G.rsum, O.rsum, a), ("max", G.rmax, O.rmax, a), ("min", G.rmin, O.rmin, a),
                           ("prod", G.rprod, O.rprod, p)]:
        A, oa = G.from_array(x, bs), O.from_array(x, bs)
        for s in starts:
            print("  %s from %s" % (red, s), flush=True)
            assert_exact(G.to_array(gf(A, list(s))), of(oa, s).val, "r%s from %s" % (red, s))


def test_viewshed_matches_oracle(G, O):
    """python/view.py's viewshed (normalised DEM, slope to the observer, rmax, visibility)
    against the oracle; every step is one IEEE operation, so bit-exact."""
    print("case viewshed 520x390 f32, blocks 128x128", flush=True)
    w, h, bs, obs, oh = 520, 390, (128, 128), (200, 150), 0.5
    dem = dem_array(w, h, 5)
    yy, xx = np.mgrid[0:h, 0:w]
    dx, dy = (obs[0] - xx).astype(np.int32), (obs[1] - yy).astype(np.int32)
    dist = np.sqrt((dx * dx + dy * dy).astype(np.float32))  # view.py distance(): sqrt(x*x + y*y + 0.0)

    def viewshed(M, D, DIST):
        norm = (D - M.zmin(D)) / (M.zmax(D) - M.zmin(D))
        slope = (norm - oh) / DIST
        return (M.rmax(slope, list(obs)) * DIST + oh) - norm

    got = G.to_array(viewshed(G, G.from_array(dem, bs), G.from_array(dist, bs)))
    ref = viewshed(O, O.from_array(dem, bs), O.from_array(dist, bs)).val
    assert_exact(got, ref, "viewshed")


@pytest.mark.parametrize("dt", [np.float32, np.int32, np.uint8, np.float64])
def test_focal_percent_matches_oracle(G, O, dt):
    """ma_focalPercent (AGE, ILE) over 3x3, 5x3 and 3x5 windows, block seams and edges, against
    the oracle bit for bit (values drawn from a few levels so equal neighbours are common)."""
    print("case focal percent", np.dtype(dt).name, flush=True)
    w, h, bs = 130, 90, (64, 32)
    a = np.random.default_rng(21).integers(0, 4, (h, w)).astype(dt)
    A, oa = G.from_array(a, bs), O.from_array(a, bs)
    for mask in ([[1, 1, 1]] * 3, [[1] * 5] * 3, [[1, 1, 1]] * 5):
        for kind in (G.AGE, G.ILE):
            assert_exact(G.to_array(G.focal_percent(A, mask, kind)), O.focal_percent(oa, mask, kind).val,
                         "percent %d %dx%d" % (kind, len(mask[0]), len(mask)))


@pytest.mark.parametrize("dt,n", [(np.int32, 4), (np.uint8, 5), (np.float32, 3), (np.int16, 2), (np.float64, 4)])
def test_diversity_matches_oracle(G, O, dt, n):
    """ma_diversity (VARI, MAJO, MINO, MEAN) over n rasters of a few levels (NaNs among the
    floats) against the oracle bit for bit; U8 means wrap as in C."""
    print("case diversity", np.dtype(dt).name, n, flush=True)
    w, h, bs = 70, 45, (32, 32)
    rng = np.random.default_rng(22 + n)
    xs = []
    for i in range(n):
        a = rng.integers(0, 4, (h, w)).astype(dt) * (60 if dt == np.uint8 else 1)
        if np.issubdtype(dt, np.floating):
            a[rng.random((h, w)) < 0.05] = np.nan
        xs.append(a)
    G_in, O_in = [G.from_array(a, bs) for a in xs], [O.from_array(a, bs) for a in xs]
    for kind in (G.VARI, G.MAJO, G.MINO, G.MEAN):
        got = G.to_array(G.diversity(G_in, kind))
        want = O.diversity(O_in, kind).val
        assert got.dtype == want.dtype
        assert_exact(got, want, "diversity %d" % kind)


def _d8_cases(w, h, seed):
    """Steepest-descent directions of a smooth DEM (long flow paths, pits) and random codes
    (invalid codes and direction cycles among them)."""
    rng = np.random.default_rng(seed)
    codes = np.array([0, 1, 2, 4, 8, 16, 32, 64, 128, 3, 255], np.uint8)
    return [scripts.d8_directions(dem_array(w, h, seed).astype(np.float64)), rng.choice(codes, (h, w))]


def _values(rng, dt, shape):
    if np.issubdtype(dt, np.floating):
        return rng.uniform(0, 1, shape).astype(dt)
    return rng.integers(0, 100, shape).astype(dt)


@pytest.mark.parametrize("dt", [np.float32, np.int32, np.uint8, np.float64, np.int64, np.uint16])
def test_spread_scan_matches_oracle(G, O, dt):
    """ma_spreadScan (flow accumulation along D8 directions) on a ragged 700x500 raster in
    256x128 blocks, against the oracle bit for bit: integer totals wrap in the type, FP
    totals add upstream totals in the fixed neighbour order."""
    print("case spread scan", np.dtype(dt).name, flush=True)
    w, h, bs = 700, 500, (256, 128)
    rng = np.random.default_rng(41)
    for k, d in enumerate(_d8_cases(w, h, 11)):
        v = _values(rng, dt, (h, w))
        got = G.to_array(G.spread_scan(G.from_array(v, bs), G.from_array(d, bs)))
        want = O.spread_scan(O.from_array(v, bs), O.from_array(d, bs)).val
        assert got.dtype == want.dtype
        assert_exact(got, want, "spread scan %s case %d" % (np.dtype(dt).name, k))


def test_flow_accumulation_long_paths(G, O):
    """A 2048x1536 DEM in 512x512 blocks: every cell's S32 total equals the oracle's (paths
    thousands of cells long cross many block seams)."""
    print("case flow accumulation 2048x1536", flush=True)
    w, h, bs = 2048, 1536, (512, 512)
    d = scripts.d8_directions(dem_array(w, h, 5).astype(np.float64))
    v = np.ones((h, w), np.int32)
    got = G.to_array(G.flow_accumulation(G.from_array(v, bs), G.from_array(d, bs)))
    want = O.flow_accumulation(O.from_array(v, bs), O.from_array(d, bs)).val
    assert_exact(got, want, "flow accumulation")
    assert got.max() > 1000


@pytest.mark.parametrize("dt", [np.float32, np.int32, np.uint8, np.float64])
def test_spread_neighbor_matches_oracle(G, O, dt):
    """ma_spreadNeighbor with SUM / MAX / MIN (and PROD for integers) on a ragged raster with
    block seams and edges (the border rule), against the oracle bit for bit; the same step
    written with neighbour and conditional ops gives the same bits."""
    print("case spread neighbor", np.dtype(dt).name, flush=True)
    w, h, bs = 300, 220, (128, 64)
    rng = np.random.default_rng(43)
    for k, d in enumerate(_d8_cases(w, h, 12)):
        v = _values(rng, dt, (h, w))
        V, D = G.from_array(v, bs), G.from_array(d, bs)
        ov, od = O.from_array(v, bs), O.from_array(d, bs)
        reds = [G.SUM, G.MAX, G.MIN] + ([G.PROD] if not np.issubdtype(dt, np.floating) else [])
        for red in reds:
            assert_exact(G.to_array(G.spread_neighbor(V, D, red)), O.spread_neighbor(ov, od, red).val,
                         "spread neighbor %d case %d" % (red, k))
        assert_exact(G.to_array(V(D)), G.to_array(scripts.flow_gather(G, V, D)), "gather vs neighbour ops")


def test_stats_barrier_access_lhs_access(G, O):
    """Stats (values + max/min in one evaluation), Barrier (values; the groups on its two sides
    stay apart), Access (one cell as a D0 value, across blocks) and LhsAccess (one cell
    replaced) against the oracle."""
    print("case stats/barrier/access", flush=True)
    w, h, bs = 130, 90, (64, 32)
    a = np.random.default_rng(44).normal(size=(h, w)).astype(np.float32)
    A, oa = G.from_array(a, bs), O.from_array(a, bs)
    S = G.stats(A * 2.0)
    G.eval(S)
    os_ = O.stats(oa * 2.0)
    assert_exact(G.to_array(S), os_.val, "stats values")
    assert G.stats_range(S) == O.stats_range(os_)
    B = G.barrier(A + 1.0) * 3.0
    assert len([g for g in G.plan(B)["groups"] if "kernel" in g]) == 2
    assert_exact(G.to_array(B), (O.barrier(oa + 1.0) * 3.0).val, "barrier")
    for cell in ([0, 0], [129, 89], [64, 32], [63, 31], [100, 5]):
        v = (A * 2.0)[cell]
        G.eval(v)
        assert G.value(v) == O.value((oa * 2.0)[cell]), cell
    C = A * 1.0
    C[[70, 40]] = 7.25
    C[[0, 89]] = -1
    oc = oa * 1.0
    oc[[70, 40]] = 7.25
    oc[[0, 89]] = -1
    assert_exact(G.to_array(C), oc.val, "lhs access")


def test_hillshade_4096_error_per_output(G, O):
    """The plain hillshade workload at its benchmark size (4096x4096 F32 DEM of seeded Gaussians
    + fBm, 512x512 blocks): slope, aspect and hillshade from the GPU against the FP64 oracle, and
    the FP32-semantics oracle against the same FP64 values, reported per output. The GPU may not
    be further from FP64 than twice the FP32 evaluation of the script itself is."""
    from map_amd import workloads as W
    print("case hillshade 4096x4096 error per output", flush=True)
    side, bs = 4096, (512, 512)
    rng = np.random.Generator(np.random.PCG64(0))
    D = G.gen_dem((side, side), bs, W.dem_gaussians(rng, side, side), 0)
    dem = G.to_array(D)
    outs = {"slope": W.slope, "aspect": W.aspect, "hillshade": W.hillshade}
    got = {k: G.to_array(f(G, D)).astype(np.float64) for k, f in outs.items()}
    ref32 = {k: f(O, O.from_array(dem, bs)).val.astype(np.float64) for k, f in outs.items()}
    O.set_mode("f64")
    od = O.from_array(dem, bs)
    ref64 = {k: np.asarray(f(O, od).val, np.float64) for k, f in outs.items()}
    t1, t2 = W.hillshade_terms(O, od)
    O.set_mode("f32")
    scale = {"slope": None, "aspect": None, "hillshade": np.abs(t1.val) + np.abs(t2.val)}
    for k in outs:
        rg = fp32_report(got[k], ref64[k], scale[k])
        ro = fp32_report(ref32[k], ref64[k], scale[k])
        d32 = np.abs(got[k] - ref32[k])
        print("  %-9s gpu vs fp64: max %.3g abs, %.3g ulp, %d cells > 4 ulp | fp32 oracle vs fp64: max %.3g abs, "
              "%.3g ulp, %d cells > 4 ulp | gpu vs fp32 oracle: max %.3g abs"
              % (k, rg["max_abs"], rg["max_ulp"], rg["n_over_4ulp"], ro["max_abs"], ro["max_ulp"], ro["n_over_4ulp"],
                 d32.max()), flush=True)
        assert np.all(np.isfinite(got[k]))
        assert rg["max_abs"] <= 2 * ro["max_abs"] + 1e-6, k


def test_serve_mix_with_scans_and_cell_reads(G, O):
    """Served requests that run precompiled scans (flow accumulation, radial max), cell reads
    (a D0 value from a raster) and Stats beside generated-kernel requests, in waves under a
    small HBM budget: every result equals the same request evaluated alone, and the scans equal
    the oracle."""
    bs = (128, 128)
    sizes = [(300, 200), (257, 389), (500, 310)]
    dem_np = [dem_array(w, h, 60 + i) for i, (w, h) in enumerate(sizes)]
    dir_np = [scripts.d8_directions(d.astype(np.float64)) for d in dem_np]

    def request(kind, di, D, R):
        if kind == "flow":
            return (G.flow_accumulation(G.ones_like(D) * 1.0, R),)
        if kind == "rmax":
            return (G.rmax(D * 1.0, [sizes[di][0] // 3, sizes[di][1] // 2]),)
        if kind == "cell":
            return ((D * 2.0)[[sizes[di][0] - 1, sizes[di][1] // 2]],)
        if kind == "stats":
            return (G.stats(D + 1.0),)
        return (scripts.hillshade(G, D),)

    def results(outs):
        return [G.value(o) if o.numdim() == G.D0 else G.to_array(o) for o in outs]

    rng = np.random.Generator(np.random.PCG64(61))
    kinds = ["flow", "rmax", "cell", "stats", "hill"]
    reqs = [(kinds[int(rng.integers(5))], int(rng.integers(len(sizes)))) for _ in range(15)]
    print("case serve with scans", reqs, flush=True)
    dems = [G.from_array(d, bs) for d in dem_np]
    dirs = [G.from_array(d, bs) for d in dir_np]
    srv = G.Server(hbm_budget=4 << 20, max_inflight=4)
    got, ids = {}, []
    try:
        for wave in range(0, len(reqs), 4):
            for kind, di in reqs[wave:wave + 4]:
                ids.append(srv.submit(*request(kind, di, dems[di], dirs[di])))
            for rid, *_ in srv.step(0):
                got[rid] = results(srv.result(rid))
        while len(got) < len(ids):
            for rid, *_ in srv.step(2000):
                got[rid] = results(srv.result(rid))
    finally:
        srv.close()
    for i, (kind, di) in enumerate(reqs):
        outs = request(kind, di, G.from_array(dem_np[di], bs), G.from_array(dir_np[di], bs))
        G.eval(*outs)
        for a, b in zip(got[ids[i]], results(outs)):
            assert_exact(np.asarray(a), np.asarray(b), "request %d (%s)" % (i, kind))
        if kind == "flow":
            want = O.flow_accumulation(O.from_array(np.ones_like(dem_np[di]), bs), O.from_array(dir_np[di], bs)).val
            assert_exact(got[ids[i]][0], want, "flow vs oracle")
