"""Benchmark of the block-task path on one MI355X.

Workload (one step): the Game of Life script of python/life.py, 200
generations over a 16384 x 16384 U8 raster of Bernoulli(p=0.3) cells
(numpy PCG64, seed 1) in 1024 x 1024 blocks, evaluated through the public API
(map_amd.map: convolve + relational/arithmetic ops -> DAG -> fusion -> one
hiprtc-compiled kernel per fused group). The input is resident in HBM before
timing; each step re-evaluates the 200-generation DAG from that input.

Prints one JSON line: value = cell-generations per second (whole job), with
the roofline of the dominant kernel (one generation over all blocks, 2 bytes
of algorithmic HBM traffic per cell: 1 read + 1 write) and a CPU baseline
(the oracle's restatement, timed on a bounded sample on the host cores).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def life_script(M, state, gens):
    mask = [[1, 1, 1], [1, 0, 1], [1, 1, 1]]
    for _ in range(gens):
        nbh = M.convolve(state, mask)
        state = (nbh == 3) + (nbh == 2) * state
    return state


def cpu_baseline(s0_full, bs, budget_s=12.0):
    """The oracle's Life (C, OpenMP over all host cores) on a bounded sample."""
    import numpy as np
    from oracle import ref as O
    cores = os.cpu_count() or 1
    side = min(s0_full.shape[0], 8192)
    sample = np.ascontiguousarray(s0_full[:side, :side])
    t0 = time.perf_counter()
    O.life(sample, bs, 1)
    one = time.perf_counter() - t0
    gens = int(max(1, min(200, budget_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    O.life(sample, bs, gens)
    dt = time.perf_counter() - t0
    return {"value": side * side * gens / dt / 1e9, "unit": "Gcell-gen/s", "cores": cores, "kind": "port",
            "sample": "Life %dx%d U8, %d generations, blocks %dx%d (top-left corner of the same input), "
                      "oracle/oracle.c or_life with OpenMP" % (side, side, gens, bs[0], bs[1])}


def _cpu_timed(run, cells, budget_s):
    """Calls run() (one pass over `cells` cells) until budget_s seconds have
    passed (at least once); returns (Gcells/s, passes)."""
    n, t0 = 0, time.perf_counter()
    while True:
        run()
        n += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s:
            return cells * n / dt / 1e9, n


CPU_PORT_NOTE = ("the oracle (oracle/ref.py, FP32 semantics: focal taps in C with OpenMP over the host "
                 "cores, the per-cell steps numpy)")


def cpu_hillshade(full_side, side=2048, block=512, budget_s=4.0):
    """The oracle's hillshade on the top-left side^2 corner of the hillshade workload's DEM."""
    import numpy as np
    from oracle import ref as O
    from map_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(0))
    D = O.from_array(O.dem(0, 0, side, side, W.dem_gaussians(rng, full_side, full_side), 0), (block, block))
    v, n = _cpu_timed(lambda: W.hillshade(O, D), side * side, budget_s)
    return {"value": float("%.4g" % v), "unit": "Gcells/s", "cores": os.cpu_count() or 1, "kind": "port",
            "sample": "hillshade of the %dx%d top-left corner of the same DEM, blocks %dx%d, %d passes, %s"
                      % (side, side, block, block, n, CPU_PORT_NOTE)}


def cpu_hillshade_zonal(full_side, side=2048, block=1024, budget_s=4.0):
    """The oracle's hillshade + 256-zone statistics on the top-left corner of the workload's inputs."""
    import numpy as np
    from oracle import ref as O
    from map_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(2))
    gauss = W.dem_gaussians(rng, full_side, full_side, ref_side=4096.0 * full_side / 4096)
    rz = np.random.Generator(np.random.PCG64(3))
    px, py = rz.integers(0, full_side, 256), rz.integers(0, full_side, 256)
    D = O.from_array(O.dem(0, 0, side, side, gauss, 2), (block, block))
    Z = O.from_array(O.voronoi(0, 0, side, side, px, py), (block, block))
    v, n = _cpu_timed(lambda: W.hillshade_zonal(O, D, Z, 256), side * side, budget_s)
    return {"value": float("%.4g" % v), "unit": "Gcells/s", "cores": os.cpu_count() or 1, "kind": "port",
            "sample": "hillshade + 256-zone stats of the %dx%d top-left corner of the same DEM and zones, "
                      "blocks %dx%d, %d passes, %s" % (side, side, block, block, n, CPU_PORT_NOTE)}


def cpu_batched_tiles(tiles, block, budget_s=5.0):
    """The oracle's per-tile pipeline over the first tiles of the same population
    ((w, h, gaussians, dem seed, zone points, weights) per tile), until the budget."""
    import numpy as np
    from oracle import ref as O
    from map_amd import workloads as W
    cells, k, dt = 0, 0, 0.0
    bs = (block, block)
    for w, h, gauss, dseed, px, py, wts in tiles:  # inputs made outside the timed part
        d = O.from_array(O.dem(0, 0, w, h, gauss, dseed), bs)
        z = O.from_array(O.voronoi(0, 0, w, h, px, py).astype(np.uint8), bs)
        t0 = time.perf_counter()
        W.batch_tile_sums(O, d, z, wts)
        dt += time.perf_counter() - t0
        cells += w * h
        k += 1
        if dt >= budget_s:
            break
    return {"value": float("%.4g" % (cells / dt / 1e9)), "unit": "Gcells/s", "cores": os.cpu_count() or 1,
            "kind": "port",
            "sample": "the first %d tiles of the same population (%d cells), %s" % (k, cells, CPU_PORT_NOTE)}


def cpu_serve_mix(block=1024, side=1024, budget_s=2.0):
    """The oracle on each request kind of the served mix over the side^2 top-left corner
    of DEM id 0 (zones for the zonal kind); the mix's rate weights each kind's time by
    its share of the requests (every kind covers its DEM's cells)."""
    import numpy as np
    from oracle import ref as O
    from map_amd import workloads as W
    r = np.random.Generator(np.random.PCG64([7, 0]))
    w, h = (int(round(np.exp(r.uniform(np.log(512), np.log(16384))))) for _ in range(2))
    r = np.random.Generator(np.random.PCG64([8, 0]))
    gauss = W.dem_gaussians(r, w, h)
    px, py = r.integers(0, w, 64), r.integers(0, h, 64)
    sw, sh = min(side, w), min(side, h)
    D = O.from_array(O.dem(0, 0, sw, sh, gauss, 5000), (block, block))
    Z = O.from_array(O.voronoi(0, 0, sw, sh, px, py), (block, block))
    per = {}
    for kind in SERVE_KINDS:
        v, _ = _cpu_timed(lambda: W.serve_request(O, kind, D, Z if kind == "zstats" else None), sw * sh, budget_s)
        per[kind] = float("%.4g" % v)
    mix = 1.0 / sum(p / per[k] for k, p in zip(SERVE_KINDS, SERVE_MIX))
    return {"value": float("%.4g" % mix), "unit": "Gcells/s", "cores": os.cpu_count() or 1, "kind": "port",
            "per_kind_gcells_s": per,
            "sample": "each request kind on the %dx%d top-left corner of DEM id 0, blocks %dx%d, the mix's "
                      "shares as weights, %s" % (sw, sh, block, block, CPU_PORT_NOTE)}


def _progress(msg):
    """A line on stderr per stage of the long sections."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def _kernel_profile(M, fn):
    """Runs fn() once with every launch timed; returns {kernel: record}."""
    M.set_option("profile", 1)
    M.kernel_times()
    fn()
    kt = M.kernel_times()
    M.set_option("profile", 0)
    return kt


def hillshade_zonal_bench(M, side=32768, block=1024, reps=3, cpu=True):
    """Hillshade + per-zone statistics over 256 Voronoi zones on a side^2 FP32
    DEM (seeded Gaussians + fBm) and an S32 zone raster, fused into one kernel
    (focal_zonal) and in the reference's grouping (hillshade materialised)."""
    import numpy as np
    from map_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(2))
    dem = M.gen_dem((side, side), (block, block), W.dem_gaussians(rng, side, side, ref_side=4096.0 * side / 4096), 2)
    rz = np.random.Generator(np.random.PCG64(3))
    zones = M.gen_zones((side, side), (block, block), rz.integers(0, side, 256), rz.integers(0, side, 256))
    cells = side * side
    # tile shapes tuned on the workload itself (a smaller raster of the same script
    # ranks the shapes differently: fewer blocks to spread over the CUs)
    out = {"tuned_tiles": {}}
    for name, fz in (("fused", 1), ("unfused", 0)):
        M.set_option("focal_zonal", fz)
        _progress("hillshade_zonal: autotuning the %s form" % name)
        out["tuned_tiles"][name] = M.autotune(W.hillshade_zonal(M, dem, zones, 256), reps=2)["chosen"]
    for name, fz in (("fused", 1), ("unfused", 0)):
        M.set_option("focal_zonal", fz)
        zs = W.hillshade_zonal(M, dem, zones, 256)
        M.eval(zs)  # compile, warm
        walls = []
        for _ in range(reps):
            M.release(zs)
            t0 = time.perf_counter()
            M.eval(zs)
            walls.append(time.perf_counter() - t0)
        kt = _kernel_profile(M, lambda: (M.release(zs), M.eval(zs)))
        main = max(kt.items(), key=lambda kv: kv[1]["total_ms"])
        dev_ms = sum(v["total_ms"] for v in kt.values())
        t = zs.table()
        out[name] = {"wall_ms": round(min(walls) * 1e3, 3), "device_ms": round(dev_ms, 3),
                     "gcells_s": round(cells / min(walls) / 1e9, 2),
                     "kernels": {k: round(v["total_ms"] / v["count"], 4) for k, v in kt.items()},
                     "main_kernel_ms": round(main[1]["total_ms"] / main[1]["count"], 4),
                     "count_total": int(t["count"].sum())}
        del zs
    M.set_option("focal_zonal", 0)
    f = out["fused"]
    alg = 8.0 * cells  # DEM 4 B + zone 4 B per cell
    f["achieved_gbs"] = round(alg / (f["main_kernel_ms"] * 1e-3) / 1e9, 1)
    f["hbm_frac"] = round(f["achieved_gbs"] / HBM_PEAK_GBS, 4)
    if cpu:
        out["cpu_baseline"] = cpu_hillshade_zonal(side, min(side, 2048), block)
    return {"workload": "hillshade + per-zone stats (256 Voronoi zones), %dx%d f32 DEM, s32 zones, blocks %dx%d"
                        % (side, side, block, block), "algorithmic_bytes_per_cell": 8, **out}


def batched_tiles_bench(M, ntiles=2048, block=1024, seed=4, forms=("batched", "plain"), cpu=True):
    """2048 independent tiles of log-uniform sides in [256, 4096]: 5x5
    weighted focal sum -> slope -> mask slope > 0.5 -> per-zone sums of the mask
    over 16 Voronoi zones (U8). Batched (one launch per kernel over a device
    task table) against the plain per-block form, identical outputs."""
    import numpy as np
    from map_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(seed))
    tiles, cells, specs = [], 0, []
    for i in range(ntiles):
        w, h = (int(round(np.exp(rng.uniform(np.log(256), np.log(4096))))) for _ in range(2))
        trng = np.random.Generator(np.random.PCG64([seed, i]))
        gauss = W.dem_gaussians(trng, w, h)
        px, py = trng.integers(0, w, 16), trng.integers(0, h, 16)
        dem = M.gen_dem((w, h), (block, block), gauss, 1000 + i)
        zones = M.gen_zones((w, h), (block, block), px, py, M.U8)
        wts = (trng.integers(-64, 65, (5, 5)) / 64.0).tolist()
        tiles.append((dem, zones, wts))
        if i < 64:
            specs.append((w, h, gauss, 1000 + i, px, py, wts))
        cells += w * h
        if i % 256 == 255:
            _progress("batched tiles: %d tiles generated" % (i + 1))
    M.set_option("focal_zonal", 1)
    res = {}
    try:
        # tile shapes tuned on the whole population, batched as it is timed (keys:
        # computation and block size; a sample of the first tiles ranks the shapes differently)
        M.set_option("form", 1)
        sample = [W.batch_tile_sums(M, d, z, w) for d, z, w in tiles]
        _progress("batched tiles: autotuning on the %d tiles" % len(tiles))
        tuned = M.autotune(*sample, reps=2)["chosen"]
        del sample
        for name, form in (("batched", 1), ("plain", 0)):
            if name not in forms:
                continue
            M.set_option("form", form)
            M.set_option("graph", 1 if form == 1 else 0)  # steady state replayed from hipGraphs
            outs = [W.batch_tile_sums(M, d, z, w) for d, z, w in tiles]
            _progress("batched tiles: %s form, first evaluation" % name)
            t0 = time.perf_counter()
            M.eval(*outs)  # plan, compile, first run
            first = time.perf_counter() - t0
            walls, devs = [], []
            for _ in range(2):
                for o in outs:
                    M.release(o)
                h0 = M.stats()
                t0 = time.perf_counter()
                M.eval(*outs)
                walls.append(time.perf_counter() - t0)
                devs.append(M.last_eval_ms())
            h1 = M.stats()
            res[name] = {"wall_ms": round(min(walls) * 1e3, 2), "device_ms": round(min(devs), 2),
                         "first_eval_ms": round(first * 1e3, 1),
                         "gcells_s": round(cells / min(walls) / 1e9, 2),
                         "host_ms_last": {k: round(h1[k] - h0[k], 2) for k in ("ms_collect", "ms_predict", "ms_plan", "ms_fuse",
                                                                               "ms_codegen", "ms_prepare", "ms_launch")},
                         "launches_last": h1["launches"] - h0["launches"],
                         "sums": [o.table()["sum"] for o in outs]}
            if name == "batched":  # per-kernel device time of one evaluation (launches timed, no graph)
                M.set_option("graph", 0)
                kt = _kernel_profile(M, lambda: ([M.release(o) for o in outs], M.eval(*outs)))
                res[name]["kernels"] = {k: {"count": v["count"], "total_ms": round(v["total_ms"], 3)}
                                        for k, v in kt.items()}
            del outs
    finally:
        M.set_option("form", 1)
        M.set_option("graph", 0)
        M.set_option("focal_zonal", 0)
    res["tuned_tiles"] = tuned
    both = "batched" in res and "plain" in res
    same = both and all(np.array_equal(a, b) for a, b in zip(res["batched"].pop("sums"), res["plain"].pop("sums")))
    for r in res.values():
        if isinstance(r, dict):
            r.pop("sums", None)
    if cpu:
        res["cpu_baseline"] = cpu_batched_tiles(specs, block)
    return {"workload": "%d tiles, sides log-uniform [256, 4096], 5x5 weighted sum -> slope -> mask -> 16-zone sums, "
                        "blocks %dx%d" % (ntiles, block, block),
            "cells": cells, "outputs_identical": bool(same) if both else None,
            "speedup_vs_plain": round(res["plain"]["wall_ms"] / res["batched"]["wall_ms"], 2) if both else None, **res}


class DemCache:
    """The served mix's DEMs (and zone rasters) by DEM id in the runtime's resident cache
    (map.ResidentCache: least recently used entries leave HBM past the byte budget)."""

    def __init__(self, M, block, budget):
        self.M, self.block = M, block
        self.cache = M.ResidentCache(budget)

    def side(self, k):
        import numpy as np
        r = np.random.Generator(np.random.PCG64([7, k]))
        return tuple(int(round(np.exp(r.uniform(np.log(512), np.log(16384))))) for _ in range(2))

    def get(self, k, zones):
        import numpy as np
        from map_amd import workloads as W
        M = self.M
        w, h = self.side(k)
        r = np.random.Generator(np.random.PCG64([8, k]))
        gauss = W.dem_gaussians(r, w, h)
        px, py = r.integers(0, w, 64), r.integers(0, h, 64)
        dem = self.cache.get(2 * k, lambda: M.gen_dem((w, h), (self.block, self.block), gauss, 5000 + k))
        z = self.cache.get(2 * k + 1, lambda: M.gen_zones((w, h), (self.block, self.block), px, py)) if zones else None
        return dem, z, (w * h * 4) * (2 if zones else 1), w * h

    def stats(self):
        return self.cache.stats()


SERVE_KINDS = ("hillshade", "slope_aspect", "mean7", "zstats")
SERVE_MIX = (0.4, 0.3, 0.2, 0.1)
SERVE_BYTES = {"hillshade": 8, "slope_aspect": 12, "mean7": 8, "zstats": 8}


def serve_stream(n, rate, seed, ndem=500):
    """n requests: Poisson arrival times at `rate`, kinds from the mix, DEM ids Zipf(1.1) over ndem."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.cumsum(rng.exponential(1.0 / rate, n)) if rate > 0 else np.zeros(n)
    kinds = rng.choice(len(SERVE_KINDS), n, p=SERVE_MIX)
    p = 1.0 / np.arange(1, ndem + 1) ** 1.1
    ids = rng.choice(ndem, n, p=p / p.sum())
    return [(float(t[i]), SERVE_KINDS[kinds[i]], int(ids[i])) for i in range(n)]


def serve_run(M, cache, stream, max_inflight=64, drain_s=20.0):
    """Open loop: each request is submitted at its arrival time (or as soon as
    the host gets to it) and completes when the server retires it. Latency =
    completion - scheduled arrival."""
    import numpy as np
    from map_amd import workloads as W
    srv = M.Server(hbm_budget=64 << 30, max_inflight=max_inflight)
    lat, cells, nbytes = [], 0, 0
    arr = {}
    i, n = 0, len(stream)
    t_build = t_submit = t_step = 0.0  # host seconds: building request scripts, submitting, stepping
    h0 = M.stats()
    t0 = time.monotonic()
    try:
        while len(lat) < n:
            now = time.monotonic() - t0
            burst = 0  # the client and the server share this thread: the server steps between bursts
            while i < n and stream[i][0] <= now and burst < max_inflight:
                burst += 1
                ta, kind, k = stream[i]
                tb = time.perf_counter()
                dem, z, _, c = cache.get(k, kind == "zstats")
                outs = W.serve_request(M, kind, dem, z)
                ts = time.perf_counter()
                rid = srv.submit(*outs)
                t_build += ts - tb
                t_submit += time.perf_counter() - ts
                arr[rid] = ta
                cells += c
                nbytes += c * SERVE_BYTES[kind]
                i += 1
            wait = int(min(2000, max(0, (stream[i][0] - (time.monotonic() - t0)) * 1e6))) if i < n else 2000
            ts = time.perf_counter()
            done = srv.step(wait)
            t_step += time.perf_counter() - ts
            for rid, ts_, tadm, td in done:
                lat.append(td - t0 - arr.pop(rid))
                srv.result(rid)
            if i >= n and time.monotonic() - t0 > stream[-1][0] + drain_s:
                break
        wall = time.monotonic() - t0
        st = srv.stats()
        h1 = M.stats()
    finally:
        srv.close()
    per = 1e3 / max(1, n)  # ms per request
    host = {"build_script_ms": round(t_build * per, 4), "submit_ms": round(t_submit * per, 4),
            "step_ms": round(t_step * per, 4)}
    for key in ("ms_plan", "ms_fuse", "ms_codegen", "ms_prepare", "ms_launch"):
        host[key[3:] + "_ms"] = round((h1[key] - h0[key]) / max(1, n), 4)  # inside step
    lat = np.array(lat) * 1e3
    return {"requests": n, "completed": int(len(lat)), "wall_s": round(wall, 3),
            "req_s": round(len(lat) / wall, 1), "gcells_s": round(cells / wall / 1e9, 2),
            "p50_ms": round(float(np.percentile(lat, 50)), 3) if len(lat) else None,
            "p99_ms": round(float(np.percentile(lat, 99)), 3) if len(lat) else None,
            "mean_tasks_per_step": round(st["group_tasks"] / max(1, st["steps"]), 2),
            "waits_memory": st["waits_memory"], "done_without_work": st["done_held"], "host_ms_per_request": host}


def sequential_run(M, cache, stream, form):
    """The same requests one at a time (ma_eval per request) in the given form."""
    import numpy as np
    from map_amd import workloads as W
    M.set_option("form", form)
    lat = []
    cells = 0
    t0 = time.monotonic()
    try:
        for ta, kind, k in stream:
            dem, z, _, c = cache.get(k, kind == "zstats")
            outs = W.serve_request(M, kind, dem, z)
            t1 = time.monotonic()
            M.eval(*outs)
            lat.append(time.monotonic() - t1)
            cells += c
            del outs
    finally:
        M.set_option("form", 1)
    wall = time.monotonic() - t0
    lat = np.array(lat) * 1e3
    return {"requests": len(stream), "req_s": round(len(stream) / wall, 1), "gcells_s": round(cells / wall / 1e9, 2),
            "service_p50_ms": round(float(np.percentile(lat, 50)), 3),
            "service_p99_ms": round(float(np.percentile(lat, 99)), 3)}


def serve_bench(M, block=1024, seconds=5.0, seed=9, cpu=True):
    """Served mix (40% hillshade, 30% slope+aspect, 20% 7x7 mean, 10% 64-zone
    stats; DEM sides log-uniform [512, 16384], 500 DEM ids, Zipf 1.1): the
    plain (per-block launches) and batched forms one request at a time, then
    the server at Poisson rates around and above the sequential capacity."""
    import gc
    # request scripts are acyclic (reference counting frees them): the cyclic
    # collector's pauses would only add to the latencies measured
    gc.collect()
    gc.disable()
    try:
        return _serve_bench(M, block, seconds, seed, cpu)
    finally:
        gc.enable()


def _serve_bench(M, block, seconds, seed, cpu):
    cache = DemCache(M, block, 96 << 30)
    warm = serve_stream(600, 0, seed + 1)
    _progress("serve: warming the DEM cache")
    sequential_run(M, cache, warm, 1)  # DEMs of the popular ids resident, kernels compiled
    # A raster whose edge blocks are narrower than a focal halo runs a second module of the
    # kernel, compiled when first needed: every kind once on a DEM with 1-cell edge blocks
    # compiles those modules now rather than inside a timed run.
    import numpy as np
    from map_amd import workloads as W
    side = block + 1
    rng = np.random.Generator(np.random.PCG64(seed))
    dem = M.gen_dem((side, side), (block, block), W.dem_gaussians(rng, side, side), 1)
    zones = M.gen_zones((side, side), (block, block), rng.integers(0, side, 64), rng.integers(0, side, 64))
    for kind in SERVE_KINDS:
        M.eval(*W.serve_request(M, kind, dem, zones))
    del dem, zones
    probe = serve_stream(200, 0, seed + 2)
    plain = sequential_run(M, cache, probe, 0)
    seq = sequential_run(M, cache, probe, 1)
    out = {"workload": "served mix over 500 DEMs (sides log-uniform [512, 16384], Zipf 1.1), blocks %dx%d"
                       % (block, block), "sequential_plain": plain, "sequential_batched": seq, "served": []}
    for mult in (0.5, 1.0, 2.0, 4.0):
        rate = max(1.0, seq["req_s"] * mult)
        _progress("serve: %.0f req/s offered" % rate)
        n = int(min(20000, max(50, rate * seconds)))
        r = serve_run(M, cache, serve_stream(n, rate, seed + int(mult * 10)))
        r["offered_req_s"] = round(rate, 1)
        out["served"].append(r)
    cs = cache.stats()
    out["dem_cache"] = {"hits": cs["hits"], "misses": cs["misses"], "evictions": cs["evictions"],
                        "resident_gb": round(cs["bytes"] / 2**30, 1)}
    if cpu:
        out["cpu_baseline"] = cpu_serve_mix(block)
    return out


def hillshade_parts_bench(M, side=4096, block=1024, nz=256):
    """Per-zone statistics (256 Voronoi zones) of each prefix of the hillshade script, fused
    into one kernel each (focal_zonal), to attribute a kernel's VALU instructions per node
    (run under rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU): the DEM alone, the horizontal
    gradient, the squared magnitude, its square root, the slope, cos / sin+cos of the slope,
    the aspect, cos of the aspect term, the hillshade. Prints each part's kernel time."""
    import numpy as np
    from map_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(2))
    dem = M.gen_dem((side, side), (block, block), W.dem_gaussians(rng, side, side), 2)
    rz = np.random.Generator(np.random.PCG64(3))
    zones = M.gen_zones((side, side), (block, block), rz.integers(0, side, nz), rz.integers(0, side, nz))
    def mag():
        x, y = W.hori(M, dem), W.vert(M, dem)
        return x * x + y * y

    # built one at a time, so that no intermediate is held (and kept) by the caller
    parts = [("dem", lambda: dem * 1.0), ("hori", lambda: W.hori(M, dem)), ("mag", mag),
             ("sqrt", lambda: M.sqrt(mag())), ("slope", lambda: W.slope(M, dem)),
             ("cos_slope", lambda: M.cos(W.slope(M, dem)) * 1.0),
             ("sincos_slope", lambda: M.cos(W.slope(M, dem)) + M.sin(W.slope(M, dem))),
             ("aspect", lambda: W.aspect(M, dem)), ("cos_aspect", lambda: M.cos(2.356 - W.aspect(M, dem))),
             ("hillshade", lambda: W.hillshade(M, dem))]
    M.set_option("focal_zonal", 1)
    out = {}
    try:
        for name, make in parts:
            zs = M.zstats(make(), zones, nz)
            M.eval(zs)
            kt = _kernel_profile(M, lambda: (M.release(zs), M.eval(zs)))
            k, rec = max(kt.items(), key=lambda kv: kv[1]["total_ms"])
            out[name] = {"kernel": k, "ms": round(rec["total_ms"] / rec["count"], 4)}
            del zs
    finally:
        M.set_option("focal_zonal", 0)
    return {"workload": "per-zone stats (%d Voronoi zones) of prefixes of the hillshade script, %dx%d f32 DEM, "
                        "blocks %dx%d, one fused kernel each" % (nz, side, side, block, block), "parts": out}


def radial_bench(M, side=16384, block=1024, reps=3):
    """Radial scan (rmax from the centre) of a resident side^2 FP32 DEM, and the viewshed
    script of python/view.py on it. Device time of each evaluation from HIP events;
    algorithmic traffic of the scan 8 B/cell (input read + output write)."""
    import numpy as np
    from map_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(2))
    dem = M.gen_dem((side, side), (block, block), W.dem_gaussians(rng, side, side, ref_side=4096.0 * side / 4096), 2)
    M.eval(dem)
    obs = [side // 2, side // 2]
    out = {}
    scan = M.rmax(dem, obs)
    M.eval(scan)
    devs, launches = [], None
    for _ in range(reps):
        M.release(scan)
        l0 = M.stats()["launches"]
        M.eval(scan)
        devs.append(M.last_eval_ms())
        launches = M.stats()["launches"] - l0
    ms = min(devs)
    gbs = 8.0 * side * side / (ms * 1e-3) / 1e9
    out["rmax"] = {"device_ms": round(ms, 3), "gcells_s": round(side * side / (ms * 1e-3) / 1e9, 2), "launches": launches,
                   "achieved_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
    del scan
    # python/view.py: normalise, slope to the observer, rmax, visibility (the distance raster from index())
    h = 0.5
    dx = obs[0] - M.index(dem, M.D1)
    dy = obs[1] - M.index(dem, M.D2)
    dist = M.sqrt(dx * dx + dy * dy + 0.0)
    norm = (dem - M.zmin(dem)) / (M.zmax(dem) - M.zmin(dem))
    view = (M.rmax((norm - h) / dist, obs) * dist + h) - norm
    M.eval(view)
    devs = []
    for _ in range(reps):
        M.release(view)
        M.eval(view)
        devs.append(M.last_eval_ms())
    out["viewshed"] = {"device_ms": round(min(devs), 3), "gcells_s": round(side * side / (min(devs) * 1e-3) / 1e9, 2)}
    del view, dist, norm, dx, dy
    M.release(dem)
    return {"workload": "radial scan from the centre of a %dx%d f32 DEM, blocks %dx%d" % (side, side, block, block), **out}


def spread_bench(M, side=16384, block=1024, reps=3):
    """Flow accumulation (ma_spreadScan, SUM) of an F32 raster of ones along the steepest-descent
    D8 directions of a side^2 DEM (directions computed on the device by a fused focal script).
    Device time of the scan from HIP events; algorithmic traffic 9 B/cell (F32 input and
    U8 direction read, F32 result written)."""
    import numpy as np
    from map_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(6))
    dem = M.gen_dem((side, side), (block, block), W.dem_gaussians(rng, side, side, ref_side=4096.0 * side / 4096), 6)
    dirs = W.d8_script(M, dem)
    water = M.ones_like(dem)
    M.eval(dirs, water)
    acc = M.flow_accumulation(water, dirs)
    M.eval(acc)
    devs = []
    for _ in range(reps):
        M.release(acc)
        M.eval(acc)
        devs.append(M.last_eval_ms())
    ms = min(devs)
    top = M.zmax(acc)
    M.eval(top)
    gbs = 9.0 * side * side / (ms * 1e-3) / 1e9
    res = {"workload": "flow accumulation of ones along D8 steepest descent, %dx%d f32 DEM, blocks %dx%d"
                       % (side, side, block, block),
           "device_ms": round(ms, 3), "gcells_s": round(side * side / (ms * 1e-3) / 1e9, 2),
           "achieved_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
           "largest_total": M.value(top)}
    del acc, top, water, dirs
    M.release(dem)
    return res


def hillshade_bench(M, side=4096, block=512, reps=5, cpu=True):
    """The plain hillshade script (python/hill.py) on a side^2 FP32 DEM in
    block^2 blocks: one fused focal+local group, plain and batched forms."""
    import numpy as np
    from map_amd import workloads as W
    rng = np.random.Generator(np.random.PCG64(0))
    dem = M.gen_dem((side, side), (block, block), W.dem_gaussians(rng, side, side), 0)
    out = {"tuned_tiles": M.autotune(W.hillshade(M, dem), reps=3)["chosen"]}
    for name, form in (("batched", 1), ("plain", 0)):
        M.set_option("form", form)
        hs = W.hillshade(M, dem)
        M.eval(hs)
        walls = []
        for _ in range(reps):
            M.release(hs)
            t0 = time.perf_counter()
            M.eval(hs)
            walls.append(time.perf_counter() - t0)
        kt = _kernel_profile(M, lambda: (M.release(hs), M.eval(hs)))
        k = max(kt.values(), key=lambda v: v["total_ms"])
        out[name] = {"wall_ms": round(min(walls) * 1e3, 3), "gcells_s": round(side * side / min(walls) / 1e9, 2),
                     "kernel_ms_per_launch": round(k["total_ms"] / k["count"], 4), "launches": k["count"]}
        del hs
    M.set_option("form", 1)
    if cpu:
        out["cpu_baseline"] = cpu_hillshade(side, min(side, 2048), block)
    return {"workload": "hillshade %dx%d f32, blocks %dx%d" % (side, side, block, block), **out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--side", type=int, default=16384)
    ap.add_argument("--block", type=int, default=1024)
    ap.add_argument("--gens", type=int, default=200)
    ap.add_argument("--form", choices=["batched", "plain"], default="batched")
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-gens", type=int, default=8, help="generations timed once in the plain form")
    ap.add_argument("--tune", type=int, default=1, help="autotune the tile shape before timing")
    ap.add_argument("--tile", default="",
                    help="TX,TY,VX[,PF[,ST[,W]]] tile shape of the Life kernel (implies --tune 0)")
    ap.add_argument("--chunk-tiles", type=int, default=0,
                    help="tiles per chunk of the persistent kernels (0: the runtime's choice)")
    ap.add_argument("--queue", type=int, default=-1, choices=[-1, 0, 1, 2],
                    help="chunk split: 0 device queue for batched launches only (the default), 1 always, 2 never")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the Life headline (skip hillshade, hillshade+zonal, batched tiles, serving)")
    ap.add_argument("--only", choices=["hillshade", "hillshade_zonal", "batched_tiles", "serve", "radial", "spread",
                             "hillshade_parts"],
                    help="run one extra section alone (no Life headline), print its JSON")
    ap.add_argument("--c3-forms", default="batched,plain",
                    help="forms of the batched-tiles section (batched, plain), comma-separated")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a U8 copy of the input raster (a local cast kernel), for counter calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from map_amd import map as M
    M.setupDevices("", M.DEV_GPU, "")
    if args.chunk_tiles:
        M.set_option("chunk_tiles", args.chunk_tiles)
    if args.queue >= 0:
        M.set_option("queue", args.queue)
    if args.only:
        fn = {"hillshade": hillshade_bench, "hillshade_zonal": hillshade_zonal_bench,
              "batched_tiles": batched_tiles_bench, "serve": serve_bench, "radial": radial_bench,
              "hillshade_parts": hillshade_parts_bench,
              "spread": spread_bench}[args.only]
        kw = {"forms": tuple(args.c3_forms.split(","))} if args.only == "batched_tiles" else {}
        if args.only in ("hillshade", "hillshade_zonal", "batched_tiles", "serve"):
            kw["cpu"] = not args.no_cpu
        print(json.dumps({args.only: fn(M, **kw)}), flush=True)
        return
    bs = (args.block, args.block)
    from map_amd import workloads as W
    s0 = W.life_input(args.side, 1)
    S = M.from_array(s0, bs)
    M.set_option("form", 1 if args.form == "batched" else 0)
    M.set_option("graph", args.graph)

    if args.copy_calib:
        cp = M.astype(S, M.U8)  # one read and one write of every cell

        def copies():
            for _ in range(max(1, args.steps)):
                M.release(cp)
                M.eval(cp)
        copies()
        M.set_option("graph", 0)  # launches replayed from a graph are not timed one by one
        kt = _kernel_profile(M, copies)  # the same launches again, each timed
        name, rec = max(kt.items(), key=lambda kv: kv[1]["total_ms"])
        avg = rec["total_ms"] / rec["count"]
        out = {"copy_calib_bytes_read": args.side * args.side, "copy_calib_bytes_written": args.side * args.side,
               "launches": 2 * max(1, args.steps), "kernel": name, "kernel_avg_ms": round(avg, 5),
               "gbs": round(2 * args.side * args.side / avg / 1e6, 1)}
        print(json.dumps(out), flush=True)
        return

    tuned = None
    if args.tile:
        tv = [int(t) for t in args.tile.split(",")]
        for key, val in zip(("tile_tx", "tile_ty", "tile_vx", "tile_pf", "tile_st", "tile_w"), tv):
            M.set_option(key, val)
        args.tune = 0
    if args.tune:
        # the tile shape is tuned on a 10-generation script over the same raster:
        # every generation is the same fused group, so the choice carries over
        tune_out = life_script(M, S, 10)
        rep = M.autotune(tune_out, reps=5)
        tuned = rep["chosen"]
        del tune_out

    out = life_script(M, S, args.gens)

    def step():
        M.release(out)
        M.eval(out)

    # inputs to HBM and kernel compile outside the timed region
    M.eval(out)
    for _ in range(args.warmup):
        step()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        dev_ms.append(M.last_eval_ms())
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    cells = args.side * args.side
    value = cells * args.gens / (wall / args.steps) / 1e9

    # roofline of the dominant kernel (one generation over all blocks), timed
    # with HIP events around each launch on the compute stream (one extra,
    # non-graph evaluation after the timed steps)
    M.set_option("graph", 0)
    M.set_option("profile", 1)
    M.kernel_times()
    step()
    kt = M.kernel_times()
    M.set_option("profile", 0)
    M.set_option("graph", args.graph)
    kname, krec = max(kt.items(), key=lambda kv: kv[1]["total_ms"])
    # every generation is one launch of one kernel: anything else means a
    # generation ran another kernel (or none) and the roofline would not be
    # the step's
    if len(kt) != 1 or krec["count"] != args.gens:
        raise SystemExit("bench.py: the profiled step launched %s, expected %d launches of one kernel"
                         % ({k: v["count"] for k, v in kt.items()}, args.gens))
    # HBM bytes per launch of this kernel, from rocprofv3 FETCH_SIZE / WRITE_SIZE
    # passes scaled by a copy kernel of known size (profiles/life_hbm_traffic.json)
    traffic = None
    try:
        with open(os.path.join(ROOT, "profiles", "life_hbm_traffic.json")) as f:
            tr = json.load(f)
        if kname in tr.get("kernels", {}):
            traffic = tr["kernels"][kname]["bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        pass
    launches = krec["count"]
    kern_ms = krec["total_ms"] / krec["count"]
    alg_bytes = 2.0 * cells  # one U8 read + one U8 write per cell and generation
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    host = M.stats()
    res = {
        "metric": "Life cell-generations/s (fused focal+local group, 200 generations)",
        "value": round(value, 3),
        "unit": "Gcell-gen/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "game-of-life %dx%d u8 Bernoulli(0.3) seed 1, %d generations, blocks %dx%d"
                               % (args.side, args.side, args.gens, bs[0], bs[1]),
                   "form": args.form, "graph": bool(args.graph)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": kname, "kernel_avg_ms": round(kern_ms, 5), "kernel_min_ms": round(krec["min_ms"], 5),
                     "launches": launches, "algorithmic_bytes_per_launch": int(alg_bytes)},
        "device_ms_per_step": round(min(dev_ms), 3),
        "tuned_tile": tuned,
        "host_ms": {"fuse": round(host["ms_fuse"], 1), "codegen_compile": round(host["ms_codegen"], 1),
                    "launch": round(host["ms_launch"], 1)},
    }
    if args.plain_gens > 0 and args.form == "batched":
        # the reference-shaped form: one launch per fused group per block
        small = life_script(M, S, args.plain_gens)
        M.set_option("graph", 0)
        M.set_option("form", 0)
        M.eval(small)
        M.release(small)
        t0 = time.perf_counter()
        M.eval(small)
        pw = time.perf_counter() - t0
        M.set_option("form", 1)
        M.release(small)
        t0 = time.perf_counter()
        M.eval(small)
        bw = time.perf_counter() - t0
        res["forms"] = {"generations": args.plain_gens,
                        "plain_gcellgen_s": round(cells * args.plain_gens / pw / 1e9, 3),
                        "batched_gcellgen_s": round(cells * args.plain_gens / bw / 1e9, 3),
                        "plain_launches": args.plain_gens * (args.side // args.block) ** 2}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(s0, bs)
    if not args.no_extras:
        M.release(out)
        M.set_option("graph", 0)
        extras = {}
        for name, fn in (("hillshade", hillshade_bench), ("hillshade_zonal", hillshade_zonal_bench),
                         ("batched_tiles", batched_tiles_bench), ("serve", serve_bench), ("radial", radial_bench),
                         ("spread", spread_bench)):
            t0 = time.perf_counter()
            try:
                extras[name] = fn(M, cpu=not args.no_cpu) if name in ("hillshade", "hillshade_zonal", "batched_tiles",
                                                                       "serve") else fn(M)
            except Exception as e:  # reported, the headline stands
                import traceback
                traceback.print_exc(file=sys.stderr)
                extras[name] = {"error": "%s: %s" % (type(e).__name__, e)}
            print("bench: %s done in %.1fs" % (name, time.perf_counter() - t0), file=sys.stderr, flush=True)
        res["extras"] = extras
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
