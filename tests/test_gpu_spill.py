"""Resident rasters under an HBM cap (option hbm_cap): when the raster store has no room,
the least recently used rasters that no running evaluation lists leave HBM (computed and
generated ones through pinned host memory, inputs with a host or file source by dropping
them) and come back when read. The results must be the bits of the same evaluations run
without the cap; an evaluation that alone needs more than the cap fails with an error
that names it, and the runtime goes on working.
"""

import gc

import numpy as np
import pytest

from map_amd import workloads as W
from tests.parity import assert_exact

pytestmark = pytest.mark.gpu

MIB = 1 << 20


def _inputs(G, n, w, h, bs):
    rng = np.random.Generator(np.random.PCG64(81))
    dems = [G.gen_dem((w, h), bs, W.dem_gaussians(rng, w, h), 900 + i) for i in range(n)]
    host = rng.uniform(-100, 100, (h, w)).astype(np.float32)
    return dems, host


def test_rasters_spill_to_host_memory_under_an_hbm_cap(G):
    w, h, bs = 1024, 768, (256, 256)  # F32: 3 MiB per raster
    print("case spill: 8 generated DEMs %dx%d, blocks %s" % (w, h, bs), flush=True)
    gc.collect()
    dems, host = _inputs(G, 8, w, h, bs)
    X = G.from_array(host, bs)
    slopes = [W.slope(G, d) for d in dems[:4]]
    G.eval(*slopes)  # computed rasters kept referenced
    hs_ref = [G.to_array(W.hillshade(G, d)) for d in dems]
    sl_ref = [G.to_array(s) for s in slopes]
    x_ref = G.to_array(X * 2.0 + 1.0)
    # 20 MiB: room for one evaluation (a DEM and its hillshade) and a few held rasters; the
    # 8 DEMs, 4 slopes and the host input (39 MiB) and what earlier tests hold move out
    cap = 20 * MIB
    G.set_option("hbm_cap", cap)
    try:
        s0 = G.stats()
        hs = [G.to_array(W.hillshade(G, d)) for d in dems]
        sl = [G.to_array(s) for s in slopes]
        xx = G.to_array(X * 2.0 + 1.0)
        for i in range(len(dems)):
            assert_exact(hs[i], hs_ref[i], "hillshade of DEM %d under the cap" % i)
        for i in range(len(slopes)):
            assert_exact(sl[i], sl_ref[i], "held slope %d under the cap" % i)
        assert_exact(xx, x_ref, "host input under the cap")
        # evaluating a spilled raster again takes it as held (nothing recomputed or re-read)
        l0 = G.stats()["launches"]
        G.eval(*slopes)
        assert G.stats()["launches"] == l0
        assert_exact(G.to_array(slopes[1]), sl_ref[1], "held slope evaluated again under the cap")
        # the generated DEMs and the slopes read again after spilling: every one back
        again = [G.to_array(W.hillshade(G, d)) for d in reversed(dems)]
        for i, a in enumerate(reversed(again)):
            assert_exact(a, hs_ref[i], "hillshade of DEM %d, second pass" % i)
        s1 = G.stats()
        print("  spills %d, unspills %d, dropped %d, in use %d + cached %d, cap %d"
              % (s1["spills"] - s0["spills"], s1["unspills"] - s0["unspills"],
                 s1["dropped_for_space"] - s0["dropped_for_space"], s1["store_in_use"], s1["store_cached"], cap),
              flush=True)
        assert s1["spills"] > s0["spills"] and s1["unspills"] > s0["unspills"]
        assert s1["store_in_use"] + s1["store_cached"] <= cap
        # one evaluation larger than the cap: a clear error, nothing half-computed kept
        big = G.gen_dem((2048, 2048), bs, W.dem_gaussians(np.random.Generator(np.random.PCG64(5)), 2048, 2048), 7)
        with pytest.raises(G.MapError, match="HBM cap"):
            G.eval(W.hillshade(G, big * 1.0) + W.slope(G, big * 2.0) + W.aspect(G, big * 3.0) + big * 4.0)
    finally:
        G.set_option("hbm_cap", 0)
    assert_exact(G.to_array(W.hillshade(G, dems[0])), hs_ref[0], "hillshade after the cap is lifted")
    assert_exact(G.to_array(slopes[3]), sl_ref[3], "held slope after the cap is lifted")
