"""ma_genDemNoAlloc, the allocation-free variant of ma_genDem: called through the C ABI
into device memory the caller owns, twice in a row (the second raster smaller, so the
scratch keeps the first call's table), each result bit-for-bit equal to ma_genDem's; a
table larger than the scratch is refused with an error."""
import ctypes as ct

import numpy as np
import pytest

from map_amd import workloads as W

pytestmark = pytest.mark.gpu

HIP = "/opt/rocm/lib/libamdhip64.so"


def _nbytes(ds, bs):
    nb0, nb1 = -(-ds[0] // bs[0]), -(-ds[1] // bs[1])
    return nb0 * nb1 * bs[0] * bs[1] * 4


def test_no_alloc_dem_equals_allocating_entry_point(G):
    hip = ct.CDLL(HIP)
    lib = G._lib
    lib.ma_genDemNoAlloc.argtypes = [ct.c_void_p, ct.c_size_t, G.Array, G.Array, ct.POINTER(ct.c_double), ct.c_int,
                                     ct.c_ulonglong]
    lib.ma_genDemNoAlloc.restype = ct.c_int
    cases = [((1500, 1100), (512, 256), 64, 5), ((300, 200), (128, 128), 9, 6)]
    cap = max(_nbytes(ds, bs) for ds, bs, _, _ in cases)
    buf = ct.c_void_p()
    assert hip.hipMalloc(ct.byref(buf), ct.c_size_t(cap)) == 0  # the caller's memory
    try:
        for ds, bs, ng, seed in cases:
            print("case no-alloc dem", ds, bs, ng, seed, flush=True)
            rng = np.random.Generator(np.random.PCG64(seed))
            g = np.ascontiguousarray(W.dem_gaussians(rng, ds[0], ds[1], n=ng), dtype=np.float64)
            n = _nbytes(ds, bs)
            rc = lib.ma_genDemNoAlloc(buf, ct.c_size_t(n), G.Array(list(ds)), G.Array(list(bs)),
                                      g.ctypes.data_as(ct.POINTER(ct.c_double)), ng, ct.c_ulonglong(seed))
            assert rc == 0, lib.ma_lastError().decode()
            got = np.empty(n // 4, np.float32)
            assert hip.hipMemcpy(got.ctypes.data_as(ct.c_void_p), buf, ct.c_size_t(n), 2) == 0  # device to host
            ref = G.gen_dem(ds, bs, g, seed)
            want = np.empty(n // 4, np.float32)
            assert lib.ma_fetch(ref, want.ctypes.data_as(ct.c_void_p), ct.c_size_t(n)) == 0
            assert got.tobytes() == want.tobytes()
            del ref
        g = np.zeros((257, 4))
        rc = lib.ma_genDemNoAlloc(buf, ct.c_size_t(cap), G.Array([64, 64]), G.Array([64, 64]),
                                  g.ctypes.data_as(ct.POINTER(ct.c_double)), 257, ct.c_ulonglong(1))
        assert rc != 0 and "at most 256 Gaussians" in lib.ma_lastError().decode()
        rc = lib.ma_genDemNoAlloc(buf, ct.c_size_t(100), G.Array([64, 64]), G.Array([64, 64]),
                                  g.ctypes.data_as(ct.POINTER(ct.c_double)), 4, ct.c_ulonglong(1))
        assert rc != 0 and "bytes of device memory" in lib.ma_lastError().decode()
    finally:
        assert hip.hipFree(buf) == 0
