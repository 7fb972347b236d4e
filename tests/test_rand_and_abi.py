"""Rand (Philox2x32-10) oracle pins and the reference's ctypes binding block.

The oracle's Philox2x32-10 is checked against the published known-answer vectors of
Random123 (kat_vectors, philox2x32 with 10 rounds), the generator the reference's Rand
statement calls (runtime/skeleton/Skeleton.cpp:194-204); the per-type output mapping
against hand-worked cells. The library must resolve every entry point the reference's
python/map.py binds at import (tests/golden/bound_names.txt, written from
python/map.py:723-830 by tests/golden/make_bound_names.py).
"""
import ctypes as ct
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# Random123 kat_vectors, "philox2x32 10": counter (2 words), key (1 word) -> output
KAT = [((0x00000000, 0x00000000), 0x00000000, (0xff1dae59, 0x6cd10df2)),
       ((0xffffffff, 0xffffffff), 0xffffffff, (0x2c3f628b, 0xab4fd7ad)),
       ((0x243f6a88, 0x85a308d3), 0x13198a2e, (0xdd7ce038, 0xf62a4c12))]


def test_philox_known_answers():
    from oracle import ref as O
    for (c0, c1), k, want in KAT:
        r = O.philox2x32_10(np.array([c0], np.uint32), np.array([c1], np.uint32), np.array([k], np.uint32))
        assert (int(r[0][0]), int(r[1][0])) == want, (hex(c0), hex(k))


def test_philox_rounds_differ_from_nine():
    """A planted mistake (a missing round) fails the known answers: nine rounds of the same
    round function applied to the first vector's input give a different output."""
    m = 0xD256D193
    c0, c1, k = 0, 0, 0
    for _ in range(9):
        p = c0 * m
        c0, c1 = ((p >> 32) ^ k ^ c1) & 0xFFFFFFFF, p & 0xFFFFFFFF
        k = (k + 0x9E3779B9) & 0xFFFFFFFF
    assert (c0, c1) != KAT[0][2]


def test_rand_cell_mapping_by_hand():
    """Cell (0, 0) of a zero seed raster has counter {0, 0} and key 0 (the first KAT vector):
    U32 takes the first word, U8/S8 its low byte, S16 its low 16 bits (two's complement),
    U64 both words (first word low), F32 first word / UINT_MAX, F64 both words / ULONG_MAX;
    cell (1, 0) has key 1, cell (0, 1) key = width."""
    from oracle import ref as O
    seed = O.from_array(np.zeros((3, 5), np.int32), (4, 4))
    w0, w1 = KAT[0][2]
    assert int(O.rand(seed, O.U32).val[0, 0]) == w0
    assert int(O.rand(seed, O.U8).val[0, 0]) == w0 & 0xFF
    assert int(O.rand(seed, O.S8).val[0, 0]) == (w0 & 0xFF) - (256 if w0 & 0x80 else 0)
    assert int(O.rand(seed, O.S16).val[0, 0]) == (w0 & 0xFFFF) - (65536 if w0 & 0x8000 else 0)
    assert int(O.rand(seed, O.U64).val[0, 0]) == w0 | (w1 << 32)
    assert O.rand(seed, O.F32).val[0, 0] == np.float32(np.float32(w0) / np.float32(4294967295.0))
    assert O.rand(seed, O.F64).val[0, 0] == float(w0 | (w1 << 32)) / 18446744073709551615.0
    one = O.philox2x32_10(np.array([0], np.uint32), np.array([0], np.uint32), np.array([1], np.uint32))[0][0]
    five = O.philox2x32_10(np.array([0], np.uint32), np.array([0], np.uint32), np.array([5], np.uint32))[0][0]
    r = O.rand(seed, O.U32).val
    assert r[0, 1] == one and r[1, 0] == five
    seeded = O.rand(O.from_array(np.full((3, 5), 0x243f6a88, np.int64).astype(np.int32), (4, 4)), O.U32).val
    want = O.philox2x32_10(np.array([0x243f6a88], np.uint32), np.array([0], np.uint32), np.array([7], np.uint32))[0][0]
    assert seeded[1, 2] == want
    f = O.rand(seed, O.F32).val
    assert f.dtype == np.float32 and np.all((f >= 0) & (f <= 1))


def test_reference_binding_block_resolves(M):
    """Every name python/map.py binds resolves in libmapx.so, with the argument and return
    types the reference declares (python/map.py:723-830), set the way it sets them."""
    names = open(os.path.join(ROOT, "tests", "golden", "bound_names.txt")).read().split()
    assert len(names) == 35
    lib = ct.CDLL(os.path.join(ROOT, "map_amd", "libmapx.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    node = ct.c_void_p
    lib.ma_rand.argtypes = [node, ct.c_int, ct.c_int]
    lib.ma_rand.restype = node
    lib.ma_boundedNbh.argtypes = [node, node, node]
    lib.ma_boundedNbh.restype = node
    for n in ("ma_loopStart", "ma_loopBody", "ma_loopAgain", "ma_loopEnd"):
        getattr(lib, n).argtypes = []
        getattr(lib, n).restype = None
    lib.ma_loopCond.argtypes = [node]
    lib.ma_loopAssemble.restype = node
    lib.ma_loopAgainTail.argtypes = [node, ct.POINTER(ct.POINTER(node)), ct.POINTER(ct.POINTER(node)),
                                     ct.POINTER(ct.c_int)]
    lib.ma_loopAgainTail.restype = None


def test_unsupported_entry_points_fail_without_aborting(M):
    """The symbolic-loop calls and ma_boundedNbh report through ma_lastError (no abort)."""
    lib = M._lib
    a = M.from_array(np.zeros((8, 8), np.int32), (8, 8))
    lib.ma_loopStart()
    assert "symbolic loops" in lib.ma_lastError().decode()
    assert not lib.ma_loopAssemble()
    agains, tails, num = ct.POINTER(ct.c_void_p)(), ct.POINTER(ct.c_void_p)(), ct.c_int(7)
    lib.ma_loopAgainTail.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
    lib.ma_loopAgainTail(None, ct.byref(agains), ct.byref(tails), ct.byref(num))
    assert num.value == 0
    lib.ma_boundedNbh.argtypes = [M.Raster, M.Raster, M.Raster]
    lib.ma_boundedNbh.restype = ct.c_void_p
    assert not lib.ma_boundedNbh(a, a, a)
    assert "boundedNbh" in lib.ma_lastError().decode()


def test_rand_plans_and_compiles(M):
    """rand(seed) is a LOCAL node that fuses with its consumer (one kernel)."""
    seed = M.from_array(np.arange(40 * 30, dtype=np.int32).reshape(30, 40), (16, 16))
    r = M.rand(seed, dt=M.F32) * 2.0 + M.astype(M.rand(seed, dt=M.U8), M.F32)
    groups = [g for g in M.plan(r)["groups"] if "kernel" in g]
    assert len(groups) == 1 and groups[0]["kinds"].count("Rand") == 2
    assert "mx::rand_cell<float>" in groups[0]["source"]
    assert M.compile_plan(r) >= 1
