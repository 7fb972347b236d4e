"""CPU oracle evaluator of map-algebra scripts. TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py as the checker; the product path never imports it.

It mirrors the API of map_amd.map (convolve, atan, sqrt, atan2, sin, cos,
con, zsum, ... and the Raster operators) but evaluates eagerly on whole
numpy rasters with the reference's semantics:

* node types: util/promote.hpp:22-157 (restated independently below),
  relational/logical ops give U8 (runtime/dag/Binary.cpp:53-57,
  Unary.cpp:40-41), convolution type promote(input, mask type)
  (Convolution.cpp:37), mask type = type of the first weight (util/Mask.cpp:18);
* each statement in C arithmetic: operands converted to C's common type,
  the result converted to the node type (runtime/skeleton/Skeleton.cpp:179-257);
* python float constants are F32 (rounded by c_float), ints S32
  (python/map.py:111-127); float mask weights pass through "%f"
  (util/VariantType.cpp:91, skeleton/util.cpp:136-165);
* focal ops with the block-relative border rule (oracle.c, load_F of
  skeleton/util.cpp:492-506).

Two modes: "f32" (reference semantics: FP32 nodes rounded to FP32 after every
statement; transcendental functions evaluated in FP64 and rounded once, i.e.
a correctly rounded FP32 libm) and "f64" (every floating node computed in
FP64 from the same FP32 inputs and constants: the exact-arithmetic value of
the script, the FP64 oracle).
"""
import ctypes as ct
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            from . import build as _b
            _b.build()
        _LIB = ct.CDLL(path)
        i64p = ct.POINTER(ct.c_int64)
        _LIB.or_tap_index.argtypes = [ct.c_int] * 6 + [i64p]
        _LIB.or_conv_f32.argtypes = [ct.c_void_p, i64p, ct.c_int64, ct.c_int, ct.c_void_p, ct.c_void_p]
        _LIB.or_conv_f64.argtypes = [ct.c_void_p, i64p, ct.c_int64, ct.c_int, ct.c_void_p, ct.c_void_p]
        _LIB.or_conv_i64.argtypes = [ct.c_void_p, i64p, ct.c_int64, ct.c_int, ct.c_void_p, ct.c_void_p]
        _LIB.or_focal_f64.argtypes = [ct.c_void_p, i64p, ct.c_int64, ct.c_int, ct.c_int, ct.c_void_p]
        _LIB.or_life.argtypes = [ct.c_void_p, ct.c_void_p] + [ct.c_int] * 5
        _LIB.or_zstats_f32.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int, ct.c_void_p, ct.c_void_p,
                                       ct.c_void_p, ct.c_void_p]
        _LIB.or_dem.argtypes = [ct.c_void_p, ct.c_int64, ct.c_int64, ct.c_int, ct.c_int, ct.c_void_p, ct.c_int,
                                ct.c_uint64]
        _LIB.or_voronoi.argtypes = [ct.c_void_p, ct.c_int64, ct.c_int64, ct.c_int, ct.c_int, ct.c_void_p,
                                    ct.c_void_p, ct.c_int]
        _LIB.or_src.argtypes = [ct.c_int] * 8 + [ct.POINTER(ct.c_int), ct.POINTER(ct.c_int)]
    return _LIB


# type codes (python/map.py:35-36)
F32, F64, B8, U8, U16, U32, U64, S8, S16, S32, S64 = range(1, 12)
SUM, PROD, rAND, rOR, MAX, MIN = 1, 2, 3, 4, 6, 7
NP = {F32: np.float32, F64: np.float64, B8: np.uint8, U8: np.uint8, U16: np.uint16, U32: np.uint32,
      U64: np.uint64, S8: np.int8, S16: np.int16, S32: np.int32, S64: np.int64}
CODE = {np.dtype(np.float32): F32, np.dtype(np.float64): F64, np.dtype(np.uint8): U8, np.dtype(np.uint16): U16,
        np.dtype(np.uint32): U32, np.dtype(np.uint64): U64, np.dtype(np.int8): S8, np.dtype(np.int16): S16,
        np.dtype(np.int32): S32, np.dtype(np.int64): S64}
FLOATS = (F32, F64)
UNSIGNED = (B8, U8, U16, U32, U64)

_MODE = ["f32"]


def set_mode(m):
    assert m in ("f32", "f64")
    _MODE[0] = m


def mode():
    return _MODE[0]


# util/promote.hpp:22-157 as a rank table: floats dominate; otherwise the
# table's rules (restated pairwise, including its irregular entries).

_PROMOTE = {
    F32: [F32, F64, F32, F32, F32, F32, F32, F32, F32, F32, F32],
    F64: [F64] * 11,
    B8: [F32, F64, B8, U8, U16, U32, U64, S8, S16, S32, S64],
    U8: [F32, F64, U8, U8, U16, U32, U64, S8, S16, S32, S64],
    U16: [F32, F64, U16, U16, U16, U32, U64, U16, S16, S32, S64],
    U32: [F32, F64, U32, U32, U32, U32, U64, S32, S32, S32, S64],
    U64: [F32, F64, B8, U64, U64, U64, U64, S64, S64, S64, S64],
    S8: [F32, F64, S8, S8, S16, S32, S64, S8, S16, S32, S64],
    S16: [F32, F64, S16, S16, S16, S32, S64, S16, S16, S32, S64],
    S32: [F32, F64, S32, S32, S32, S32, S64, S32, S32, S32, S64],
    S64: [F32, F64, S64, S64, S64, S64, S64, S64, S64, S64, S64],
}


def promote(a, b):
    return _PROMOTE[a][b - 1]


# C's integer promotion and usual arithmetic conversions (C11 6.3.1.1, 6.3.1.8)
def _c_promoted(dt):
    if dt in (B8, U8, S8, U16, S16, S32):
        return np.int32
    return {U32: np.uint32, S64: np.int64, U64: np.uint64, F32: np.float32, F64: np.float64}[dt]


def _c_common(a, b):
    pa, pb = np.dtype(_c_promoted(a)), np.dtype(_c_promoted(b))
    if np.float64 in (pa, pb):
        return np.float64
    if np.float32 in (pa, pb):
        return np.float32
    if pa == pb:
        return pa.type
    ua, ub = pa.kind == 'u', pb.kind == 'u'
    if ua == ub:
        return (pa if pa.itemsize >= pb.itemsize else pb).type
    u, s = (pa, pb) if ua else (pb, pa)
    if u.itemsize >= s.itemsize:
        return u.type
    return s.type


def _work(np_t):
    """Storage type of a C type in the current mode (FP64 mode widens floats)."""
    if _MODE[0] == "f64" and np.dtype(np_t).kind == 'f':
        return np.float64
    return np_t


def _conv(v, dt):
    """C conversion of values v to the node type dt."""
    t = _work(NP[dt])
    v = np.asarray(v)
    if dt == B8:
        return (v != 0).astype(np.uint8)
    if np.dtype(t).kind in 'iu' and v.dtype.kind == 'f':
        with np.errstate(invalid='ignore'):
            return np.trunc(v).astype(np.int64).astype(t)
    with np.errstate(over='ignore'):
        return v.astype(t)


class R:
    """An oracle raster (2D, rows = y) or D0 value with its reference type."""

    def __init__(self, val, dt, bs=None):
        self.val = val
        self.dt = dt
        self.bs = bs

    @property
    def d2(self):
        return self.bs is not None

    def datasize(self):
        return [self.val.shape[1], self.val.shape[0]]

    def blocksize(self):
        return list(self.bs)

    def datatype(self):
        return self.dt

    def __bool__(self):
        raise TypeError("oracle rasters are not booleans")

    def __hash__(self):
        return id(self)

    def __pow__(self, o):
        return binary(self, o, 'POW')

    def __rpow__(self, o):
        return binary(o, self, 'POW')

    def astype(self, dt):
        return astype(self, dt)

    def __neg__(self):
        return unary(self, 'NEG')

    def __pos__(self):
        return unary(self, 'POS')

    def __invert__(self):
        return unary(self, 'bNOT')

    def no(self):
        return unary(self, 'NOT')


for _op, _name in [('ADD', 'add'), ('SUB', 'sub'), ('MUL', 'mul'), ('DIV', 'truediv'), ('MOD', 'mod'),
                   ('EQ', 'eq'), ('NE', 'ne'), ('LT', 'lt'), ('GT', 'gt'), ('LE', 'le'), ('GE', 'ge'),
                   ('bAND', 'and'), ('bOR', 'or'), ('bXOR', 'xor'), ('SHL', 'lshift'), ('SHR', 'rshift')]:
    def _mk(op):
        return (lambda self, o: binary(self, o, op)), (lambda self, o: binary(o, self, op))
    _f, _rf = _mk(_op)
    setattr(R, '__%s__' % _name, _f)
    setattr(R, '__r%s__' % _name, _rf)


def const(v):
    if isinstance(v, R):
        return v
    if isinstance(v, float):
        return R(_work(np.float32)(np.float32(v)), F32)
    if isinstance(v, bool):
        return R(np.uint8(v), B8)
    if isinstance(v, int):
        return R(np.int32(v), S32)
    raise TypeError(v)


def from_array(a, bs):
    a = np.asarray(a)
    dt = CODE[a.dtype]
    return R(a.astype(_work(a.dtype.type)), dt, tuple(bs))


def _shape_of(*xs):
    bs = None
    for x in xs:
        if x.d2:
            bs = x.bs
    return bs


def _cdiv(a, b):
    q = a // b
    adj = (a % b != 0) & ((a < 0) ^ (b < 0))
    return q + adj.astype(q.dtype)


def binary(l, r, op):
    l, r = const(l), const(r)
    bs = _shape_of(l, r)
    rel = op in ('EQ', 'NE', 'LT', 'GT', 'LE', 'GE', 'AND', 'OR')
    dt = U8 if rel else promote(l.dt, r.dt)
    if op in ('MAX2', 'MIN2', 'ATAN2', 'POW', 'HYPOT', 'FMOD'):
        a, b = _conv(l.val, dt), _conv(r.val, dt)
        if dt in FLOATS:
            a64, b64 = a.astype(np.float64), b.astype(np.float64)
            f = {'MAX2': np.fmax, 'MIN2': np.fmin, 'ATAN2': np.arctan2, 'POW': np.power, 'HYPOT': np.hypot,
                 'FMOD': np.fmod}[op]
            with np.errstate(all='ignore'):
                res = f(a64, b64)
            return R(_conv(res, dt), dt, bs)
        if op in ('MAX2', 'MIN2'):
            return R((np.maximum if op == 'MAX2' else np.minimum)(a, b), dt, bs)
        f = {'ATAN2': np.arctan2, 'POW': np.power, 'HYPOT': np.hypot, 'FMOD': np.fmod}[op]
        return R(_conv(f(a.astype(np.float64), b.astype(np.float64)), dt), dt, bs)
    ct_ = _work(_c_common(l.dt, r.dt))
    a = np.asarray(l.val).astype(ct_)
    b = np.asarray(r.val).astype(ct_)
    with np.errstate(all='ignore'):
        if op == 'ADD':
            res = a + b
        elif op == 'SUB':
            res = a - b
        elif op == 'MUL':
            res = a * b
        elif op == 'DIV':
            res = a / b if np.dtype(ct_).kind == 'f' else _cdiv(a, b)
        elif op == 'MOD':
            res = np.fmod(a, b) if np.dtype(ct_).kind == 'f' else a - b * _cdiv(a, b)
        elif op in ('EQ', 'NE', 'LT', 'GT', 'LE', 'GE'):
            res = {'EQ': a == b, 'NE': a != b, 'LT': a < b, 'GT': a > b, 'LE': a <= b, 'GE': a >= b}[op]
        elif op == 'AND':
            res = (a != 0) & (b != 0)
        elif op == 'OR':
            res = (a != 0) | (b != 0)
        elif op == 'bAND':
            res = a & b
        elif op == 'bOR':
            res = a | b
        elif op == 'bXOR':
            res = a ^ b
        elif op == 'SHL':
            res = a << b
        elif op == 'SHR':
            res = a >> b
        else:
            raise ValueError(op)
    if _MODE[0] == "f32" and np.dtype(ct_) == np.float32:
        res = np.asarray(res).astype(np.float32)
    return R(_conv(res, dt), dt, bs)


_UFUN = {'SIN': np.sin, 'COS': np.cos, 'TAN': np.tan, 'ASIN': np.arcsin, 'ACOS': np.arccos, 'ATAN': np.arctan,
         'SINH': np.sinh, 'COSH': np.cosh, 'TANH': np.tanh, 'ASINH': np.arcsinh, 'ACOSH': np.arccosh,
         'ATANH': np.arctanh, 'EXP': np.exp, 'EXP2': np.exp2, 'EXP10': lambda x: np.power(10.0, x),
         'LOG': np.log, 'LOG2': np.log2, 'LOG10': np.log10, 'SQRT': np.sqrt, 'CBRT': np.cbrt, 'ABS': np.abs,
         'CEIL': np.ceil, 'FLOOR': np.floor, 'TRUNC': np.trunc, 'ROUND': lambda x: np.trunc(x + np.copysign(0.5, x))}


def unary(x, op):
    x = const(x)
    dt = U8 if op == 'NOT' else x.dt
    if op in ('POS', 'NEG', 'bNOT'):
        p = np.asarray(x.val).astype(_work(_c_promoted(x.dt)))
        with np.errstate(over='ignore'):
            res = p if op == 'POS' else (-p if op == 'NEG' else ~p)
        return R(_conv(res, dt), dt, x.bs)
    if op == 'NOT':
        return R((np.asarray(x.val) == 0).astype(np.uint8), U8, x.bs)
    if op == 'ABS' and dt not in FLOATS:
        p = np.asarray(x.val).astype(_c_promoted(x.dt))
        return R(_conv(np.abs(p), dt), dt, x.bs)
    with np.errstate(all='ignore'):
        res = _UFUN[op](np.asarray(x.val).astype(np.float64))
    return R(_conv(res, dt), dt, x.bs)


def astype(x, dt):
    return R(_conv(x.val, dt), dt, x.bs)


def philox2x32_10(c0, c1, key):
    """Philox2x32-10 (Salmon, Moraes, Dror and Shaw, SC 2011; Random123, the generator the
    reference's Rand statement calls, runtime/skeleton/Skeleton.cpp:194-197) on uint32 arrays:
    ten rounds of (c0, c1) <- (mulhi(M, c0) ^ key ^ c1, mullo(M, c0)), M = 0xD256D193, key
    bumped by 0x9E3779B9 between rounds. Returns (c0, c1)."""
    m = np.uint64(0xD256D193)
    c0 = np.asarray(c0, np.uint32).astype(np.uint64)
    c1 = np.asarray(c1, np.uint32).astype(np.uint64)
    k = np.asarray(key, np.uint32).astype(np.uint64)
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        prod = c0 * m
        c0, c1 = ((prod >> np.uint64(32)) ^ k ^ c1) & mask, prod & mask
        k = (k + np.uint64(0x9E3779B9)) & mask
    return c0.astype(np.uint32), c1.astype(np.uint32)


def rand(seed, dt=None):
    """Rand (runtime/dag/Rand.cpp, Skeleton.cpp:183-206): per cell Philox2x32-10 of counter
    {seed cell, 0} and key {y * width + x}; integer types take the low bytes of the output
    (both words for 8-byte types), F32 the first word / UINT_MAX, F64 both words / ULONG_MAX."""
    dt = seed.dt if dt is None else dt
    h, w = seed.val.shape
    idx = (np.arange(h, dtype=np.uint64)[:, None] * np.uint64(w) + np.arange(w, dtype=np.uint64)[None, :])
    sv = np.asarray(seed.val).astype(np.int64).astype(np.uint64) & np.uint64(0xFFFFFFFF)
    c0, c1 = philox2x32_10(sv.astype(np.uint32), np.zeros_like(sv, np.uint32),
                           (idx & np.uint64(0xFFFFFFFF)).astype(np.uint32))
    wide = c0.astype(np.uint64) | (c1.astype(np.uint64) << np.uint64(32))
    if dt == F32:
        v = (c0.astype(np.float32) / np.float32(4294967295.0)).astype(np.float32)
    elif dt == F64:
        v = wide.astype(np.float64) / np.float64(18446744073709551615.0)
    elif NP[dt]().itemsize == 8:
        v = wide.view(np.int64).astype(NP[dt]) if dt == S64 else wide
    else:
        v = c0.astype(NP[dt])
    return R(np.asarray(v, NP[dt]), dt, seed.bs)


def con(c, l, r):
    c, l, r = const(c), const(l), const(r)
    if l.dt != r.dt:
        raise TypeError("conditional branches differ in type")
    bs = _shape_of(c, l, r)
    return R(_conv(np.where(np.asarray(c.val) != 0, l.val, r.val), l.dt), l.dt, bs)


def _quantize(w):
    """A mask weight as the reference's kernel source spells it (std::to_string, %f)."""
    if isinstance(w, float):
        return float(np.float32(float("%f" % float(np.float32(w)))))
    return w


_TAPCACHE = {}


def taps(ds, bs, h0, h1):
    key = (tuple(ds), tuple(bs), h0, h1)
    if key not in _TAPCACHE:
        nt = (2 * h0 + 1) * (2 * h1 + 1)
        idx = np.empty(ds[0] * ds[1] * nt, dtype=np.int64)
        lib().or_tap_index(ds[0], ds[1], bs[0], bs[1], h0, h1, idx.ctypes.data_as(ct.POINTER(ct.c_int64)))
        if len(_TAPCACHE) > 8:
            _TAPCACHE.clear()
        _TAPCACHE[key] = (idx, nt)
    return _TAPCACHE[key]


def convolve(x, lst):
    if not x.d2:
        raise TypeError("convolve needs a raster")
    ncol, nrow = len(lst[0]), len(lst)
    flat = [e for row in lst for e in row]
    first = flat[0]
    mdt = F32 if isinstance(first, float) else (B8 if isinstance(first, bool) else S32)
    dt = promote(x.dt, mdt)
    h0, h1 = ncol // 2, nrow // 2
    ds = x.datasize()
    idx, nt = taps(ds, x.bs, h0, h1)
    n = ds[0] * ds[1]
    L = lib()
    i64p = ct.POINTER(ct.c_int64)
    if dt in FLOATS:
        if _MODE[0] == "f64" or dt == F64:
            w = np.array([float(_quantize(e)) if mdt == F32 else float(e) for e in flat], np.float64)
            src = np.ascontiguousarray(np.asarray(x.val, np.float64).reshape(-1))
            out = np.empty(n, np.float64)
            L.or_conv_f64(src.ctypes.data, idx.ctypes.data_as(i64p), n, nt, w.ctypes.data, out.ctypes.data)
        else:
            w = np.array([_quantize(e) if mdt == F32 else e for e in flat], np.float32)
            src = np.ascontiguousarray(np.asarray(x.val).astype(np.float32).reshape(-1))
            out = np.empty(n, np.float32)
            L.or_conv_f32(src.ctypes.data, idx.ctypes.data_as(i64p), n, nt, w.ctypes.data, out.ctypes.data)
        return R(_conv(out.reshape(ds[1], ds[0]), dt), dt, x.bs)
    w = np.array([int(e) for e in flat], np.int64)
    src = np.ascontiguousarray(np.asarray(x.val).astype(np.int64).reshape(-1))
    out = np.empty(n, np.int64)
    L.or_conv_i64(src.ctypes.data, idx.ctypes.data_as(i64p), n, nt, w.ctypes.data, out.ctypes.data)
    return R(_conv(out.reshape(ds[1], ds[0]), dt), dt, x.bs)


def neighbor(x, coord):
    dx, dy = coord
    h0, h1 = abs(dx), abs(dy)
    ds = x.datasize()
    idx, nt = taps(ds, x.bs, h0, h1)
    t = (dy + h1) * (2 * h0 + 1) + (dx + h0)
    src = np.asarray(x.val).reshape(-1)
    return R(src[idx.reshape(-1, nt)[:, t]].reshape(ds[1], ds[0]), x.dt, x.bs)


def _focal(x, lst, red):
    """FocalFunc (runtime/skeleton/FocalSkeleton.cpp:338-372): from the reduction's neutral
    value, OP over the window of the mask's size (y outer, x inner, border rule) in the
    raster's type; the mask's values are not used (quirk B3). Floating MAX / MIN skip NaN
    (fmax / fmin: the reference's OpenCL max / min leave NaN undefined)."""
    ncol, nrow = len(lst[0]), len(lst)
    h0, h1 = ncol // 2, nrow // 2
    ds = x.datasize()
    idx, nt = taps(ds, x.bs, h0, h1)
    src = np.asarray(x.val).reshape(-1)
    g = src[idx.reshape(-1, nt)]
    acc = np.full(g.shape[0], _neutral(red, x.dt), dtype=src.dtype)
    for t in range(nt):
        acc = np.asarray(_red_step(acc, g[:, t], red, x.dt)).astype(src.dtype)
    return R(acc.reshape(ds[1], ds[0]), x.dt, x.bs)


def ffunc(x, lst, red):
    return _focal(x, lst, red)


AGE, ILE = 1, 2
VARI, MAJO, MINO, MEAN = 1, 2, 3, 4


def focal_percent(x, lst, kind):
    """FocalPercent (runtime/skeleton/FocalSkeleton.cpp:375-403): count, over the window of the
    mask's size (y outer, x inner, border rule), of centre == cell (AGE) or centre > cell (ILE),
    accumulated in the raster's type, then divided by the reference's nbh_size
    (skeleton/util.cpp:321-327): "(H0*2+1)*H1*2+1", i.e. (2*h0+1)*h1*2+1 for a node alone
    (C precedence; 7 for a 3x3 window)."""
    ncol, nrow = len(lst[0]), len(lst)
    h0, h1 = ncol // 2, nrow // 2
    ds = x.datasize()
    idx, nt = taps(ds, x.bs, h0, h1)
    src = np.asarray(x.val).reshape(-1)
    g = src[idx.reshape(-1, nt)]
    with np.errstate(invalid="ignore"):
        hit = (g == src[:, None]) if kind == AGE else (src[:, None] > g)
    cnt = hit.sum(axis=1)  # at most 2*h0+1 by 2*h1+1: exact in every type
    nbh = (2 * h0 + 1) * h1 * 2 + 1
    if x.dt in FLOATS:
        T = src.dtype.type
        res = (cnt.astype(T) / T(nbh)).astype(T)
    else:
        res = _conv(cnt // nbh, x.dt)  # non-negative: C division = floor
    return R(np.asarray(res).reshape(ds[1], ds[0]).astype(src.dtype), x.dt, x.bs)


def _c_int(v, dt):
    """A Python integer wrapped to the C type dt."""
    bits = np.dtype(NP[dt]).itemsize * 8
    v &= (1 << bits) - 1
    if dt in (S8, S16, S32, S64) and v >= 1 << (bits - 1):
        v -= 1 << bits
    return v


def diversity(xs, kind):
    """Diversity (runtime/skeleton/Skeleton.cpp:259-288, helpers skeleton/util.cpp:439-481) per
    cell over the rasters in order: distinct values in order of first appearance (== equality)
    with counts; VARI = their number (S8), MAJO / MINO = the first most / least frequent,
    MEAN = (sum of value * count, accumulated in the type with C conversions) / number of
    distinct values. Per-cell Python loop: small rasters only."""
    vals = [np.asarray(x.val).reshape(-1) for x in xs]
    dt = xs[0].dt
    T = vals[0].dtype.type
    n = vals[0].size
    out = np.empty(n, np.int8 if kind == VARI else vals[0].dtype)
    fl = dt in FLOATS
    for i in range(n):
        el, cnt = [], []
        for v in vals:
            e = v[i]
            for k in range(len(el)):
                if e == el[k]:
                    cnt[k] += 1
                    break
            else:
                el.append(e)
                cnt.append(1)
        if kind == VARI:
            out[i] = len(el)
        elif kind in (MAJO, MINO):
            m = 0
            for k in range(1, len(el)):
                if (cnt[k] > cnt[m]) if kind == MAJO else (cnt[k] < cnt[m]):
                    m = k
            out[i] = el[m]
        elif fl:
            mean = T(0)
            for e, c in zip(el, cnt):
                mean = T(mean + T(e * T(c)))
            out[i] = T(mean / T(len(el)))
        else:
            mean = 0
            for e, c in zip(el, cnt):
                mean = _c_int(mean + int(e) * c, dt)
            q = abs(mean) // len(el)
            out[i] = _c_int(q if mean >= 0 else -q, dt)
    return R(out.reshape(np.asarray(xs[0].val).shape), S8 if kind == VARI else dt, xs[0].bs)


def fmin(x, lst):
    return _focal(x, lst, MIN)


def fmax(x, lst):
    return _focal(x, lst, MAX)


def fsum(x):
    return convolve(x, [[1, 1, 1], [1, 1, 1], [1, 1, 1]])


def _zred(x, red):
    v = np.asarray(x.val)
    dt = x.dt
    if red in (MAX, MIN):
        r = v.max() if red == MAX else v.min()
        return R(_conv(r, dt), dt)
    if dt in FLOATS:
        a = v.astype(np.float64).reshape(-1)
        r = math.fsum(a) if red == SUM else float(np.prod(a))
        return R(_conv(np.float64(r), dt), dt)
    a = v.astype(np.int64).reshape(-1)
    if red == SUM:
        r = int(a.sum(dtype=np.int64))
    elif red == PROD:
        r = 1
        for e in a:
            r = (r * int(e)) & ((1 << 64) - 1)
    elif red == rAND:
        # the fold starts from the reduction's neutral (ZonalTask preCompute fills the slot
        # with it): for signed types that is the type's maximum (ReductionType.cpp:80-90),
        # not an identity of &, so the sign bit is always cleared
        r = int(np.bitwise_and.reduce(a)) & int(_neutral(rAND, dt))
    else:
        # likewise the rOR neutral of a signed type is its minimum: the sign bit is always set
        r = int(np.bitwise_or.reduce(a)) | int(_neutral(rOR, dt))
    bits = np.dtype(NP[dt]).itemsize * 8
    r &= (1 << bits) - 1
    return R(np.array(r, dtype=np.uint64).astype(NP[dt]), dt)


# ---- spreading (D8 flow directions) -----------------------------------------------------
# The code of the neighbour at offset (i0, i1) of a cell when that neighbour flows into the
# cell, indexed [i1 + 1][i0 + 1] (dircod of runtime/skeleton/util.cpp:628 and
# front/Raster.cpp:586-589); the downstream offset of each code follows from it.
D8_INFLOW = [[2, 4, 8], [1, 0, 16], [128, 64, 32]]
D8_OUT = {D8_INFLOW[i1 + 1][i0 + 1]: (-i0, -i1) for i1 in (-1, 0, 1) for i0 in (-1, 0, 1) if i0 or i1}


def _neutral(red, dt):
    t = _work(NP[dt])
    if red == SUM:
        return t(0)
    if red == PROD:
        return t(1)
    if dt in FLOATS:
        return t(-np.inf) if red == MAX else t(np.inf)
    info = np.iinfo(NP[dt]) if dt != B8 else np.iinfo(np.uint8)
    if red in (MAX, rOR):
        return t(info.min if dt != B8 else 0)
    return t(info.max if dt != B8 else 1)  # MIN, rAND


def _red_step(acc, v, red, dt):
    """acc = OP(acc, v) in the node's C type (the statements of runtime/skeleton/Skeleton.cpp)."""
    if dt in FLOATS:
        if red == MAX:
            return np.fmax(acc, v)
        if red == MIN:
            return np.fmin(acc, v)
        return (acc + v) if red == SUM else (acc * v)
    if red == MAX:
        return np.maximum(acc, v)
    if red == MIN:
        return np.minimum(acc, v)
    wide = np.uint64 if dt == U64 else np.int64
    a, b = np.asarray(acc).astype(wide), np.asarray(v).astype(wide)
    with np.errstate(over='ignore'):
        r = a + b if red == SUM else a * b if red == PROD else a & b if red == rAND else a | b
    return _conv(r, dt)


def spread_neighbor(x, d, red=SUM):
    """SpreadNeighbor (runtime/dag/SpreadNeighbor.cpp:31-47): per cell, red over the 3x3
    neighbours (y outer, x inner, centre skipped) whose code in d points at the cell, from the
    reduction's neutral value, with the focal border rule; the gather of flowAccuGather
    (front/Raster.cpp:583-606) as one node."""
    ds = x.datasize()
    idx, nt = taps(ds, x.bs, 1, 1)
    g = idx.reshape(-1, nt)
    src = np.asarray(x.val).reshape(-1)
    dsrc = np.asarray(d.val).reshape(-1)
    acc = np.full(src.shape, _neutral(red, x.dt), dtype=src.dtype)
    for t in range(9):
        i1, i0 = t // 3 - 1, t % 3 - 1
        if i0 == 0 and i1 == 0:
            continue
        sel = dsrc[g[:, t]] == D8_INFLOW[i1 + 1][i0 + 1]
        acc = np.where(sel, _red_step(acc, src[g[:, t]], red, x.dt), acc).astype(src.dtype)
    return R(acc.reshape(ds[1], ds[0]), x.dt, x.bs)


def _d8_downstream(d):
    """Downstream cell of every cell of a direction raster (rows = y) and whether it has one."""
    h, w = d.shape
    yy, xx = np.mgrid[0:h, 0:w]
    tx, ty = xx.copy(), yy.copy()
    has = np.zeros((h, w), bool)
    for code, (ox, oy) in D8_OUT.items():
        m = d == code
        tx[m] += ox
        ty[m] += oy
        has |= m
    has &= (tx >= 0) & (tx < w) & (ty >= 0) & (ty < h)
    return tx, ty, has


def spread_scan(x, d, red=SUM):
    """SpreadScan with SUM (runtime/dag/SpreadScan.cpp; SpreadingSkeleton.cpp:288-388): with
    T(c) = in(c) + the sum of T(u) over the upstream neighbours u of c (the neighbours whose
    code points at c, y outer, x inner; in the raster's C type), a cell with a downstream cell
    inside the raster gets T(c) (what leaves it), any other cell 0; cells on a direction cycle
    never have all their upstream totals and get 0 (the reference's iteration does not end
    for them). Evaluated in topological order (Kahn's algorithm), one
    front of ready cells at a time."""
    if red != SUM:
        raise NotImplementedError("spread_scan: SUM only")
    dv = np.asarray(d.val)
    h, w = dv.shape
    v = np.asarray(x.val)
    tx, ty, has = _d8_downstream(dv)
    left = np.zeros((h, w), np.int64)
    np.add.at(left, (ty[has], tx[has]), 1)
    tot = np.zeros_like(v)
    done = np.zeros((h, w), bool)
    ry, rx = np.nonzero(left == 0)
    while ry.size:
        acc = v[ry, rx].copy()
        for i1 in (-1, 0, 1):
            for i0 in (-1, 0, 1):
                if i0 == 0 and i1 == 0:
                    continue
                ny, nx = ry + i1, rx + i0
                ok = (nx >= 0) & (nx < w) & (ny >= 0) & (ny < h)
                nyc, nxc = np.clip(ny, 0, h - 1), np.clip(nx, 0, w - 1)
                ok &= dv[nyc, nxc] == D8_INFLOW[i1 + 1][i0 + 1]
                acc = np.where(ok, _red_step(acc, tot[nyc, nxc], SUM, x.dt), acc).astype(v.dtype)
        tot[ry, rx] = acc
        done[ry, rx] = True
        go = has[ry, rx]
        gy, gx = ty[ry, rx][go], tx[ry, rx][go]
        np.subtract.at(left, (gy, gx), 1)
        nxt = np.unique(gy * w + gx)
        nxt = nxt[left.reshape(-1)[nxt] == 0]
        ry, rx = nxt // w, nxt % w
    out = np.where(has & done, tot, np.zeros_like(tot)).astype(v.dtype)
    return R(out, x.dt, x.bs)


def flow_accumulation(x, d):
    return spread_scan(x, d, SUM)


def stats(x):
    """Stats (runtime/dag/Stats.cpp): the raster's values, with its maximum and minimum."""
    r = R(np.asarray(x.val).copy(), x.dt, x.bs)
    r.range = (zmax(x), zmin(x))
    return r


def stats_range(s):
    return value(s.range[0]), value(s.range[1])


def barrier(x):
    return R(np.asarray(x.val).copy(), x.dt, x.bs)


def access(x, cell):
    """Access (runtime/dag/Access.cpp): the value of cell [x, y]."""
    return R(np.asarray(x.val)[cell[1], cell[0]].copy(), x.dt)


def lhs_access(x, v, cell):
    """LhsAccess (Skeleton.cpp:290-297): x with cell [x, y] set to v in x's type."""
    out = np.asarray(x.val).copy()
    out[cell[1], cell[0]] = np.asarray(_conv(np.asarray(const(v).val), x.dt))
    return R(out, x.dt, x.bs)


def zeros_like(x, dt=None):
    dt = x.dt if dt is None else dt
    return R(np.zeros(np.asarray(x.val).shape, _work(NP[dt])), dt, x.bs)


def full_like(c, x, dt=None):
    """A raster of x's shape holding the constant c converted to dt (x's type by default), as
    python/map.py's _like makes it (Variant(c).convert(dt))."""
    dt = x.dt if dt is None else dt
    v = _conv(np.asarray(const(c).val), dt)
    return R(np.full(np.asarray(x.val).shape, v, _work(NP[dt])), dt, x.bs)


def ones_like(x, dt=None):
    return full_like(1, x, dt)


def _r_getitem(self, key):
    if isinstance(key, (list, tuple)):
        return access(self, key)
    return con(key, self, const(0).astype(self.dt))


def _r_setitem(self, key, other):
    other = const(other).astype(self.dt)
    new = lhs_access(self, other, key) if isinstance(key, (list, tuple)) else con(key, other, self)
    self.val, self.dt = new.val, new.dt


def _r_call(self, ngb, red=None):
    if isinstance(ngb, R):
        return spread_neighbor(self, ngb, SUM if red is None else red)
    return neighbor(self, ngb)


R.__getitem__, R.__setitem__, R.__call__ = _r_getitem, _r_setitem, _r_call


def zsum(x):
    return _zred(x, SUM)


def zprod(x):
    return _zred(x, PROD)


def zmax(x):
    return _zred(x, MAX)


def zmin(x):
    return _zred(x, MIN)


# Octants of a radial scan in the reference's RadialCase order (runtime/Direction.hpp:68-97):
# (first direction, second direction) as unit vectors (fx, fy, sx, sy), x east, y south.
RADIAL_OCTANTS = [(0, -1, 1, 0), (1, 0, 0, -1), (1, 0, 0, 1), (0, 1, 1, 0),
                  (0, 1, -1, 0), (-1, 0, 0, 1), (-1, 0, 0, -1), (0, -1, -1, 0)]


def _reach(s, n, u):
    return n - 1 - s if u > 0 else s


def _radial(x, red, coord):
    """Radial scan from the cell coord = (x, y) (RadialScan, runtime/skeleton/RadiatingSkeleton.cpp:
    188-247 and the radial helpers of runtime/skeleton/util.cpp:543-616), as the recurrence that
    kernel computes per octant (first direction F, second S; cells with d0 = (p-s).F >= d1 =
    (p-s).S >= 0), column d0 by column:
      base = neutral at the start cell; R(p-F) on the F axis (d1 = 0); R(p-F-S) on the diagonal
      (d1 = d0); otherwise a*w + b*(1-w) with a = R(p-F), b = R(p-F-S),
      w = d1/d0*(d0-1) - (d1-1) (linear_interpolation, util.cpp:564);
      R(p) = OP(base, in(p)) when in(p) is finite, else base (RadiatingSkeleton.cpp:305-316).
    Every operation rounds in the raster's type (numpy element-wise ops). Cells on an axis or a
    diagonal are computed by both of their octants; the two results are checked to be equal."""
    v = np.asarray(x.val)
    T = v.dtype.type
    H, W = v.shape
    s0, s1 = int(coord[0]), int(coord[1])
    if not (0 <= s0 < W and 0 <= s1 < H):
        raise ValueError("radial: start cell outside the raster")
    neutral = {SUM: 0.0, PROD: 1.0, MAX: -np.inf, MIN: np.inf}[red]
    op = {SUM: np.add, PROD: np.multiply, MAX: np.fmax, MIN: np.fmin}[red]
    out = np.empty_like(v)
    seen = np.zeros(v.shape, bool)
    one = T(1)
    for fx, fy, sx, sy in RADIAL_OCTANTS:
        L0 = _reach(s0, W, fx) if fx else _reach(s1, H, fy)
        L1 = _reach(s0, W, sx) if sx else _reach(s1, H, sy)
        prev = None
        for c in range(L0 + 1):
            n = min(c, L1) + 1
            r = np.arange(n)
            xs, ys = s0 + c * fx + r * sx, s1 + c * fy + r * sy
            xin = v[ys, xs]
            base = np.empty(n, T)
            if c == 0:
                base[0] = T(neutral)
            else:
                base[0] = prev[0]
                m = min(c - 1, L1)  # rows 0 < d1 < d0 with both predecessors in column c-1
                if m >= 1:
                    d1 = np.arange(1, m + 1)
                    w = d1.astype(T) / T(c) * T(c - 1) - (d1 - 1).astype(T)
                    base[1:m + 1] = prev[1:m + 1] * w + prev[0:m] * (one - w)
                if c <= L1:
                    base[c] = prev[c - 1]
            with np.errstate(all="ignore"):
                val = np.where(np.isfinite(xin), op(base, xin), base).astype(T)
            old = seen[ys, xs]
            if old.any():
                a, b = out[ys, xs][old], val[old]
                if not np.array_equal(a.view(np.uint8), b.view(np.uint8)):
                    raise AssertionError("radial: octants disagree on a shared cell")
            out[ys, xs] = val
            seen[ys, xs] = True
            prev = val
    return R(out, x.dt, x.bs)


def rsum(x, coord):
    return _radial(x, SUM, coord)


def rprod(x, coord):
    return _radial(x, PROD, coord)


def rmax(x, coord):
    return _radial(x, MAX, coord)


def rmin(x, coord):
    return _radial(x, MIN, coord)


def zstats(x, zones, nz):
    """Per-zone (sum FP64, min, max, count, mean) of x over integer zones."""
    v = np.asarray(x.val).astype(np.float64).reshape(-1)
    z = np.asarray(zones.val).astype(np.int64).reshape(-1)
    ok = (z >= 0) & (z < nz)
    v, z = v[ok], z[ok]
    cnt = np.bincount(z, minlength=nz).astype(np.int64)
    s = np.zeros(nz, np.float64)
    # exact-ish FP64 sums per zone (sorted by zone, fsum per group)
    order = np.argsort(z, kind='stable')
    zs, vs = z[order], v[order]
    bounds = np.searchsorted(zs, np.arange(nz + 1))
    mn = np.full(nz, np.inf)
    mx = np.full(nz, -np.inf)
    for k in range(nz):
        seg = vs[bounds[k]:bounds[k + 1]]
        if len(seg):
            s[k] = math.fsum(seg)
            mn[k] = seg.min()
            mx[k] = seg.max()
    with np.errstate(invalid='ignore', divide='ignore'):
        mean = np.where(cnt > 0, s / np.maximum(cnt, 1), np.nan)
    return {'sum': s, 'min': mn, 'max': mx, 'count': cnt, 'mean': mean}


def value(x):
    return np.asarray(x.val).item()


def _def_u(name, op):
    def f(x):
        if isinstance(x, R):
            return unary(x, op)
        return getattr(math, name)(x)
    f.__name__ = name
    globals()[name] = f


for _name, _op in [('sin', 'SIN'), ('cos', 'COS'), ('tan', 'TAN'), ('asin', 'ASIN'), ('acos', 'ACOS'),
                   ('atan', 'ATAN'), ('sinh', 'SINH'), ('cosh', 'COSH'), ('tanh', 'TANH'), ('exp', 'EXP'),
                   ('exp2', 'EXP2'), ('log', 'LOG'), ('log2', 'LOG2'), ('log10', 'LOG10'), ('sqrt', 'SQRT'),
                   ('cbrt', 'CBRT'), ('ceil', 'CEIL'), ('floor', 'FLOOR'), ('trunc', 'TRUNC')]:
    _def_u(_name, _op)


def _def_b(name, op, scalar):
    def f(a, b):
        if not isinstance(a, R) and not isinstance(b, R):
            return scalar(a, b)
        return binary(a, b, op)
    f.__name__ = name
    globals()[name] = f


_def_b('atan2', 'ATAN2', math.atan2)
_def_b('hypot', 'HYPOT', math.hypot)
_def_b('fmod', 'FMOD', math.fmod)
_def_b('max', 'MAX2', max)
_def_b('min', 'MIN2', min)


def life(state_u8, bs, gens):
    """`gens` Life generations of a U8 raster (rows = y) with the reference semantics."""
    a = np.ascontiguousarray(state_u8, dtype=np.uint8).copy()
    b = np.empty_like(a)
    lib().or_life(a.ctypes.data, b.ctypes.data, a.shape[1], a.shape[0], bs[0], bs[1], gens)
    return a if gens % 2 == 0 else b


def dem(x0, y0, w, h, gauss, seed):
    g = np.ascontiguousarray(np.asarray(gauss, np.float64).reshape(-1, 4))
    out = np.empty((h, w), np.float32)
    lib().or_dem(out.ctypes.data, x0, y0, w, h, g.ctypes.data, len(g), ct.c_uint64(seed))
    return out


def voronoi(x0, y0, w, h, px, py):
    px = np.ascontiguousarray(px, np.int32)
    py = np.ascontiguousarray(py, np.int32)
    out = np.empty((h, w), np.int32)
    lib().or_voronoi(out.ctypes.data, x0, y0, w, h, px.ctypes.data, py.ctypes.data, len(px))
    return out


def src_cell(ds, bs, bx, by, lx, ly):
    gx, gy = ct.c_int(), ct.c_int()
    lib().or_src(ds[0], ds[1], bs[0], bs[1], bx, by, lx, ly, ct.byref(gx), ct.byref(gy))
    return gx.value, gy.value
