"""Python 3 front end of the map-algebra API over libmapx.so.

Mirrors the reference's python/map.py (names, enum values, operator
overloading, constant typing: Python float -> F32 rounded by c_float, int ->
S32, bool -> B8, python/map.py:111-127) so scripts such as hill.py and
life.py run unchanged apart from Python 3 syntax. Calls go through ctypes to
the C ABI declared in include/mapx.h.

Differences from the reference: failures raise MapError instead of aborting
the process, and a `while` over rasters runs as ordinary Python with the D0
condition evaluated on the device each iteration (Raster.__bool__) instead of
being rebuilt from the loop's source into a Loop node (python/map.py:234-327).
Extensions: zstats/zone_table (per-zone statistics), from_array/to_array
(in-memory rasters), gen_dem/gen_zones (synthetic inputs in HBM),
set_option/get_option, plan, compile_plan, stats.
"""
import builtins as _builtins
import ctypes as ct
import json as _json
import math
import operator
import os
from functools import reduce

import numpy as np

from . import binfile as _binfile

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "libmapx.so")


class MapError(RuntimeError):
    """A failed call into the backend (the reference aborts the process instead)."""


def _load():
    if not os.path.exists(_LIBPATH):
        raise ImportError("libmapx.so is not built: run `python -m map_amd.build` (%s missing)" % _LIBPATH)
    return ct.CDLL(_LIBPATH)


_lib = _load()

# ---- enumerations (values of python/map.py:35-83) ---------------------------------

DataTypeId = ['NONE_DATATYPE', 'F32', 'F64', 'B8', 'U8', 'U16', 'U32', 'U64', 'S8', 'S16', 'S32', 'S64', 'N_DATATYPE']
DataTypeVal = list(range(len(DataTypeId)))
NumDimId = ['NONE_NUMDIM', 'TIME', 'D0', 'D1', 'D2', 'D3', 'N_NUMDIM']
NumDimVal = [0x00, 0x01, 0x02, 0x04, 0x06, 0x08, 0x0A]
MemOrderId = ['NONE_MEMORDER', 'BLK', 'ROW', 'COL', 'SFC', 'N_MEMORDER']
MemOrderVal = [0x00, 0x01, 0x02, 0x04, 0x06, 0x08]
DeviceTypeId = ['DEV_DEF', 'DEV_CPU', 'DEV_GPU', 'DEV_ACC', 'DEV_CUS', 'DEV_ALL']
DeviceTypeVal = [0x01, 0x02, 0x04, 0x08, 0x16, 0xFFFFFFFF]
UnaryTypeId = ['NONE_UNARY', 'POS', 'NEG', 'NOT', 'bNOT', 'MARK_UNARY',
               'SIN', 'COS', 'TAN', 'ASIN', 'ACOS', 'ATAN', 'SINH', 'COSH', 'TANH', 'ASINH', 'ACOSH', 'ATANH',
               'EXP', 'EXP2', 'EXP10', 'LOG', 'LOG2', 'LOG10', 'SQRT', 'CBRT', 'ABS',
               'CEIL', 'FLOOR', 'TRUNC', 'ROUND', 'N_UNARY']
UnaryTypeSym = ['', '__pos__', '__neg__', '__not__', '__invert__', '',
                'sin', 'cos', 'tan', 'asin', 'acos', 'atan', 'sinh', 'cosh', 'tanh', 'asinh', 'acosh', 'atanh',
                'exp', 'exp2', 'exp10', 'log', 'log2', 'log10', 'sqrt', 'cbrt', 'abs',
                'ceil', 'floor', 'trunc', 'round', '']
UnaryTypeVal = list(range(len(UnaryTypeId)))
BinaryTypeId = ['NONE_BINARY', 'ADD', 'SUB', 'MUL', 'DIV', 'MOD', 'EQ', 'NE', 'LT', 'GT', 'LE', 'GE', 'AND', 'OR',
                'bAND', 'bOR', 'bXOR', 'SHL', 'SHR', 'MARK_BINARY', 'MAX2', 'MIN2', 'ATAN2', 'POW', 'HYPOT', 'FMOD',
                'N_BINARY']
BinaryTypeSym = ['', '__add__', '__sub__', '__mul__', '__truediv__', '__mod__',
                 '__eq__', '__ne__', '__lt__', '__gt__', '__le__', '__ge__', '', '',
                 '__and__', '__or__', '__xor__', '__lshift__', '__rshift__', '',
                 'max', 'min', 'atan2', 'pow', 'hypot', 'fmod', '']
BinaryTypeVal = list(range(len(BinaryTypeId)))
ReductionTypeId = ['NONE_REDUCTION', 'SUM', 'PROD', 'rAND', 'rOR', 'MARK_REDUCTION', 'MAX', 'MIN', 'N_REDUCTION']
ReductionTypeVal = list(range(len(ReductionTypeId)))

for _ids, _vals in zip([DataTypeId, NumDimId, MemOrderId, DeviceTypeId, UnaryTypeId, BinaryTypeId, ReductionTypeId],
                       [DataTypeVal, NumDimVal, MemOrderVal, DeviceTypeVal, UnaryTypeVal, BinaryTypeVal,
                        ReductionTypeVal]):
    for _i, _v in zip(_ids, _vals):
        globals()[_i] = _v

_NP = {F32: np.float32, F64: np.float64, B8: np.uint8, U8: np.uint8, U16: np.uint16, U32: np.uint32,
       U64: np.uint64, S8: np.int8, S16: np.int16, S32: np.int32, S64: np.int64}
_CODE = {np.dtype(v): k for k, v in _NP.items() if k != B8}

# ---- by-value ABI structs ------------------------------------------------------------


class Variant(ct.Structure):
    class Union(ct.Union):
        _fields_ = [('b8', ct.c_bool), ('s8', ct.c_int8), ('s16', ct.c_int16), ('s32', ct.c_int32),
                    ('s64', ct.c_int64), ('u8', ct.c_uint8), ('u16', ct.c_uint16), ('u32', ct.c_uint32),
                    ('u64', ct.c_uint64), ('f32', ct.c_float), ('f64', ct.c_double)]
    _fields_ = [('union', Union), ('dtype', ct.c_int)]

    def __init__(self, val=None, dtype=None):
        super().__init__()
        if val is None:
            return
        if dtype is not None:
            self._set(val, dtype)
        elif isinstance(val, float):
            self._set(val, F32)
        elif isinstance(val, bool):
            self._set(val, B8)
        elif isinstance(val, int):
            self._set(val, S32)
        elif isinstance(val, ct.c_uint8):
            self._set(val.value, U8)
        else:
            raise TypeError("unsupported constant %r" % (val,))

    _FIELD = {F32: 'f32', F64: 'f64', B8: 'b8', U8: 'u8', U16: 'u16', U32: 'u32', U64: 'u64', S8: 's8',
              S16: 's16', S32: 's32', S64: 's64'}

    def _set(self, val, dt):
        f = self._FIELD[dt]
        if dt in (F32, F64):
            val = float(val)
        elif dt == B8:
            val = bool(val)
        else:
            val = int(val)
        setattr(self.union, f, val)
        self.dtype = dt

    def datatype(self):
        return self.dtype

    def value(self):
        if self.dtype not in self._FIELD:
            raise MapError("empty value")
        return getattr(self.union, self._FIELD[self.dtype])

    def convert(self, new_dt):
        return Variant(self.value(), new_dt)


class Array(ct.Structure):
    N = 2
    T = ct.c_int * N
    _fields_ = [('arr', T), ('n', ct.c_int)]

    def __init__(self, arg=()):
        if arg is None or (isinstance(arg, (list, tuple)) and len(arg) == 0):
            super().__init__(self.T(-1, -1), 0)
        elif isinstance(arg, (list, tuple)):
            if len(arg) != self.N:
                raise ValueError("a size or coordinate needs %d values" % self.N)
            super().__init__(self.T(*[int(x) for x in arg]), self.N)
        elif isinstance(arg, Array):
            super().__init__(arg.arr, arg.n)
        else:
            raise TypeError("bad array argument %r" % (arg,))

    def isNone(self):
        return all(e == -1 for e in self.arr)

    def toList(self):
        return [x for x in self.arr[:self.n]]


class Node(ct.c_void_p):
    pass


def _err(ptr):
    msg = _lib.ma_lastError().decode()
    raise MapError(msg or "backend call failed")


def _node(ptr):
    if not ptr:
        _err(ptr)
    return ptr

# ---- Raster ------------------------------------------------------------------------


class Raster:
    """Handle of one lazy DAG node (a raster, or a D0 value)."""
    __slots__ = ('_node', '__weakref__')

    def __init__(self, node=None):
        if isinstance(node, Node):
            if not node:
                _err(node)
            self._node = node
            _fast.ma_increaseRef(node)
        elif node is None:
            self._node = Node(None)
        else:
            raise TypeError("Raster wraps a Node")

    def __del__(self):
        try:
            if _fast is not None and self._node:
                _fast.ma_decreaseRef(self._node)
        except Exception:
            pass

    @classmethod
    def from_param(cls, ins):
        return ins._node

    def __bool__(self):
        """Truth value of a D0 raster, evaluated on the device, so a `while` over rasters
        (`while zor(change): ...`) runs unmodified as ordinary Python. The reference instead
        rebuilds the loop from its source into a Loop node (python/map.py:234-327,
        runtime/Runtime.cpp:156-311); both read the condition back once per iteration. Each
        iteration's body is a new lazy DAG: its kernels come from the compiled-kernel cache
        and the previous iterate stays resident while the new one links to it. A D2 raster
        has no truth value."""
        if self.numdim() != D0:
            raise MapError("the truth value of a D2 raster is ambiguous (reduce it first, e.g. zor/zmax)")
        eval(self)
        return bool(value(self))

    def __hash__(self):
        return id(self)

    def __getitem__(self, key):
        """raster[x, y]: the D0 value of one cell (python/map.py:205-211); raster[mask]:
        the masked raster, con(mask, raster, 0)."""
        if isinstance(key, (list, tuple)):
            return Raster(_lib.ma_access(self, Array(list(key))))
        if isinstance(key, Raster):
            return Raster(_lib.ma_conditional(key, self, zeros_like(self)))
        raise MapError("a raster is indexed by a cell [x, y] or by a boolean raster")

    def __setitem__(self, key, other):
        """raster[x, y] = v replaces one cell; raster[mask] = v replaces the masked cells
        (python/map.py:213-225)."""
        other = _const_to_raster(other).astype(self.datatype())
        if isinstance(key, (list, tuple)):
            new = _node(_lib.ma_lhsAccess(self, other, Array(list(key))))
        elif isinstance(key, Raster):
            new = _node(_lib.ma_conditional(key, other, self))
        else:
            raise MapError("a raster is indexed by a cell [x, y] or by a boolean raster")
        old = Raster()
        old._node, self._node = self._node, new
        _lib.ma_increaseRef(self)  # `old` drops the previous node's reference when it goes

    def __call__(self, ngb, red=None):
        """raster([dx, dy]): the neighbour at a fixed offset (python/map.py:227-233);
        raster(dir, red): the reduction of the neighbours whose D8 direction in the raster
        `dir` points at each cell (front/Raster.cpp:137-139, ma_spreadNeighbor)."""
        if isinstance(ngb, Raster):
            return Raster(_lib.ma_spreadNeighbor(self, ngb, SUM if red is None else red))
        if isinstance(ngb, (list, tuple)) and len(ngb) == 2 and not isinstance(ngb[0], Raster):
            return Raster(_lib.ma_neighbor(self, Array(list(ngb))))
        raise MapError("bounded neighbourhoods are not supported: the reference has no kernel for them "
                       "(runtime/skeleton has no BoundedNbh statement)")

    def nodeid(self):
        return _lib.ma_nodeid(self)

    def datatype(self):
        return _lib.ma_datatype(self)

    def numdim(self):
        return _lib.ma_numdim(self)

    def memorder(self):
        return _lib.ma_memorder(self)

    def datasize(self):
        return _lib.ma_datasize(self).toList()

    def blocksize(self):
        return _lib.ma_blocksize(self).toList()

    def numblock(self):
        ds, bs = self.datasize(), self.blocksize()
        return [(d + b - 1) // b for d, b in zip(ds, bs)]

    def __pow__(self, other):
        return Raster(_lib.ma_binary(self, _const_to_raster(other), POW))

    def __rpow__(self, other):
        return Raster(_lib.ma_binary(_const_to_raster(other), self, POW))

    def no(self):
        return Raster(_lib.ma_unary(self, NOT))

    def astype(self, dt):
        return astype(self, dt)




# Scalar operands become D0 constant nodes; the backend's value numbering returns
# the same node for the same constant, so the front end keeps one Raster per
# constant instead of building a Variant and calling ma_constant each time (the
# per-request cost of a served script is mostly these calls). Bounded.
_CONSTS = {}


def _const_to_raster(arg):
    if isinstance(arg, Raster):
        return arg
    if isinstance(arg, (bool, int, float)):
        key = repr(arg)  # tells apart what == does not: 1 / 1.0 / True, 0.0 / -0.0; NaN equals itself
        r = _CONSTS.get(key)
        if r is None:
            var = Variant(arg)
            r = Raster(_lib.ma_constant(var, Array(), var.datatype(), ROW + BLK, Array()))
            if len(_CONSTS) >= 4096:
                _CONSTS.clear()
            _CONSTS[key] = r
        return r
    raise TypeError("cannot use %r as a raster" % (arg,))


def _def_unary_op(symbol, utype):
    def uop(self):
        return Raster(_lib.ma_unary(self, utype))
    setattr(Raster, symbol, uop)


_SCALAR_FUN = {'exp10': lambda x: 10.0 ** x, 'abs': _builtins.abs, 'round': _builtins.round,
               'max': _builtins.max, 'min': _builtins.min, 'pow': _builtins.pow}


def _scalar_fun(symbol):
    if symbol in _SCALAR_FUN:
        return _SCALAR_FUN[symbol]
    return getattr(math, symbol)


def _def_unary_fun(symbol, utype):
    def ufun(raster):
        if isinstance(raster, Raster):
            return Raster(_fast.ma_unary(raster._node, utype))
        return _scalar_fun(symbol)(raster)
    ufun.__name__ = symbol
    globals()[symbol] = ufun


for _sym, _ut in zip(UnaryTypeSym, UnaryTypeVal):
    if _sym and _ut < MARK_UNARY:
        _def_unary_op(_sym, _ut)
    elif _sym and _ut > MARK_UNARY:
        _def_unary_fun(_sym, _ut)


def _def_binary_op(symbol, btype):
    def biop(self, rhs):
        return Raster(_fast.ma_binary(self._node, _const_to_raster(rhs)._node, btype))

    def rbiop(self, lhs):
        return Raster(_fast.ma_binary(_const_to_raster(lhs)._node, self._node, btype))
    setattr(Raster, symbol, biop)
    setattr(Raster, symbol[:2] + 'r' + symbol[2:], rbiop)


def _def_binary_fun(symbol, btype):
    def bifun(lhs, rhs):
        if not isinstance(lhs, Raster) and not isinstance(rhs, Raster):
            return _scalar_fun(symbol)(lhs, rhs)
        return Raster(_lib.ma_binary(_const_to_raster(lhs), _const_to_raster(rhs), btype))
    bifun.__name__ = symbol
    globals()[symbol] = bifun


for _sym, _bt in zip(BinaryTypeSym, BinaryTypeVal):
    if _sym and _bt < MARK_BINARY:
        _def_binary_op(_sym, _bt)
    elif _sym and _bt > MARK_BINARY:
        _def_binary_fun(_sym, _bt)
Raster.__div__ = Raster.__truediv__

# ---- module functions (python/map.py:451-576) -------------------------------------------


def setupDevices(plat_name, dev_type, dev_name):
    _lib.ma_setupDevices(plat_name.encode(), dev_type, dev_name.encode())
    if _lib.ma_lastError():
        _err(None)


def eval(*args):
    cond = [isinstance(a, Raster) for a in args]
    if len(args) and not all(cond):
        if not any(cond):
            return _builtins.eval(*args)
        raise TypeError("eval takes only rasters, or only the arguments of the built-in eval")
    arr = (Node * len(args))(*[r._node for r in args])
    _lib.ma_eval(ct.cast(arr, ct.POINTER(Node)), len(args))
    if _lib.ma_lastError():
        _err(None)


def value(arg):
    var = _lib.ma_value(arg)
    if var.dtype == 0:
        _err(None)
    return var.value()


def read(file):
    return Raster(_lib.ma_read(file.encode() if isinstance(file, str) else file))


def write(raster, file):
    r = _lib.ma_write(raster, file.encode() if isinstance(file, str) else file)
    if r != 0:
        _err(None)
    return r


def _cnst(cnst, ds, dt, mo, bs):
    var = Variant(cnst).convert(dt)
    return Raster(_lib.ma_constant(var, Array(ds), dt, mo, Array(bs)))


def zeros(ds=(), dt=F32, mo=ROW + BLK, bs=()):
    return _cnst(0, ds, dt, mo, bs)


def ones(ds=(), dt=F32, mo=ROW + BLK, bs=()):
    return _cnst(1, ds, dt, mo, bs)


def trues(ds=(), dt=B8, mo=ROW + BLK, bs=()):
    return _cnst(True, ds, dt, mo, bs)


def falses(ds=(), dt=B8, mo=ROW + BLK, bs=()):
    return _cnst(False, ds, dt, mo, bs)


def full(cnst, ds=(), dt=F32, mo=ROW + BLK, bs=()):
    return _cnst(cnst, ds, dt, mo, bs)


def _like(cnst, raster, dt, mo):
    dt = raster.datatype() if dt == NONE_DATATYPE else dt
    mo = raster.memorder() if mo == NONE_MEMORDER else mo
    var = Variant(cnst).convert(dt)
    return Raster(_lib.ma_constant(var, Array(raster.datasize()), dt, mo, Array(raster.blocksize())))


def zeros_like(raster, dt=NONE_DATATYPE, mo=NONE_MEMORDER):
    return _like(0, raster, dt, mo)


def ones_like(raster, dt=NONE_DATATYPE, mo=NONE_MEMORDER):
    return _like(1, raster, dt, mo)


def full_like(cnst, raster, dt=NONE_DATATYPE, mo=NONE_MEMORDER):
    return _like(cnst, raster, dt, mo)


def index(raster, dim):
    return Raster(_lib.ma_index(raster, dim))


def rand(seed, ds=(), dt=NONE_DATATYPE, mo=NONE_MEMORDER, bs=()):
    """Random raster (python/map.py:523-531): Philox2x32-10 of counter {seed cell, 0} and key
    {cell index}; `seed` is an integer raster, or an int made a constant raster of size ds."""
    if not isinstance(seed, Raster):
        dt = S32 if dt == NONE_DATATYPE else dt
        seed = Raster(_lib.ma_constant(Variant(seed).convert(S32), Array(list(ds)), S32,
                                       BLK + ROW if mo == NONE_MEMORDER else mo, Array(list(bs))))
    return Raster(_lib.ma_rand(seed, dt, mo))


def astype(raster, dt):
    return Raster(_lib.ma_cast(raster, dt))


def con(cond, lhs, rhs):
    return Raster(_lib.ma_conditional(_const_to_raster(cond), _const_to_raster(lhs), _const_to_raster(rhs)))


_MASKS = {}  # mask values -> the ctypes arguments (read only by the backend), bounded


def _mask(lst):
    if not lst or not all(len(e) == len(lst[0]) for e in lst):
        raise ValueError("a mask is a non-empty list of equal-length rows")
    key = repr(lst)  # shape and values, 1 / 1.0 / True and 0.0 / -0.0 apart
    m = _MASKS.get(key)
    if m is None:
        flat = [e for row in lst for e in row]
        ds = Array([len(lst[0]), len(lst)])
        arr = (Variant * len(flat))(*[Variant(e) for e in flat])
        m = (ct.cast(arr, ct.POINTER(Variant)), len(flat), ds, arr)
        if len(_MASKS) >= 1024:
            _MASKS.clear()
        _MASKS[key] = m
    return m


def convolve(raster, lst):
    ptr, n, ds, keep = _mask(lst)
    return Raster(_fast.ma_convolution(raster._node, ptr, n, ds))


def fsum(raster):
    return convolve(raster, [[1, 1, 1], [1, 1, 1], [1, 1, 1]])


def _focal(raster, lst, red):
    ptr, n, ds, keep = _mask(lst)
    return Raster(_lib.ma_focalFunc(raster, ptr, n, ds, red))


AGE, ILE = 1, 2                      # PercentEnum (util/PercentType.hpp:18)
VARI, MAJO, MINO, MEAN = 1, 2, 3, 4  # DiversityEnum (util/DiversityType.hpp:19-21)


def focal_percent(raster, lst, kind):
    """Share of the window (its size from the mask lst) equal to the centre (AGE) or below
    it (ILE); the reference's ma_focalPercent (not exposed by its map.py)."""
    ptr, n, ds, _keep = _mask(lst)
    return Raster(_lib.ma_focalPercent(raster, ptr, n, ds, kind))


def diversity(rasters, kind):
    """Per-cell VARI / MAJO / MINO / MEAN over rasters of one shape and type; the reference's
    ma_diversity (not exposed by its map.py)."""
    rs = list(rasters)
    arr = (Node * len(rs))(*[r._node for r in rs])
    return Raster(_lib.ma_diversity(ct.cast(arr, ct.POINTER(Node)), len(rs), kind))


def fmin(raster, lst):
    return _focal(raster, lst, MIN)


def fmax(raster, lst):
    return _focal(raster, lst, MAX)


def rsum(raster, coord):
    """Radial scans from the cell coord = [x, y] (python/map.py:578-588)."""
    return Raster(_lib.ma_radialScan(raster, SUM, Array(list(coord))))


def rprod(raster, coord):
    return Raster(_lib.ma_radialScan(raster, PROD, Array(list(coord))))


def rmax(raster, coord):
    return Raster(_lib.ma_radialScan(raster, MAX, Array(list(coord))))


def rmin(raster, coord):
    return Raster(_lib.ma_radialScan(raster, MIN, Array(list(coord))))


def spread_neighbor(raster, dir, red=SUM):
    """One spreading step: per cell, red over the 3x3 neighbours whose D8 code in `dir`
    points at it (ma_spreadNeighbor; the gather of front/Raster.cpp:583-606)."""
    return Raster(_lib.ma_spreadNeighbor(raster, dir, red))


def spread_scan(raster, dir, red=SUM):
    """Accumulation along the U8 D8 directions `dir` (ma_spreadScan, SUM only): the total
    leaving each cell towards its downstream cell (front/Raster.cpp:570-577 flowAccumulation)."""
    return Raster(_lib.ma_spreadScan(raster, dir, red))


def flow_accumulation(raster, dir):
    return spread_scan(raster, dir, SUM)


def barrier(raster):
    """The raster's values; fusion never crosses it (ma_barrier)."""
    return Raster(_lib.ma_barrier(raster))


def stats_range(s):
    """(max, min) of an evaluated stats(raster) node."""
    mx, mn = Variant(), Variant()
    if _lib.ma_statsRange(s, ct.byref(mx), ct.byref(mn)) != 0:
        _err(None)
    return mx.value(), mn.value()


def zsum(raster):
    return Raster(_lib.ma_zonalReduc(raster, SUM))


def zprod(raster):
    return Raster(_lib.ma_zonalReduc(raster, PROD))


def zmax(raster):
    return Raster(_lib.ma_zonalReduc(raster, MAX))


def zmin(raster):
    return Raster(_lib.ma_zonalReduc(raster, MIN))


def lmax(lst):
    return reduce(lambda a, b: max(a, b), lst)


def lmin(lst):
    return reduce(lambda a, b: min(a, b), lst)


def lor(lst):
    return reduce(lambda a, b: a + b, lst)


def pick(raster, lst):
    acu = _const_to_raster(lst[0])
    for i in range(len(lst)):
        acu = con(raster == i, lst[i], acu)
    return acu


def degrees(raster):
    return raster / (2 * math.pi) * 360


def radians(raster):
    return raster / 360 * (2 * math.pi)


def mean(raster):
    return zsum(raster) / prod(raster.datasize())


def gradient(raster):
    ix = convolve(raster, [[-1, 0, 1], [-1, 0, 1], [-1, 0, 1]])
    iy = convolve(raster, [[-1, -1, -1], [0, 0, 0], [1, 1, 1]])
    return ix, iy


def blur(raster, size=3):
    if size == 3:
        return convolve(raster, [[1, 2, 1], [2, 4, 2], [1, 2, 1]]) / 16
    if size == 5:
        k = [[1, 4, 7, 4, 1], [4, 16, 26, 16, 4], [7, 26, 41, 26, 7], [4, 16, 26, 16, 4], [1, 4, 7, 4, 1]]
        return convolve(raster, k) / 276
    raise ValueError("blur size must be 3 or 5")


def harris(raster):
    ix, iy = gradient(raster)
    xx, yy, xy = blur(ix * ix), blur(iy * iy), blur(ix * iy)
    return (xx * yy - xy) - 0.04 * (xx + yy) ** 2


def prod(arg):
    if isinstance(arg, (list, tuple)):
        return reduce(operator.mul, arg, 1)
    if isinstance(arg, Array):
        return arg.arr[0] * arg.arr[1]
    raise TypeError("prod of a list or Array")

# ---- extensions ------------------------------------------------------------------------


class ZoneStats(Raster):
    """Per-zone statistics node: after eval, table() gives sum/min/max/count/mean."""

    def __init__(self, node, nzones):
        super().__init__(node)
        self.nzones = nzones

    def table(self):
        n = self.nzones
        s, mn, mx = (np.zeros(n, np.float64) for _ in range(3))
        c = np.zeros(n, np.int64)
        dp = ct.POINTER(ct.c_double)
        r = _lib.ma_zonalTable(self, s.ctypes.data_as(dp), mn.ctypes.data_as(dp), mx.ctypes.data_as(dp),
                               c.ctypes.data_as(ct.POINTER(ct.c_longlong)), n)
        if r != 0:
            _err(None)
        with np.errstate(invalid='ignore', divide='ignore'):
            mean_ = np.where(c > 0, s / np.maximum(c, 1), np.nan)
        return {'sum': s, 'sum_f32': s.astype(np.float32), 'min': mn, 'max': mx, 'count': c, 'mean': mean_}


def zstats(raster, zones, nzones):
    return ZoneStats(_lib.ma_zonalStats(raster, zones, int(nzones)), int(nzones))


def zone_table(zs):
    return zs.table()


def from_array(a, bs, dt=None):
    """In-memory input raster from a 2D numpy array (rows = y)."""
    a = np.ascontiguousarray(a)
    code = dt if dt is not None else _CODE[a.dtype]
    flat = _binfile.to_blocks(a.astype(_NP[code], copy=False), bs)
    return Raster(_lib.ma_input(flat.ctypes.data_as(ct.c_void_p), ct.c_size_t(flat.nbytes),
                                Array([a.shape[1], a.shape[0]]), code, Array(list(bs))))


def to_array(raster):
    """Evaluated raster -> 2D numpy array of its data size (evaluates it if needed)."""
    ds, bs, dt = raster.datasize(), raster.blocksize(), raster.datatype()
    nb = raster.numblock()
    flat = np.empty(nb[0] * nb[1] * bs[0] * bs[1], dtype=_NP[dt])
    if _lib.ma_fetch(raster, flat.ctypes.data_as(ct.c_void_p), ct.c_size_t(flat.nbytes)) != 0:
        eval(raster)
        if _lib.ma_fetch(raster, flat.ctypes.data_as(ct.c_void_p), ct.c_size_t(flat.nbytes)) != 0:
            _err(None)
    return _binfile.from_blocks(flat, ds, bs)


def file_load(path):
    """Host-only read of a .bin or tiled .tif/.tiff raster file -> 2D numpy array of its
    data size (no device needed)."""
    r = Raster(_lib.ma_read(path.encode()))
    ds, bs, dt = r.datasize(), r.blocksize(), r.datatype()
    nb = r.numblock()
    del r
    flat = np.empty(nb[0] * nb[1] * bs[0] * bs[1], dtype=_NP[dt])
    if _lib.ma_fileLoad(path.encode(), flat.ctypes.data_as(ct.c_void_p), ct.c_size_t(flat.nbytes)) != 0:
        _err(None)
    return _binfile.from_blocks(flat, ds, bs)


def file_store(path, a, bs, dt=None):
    """Host-only write of a 2D numpy array as a .bin or tiled BigTIFF (.tif/.tiff) raster file
    with blocks (= TIFF tiles) of size bs."""
    a = np.ascontiguousarray(a)
    code = dt if dt is not None else _CODE[a.dtype]
    flat = _binfile.to_blocks(a.astype(_NP[code], copy=False), bs)
    if _lib.ma_fileStore(path.encode(), flat.ctypes.data_as(ct.c_void_p), ct.c_size_t(flat.nbytes),
                         Array([a.shape[1], a.shape[0]]), code, Array(list(bs))) != 0:
        _err(None)


def release(raster):
    if _lib.ma_release(raster) != 0:
        _err(None)


def gen_dem(ds, bs, gauss, noise_seed):
    """Synthetic DEM in HBM: `gauss` is an (n, 4) array of (cx, cy, sigma, amplitude)."""
    g = np.ascontiguousarray(np.asarray(gauss, dtype=np.float64).reshape(-1, 4))
    return Raster(_lib.ma_genDem(Array(list(ds)), Array(list(bs)), g.ctypes.data_as(ct.POINTER(ct.c_double)),
                                 len(g), ct.c_ulonglong(noise_seed)))


def gen_zones(ds, bs, px, py, dt=S32):
    px = np.ascontiguousarray(px, dtype=np.int32)
    py = np.ascontiguousarray(py, dtype=np.int32)
    ip = ct.POINTER(ct.c_int)
    return Raster(_lib.ma_genZones(Array(list(ds)), Array(list(bs)), dt, px.ctypes.data_as(ip),
                                   py.ctypes.data_as(ip), len(px)))


def set_option(key, val):
    if _lib.ma_setOption(key.encode(), int(val)) != 0:
        _err(None)


def get_option(key):
    return _lib.ma_getOption(key.encode())


def _nodes(args):
    arr = (Node * len(args))(*[r._node for r in args])
    return ct.cast(arr, ct.POINTER(Node)), len(args), arr


def plan(*args):
    ptr, n, keep = _nodes(args)
    s = _lib.ma_plan(ptr, n)
    if not s:
        _err(None)
    return _json.loads(s.decode())


def compile_plan(*args):
    ptr, n, keep = _nodes(args)
    r = _lib.ma_compilePlan(ptr, n)
    if r < 0:
        _err(None)
    return r


def autotune(*args, reps=3):
    """Times candidate tile shapes for the groups of an evaluation and keeps the fastest."""
    ptr, n, keep = _nodes(args)
    s = _lib.ma_autotune(ptr, n, int(reps))
    if not s:
        _err(None)
    return _json.loads(s.decode())


def tune_file(path, save=False):
    r = _lib.ma_tuneFile(path.encode(), 1 if save else 0)
    if r < 0:
        _err(None)
    return r


def stats(raster=None):
    """stats(raster): the raster with its maximum and minimum computed alongside
    (ma_stats; read them with stats_range). stats(): the runtime's counters and
    device-memory figures as a dict (ma_runStats)."""
    if raster is not None:
        return Raster(_lib.ma_stats(raster))
    return _json.loads(_lib.ma_runStats().decode())


def last_eval_ms():
    return _lib.ma_lastEvalMs()


def kernel_times():
    """{kernel: {count, total_ms, min_ms, max_ms}} of launches timed since the last call ("profile" option)."""
    s = _lib.ma_kernelTimes()
    if s is None:
        _err(None)
    return _json.loads(s.decode())

class Server:
    """Serving loop over the backend: requests (tuples of result rasters /
    zonal nodes) are submitted as they arrive and evaluated with continuous
    batching; step() advances the loop and returns the retired requests as
    (id, t_submit, t_admit, t_done) with CLOCK_MONOTONIC seconds
    (time.monotonic())."""

    def __init__(self, hbm_budget=0, max_inflight=64):
        if _lib.ma_serveOpen(int(hbm_budget), int(max_inflight)) != 0:
            _err(None)
        self._held = {}

    def submit(self, *targets):
        ptr, n, keep = _nodes(targets)
        rid = _lib.ma_serveSubmit(ptr, n)
        if rid < 0:
            _err(None)
        self._held[rid] = targets
        return rid

    def step(self, wait_us=0):
        if _lib.ma_serveStep(int(wait_us)) < 0:
            _err(None)
        return self.take()

    def take(self, cap=4096):
        ids = (ct.c_longlong * cap)()
        ts, ta, td = (ct.c_double * cap)(), (ct.c_double * cap)(), (ct.c_double * cap)()
        k = _lib.ma_serveTake(ids, ts, ta, td, cap)
        if k < 0:
            _err(None)
        out = [(ids[i], ts[i], ta[i], td[i]) for i in range(k)]
        return out

    def result(self, rid):
        """The targets of a retired request (the caller drops them when done)."""
        return self._held.pop(rid)

    def stats(self):
        s = _lib.ma_serveStats()
        if s is None:
            _err(None)
        return _json.loads(s.decode())

    def close(self):
        if _lib.ma_serveClose() != 0:
            _err(None)


class ResidentCache:
    """Rasters kept resident in HBM under caller keys (ints; e.g. DEM ids shared by served
    requests), least recently used ones released past a byte budget (0: no limit). The
    cache is the runtime's (one per process): get(key, make) returns the cached raster, or
    calls make() on a miss and caches its result."""

    def __init__(self, budget_bytes=0):
        self.budget(budget_bytes)

    @staticmethod
    def budget(nbytes):
        if _lib.ma_cacheBudget(int(nbytes)) != 0:
            _err(None)

    @staticmethod
    def get(key, make=None):
        p = _lib.ma_cacheGet(int(key))
        if p:
            return Raster(p)
        if make is None:
            return None
        r = make()
        if _lib.ma_cachePut(int(key), r) != 0:
            _err(None)
        return r

    @staticmethod
    def drop(key):
        if _lib.ma_cacheDrop(int(key)) != 0:
            _err(None)

    @staticmethod
    def stats():
        s = _lib.ma_cacheStats()
        if s is None:
            _err(None)
        return _json.loads(s.decode())


def math_check(fn, lo=0, hi=1 << 32):
    """Accuracy of an FP32 fast path of the generated kernels (sqrt, sin, cos, sincos, atan2,
    atan2_near) over FP32 bit patterns [lo, hi) against FP64 on the device:
    {max_ulp, n_over_4ulp, n, worst_input_bits}."""
    mu, no, n, wi = ct.c_double(), ct.c_ulonglong(), ct.c_ulonglong(), ct.c_uint()
    if _lib.ma_mathCheck(fn.encode(), ct.c_ulonglong(lo), ct.c_ulonglong(hi), ct.byref(mu), ct.byref(no),
                         ct.byref(n), ct.byref(wi)) != 0:
        _err(None)
    return {"max_ulp": mu.value, "n_over_4ulp": no.value, "n": n.value, "worst_input_bits": wi.value}


def kernel_source(name):
    """Generated HIP source of a kernel built in this process."""
    s = _lib.ma_kernelSource(name.encode())
    if s is None:
        _err(None)
    return s.decode()

# ---- ctypes prototypes (python/map.py:723-830) ------------------------------------------


def _proto(name, args, res):
    f = getattr(_lib, name)
    f.argtypes = args
    f.restype = res


# The node factories and reference counts the scripts call most, through a second
# handle of the library whose argument types are the Node pointers themselves
# (converted by ctypes without a Python call per argument).
_fast = ct.CDLL(_LIBPATH)
for _name, _args in (('ma_increaseRef', [Node]), ('ma_decreaseRef', [Node]),
                     ('ma_binary', [Node, Node, ct.c_int]), ('ma_unary', [Node, ct.c_int]),
                     ('ma_convolution', [Node, ct.POINTER(Variant), ct.c_int, Array])):
    getattr(_fast, _name).argtypes = _args
    getattr(_fast, _name).restype = None if _name.endswith('Ref') else Node


_proto('ma_lastError', [], ct.c_char_p)
_proto('ma_setupDevices', [ct.c_char_p, ct.c_int, ct.c_char_p], None)
_proto('ma_increaseRef', [Raster], None)
_proto('ma_decreaseRef', [Raster], None)
_proto('ma_eval', [ct.POINTER(Node), ct.c_int], None)
_proto('ma_nodeid', [Raster], ct.c_int)
_proto('ma_value', [Raster], Variant)
_proto('ma_datatype', [Raster], ct.c_int)
_proto('ma_numdim', [Raster], ct.c_int)
_proto('ma_memorder', [Raster], ct.c_int)
_proto('ma_datasize', [Raster], Array)
_proto('ma_blocksize', [Raster], Array)
_proto('ma_read', [ct.c_char_p], Node)
_proto('ma_write', [Raster, ct.c_char_p], ct.c_int)
_proto('ma_index', [Raster, ct.c_int], Node)
_proto('ma_rand', [Raster, ct.c_int, ct.c_int], Node)
_proto('ma_cast', [Raster, ct.c_int], Node)
_proto('ma_conditional', [Raster, Raster, Raster], Node)
_proto('ma_unary', [Raster, ct.c_int], Node)
_proto('ma_binary', [Raster, Raster, ct.c_int], Node)
_proto('ma_constant', [Variant, Array, ct.c_int, ct.c_int, Array], Node)
_proto('ma_neighbor', [Raster, Array], Node)
_proto('ma_convolution', [Raster, ct.POINTER(Variant), ct.c_int, Array], Node)
_proto('ma_focalFunc', [Raster, ct.POINTER(Variant), ct.c_int, Array, ct.c_int], Node)
_proto('ma_zonalReduc', [Raster, ct.c_int], Node)
_proto('ma_radialScan', [Raster, ct.c_int, Array], Node)
_proto('ma_focalPercent', [Raster, ct.POINTER(Variant), ct.c_int, Array, ct.c_int], Node)
_proto('ma_diversity', [ct.POINTER(Node), ct.c_int, ct.c_int], Node)
_proto('ma_zonalStats', [Raster, Raster, ct.c_int], Node)
_proto('ma_zonalTable', [Raster, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double), ct.POINTER(ct.c_double),
                         ct.POINTER(ct.c_longlong), ct.c_int], ct.c_int)
_proto('ma_input', [ct.c_void_p, ct.c_size_t, Array, ct.c_int, Array], Node)
_proto('ma_fetch', [Raster, ct.c_void_p, ct.c_size_t], ct.c_int)
_proto('ma_fileLoad', [ct.c_char_p, ct.c_void_p, ct.c_size_t], ct.c_int)
_proto('ma_fileStore', [ct.c_char_p, ct.c_void_p, ct.c_size_t, Array, ct.c_int, Array], ct.c_int)
_proto('ma_release', [Raster], ct.c_int)
_proto('ma_setOption', [ct.c_char_p, ct.c_longlong], ct.c_int)
_proto('ma_getOption', [ct.c_char_p], ct.c_longlong)
_proto('ma_plan', [ct.POINTER(Node), ct.c_int], ct.c_char_p)
_proto('ma_compilePlan', [ct.POINTER(Node), ct.c_int], ct.c_int)
_proto('ma_runStats', [], ct.c_char_p)
_proto('ma_stats', [Raster], Node)
_proto('ma_statsRange', [Raster, ct.POINTER(Variant), ct.POINTER(Variant)], ct.c_int)
_proto('ma_barrier', [Raster], Node)
_proto('ma_access', [Raster, Array], Node)
_proto('ma_lhsAccess', [Raster, Raster, Array], Node)
_proto('ma_spreadNeighbor', [Raster, Raster, ct.c_int], Node)
_proto('ma_spreadScan', [Raster, Raster, ct.c_int], Node)
_proto('ma_lastEvalMs', [], ct.c_double)
_proto('ma_kernelTimes', [], ct.c_char_p)
_proto('ma_kernelSource', [ct.c_char_p], ct.c_char_p)
_proto('ma_cacheBudget', [ct.c_longlong], ct.c_int)
_proto('ma_cacheGet', [ct.c_longlong], Node)
_proto('ma_cachePut', [ct.c_longlong, Raster], ct.c_int)
_proto('ma_cacheDrop', [ct.c_longlong], ct.c_int)
_proto('ma_cacheStats', [], ct.c_char_p)
_proto('ma_mathCheck', [ct.c_char_p, ct.c_ulonglong, ct.c_ulonglong, ct.POINTER(ct.c_double), ct.POINTER(ct.c_ulonglong),
                        ct.POINTER(ct.c_ulonglong), ct.POINTER(ct.c_uint)], ct.c_int)
_proto('ma_serveOpen', [ct.c_longlong, ct.c_int], ct.c_int)
_proto('ma_serveSubmit', [ct.POINTER(Node), ct.c_int], ct.c_longlong)
_proto('ma_serveStep', [ct.c_int], ct.c_int)
_proto('ma_serveTake', [ct.POINTER(ct.c_longlong), ct.POINTER(ct.c_double), ct.POINTER(ct.c_double),
                        ct.POINTER(ct.c_double), ct.c_int], ct.c_int)
_proto('ma_serveStats', [], ct.c_char_p)
_proto('ma_serveClose', [], ct.c_int)
_proto('ma_autotune', [ct.POINTER(Node), ct.c_int, ct.c_int], ct.c_char_p)
_proto('ma_tuneFile', [ct.c_char_p, ct.c_int], ct.c_int)
_proto('ma_genDem', [Array, Array, ct.POINTER(ct.c_double), ct.c_int, ct.c_ulonglong], Node)
_proto('ma_genZones', [Array, Array, ct.c_int, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), ct.c_int], Node)
