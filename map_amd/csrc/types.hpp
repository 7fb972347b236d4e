// Type system of the Parallel Map Algebra front end: enum values, by-value ABI
// structs, the promotion table, operator classification and reduction neutrals.
//
// Enum values are part of the C ABI that map.py binds through ctypes, so they
// keep the reference's numbering:
//   DataTypeEnum    util/DataType.hpp:21        (python/map.py:35-36)
//   NumDimEnum      util/NumDim.hpp             (python/map.py:38-39)
//   MemOrderEnum    util/MemOrder.hpp           (python/map.py:41-42)
//   UnaryEnum       util/UnaryType.hpp          (python/map.py:47-66)
//   BinaryEnum      util/BinaryType.hpp         (python/map.py:68-80)
//   ReductionEnum   util/ReductionType.hpp      (python/map.py:82-83)
//   PatternEnum     runtime/Pattern.hpp:20
#pragma once
#include <cstdint>
#include <cstring>
#include <string>
#include <stdexcept>

namespace mapx {

enum DataTypeEnum : int {
  NONE_DATATYPE = 0, F32, F64, B8, U8, U16, U32, U64, S8, S16, S32, S64, N_DATATYPE
};
enum NumDimEnum : int { NONE_NUMDIM = 0x00, TIME = 0x01, D0 = 0x02, D1 = 0x04, D2 = 0x06, D3 = 0x08 };
enum MemOrderEnum : int { NONE_MEMORDER = 0x00, BLK = 0x01, ROW = 0x02, COL = 0x04, SFC = 0x06 };
enum DeviceType : unsigned { DEV_DEF = 0x01, DEV_CPU = 0x02, DEV_GPU = 0x04, DEV_ACC = 0x08 };

enum UnaryEnum : int {
  NONE_UNARY = 0, POS, NEG, NOT, bNOT, MARK_UNARY,
  SIN, COS, TAN, ASIN, ACOS, ATAN, SINH, COSH, TANH, ASINH, ACOSH, ATANH,
  EXP, EXP2, EXP10, LOG, LOG2, LOG10, SQRT, CBRT, ABS, CEIL, FLOOR, TRUNC, ROUND, N_UNARY
};
enum BinaryEnum : int {
  NONE_BINARY = 0, ADD, SUB, MUL, DIV, MOD, EQ, NE, LT, GT, LE, GE, AND, OR,
  bAND, bOR, bXOR, SHL, SHR, MARK_BINARY, MAX2, MIN2, ATAN2, POW, HYPOT, FMOD, N_BINARY
};
// COUNT is an extension used only by the per-zone statistics operator; it lies
// past the reference's N_REDUCTION so no reference value changes meaning.
enum ReductionEnum : int {
  NONE_REDUCTION = 0, SUM, PROD, rAND, rOR, MARK_REDUCTION, MAX, MIN, N_REDUCTION, COUNT = 16
};
enum PatternEnum : int {
  NONE_PAT = 0x00, FREE = 0x01, LOCAL = 0x02, FOCAL = 0x04, ZONAL = 0x08,
  RADIAL = 0x10, SPREAD = 0x20, SPECIAL = 0x40, BARRIER = 0x80
};

// By-value struct of the C ABI: an 8-byte union plus the dtype tag
// (python/map.py:91-109). 16 bytes with padding on x86-64.
struct Variant {
  union {
    bool b8; int8_t s8; int16_t s16; int32_t s32; int64_t s64;
    uint8_t u8; uint16_t u16; uint32_t u32; uint64_t u64;
    float f32; double f64; uint64_t raw;
  } u;
  int dtype;
};
static_assert(sizeof(Variant) == 16, "Variant ABI layout");

// Array4<int> with N=2 as seen across the ABI: {int arr[2]; int n} (python/map.py:157-161).
struct Dim2 {
  int arr[2];
  int n;
};
static_assert(sizeof(Dim2) == 12, "Dim2 ABI layout");

inline Dim2 dim2_none() { return Dim2{{-1, -1}, 0}; }
inline Dim2 dim2(int a, int b) { return Dim2{{a, b}, 2}; }
inline bool dim2_eq(const Dim2& a, const Dim2& b) {
  if (a.n != b.n) return false;
  for (int i = 0; i < a.n; i++) if (a.arr[i] != b.arr[i]) return false;
  return true;
}

inline int sizeof_dt(int dt) {
  switch (dt) {
    case F32: return 4; case F64: return 8; case B8: return 1; case U8: return 1;
    case U16: return 2; case U32: return 4; case U64: return 8; case S8: return 1;
    case S16: return 2; case S32: return 4; case S64: return 8;
    default: throw std::runtime_error("sizeof_dt: bad data type");
  }
}
inline bool is_float_dt(int dt) { return dt == F32 || dt == F64; }
inline bool is_unsigned_dt(int dt) { return dt == B8 || dt == U8 || dt == U16 || dt == U32 || dt == U64; }
inline bool is_signed_dt(int dt) { return dt == S8 || dt == S16 || dt == S32 || dt == S64; }

// Device C type used in generated HIP source. B8 is stored as one byte
// (unsigned char), the reference's OpenCL `bool`-as-uchar convention.
inline const char* ctype_dt(int dt) {
  switch (dt) {
    case F32: return "float"; case F64: return "double"; case B8: return "unsigned char";
    case U8: return "unsigned char"; case U16: return "unsigned short"; case U32: return "unsigned int";
    case U64: return "unsigned long long"; case S8: return "signed char"; case S16: return "short";
    case S32: return "int"; case S64: return "long long";
    default: throw std::runtime_error("ctype_dt: bad data type");
  }
}
inline const char* name_dt(int dt) {
  static const char* n[] = {"NONE", "F32", "F64", "B8", "U8", "U16", "U32", "U64", "S8", "S16", "S32", "S64"};
  if (dt < 0 || dt >= N_DATATYPE) return "BAD";
  return n[dt];
}

// Promotion of a pair of types: restates the table of util/promote.hpp:22-157
// row by row (the table is a contract, so its entries are identical).
inline int promote(int a, int b) {
  static const int T[N_DATATYPE][N_DATATYPE] = {
      //        NONE F32  F64  B8   U8   U16  U32  U64  S8   S16  S32  S64
      /*NONE*/ {0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0,   0},
      /*F32 */ {0,   F32, F64, F32, F32, F32, F32, F32, F32, F32, F32, F32},
      /*F64 */ {0,   F64, F64, F64, F64, F64, F64, F64, F64, F64, F64, F64},
      /*B8  */ {0,   F32, F64, B8,  U8,  U16, U32, U64, S8,  S16, S32, S64},
      /*U8  */ {0,   F32, F64, U8,  U8,  U16, U32, U64, S8,  S16, S32, S64},
      /*U16 */ {0,   F32, F64, U16, U16, U16, U32, U64, U16, S16, S32, S64},
      /*U32 */ {0,   F32, F64, U32, U32, U32, U32, U64, S32, S32, S32, S64},
      /*U64 */ {0,   F32, F64, B8,  U64, U64, U64, U64, S64, S64, S64, S64},
      /*S8  */ {0,   F32, F64, S8,  S8,  S16, S32, S64, S8,  S16, S32, S64},
      /*S16 */ {0,   F32, F64, S16, S16, S16, S32, S64, S16, S16, S32, S64},
      /*S32 */ {0,   F32, F64, S32, S32, S32, S32, S64, S32, S32, S32, S64},
      /*S64 */ {0,   F32, F64, S64, S64, S64, S64, S64, S64, S64, S64, S64},
  };
  if (a <= 0 || a >= N_DATATYPE || b <= 0 || b >= N_DATATYPE) throw std::runtime_error("promote: bad type");
  return T[a][b];
}

// Operator classification (util/UnaryType.cpp:33-48, util/BinaryType.cpp:33-47).
inline bool unary_is_operator(int t) { return t > NONE_UNARY && t < MARK_UNARY; }
inline bool unary_is_relational(int t) { return t == NOT; }
inline bool unary_is_bitwise(int t) { return t == bNOT; }
inline bool binary_is_operator(int t) { return t > NONE_BINARY && t < MARK_BINARY; }
inline bool binary_is_relational(int t) { return t >= EQ && t <= OR; }
inline bool binary_is_bitwise(int t) { return t >= bAND && t <= SHR; }

// Spelling of each operator in C (util/UnaryType.cpp:87-120, util/BinaryType.cpp:80-106).
inline const char* unary_code(int t) {
  switch (t) {
    case POS: return "+"; case NEG: return "-"; case NOT: return "!"; case bNOT: return "~";
    case SIN: return "sin"; case COS: return "cos"; case TAN: return "tan"; case ASIN: return "asin";
    case ACOS: return "acos"; case ATAN: return "atan"; case SINH: return "sinh"; case COSH: return "cosh";
    case TANH: return "tanh"; case ASINH: return "asinh"; case ACOSH: return "acosh"; case ATANH: return "atanh";
    case EXP: return "exp"; case EXP2: return "exp2"; case EXP10: return "exp10"; case LOG: return "log";
    case LOG2: return "log2"; case LOG10: return "log10"; case SQRT: return "sqrt"; case CBRT: return "cbrt";
    case ABS: return "fabs"; case CEIL: return "ceil"; case FLOOR: return "floor"; case TRUNC: return "trunc";
    case ROUND: return "round";
    default: throw std::runtime_error("unary_code: bad op");
  }
}
inline const char* binary_code(int t) {
  switch (t) {
    case ADD: return "+"; case SUB: return "-"; case MUL: return "*"; case DIV: return "/"; case MOD: return "%";
    case EQ: return "=="; case NE: return "!="; case LT: return "<"; case GT: return ">"; case LE: return "<=";
    case GE: return ">="; case AND: return "&&"; case OR: return "||"; case bAND: return "&"; case bOR: return "|";
    case bXOR: return "^"; case SHL: return "<<"; case SHR: return ">>"; case MAX2: return "max";
    case MIN2: return "min"; case ATAN2: return "atan2"; case POW: return "pow"; case HYPOT: return "hypot";
    case FMOD: return "fmod";
    default: throw std::runtime_error("binary_code: bad op");
  }
}

inline bool reduction_valid(int r) { return r == SUM || r == PROD || r == rAND || r == rOR || r == MAX || r == MIN; }

// ---- Variant helpers --------------------------------------------------------
inline Variant variant_zero(int dt) { Variant v; v.u.raw = 0; v.dtype = dt; return v; }

template <typename T> inline T variant_as(const Variant& v) {
  switch (v.dtype) {
    case F32: return (T)v.u.f32; case F64: return (T)v.u.f64; case B8: return (T)v.u.b8;
    case U8: return (T)v.u.u8; case U16: return (T)v.u.u16; case U32: return (T)v.u.u32;
    case U64: return (T)v.u.u64; case S8: return (T)v.u.s8; case S16: return (T)v.u.s16;
    case S32: return (T)v.u.s32; case S64: return (T)v.u.s64;
    default: throw std::runtime_error("variant_as: empty variant");
  }
}
// C conversion of a value to another type (the conversion a C assignment does).
inline Variant variant_convert(const Variant& v, int dt) {
  Variant r = variant_zero(dt);
  switch (dt) {
    case F32: r.u.f32 = variant_as<float>(v); break;
    case F64: r.u.f64 = variant_as<double>(v); break;
    case B8: r.u.b8 = variant_as<double>(v) != 0; break;
    case U8: r.u.u8 = is_float_dt(v.dtype) ? (uint8_t)(int64_t)variant_as<double>(v) : (uint8_t)variant_as<int64_t>(v); break;
    case U16: r.u.u16 = is_float_dt(v.dtype) ? (uint16_t)(int64_t)variant_as<double>(v) : (uint16_t)variant_as<int64_t>(v); break;
    case U32: r.u.u32 = is_float_dt(v.dtype) ? (uint32_t)(int64_t)variant_as<double>(v) : (uint32_t)variant_as<int64_t>(v); break;
    case U64: r.u.u64 = variant_as<uint64_t>(v); break;
    case S8: r.u.s8 = (int8_t)variant_as<int64_t>(v); break;
    case S16: r.u.s16 = (int16_t)variant_as<int64_t>(v); break;
    case S32: r.u.s32 = (int32_t)variant_as<int64_t>(v); break;
    case S64: r.u.s64 = variant_as<int64_t>(v); break;
    default: throw std::runtime_error("variant_convert: bad type");
  }
  return r;
}
// Bitwise equality on the stored bytes of the type (GVN key equality).
inline bool variant_equal(const Variant& a, const Variant& b) {
  if (a.dtype != b.dtype) return false;
  if (a.dtype <= 0 || a.dtype >= N_DATATYPE) return a.dtype == b.dtype;
  return std::memcmp(&a.u, &b.u, sizeof_dt(a.dtype)) == 0;
}
inline uint64_t variant_bits(const Variant& v) {
  uint64_t r = 0;
  if (v.dtype > 0 && v.dtype < N_DATATYPE) std::memcpy(&r, &v.u, sizeof_dt(v.dtype));
  return r;
}

// Exact C literal for a typed value, as used inside generated device code.
// Floating values are printed as hexadecimal floats so no rounding happens.
std::string variant_literal(const Variant& v);

// Reduction neutral values (util/ReductionType.tpl:40-65): 0 for SUM, 1 for
// PROD, all-ones for rAND, zero for rOR, the type's lowest value (or -inf)
// for MAX and the highest value (or +inf) for MIN.
Variant reduction_neutral(int red, int dt);

}  // namespace mapx
