// Node factories (type and shape inference) and value numbering.
#include "dag.hpp"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <set>
#include <sstream>

namespace mapx {

const char* kind_name(Kind k) {
  switch (k) {
    case Kind::Constant: return "Constant"; case Kind::Index: return "Index"; case Kind::Cast: return "Cast";
    case Kind::Unary: return "Unary"; case Kind::Binary: return "Binary"; case Kind::Conditional: return "Conditional";
    case Kind::Neighbor: return "Neighbor"; case Kind::Convolution: return "Convolution";
    case Kind::FocalFunc: return "FocalFunc"; case Kind::ZonalReduc: return "ZonalReduc";
    case Kind::ZonalStats: return "ZonalStats"; case Kind::Read: return "Read"; case Kind::Write: return "Write";
    case Kind::Scalar: return "Scalar"; case Kind::RadialScan: return "RadialScan";
    case Kind::FocalPercent: return "FocalPercent"; case Kind::Diversity: return "Diversity";
    case Kind::SpreadNeighbor: return "SpreadNeighbor"; case Kind::SpreadScan: return "SpreadScan";
    case Kind::Stats: return "Stats"; case Kind::Barrier: return "Barrier";
    case Kind::Access: return "Access"; case Kind::LhsAccess: return "LhsAccess";
    case Kind::Rand: return "Rand";
  }
  return "?";
}

static void fail(const std::string& m) { throw std::runtime_error(m); }

std::string variant_literal(const Variant& v) {
  char buf[96];
  switch (v.dtype) {
    case F32:
      if (std::isinf(v.u.f32)) return v.u.f32 > 0 ? "__builtin_inff()" : "(-__builtin_inff())";
      if (std::isnan(v.u.f32)) return "__builtin_nanf(\"\")";
      std::snprintf(buf, sizeof buf, "%af", (double)v.u.f32); return buf;
    case F64:
      if (std::isinf(v.u.f64)) return v.u.f64 > 0 ? "__builtin_inf()" : "(-__builtin_inf())";
      if (std::isnan(v.u.f64)) return "__builtin_nan(\"\")";
      std::snprintf(buf, sizeof buf, "%a", v.u.f64); return buf;
    case B8: return v.u.b8 ? "((unsigned char)1)" : "((unsigned char)0)";
    case U8: std::snprintf(buf, sizeof buf, "((unsigned char)%uu)", (unsigned)v.u.u8); return buf;
    case U16: std::snprintf(buf, sizeof buf, "((unsigned short)%uu)", (unsigned)v.u.u16); return buf;
    case U32: std::snprintf(buf, sizeof buf, "%uu", v.u.u32); return buf;
    case U64: std::snprintf(buf, sizeof buf, "%lluull", (unsigned long long)v.u.u64); return buf;
    case S8: std::snprintf(buf, sizeof buf, "((signed char)%d)", (int)v.u.s8); return buf;
    case S16: std::snprintf(buf, sizeof buf, "((short)%d)", (int)v.u.s16); return buf;
    case S32:
      if (v.u.s32 == INT32_MIN) return "(-2147483647-1)";
      std::snprintf(buf, sizeof buf, "%d", v.u.s32); return buf;
    case S64:
      if (v.u.s64 == INT64_MIN) return "(-9223372036854775807ll-1)";
      std::snprintf(buf, sizeof buf, "%lldll", (long long)v.u.s64); return buf;
    default: fail("variant_literal: empty variant");
  }
  return "";
}

Variant reduction_neutral(int red, int dt) {
  Variant v = variant_zero(dt);
  auto set_max = [&]() {
    switch (dt) {
      case F32: v.u.f32 = INFINITY; break; case F64: v.u.f64 = INFINITY; break;
      case B8: v.u.b8 = true; break; case U8: v.u.u8 = UINT8_MAX; break; case U16: v.u.u16 = UINT16_MAX; break;
      case U32: v.u.u32 = UINT32_MAX; break; case U64: v.u.u64 = UINT64_MAX; break;
      case S8: v.u.s8 = INT8_MAX; break; case S16: v.u.s16 = INT16_MAX; break;
      case S32: v.u.s32 = INT32_MAX; break; case S64: v.u.s64 = INT64_MAX; break;
    }
  };
  auto set_min = [&]() {
    switch (dt) {
      case F32: v.u.f32 = -INFINITY; break; case F64: v.u.f64 = -INFINITY; break;
      case S8: v.u.s8 = INT8_MIN; break; case S16: v.u.s16 = INT16_MIN; break;
      case S32: v.u.s32 = INT32_MIN; break; case S64: v.u.s64 = INT64_MIN; break;
      default: v.u.raw = 0; break;  // unsigned and B8: lowest is 0
    }
  };
  switch (red) {
    case SUM: break;
    case PROD: v = variant_convert(Variant{{.s32 = 1}, S32}, dt); break;
    case rAND: set_max(); break;
    case rOR: set_min(); break;
    case MAX: set_min(); break;
    case MIN: set_max(); break;
    default: fail("reduction_neutral: bad reduction");
  }
  return v;
}

// ---- Node --------------------------------------------------------------------

int Node::pattern() const {
  if (cut) return FREE;
  switch (kind) {
    case Kind::Cast: case Kind::Unary: case Kind::Binary: case Kind::Conditional: case Kind::Diversity:
    case Kind::Access: case Kind::LhsAccess: return LOCAL;  // Access.hpp:44, LhsAccess.hpp:45
    case Kind::Rand: return LOCAL;  // Rand.hpp:41
    case Kind::Neighbor: case Kind::Convolution: case Kind::FocalFunc: case Kind::FocalPercent:
    case Kind::SpreadNeighbor: return FOCAL;  // SpreadNeighbor.hpp:42
    case Kind::ZonalReduc: case Kind::ZonalStats: return ZONAL;
    case Kind::RadialScan: return RADIAL;
    case Kind::SpreadScan: return SPREAD;  // SpreadScan.hpp:48
    case Kind::Stats: return LOCAL;        // an identity statement here (the reference: SPECIAL, Stats.hpp:43)
    case Kind::Barrier: return BARRIER;    // Barrier.hpp:39
    default: return FREE;  // Constant, Index, Read, Write, Scalar (runtime/dag/Node.cpp:98)
  }
}

int Node::halo(int i) const {
  switch (kind) {
    case Kind::Neighbor: return std::abs(coord.arr[i]);       // Neighbor.cpp:86
    case Kind::Convolution: case Kind::FocalFunc: case Kind::FocalPercent: return mask.halo(i);  // Convolution.cpp:84
    case Kind::SpreadNeighbor: return 1;  // SpreadNeighbor.cpp:88-90
    default: return 0;
  }
}

std::string Node::signature(bool float_weights) const {
  std::ostringstream s;
  s << kind_name(kind)[0] << kind_name(kind)[1] << meta.nd << name_dt(meta.dt);
  switch (kind) {
    case Kind::Constant: s << "c" << variant_bits(cnst); break;
    case Kind::Index: case Kind::Cast: case Kind::Unary: case Kind::Binary:
    case Kind::ZonalReduc: s << "o" << op; break;
    case Kind::Neighbor: s << "n" << coord.arr[0] << "," << coord.arr[1]; break;
    case Kind::Convolution: case Kind::FocalFunc:
      s << "m" << op << ":" << mask.ds.arr[0] << "x" << mask.ds.arr[1] << name_dt(mask.dtype());
      if (float_weights || !is_float_dt(mask.dtype()))
        for (auto& w : mask.v) s << "," << variant_bits(w);
      break;
    case Kind::ZonalStats: s << "z" << nzones; break;
    case Kind::FocalPercent: s << "p" << op << ":" << mask.ds.arr[0] << "x" << mask.ds.arr[1]; break;
    case Kind::Diversity: s << "d" << op << ":" << prev.size() << name_dt(prev.empty() ? 0 : prev[0]->meta.dt); break;
    case Kind::RadialScan: s << "r" << op << "@" << coord.arr[0] << "," << coord.arr[1]; break;
    case Kind::SpreadNeighbor: s << "sn" << op << name_dt(prev[1]->meta.dt); break;
    case Kind::SpreadScan: s << "ss" << op << name_dt(prev[1]->meta.dt); break;
    case Kind::Access: case Kind::LhsAccess: s << "a" << coord.arr[0] << "," << coord.arr[1]; break;
    default: break;
  }
  return s.str();
}

// ---- Dag ---------------------------------------------------------------------

// Value-numbering key: the node's kind, inputs (by id), type, geometry and
// parameters as raw bytes (every field a node kind uses; the others keep
// their defaults, so they never tell equal nodes apart).
std::string Dag::key(const Node* n) const {
  std::string k;
  k.reserve(48 + 4 * n->prev.size() + 8 * n->mask.v.size() + n->path.size());
  auto i32 = [&](int v) { k.append((const char*)&v, sizeof v); };
  auto u64 = [&](uint64_t v) { k.append((const char*)&v, sizeof v); };
  i32((int)n->kind);
  i32((int)n->prev.size());
  for (auto* p : n->prev) i32(p->id);
  i32(n->meta.dt); i32(n->meta.nd);
  i32(n->meta.ds.arr[0]); i32(n->meta.ds.arr[1]); i32(n->meta.bs.arr[0]); i32(n->meta.bs.arr[1]);
  switch (n->kind) {
    case Kind::Constant: i32(n->cnst.dtype); u64(variant_bits(n->cnst)); break;
    case Kind::Read:
      if (n->host) { k.push_back('m'); u64(n->host->uid); }
      else { k.push_back('f'); k += n->path; }
      break;
    case Kind::Write: k += n->path; break;
    default:
      i32(n->op); i32(n->coord.arr[0]); i32(n->coord.arr[1]); i32(n->coord.n); i32(n->nzones);
      i32(n->mask.ds.arr[0]); i32(n->mask.ds.arr[1]); i32((int)n->mask.v.size());
      for (auto& w : n->mask.v) { i32(w.dtype); u64(variant_bits(w)); }
      break;
  }
  return k;
}

Node* Dag::add(std::unique_ptr<Node> n) {
  std::string k = key(n.get());
  auto it = gvn_.find(k);
  if (it != gvn_.end()) return it->second;  // SimplifierOnline.tpl:26-34
  // Output nodes share the id of the node they back (runtime/dag/IO.cpp: OutputNode)
  if ((n->kind == Kind::Write || n->kind == Kind::Scalar) && !n->prev.empty()) n->id = n->prev[0]->id;
  else n->id = id_count_++;
  Node* raw = n.get();
  for (auto* p : raw->prev) { p->next.push_back(raw); p->ref++; }  // Node::addNext (Node.cpp:83-86)
  for (auto* a : raw->aux) a->ref++;
  raw->gkey = k;
  nodes_.push_back(std::move(n));
  gvn_.emplace(std::move(k), raw);
  return raw;
}

static Meta meta_of(Dim2 ds, int dt, int mo, Dim2 bs) {
  Meta m;
  m.ds = ds; m.dt = dt; m.mo = mo; m.bs = bs;
  m.nd = (ds.n == 0) ? D0 : (ds.n == 1 ? D1 : D2);
  return m;
}

static void check_d2(const Meta& m, const char* who) {
  if (m.nd != D2) return;
  if (m.ds.arr[0] <= 0 || m.ds.arr[1] <= 0 || m.bs.arr[0] <= 0 || m.bs.arr[1] <= 0)
    fail(std::string(who) + ": raster needs positive data and block sizes");
}

// Constant (runtime/dag/Constant.cpp:29-33). The value must carry the declared type.
Node* Dag::constant(Variant v, Dim2 ds, int dt, int mo, Dim2 bs) {
  if (v.dtype != dt) fail("constant: value type differs from the declared type");
  auto n = std::make_unique<Node>();
  n->kind = Kind::Constant;
  n->meta = meta_of(ds, dt, mo, bs);
  check_d2(n->meta, "constant");
  n->cnst = v;
  if (n->meta.nd == D0) { n->value = v; n->evaluated = true; }
  return add(std::move(n));
}

// Index (runtime/dag/Index.cpp): the global coordinate along D1 (x) or D2 (y), S32.
Node* Dag::index(Node* p, int dim) {
  if (!p || p->meta.nd != D2) fail("index: needs a D2 raster");
  if (dim != D1 && dim != D2) fail("index: dim must be D1 or D2");
  auto n = std::make_unique<Node>();
  n->kind = Kind::Index;
  n->meta = p->meta; n->meta.dt = S32;
  n->op = dim;
  n->prev = {p};
  return add(std::move(n));
}

// Rand (runtime/dag/Rand.cpp:78-86, Skeleton.cpp:183-206): the seed raster's
// shape, type dt (the seed's type when NONE_DATATYPE); each cell is a
// Philox2x32-10 output of counter {seed cell, 0} and key {the cell's linear
// index y * width + x}, integers taking its low bytes and floats the first
// word divided by UINT_MAX (F64: both words over ULONG_MAX).
Node* Dag::rand(Node* seed, int dt, int mo) {
  if (!seed || seed->meta.nd != D2) fail("rand: needs a D2 seed raster");
  if (is_float_dt(seed->meta.dt) || seed->meta.dt == B8) fail("rand: the seed raster must have an integer type");
  if (dt == NONE_DATATYPE) dt = seed->meta.dt;
  if (dt <= NONE_DATATYPE || dt == B8 || dt >= N_DATATYPE) fail("rand: bad type");
  auto n = std::make_unique<Node>();
  n->kind = Kind::Rand;
  n->meta = seed->meta;
  n->meta.dt = dt;
  if (mo != NONE_MEMORDER) n->meta.mo = mo;
  n->prev = {seed};
  return add(std::move(n));
}

// Cast (runtime/dag/Cast.cpp): same shape, new type.
Node* Dag::cast(Node* p, int dt) {
  if (!p) fail("cast: null node");
  if (dt <= NONE_DATATYPE || dt >= N_DATATYPE) fail("cast: bad type");
  auto n = std::make_unique<Node>();
  n->kind = Kind::Cast;
  n->meta = p->meta; n->meta.dt = dt;
  n->op = dt;
  n->prev = {p};
  return add(std::move(n));
}

// Unary (runtime/dag/Unary.cpp:30-49): type kept, U8 for the relational NOT,
// bitwise NOT only on unsigned types.
Node* Dag::unary(Node* p, int op) {
  if (!p) fail("unary: null node");
  if (op <= NONE_UNARY || op >= N_UNARY || op == MARK_UNARY) fail("unary: bad op");
  if (unary_is_bitwise(op) && !is_unsigned_dt(p->meta.dt)) fail("unary: bitwise op needs an unsigned type");
  auto n = std::make_unique<Node>();
  n->kind = Kind::Unary;
  n->meta = p->meta;
  if (unary_is_relational(op)) n->meta.dt = U8;
  n->op = op;
  n->prev = {p};
  return add(std::move(n));
}

// Binary (runtime/dag/Binary.cpp:31-63): shape of the non-D0 side, promoted
// type, U8 for relational and logical ops.
Node* Dag::binary(Node* l, Node* r, int op) {
  if (!l || !r) fail("binary: null node");
  if (op <= NONE_BINARY || op >= N_BINARY || op == MARK_BINARY) fail("binary: bad op");
  if (binary_is_bitwise(op) && !(is_unsigned_dt(l->meta.dt) && is_unsigned_dt(r->meta.dt)))
    fail("binary: bitwise op needs unsigned types");
  auto n = std::make_unique<Node>();
  n->kind = Kind::Binary;
  if (l->meta.nd == D0) n->meta = r->meta;
  else if (r->meta.nd == D0) n->meta = l->meta;
  else {
    if (l->meta.nd != r->meta.nd || !dim2_eq(l->meta.ds, r->meta.ds) || l->meta.mo != r->meta.mo ||
        !dim2_eq(l->meta.bs, r->meta.bs))
      fail("binary: operands differ in shape, memory order or block size");
    n->meta = l->meta;
  }
  n->meta.dt = binary_is_relational(op) ? U8 : promote(l->meta.dt, r->meta.dt);
  n->op = op;
  n->prev = {l, r};
  return add(std::move(n));
}

// Conditional (runtime/dag/Conditional.cpp:30-90): shape of the non-D0
// operands, both branches of one type.
Node* Dag::conditional(Node* c, Node* l, Node* r) {
  if (!c || !l || !r) fail("conditional: null node");
  if (l->meta.dt != r->meta.dt) fail("conditional: branches differ in type");
  auto n = std::make_unique<Node>();
  n->kind = Kind::Conditional;
  const Node* shape = nullptr;
  for (Node* x : {c, l, r})
    if (x->meta.nd != D0) {
      if (shape && (!dim2_eq(shape->meta.ds, x->meta.ds) || !dim2_eq(shape->meta.bs, x->meta.bs)))
        fail("conditional: operands differ in shape");
      if (!shape) shape = x;
    }
  n->meta = shape ? shape->meta : l->meta;
  n->meta.dt = l->meta.dt;
  n->prev = {c, l, r};
  return add(std::move(n));
}

// Neighbor (runtime/dag/Neighbor.cpp): value of the cell at +coord.
Node* Dag::neighbor(Node* p, Dim2 coord) {
  if (!p || p->meta.nd != D2) fail("neighbor: needs a D2 raster");
  if (coord.n != 2) fail("neighbor: coordinate needs 2 values");
  if (std::abs(coord.arr[0]) > p->meta.bs.arr[0] || std::abs(coord.arr[1]) > p->meta.bs.arr[1])
    fail("neighbor: offset larger than the block size");
  auto n = std::make_unique<Node>();
  n->kind = Kind::Neighbor;
  n->meta = p->meta;
  n->coord = coord;
  n->prev = {p};
  return add(std::move(n));
}

static void check_mask(const Mask& m, const Node* p) {
  if (m.ds.n != 2 || m.ds.arr[0] <= 0 || m.ds.arr[1] <= 0) fail("mask: needs 2 positive dimensions");
  if ((m.ds.arr[0] % 2) == 0 || (m.ds.arr[1] % 2) == 0) fail("mask: dimensions must be odd");
  if ((int)m.v.size() != m.ds.arr[0] * m.ds.arr[1]) fail("mask: value count differs from its dimensions");
  for (auto& w : m.v) if (w.dtype <= NONE_DATATYPE || w.dtype >= N_DATATYPE) fail("mask: untyped value");
  if (m.halo(0) > p->meta.bs.arr[0] || m.halo(1) > p->meta.bs.arr[1]) fail("mask: halo larger than the block");
}

// Convolution (runtime/dag/Convolution.cpp:30-43): type promote(input, mask).
Node* Dag::convolution(Node* p, const Mask& m) {
  if (!p || p->meta.nd != D2) fail("convolution: needs a D2 raster");
  check_mask(m, p);
  auto n = std::make_unique<Node>();
  n->kind = Kind::Convolution;
  n->meta = p->meta;
  n->meta.dt = promote(p->meta.dt, m.dtype());
  n->mask = m;
  n->prev = {p};
  return add(std::move(n));
}

// FocalFunc (runtime/dag/FocalFunc.cpp): reduction over the mask's extent in
// the input type. The GPU skeleton of the reference ignores the weights
// (FocalSkeleton.cpp:338-372); so does this backend.
Node* Dag::focalFunc(Node* p, const Mask& m, int red) {
  if (!p || p->meta.nd != D2) fail("focalFunc: needs a D2 raster");
  if (!reduction_valid(red)) fail("focalFunc: bad reduction");
  check_mask(m, p);
  auto n = std::make_unique<Node>();
  n->kind = Kind::FocalFunc;
  n->meta = p->meta;
  n->mask = m;
  n->op = red;
  n->prev = {p};
  return add(std::move(n));
}

// ZonalReduc (runtime/dag/ZonalReduc.cpp:30-52): whole-raster reduction to a
// D0 value of the input type.
Node* Dag::zonalReduc(Node* p, int red) {
  if (!p || p->meta.nd != D2) fail("zonalReduc: needs a D2 raster");
  if (!reduction_valid(red)) fail("zonalReduc: bad reduction");
  auto n = std::make_unique<Node>();
  n->kind = Kind::ZonalReduc;
  n->meta = meta_of(dim2_none(), p->meta.dt, p->meta.mo, dim2_none());
  n->meta.ds.n = 0;
  n->op = red;
  n->prev = {p};
  n->value = reduction_neutral(red, p->meta.dt);
  return add(std::move(n));
}

// FocalPercent (runtime/dag/FocalPercent.cpp:29-41): the share of the mask's
// window (its values are not used, only its size) that equals the centre
// (AGE) or that the centre exceeds (ILE); the input's type.
Node* Dag::focalPercent(Node* p, const Mask& m, int type) {
  if (!p || p->meta.nd != D2) fail("focalPercent: needs a D2 raster");
  if (type != 1 && type != 2) fail("focalPercent: type must be AGE (1) or ILE (2)");
  check_mask(m, p);
  auto n = std::make_unique<Node>();
  n->kind = Kind::FocalPercent;
  n->meta = p->meta;
  n->mask = m;
  n->op = type;
  n->prev = {p};
  return add(std::move(n));
}

// Diversity (runtime/dag/Diversity.cpp:40-63) of two or more rasters of one
// shape and type: per cell, the number of distinct values (VARI, an S8
// raster), the most (MAJO) or least (MINO) frequent value, or MEAN; the
// inputs' type otherwise.
Node* Dag::diversity(const std::vector<Node*>& ps, int type) {
  if (ps.size() < 2) fail("diversity: needs at least two rasters");
  if (ps.size() > 16) fail("diversity: at most 16 rasters");
  if (type < 1 || type > 4) fail("diversity: type must be VARI, MAJO, MINO or MEAN (1-4)");
  for (Node* p : ps) {
    if (!p || p->meta.nd != D2) fail("diversity: needs D2 rasters");
    if (!dim2_eq(p->meta.ds, ps[0]->meta.ds) || !dim2_eq(p->meta.bs, ps[0]->meta.bs) || p->meta.dt != ps[0]->meta.dt)
      fail("diversity: rasters differ in shape or type");
  }
  auto n = std::make_unique<Node>();
  n->kind = Kind::Diversity;
  n->meta = ps[0]->meta;
  if (type == 1) n->meta.dt = S8;  // variety is always an integer (Diversity.cpp:56-58)
  n->op = type;
  n->prev = ps;
  return add(std::move(n));
}

// SpreadNeighbor (runtime/dag/SpreadNeighbor.cpp:31-47): per cell, the
// reduction of the neighbours' values (3x3 window, y outer, x inner, centre
// excluded) whose D8 direction code in `dir` points at the cell, starting from
// the reduction's neutral value; the input's type. It is the gather step of the
// reference's flowAccuGather (front/Raster.cpp:583-606) as one focal node.
Node* Dag::spreadNeighbor(Node* p, Node* dir, int red) {
  if (!p || !dir || p->meta.nd != D2 || dir->meta.nd != D2) fail("spreadNeighbor: needs two D2 rasters");
  if (!dim2_eq(p->meta.ds, dir->meta.ds) || !dim2_eq(p->meta.bs, dir->meta.bs))
    fail("spreadNeighbor: the value and direction rasters differ in shape");
  if (!reduction_valid(red)) fail("spreadNeighbor: bad reduction");
  if (p->meta.bs.arr[0] < 1 || p->meta.bs.arr[1] < 1) fail("spreadNeighbor: bad block size");
  auto n = std::make_unique<Node>();
  n->kind = Kind::SpreadNeighbor;
  n->meta = p->meta;
  n->op = red;
  n->prev = {p, dir};
  return add(std::move(n));
}

// SpreadScan (runtime/dag/SpreadScan.cpp:31-45): accumulation of `p` along the
// D8 flow directions of `dir` (U8). Each cell's result is the total that leaves
// it towards its downstream cell: its own value plus everything that reaches it
// from upstream; a cell without a downstream cell inside the raster keeps the
// neutral value. SUM only: the reference's kernel applies `+=` whatever the
// reduction (SpreadingSkeleton.cpp:318-320).
Node* Dag::spreadScan(Node* p, Node* dir, int red) {
  if (!p || !dir || p->meta.nd != D2 || dir->meta.nd != D2) fail("spreadScan: needs two D2 rasters");
  if (!dim2_eq(p->meta.ds, dir->meta.ds) || !dim2_eq(p->meta.bs, dir->meta.bs))
    fail("spreadScan: the value and direction rasters differ in shape");
  if (dir->meta.dt != U8) fail("spreadScan: directions must be U8 (SpreadScan.cpp:37)");
  if (red != SUM) fail("spreadScan: only SUM is supported (the reference accumulates with +=)");
  if (p->meta.dt == B8) fail("spreadScan: needs a numeric raster");
  // the scan's walk lists and cell indices are 32-bit element indices of the
  // block-tiled raster (edge blocks padded to full size)
  const uint64_t cells = (uint64_t)p->meta.nb(0) * p->meta.nb(1) * p->meta.bs.arr[0] * p->meta.bs.arr[1];
  if (cells >= (uint64_t(1) << 32))
    fail("spreadScan: rasters of 2^32 or more cells (edge blocks padded to whole blocks) are not supported");
  auto n = std::make_unique<Node>();
  n->kind = Kind::SpreadScan;
  n->meta = p->meta;
  n->op = red;
  n->prev = {p, dir};
  return add(std::move(n));
}

// Stats (runtime/dag/Stats.cpp:30-61): the input's values, with its global
// maximum and minimum (ZonalReduc MAX / MIN of the input) computed whenever
// the Stats node is evaluated: the node holds them (aux) as the reference's
// prev_list does, without making them inputs of its own statement.
Node* Dag::stats(Node* p) {
  if (!p || p->meta.nd != D2) fail("stats: needs a D2 raster");
  Node* mx = zonalReduc(p, MAX);
  Node* mn = zonalReduc(p, MIN);
  auto n = std::make_unique<Node>();
  n->kind = Kind::Stats;
  n->meta = p->meta;
  n->prev = {p};
  n->aux = {mx, mn};
  return add(std::move(n));
}

// Barrier (runtime/dag/Barrier.cpp): the input's values; its BARRIER pattern
// keeps the groups on either side apart (Pattern.cpp:65-119).
Node* Dag::barrier(Node* p) {
  if (!p) fail("barrier: null node");
  auto n = std::make_unique<Node>();
  n->kind = Kind::Barrier;
  n->meta = p->meta;
  n->prev = {p};
  return add(std::move(n));
}

// Access (runtime/dag/Access.cpp:29-40): the value of one cell of a D2 raster,
// a D0 value of the raster's type.
Node* Dag::access(Node* p, Dim2 cell) {
  if (!p || p->meta.nd != D2) fail("access: needs a D2 raster");
  if (cell.n != 2 || cell.arr[0] < 0 || cell.arr[1] < 0 || cell.arr[0] >= p->meta.ds.arr[0] ||
      cell.arr[1] >= p->meta.ds.arr[1])
    fail("access: the cell must lie inside the raster");
  auto n = std::make_unique<Node>();
  n->kind = Kind::Access;
  n->meta.dt = p->meta.dt;
  n->meta.nd = D0;
  n->meta.mo = p->meta.mo;
  n->coord = cell;
  n->prev = {p};
  return add(std::move(n));
}

// LhsAccess (runtime/dag/LhsAccess.cpp; Skeleton.cpp:290-297): `lhs` with the
// one cell replaced by the D0 value `rhs` (converted to lhs's type by the
// caller, python/map.py:213).
Node* Dag::lhsAccess(Node* lhs, Node* rhs, Dim2 cell) {
  if (!lhs || !rhs || lhs->meta.nd != D2 || rhs->meta.nd != D0) fail("lhsAccess: needs a D2 raster and a D0 value");
  if (cell.n != 2 || cell.arr[0] < 0 || cell.arr[1] < 0 || cell.arr[0] >= lhs->meta.ds.arr[0] ||
      cell.arr[1] >= lhs->meta.ds.arr[1])
    fail("lhsAccess: the cell must lie inside the raster");
  auto n = std::make_unique<Node>();
  n->kind = Kind::LhsAccess;
  n->meta = lhs->meta;
  n->coord = cell;
  n->prev = {lhs, rhs};
  return add(std::move(n));
}

// Radial scan from the start cell (RadialScan::Factory, runtime/dag/RadialScan.cpp:34-45): a
// raster of the input's shape and type. Floating inputs only: the reference's
// interpolation weight is meaningless for integers (RadiatingSkeleton's
// linear_interpolation truncates it to 1 - dif1).
Node* Dag::radialScan(Node* p, int red, Dim2 start) {
  if (!p || p->meta.nd != D2) fail("radialScan: needs a D2 raster");
  if (!is_float_dt(p->meta.dt)) fail("radialScan: needs a floating raster (F32 or F64)");
  if (red != SUM && red != PROD && red != MAX && red != MIN) fail("radialScan: reduction must be SUM, PROD, MAX or MIN");
  if (start.n != 2 || start.arr[0] < 0 || start.arr[1] < 0 || start.arr[0] >= p->meta.ds.arr[0] ||
      start.arr[1] >= p->meta.ds.arr[1])
    fail("radialScan: the start cell must lie inside the raster");
  auto n = std::make_unique<Node>();
  n->kind = Kind::RadialScan;
  n->meta = p->meta;
  n->op = red;
  n->coord = start;
  n->prev = {p};
  return add(std::move(n));
}

// Per-zone statistics (no reference counterpart): sum, min, max and count of
// `data` per value of the integer `zone` raster in [0, nzones).
Node* Dag::zonalStats(Node* data, Node* zone, int nzones) {
  if (!data || !zone || data->meta.nd != D2 || zone->meta.nd != D2) fail("zonalStats: needs D2 rasters");
  if (!dim2_eq(data->meta.ds, zone->meta.ds) || !dim2_eq(data->meta.bs, zone->meta.bs))
    fail("zonalStats: data and zone rasters differ in shape");
  if (is_float_dt(zone->meta.dt) || zone->meta.dt == B8) fail("zonalStats: zone raster must be an integer type");
  if (nzones <= 0 || nzones > 4096) fail("zonalStats: number of zones must be in [1, 4096]");
  auto n = std::make_unique<Node>();
  n->kind = Kind::ZonalStats;
  n->meta = meta_of(dim2(nzones, 1), data->meta.dt, data->meta.mo, dim2(nzones, 1));
  n->meta.nd = D1;
  n->nzones = nzones;
  n->prev = {data, zone};
  return add(std::move(n));
}

Node* Dag::input(std::shared_ptr<HostRaster> r) {
  auto n = std::make_unique<Node>();
  n->kind = Kind::Read;
  n->meta = r->meta;
  check_d2(n->meta, "input");
  n->host = std::move(r);
  return add(std::move(n));
}

Node* Dag::write(Node* p, const std::string& path) {
  if (!p || p->meta.nd != D2) fail("write: needs a D2 raster");
  if (path.empty()) fail("write: empty path");
  auto n = std::make_unique<Node>();
  n->kind = Kind::Write;
  n->meta = p->meta;
  n->path = path;
  n->prev = {p};
  return add(std::move(n));
}

Node* Dag::scalar(Node* p) {
  if (!p || p->meta.nd != D0) fail("scalar: needs a D0 node");
  auto n = std::make_unique<Node>();
  n->kind = Kind::Scalar;
  n->meta = p->meta;
  n->prev = {p};
  return add(std::move(n));
}

void Dag::collect(const std::function<void(Node*)>& on_delete) {
  // Reverse id order: unlinking a node may free its prevs (Runtime.cpp:375-393).
  bool any = false;
  for (auto it = nodes_.rbegin(); it != nodes_.rend(); ++it) {
    Node* n = it->get();
    if (n->ref > 0) continue;
    any = true;
    for (Node* p : n->prev) {
      auto f = std::find(p->next.begin(), p->next.end(), n);
      if (f != p->next.end()) { p->next.erase(f); p->ref--; }
    }
    n->prev.clear();
    for (Node* a : n->aux) a->ref--;
    n->aux.clear();
  }
  if (!any) return;
  std::vector<std::unique_ptr<Node>> keep;
  keep.reserve(nodes_.size());
  for (auto& n : nodes_) {
    if (n->ref > 0) { keep.push_back(std::move(n)); continue; }
    auto g = gvn_.find(n->gkey);
    if (g != gvn_.end() && g->second == n.get()) gvn_.erase(g);
    if (on_delete) on_delete(n.get());
  }
  nodes_.swap(keep);
}

void Dag::collect_grown(const std::function<void(Node*)>& on_delete) {
  if (nodes_.size() < std::max<size_t>(4096, 2 * live_after_collect_)) return;
  collect(on_delete);
  live_after_collect_ = nodes_.size();
}

std::vector<Node*> list_upwards(const std::vector<Node*>& roots) {
  std::set<Node*> seen;
  std::vector<Node*> stack(roots.begin(), roots.end()), out;
  while (!stack.empty()) {
    Node* n = stack.back(); stack.pop_back();
    if (!seen.insert(n).second) continue;
    out.push_back(n);
    for (Node* p : n->prev) stack.push_back(p);
  }
  std::sort(out.begin(), out.end(), [](Node* a, Node* b) {
    if (a->id != b->id) return a->id < b->id;
    return (int)a->isOutput() < (int)b->isOutput();
  });
  return out;
}

}  // namespace mapx
