#include "codegen.hpp"
#include "dev.hpp"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <map>
#include <set>
#include <sstream>
#include <unordered_map>

namespace mapx {

namespace {

// Group-local names: generated source must not depend on global node ids, so
// that equal groups (e.g. every generation of an iterated stencil) share one
// compiled kernel.
thread_local const std::map<const Node*, int>* g_local = nullptr;
std::string LN(const Node* n) {
  auto it = g_local->find(n);
  if (it == g_local->end()) throw std::runtime_error("codegen: node outside its group");
  return std::to_string(it->second);
}
std::string V(const Node* n) { return "v" + LN(n); }
const char* CT(int dt) { return ctype_dt(dt); }

bool contains(const std::vector<Node*>& v, const Node* n) { return std::find(v.begin(), v.end(), n) != v.end(); }

std::string hex_hash(const std::string& s) {
  // FNV-1a 64: stable across processes (std::hash is not required to be).
  uint64_t h = 1469598103934665603ull;
  for (unsigned char c : s) { h ^= c; h *= 1099511628211ull; }
  char buf[24];
  std::snprintf(buf, sizeof buf, "%016llx", (unsigned long long)h);
  return buf;
}

const char* red_op(int red) {
  switch (red) {
    case SUM: return "mx::OpSum"; case PROD: return "mx::OpProd"; case rAND: return "mx::OpAnd";
    case rOR: return "mx::OpOr"; case MAX: return "mx::OpMax"; case MIN: return "mx::OpMin";
  }
  throw std::runtime_error("codegen: bad reduction");
}

// Accumulator type: floating SUM and PROD accumulate in FP64 (a fixed-order
// sum of FP64 partials); every other reduction runs in the data type itself.
int acc_type(int red, int dt) { return (is_float_dt(dt) && (red == SUM || red == PROD)) ? F64 : dt; }

std::string lit(const Variant& v) { return variant_literal(v); }

// `fast`: FP32 sqrt goes through mx::sqrt_f32 (v_sqrt_f32 on its normal
// range, within 1 ulp) instead of the correctly rounded library sqrtf.
std::string unary_stmt(const Node* n, const std::string& p, int pdt, bool fast) {
  const int op = n->op, dt = n->meta.dt;
  if (unary_is_operator(op)) return std::string("(") + CT(dt) + ")(" + unary_code(op) + p + ")";
  if (op == ABS && !is_float_dt(dt)) return std::string("(") + CT(dt) + ")(" + p + " < 0 ? -" + p + " : " + p + ")";
  if (fast && dt == F32 && op == SQRT) return "mx::sqrt_f32(" + p + ")";
  std::string fn = op == ABS ? "fabs" : unary_code(op);
  if (dt == F32) return fn + "f(" + p + ")";
  if (dt == F64) return fn + "(" + p + ")";
  (void)pdt;
  return std::string("(") + CT(dt) + ")" + fn + "((double)" + p + ")";
}

// `fast`: FP32 atan2 goes through mx::atan2_f32 (within 4 ulp; the library
// atan2f outside its fast domain).
std::string binary_stmt(const Node* n, const std::string& l, const std::string& r, bool fast) {
  const int op = n->op, dt = n->meta.dt;
  const int ldt = n->prev[0]->meta.dt, rdt = n->prev[1]->meta.dt;
  if (op == MOD && (is_float_dt(ldt) || is_float_dt(rdt))) {
    int t = promote(ldt, rdt);
    std::string f = t == F32 ? "fmodf" : "fmod";
    return std::string("(") + CT(dt) + ")" + f + "((" + CT(t) + ")" + l + ", (" + CT(t) + ")" + r + ")";
  }
  if (op == DIV && n->prev[1]->meta.nd == D0 && (is_float_dt(ldt) || is_float_dt(rdt))) {
    // floating division by a work-group-uniform divisor (C++ operand conversions)
    const int t = (ldt == F64 || rdt == F64) ? F64 : F32;
    return std::string("(") + CT(dt) + ")mx::div_uniform((" + CT(t) + ")" + l + ", (" + CT(t) + ")" + r + ")";
  }
  if (binary_is_operator(op)) return std::string("(") + CT(dt) + ")(" + l + " " + binary_code(op) + " " + r + ")";
  std::string L = std::string("((") + CT(dt) + ")" + l + ")", R = std::string("((") + CT(dt) + ")" + r + ")";
  if (op == MAX2 || op == MIN2) {
    if (dt == F32) return std::string(op == MAX2 ? "fmaxf(" : "fminf(") + L + ", " + R + ")";
    if (dt == F64) return std::string(op == MAX2 ? "fmax(" : "fmin(") + L + ", " + R + ")";
    return std::string("(") + L + (op == MAX2 ? " > " : " < ") + R + " ? " + L + " : " + R + ")";
  }
  std::string fn = binary_code(op);
  if (fast && dt == F32 && op == ATAN2) return "mx::atan2_f32(" + L + ", " + R + ")";
  if (dt == F32) return fn + "f(" + L + ", " + R + ")";
  if (dt == F64) return fn + "(" + L + ", " + R + ")";
  return std::string("(") + CT(dt) + ")" + fn + "((double)" + l + ", (double)" + r + ")";
}

// ---- byte-lane form ---------------------------------------------------------------
//
// Exact value range of a node of a focal group over one U8 raster `s`, for staged
// values in [0, R]. Uniform nodes (D0 inputs, constants) are taken in [0, 127]; the
// kernel checks its D0 arguments against that before it uses the lanes.
struct LaneRange {
  int lo = 0, hi = 0;
  bool uni = false;
};

bool lane_dt_ok(int dt, int lo, int hi) {
  if (is_float_dt(dt)) return false;
  if (dt == S8) return lo >= -128 && hi <= 127;
  return lo >= 0;  // every other integer type holds [0, 255]
}

// Ranges of the nodes of `all` (kernel order), or false when some node has no exact
// byte-lane statement: 0/1-weighted convolutions and neighbours of `s`, casts, sums,
// differences that stay non-negative, products with a 0/1 operand, comparisons of
// operands in [0, 127], logical and bitwise and/or/xor, conditionals on a 0/1
// condition, with every value in [0, 255] (so the C conversions are identities).
bool lane_ranges(const std::vector<Node*>& all, const std::set<Node*>& in_set, const Node* s, int R,
                 std::map<const Node*, LaneRange>& rg) {
  rg.clear();
  for (Node* n : all) {
    LaneRange r;
    if (in_set.count(n)) {
      if (is_float_dt(n->meta.dt)) return false;
      if (n->meta.nd == D2) {
        if (n != s) return false;
        r.hi = R;
      } else {
        r.hi = 127;
        r.uni = true;
      }
      rg[n] = r;
      continue;
    }
    if (is_float_dt(n->meta.dt)) return false;
    for (Node* p : n->prev)
      if (!rg.count(p)) return false;
    auto arg = [&](int i) { return rg.at(n->prev[i]); };
    switch (n->kind) {
      case Kind::Constant: {
        if (is_float_dt(n->cnst.dtype)) return false;
        const long long v = variant_as<long long>(n->cnst);
        if (v < 0 || v > 127) return false;
        r.lo = r.hi = (int)v;
        r.uni = true;
        break;
      }
      case Kind::Neighbor:
        if (n->prev[0] != s) return false;
        r.hi = R;
        break;
      case Kind::Convolution: {
        if (n->prev[0] != s || is_float_dt(n->mask.dtype())) return false;
        int w = 0;
        for (int i1 = 0; i1 < n->mask.ds.arr[1]; i1++)
          for (int i0 = 0; i0 < n->mask.ds.arr[0]; i0++) {
            const long long m = variant_as<long long>(n->mask.at(i0, i1));
            if (m < 0 || m > 1) return false;
            w += (int)m;
          }
        r.hi = R * w;
        break;
      }
      case Kind::Cast:
        r = arg(0);
        break;
      case Kind::Unary: {
        const LaneRange a = arg(0);
        if (n->op == POS) r = a;
        else if (n->op == NOT && !a.uni && a.lo >= 0 && a.hi <= 1) r.hi = 1;
        else return false;
        break;
      }
      case Kind::Binary: {
        const LaneRange a = arg(0), b = arg(1);
        if (a.uni && b.uni) return false;
        if (is_float_dt(n->prev[0]->meta.dt) || is_float_dt(n->prev[1]->meta.dt)) return false;
        const bool a01 = !a.uni && a.lo >= 0 && a.hi <= 1, b01 = !b.uni && b.lo >= 0 && b.hi <= 1;
        switch (n->op) {
          case ADD: r.lo = a.lo + b.lo; r.hi = a.hi + b.hi; break;
          case SUB:
            if (a.uni || b.uni) return false;
            r.lo = a.lo - b.hi; r.hi = a.hi - b.lo;
            break;
          case MUL:
            if (a.uni || b.uni || !(a01 || b01)) return false;
            r.lo = a.lo * b.lo; r.hi = a.hi * b.hi;
            break;
          case EQ: case NE: case LT: case GT: case LE: case GE:
            if (a.lo < 0 || b.lo < 0 || a.hi > 127 || b.hi > 127) return false;
            r.hi = 1;
            break;
          case AND: case OR:
            if (!a01 || !b01) return false;
            r.hi = 1;
            break;
          case bAND: r.hi = std::min(a.hi, b.hi); break;
          case bOR: case bXOR: {
            int p = 1;
            while (p <= std::max(a.hi, b.hi)) p <<= 1;
            r.hi = p - 1;
            break;
          }
          default: return false;
        }
        break;
      }
      case Kind::Conditional: {
        const LaneRange c = arg(0), a = arg(1), b = arg(2);
        if (c.uni || c.lo < 0 || c.hi > 1) return false;
        r.lo = std::min(a.lo, b.lo);
        r.hi = std::max(a.hi, b.hi);
        break;
      }
      default: return false;
    }
    if (r.lo < 0 || r.hi > 255 || !lane_dt_ok(n->meta.dt, r.lo, r.hi)) return false;
    rg[n] = r;
  }
  return true;
}

}  // namespace

Variant quantize_literal(const Variant& v) {
  if (v.dtype == F32) {
    std::string s = std::to_string((double)v.u.f32);
    Variant r = v; r.u.f32 = (float)std::strtod(s.c_str(), nullptr); return r;
  }
  if (v.dtype == F64) {
    std::string s = std::to_string(v.u.f64);
    Variant r = v; r.u.f64 = std::strtod(s.c_str(), nullptr); return r;
  }
  return v;
}

TileShape default_tile(const Group& g) {
  // cells per thread from the narrowest raster read or written (intermediates live in registers)
  int minsz = 8;
  for (Node* n : g.ins)
    if (n->meta.nd == D2) minsz = std::min(minsz, sizeof_dt(n->meta.dt));
  for (Node* n : g.nodes)
    if (n->meta.nd == D2 && (std::find(g.outs.begin(), g.outs.end(), n) != g.outs.end() || n->next.empty()))
      minsz = std::min(minsz, sizeof_dt(n->meta.dt));
  if (minsz == 8) minsz = 4;
  TileShape t;
  // 64 threads per tile row, each moving one vector of the narrowest type:
  // 8 cells of 1-byte types, 4 cells of 2- and 4-byte types, 2 of 8-byte types
  if (minsz >= 8) { t.VX = 2; t.TX = 128; t.TY = 16; }
  else if (minsz == 1) { t.VX = 8; t.TX = 512; t.TY = 16; }
  else { t.VX = 4; t.TX = 256; t.TY = 16; }
  return t;
}

// `acc = OP(acc, v)` in the node's C type: fmax/fmin for floating MAX/MIN (the
// reference's max/min builtins, ReductionType.cpp:66-78), a comparison for
// integers, the operator otherwise.
static std::string reduce_stmt(int op, int dt, const std::string& acc, const std::string& v) {
  const std::string T = CT(dt);
  std::ostringstream o;
  if (op == MAX || op == MIN) {
    if (dt == F32) o << acc << " = " << (op == MAX ? "fmaxf(" : "fminf(") << acc << ", " << v << ");";
    else if (dt == F64) o << acc << " = " << (op == MAX ? "fmax(" : "fmin(") << acc << ", " << v << ");";
    else o << "{ const " << T << " t_ = " << v << "; " << acc << " = (t_ " << (op == MAX ? ">" : "<") << " " << acc << ") ? t_ : " << acc << "; }";
  } else {
    const char* c = op == SUM ? "+" : op == PROD ? "*" : op == rAND ? "&" : "|";
    o << acc << " = (" << T << ")(" << acc << " " << c << " " << v << ");";
  }
  return o.str();
}

// inputs first (D2 then D0, each by id), then the nodes by id: the local
// numbering is then the same for groups of equal structure
std::vector<Node*> kernel_order(const Group& g) {
  std::vector<Node*> all(g.ins.begin(), g.ins.end()), body(g.nodes.begin(), g.nodes.end());
  std::sort(all.begin(), all.end(), [](Node* a, Node* b) {
    if ((a->meta.nd == D0) != (b->meta.nd == D0)) return a->meta.nd != D0;
    return a->id < b->id;
  });
  std::sort(body.begin(), body.end(), [](Node* a, Node* b) { return a->id < b->id; });
  all.insert(all.end(), body.begin(), body.end());
  return all;
}

std::string kernel_shape_key(const Group& g, const std::vector<Node*>& d2_outs, TileShape tile, bool libm) {
  const std::vector<Node*> all = kernel_order(g);
  std::unordered_map<const Node*, int> local;
  for (size_t i = 0; i < all.size(); i++) local[all[i]] = (int)i;
  std::set<const Node*> in_set(g.ins.begin(), g.ins.end());
  std::ostringstream k;
  k << tile.TX << "," << tile.TY << "," << tile.VX << "," << tile.PF << "," << tile.ST << "," << tile.W << "," << libm << "|";
  for (Node* n : all) {
    if (in_set.count(n)) { k << "I" << n->meta.nd << "." << n->meta.dt << ";"; continue; }
    k << n->signature(false) << "(";
    for (Node* p : n->prev) {
      auto it = local.find(p);
      k << (it == local.end() ? -1 : it->second) << ",";
    }
    k << ");";
  }
  k << "|";
  for (Node* o : d2_outs) {
    auto it = local.find(o);
    k << (it == local.end() ? -1 : it->second) << ",";
  }
  return k.str();
}

KernelSpec generate_kernel(const Group& g, const std::vector<Node*>& d2_outs, TileShape tile, bool libm) {
  const bool fast = !libm;
  KernelSpec ks;
  ks.tile = tile;
  ks.d2_outs = d2_outs;
  if (tile.TX % tile.VX) throw std::runtime_error("codegen: TX must be a multiple of VX");
  const int TWX = tile.TX / tile.VX;
  if (256 % TWX || tile.TY % (256 / TWX)) throw std::runtime_error("codegen: tile does not fit 256 threads");

  std::vector<Node*> all = kernel_order(g), body(g.nodes.begin(), g.nodes.end());
  std::sort(body.begin(), body.end(), [](Node* a, Node* b) { return a->id < b->id; });
  std::set<Node*> in_set(g.ins.begin(), g.ins.end());
  std::set<Node*> member(all.begin(), all.end());
  std::map<const Node*, int> local;
  for (size_t i = 0; i < all.size(); i++) local[all[i]] = (int)i;
  for (Node* o : d2_outs)
    if (!local.count(o)) throw std::runtime_error("codegen: output outside its group");
  g_local = &local;

  // ---- classification
  std::vector<Node*> focal, staged, zonal;
  for (Node* n : g.nodes) {
    if (n->pattern() == FOCAL) focal.push_back(n);
    if (n->kind == Kind::ZonalReduc || n->kind == Kind::ZonalStats) zonal.push_back(n);
  }
  for (Node* f : focal) {
    // the windowed sources of a focal node: its input (and the direction raster of a SpreadNeighbor)
    const size_t ns = f->kind == Kind::SpreadNeighbor ? 2 : 1;
    for (size_t k = 0; k < ns; k++)
      if (!contains(staged, f->prev[k])) staged.push_back(f->prev[k]);
  }
  std::sort(staged.begin(), staged.end(), [](Node* a, Node* b) { return a->id < b->id; });
  for (Node* s : staged)
    if (!member.count(s)) throw std::runtime_error("codegen: focal source outside its group");
  // halo nodes: staged nodes and their ancestors inside the group
  std::set<Node*> halo;
  std::function<void(Node*)> up = [&](Node* n) {
    if (!member.count(n) || !halo.insert(n).second) return;
    if (in_set.count(n)) return;
    for (Node* p : n->prev) up(p);
  };
  for (Node* s : staged) up(s);
  for (Node* n : halo)  // inputs are materialised rasters, whatever produced them
    if (n->pattern() == FOCAL && !in_set.count(n))
      throw std::runtime_error("codegen: focal node feeding a focal node in one group");

  ks.form = focal.empty() ? Form::Local : Form::Focal;
  for (Node* f : focal) { ks.H0 = std::max(ks.H0, f->halo(0)); ks.H1 = std::max(ks.H1, f->halo(1)); }

  // cell nodes: everything computed per output cell
  std::vector<Node*> cell;
  for (Node* n : g.nodes) if (!halo.count(n)) cell.push_back(n);
  // halo values needed at the centre cell (by non-focal cell nodes or as outputs)
  std::set<Node*> centre;
  for (Node* n : cell) {
    if (n->pattern() == FOCAL) continue;
    for (Node* p : n->prev) if (halo.count(p)) centre.insert(p);
  }
  for (Node* o : d2_outs) if (halo.count(o)) centre.insert(o);
  {
    std::vector<Node*> stack(centre.begin(), centre.end());
    while (!stack.empty()) {
      Node* n = stack.back(); stack.pop_back();
      if (in_set.count(n) || contains(staged, n)) continue;
      for (Node* p : n->prev)
        if (halo.count(p) && centre.insert(p).second) stack.push_back(p);
    }
  }
  // inputs read by cell nodes directly
  std::set<Node*> cell_inputs;
  for (Node* n : cell)
    for (Node* p : n->prev)
      if (in_set.count(p) && !halo.count(p)) cell_inputs.insert(p);
  for (Node* o : d2_outs) if (in_set.count(o) && !halo.count(o)) cell_inputs.insert(o);

  // ---- parameters
  for (Node* n : all) {
    if (!in_set.count(n)) continue;
    Param p;
    p.kind = n->meta.nd == D0 ? Param::InD0 : Param::InD2;
    p.node = n; p.dt = n->meta.dt;
    ks.params.push_back(p);
  }
  for (Node* n : d2_outs) { Param p; p.kind = Param::OutD2; p.node = n; p.dt = n->meta.dt; ks.params.push_back(p); }
  for (Node* n : body)
    if (n->kind == Kind::Convolution && is_float_dt(n->mask.dtype())) {
      Param p; p.kind = Param::Mask; p.node = n; p.dt = n->mask.dtype(); ks.params.push_back(p);
    }
  // layout of P::Args: members in parameter order at their natural alignment
  // (In<T> = {const T* p})
  {
    size_t off = 0, al_max = 1;
    for (Param& p : ks.params) {
      size_t sz = 8, al = 8;
      if (p.kind == Param::InD0) sz = al = (size_t)sizeof_dt(p.dt);
      if (p.kind == Param::Mask) { al = (size_t)sizeof_dt(p.dt); sz = al * p.node->mask.v.size(); }
      off = (off + al - 1) / al * al;
      p.off = off;
      p.size = sz;
      off += sz;
      al_max = std::max(al_max, al);
    }
    if (ks.params.empty()) { off = 4; al_max = 4; }  // Args { int unused; }
    ks.args_size = (off + al_max - 1) / al_max * al_max;
    const size_t lal = std::max<size_t>(al_max, 4);
    ks.task_launch_off = (ks.args_size + 3) / 4 * 4;
    ks.task_size = (ks.task_launch_off + 13 * 4 + lal - 1) / lal * lal;
  }
  size_t part_bytes = 0;
  std::vector<size_t> part_off;
  for (Node* z : zonal) {
    ZonalOut zo{};
    zo.node = z;
    zo.dt = z->prev[0]->meta.dt;
    if (z->kind == Kind::ZonalReduc) {
      zo.red = z->op;
      zo.acc_dt = acc_type(z->op, zo.dt);
      part_off.push_back(part_bytes);
      zo.result_off = part_bytes;
      part_bytes += 8;
    } else {
      if (!(zo.dt == F32 || zo.dt == S32 || zo.dt == U32 || zo.dt == U8 || zo.dt == U16 || zo.dt == S8 || zo.dt == S16))
        throw std::runtime_error("codegen: per-zone statistics support 8/16/32-bit data");
      zo.red = 0;
      zo.acc_dt = F64;
      zo.nzones = z->nzones;
      part_off.push_back(part_bytes);
      zo.result_off = part_bytes;
      part_bytes += (size_t)z->nzones * 24;
    }
    ks.zouts.push_back(zo);
  }
  ks.part_bytes = part_bytes;

  // the kernel is named after the hash of its own source (with the name
  // replaced by a placeholder), which is also the compile-cache key
  ks.name = "KRN_NAME";
  ks.fin_name = zonal.empty() ? "" : "KRN_NAME_fin";

  // ---- source
  std::ostringstream s;
  const int VX = tile.VX;
  const int PAD = ks.H0 == 0 ? 0 : VX * ((ks.H0 + VX - 1) / VX);  // halo rounded up to whole vectors
  const int SW = tile.TX + 2 * PAD, HH = tile.TY + 2 * ks.H1, WW = VX + 2 * PAD, WR = 2 * ks.H1 + 1;
  // streaming form: every staged node is an input raster, read per row as one
  // dword-aligned span of VX + 2*A cells (A: the halo rounded up to 4 bytes)
  std::map<const Node*, int> span_pad;
  {
    bool ok = ks.form == Form::Focal && tile.ST && !staged.empty();
    for (Node* st : staged) {
      const int sz = sizeof_dt(st->meta.dt);
      if (!in_set.count(st) || st->meta.nd != D2 || (VX * sz) % 4) { ok = false; break; }
      const int unit = sz >= 4 ? 1 : 4 / sz;
      span_pad[st] = unit * ((ks.H0 + unit - 1) / unit);
    }
    ks.stream = ok;
  }
  // byte-lane form: a staged focal group over one U8 input raster with one 1-byte output,
  // every node exact in byte lanes for staged values in [0, R]. R = 255 needs no check of
  // the data; otherwise R = 1 (0/1 rasters, where products of 0/1 values are ANDs). A
  // group exact for some R is exact for every smaller one, so there is no other choice.
  std::map<const Node*, LaneRange> lane_rg;
  Node* lane_s = staged.size() == 1 ? staged[0] : nullptr;
  Node* lane_out = d2_outs.size() == 1 ? d2_outs[0] : nullptr;
  if (ks.form == Form::Focal && !ks.stream && zonal.empty() && lane_s && lane_out && in_set.count(staged[0]) &&
      lane_s->meta.nd == D2 && (lane_s->meta.dt == U8 || lane_s->meta.dt == B8) && VX % 4 == 0 &&
      !in_set.count(lane_out) && sizeof_dt(lane_out->meta.dt) == 1 &&
      std::none_of(cell_inputs.begin(), cell_inputs.end(), [](Node* n) { return n->meta.nd == D2; })) {
    for (int R : {255, 1})
      if (lane_ranges(all, in_set, lane_s, R, lane_rg) && !lane_rg.at(lane_out).uni) {  // the output is per cell
        ks.lane_r = (unsigned)R;
        break;
      }
    if (!ks.lane_r) lane_rg.clear();
  }
  s << "#include \"mapx_dev.hiph\"\n";
  // the module for rasters with an edge block narrower than the halo is the
  // same source compiled with MX_NARROW 1 (narrow_source)
  s << "#ifndef MX_NARROW\n#define MX_NARROW 0\n#endif\n";
  s << "namespace {\nstruct P {\n";
  s << "  static constexpr int H0 = " << ks.H0 << ", H1 = " << ks.H1 << ", PAD = " << PAD << ", SW = " << SW << ";\n";
  s << "  struct Args {\n";
  for (const Param& p : ks.params) {
    if (p.kind == Param::InD2) s << "    mx::In<" << CT(p.dt) << "> i" << LN(p.node) << ";\n";
    if (p.kind == Param::InD0) s << "    " << CT(p.dt) << " s" << LN(p.node) << ";\n";
    if (p.kind == Param::OutD2) s << "    " << CT(p.dt) << "* __restrict__ o" << LN(p.node) << ";\n";
    if (p.kind == Param::Mask) s << "    " << CT(p.dt) << " m" << LN(p.node) << "[" << p.node->mask.v.size() << "];\n";
  }
  if (ks.params.empty()) s << "    int unused;\n";
  s << "  };\n";
  s << "  struct Smem {\n";
  for (Node* st : staged) s << "    alignas(16) " << CT(st->meta.dt) << " sh" << LN(st) << "[" << SW * HH << "];\n";
  if (ks.lane_r) s << "    int lane_bad;  // the last tile with a staged value above LANE_R\n";
  if (staged.empty()) s << "    int unused;\n";
  s << "  };\n";
  // per-thread register window of every staged node
  s << "  struct Win {\n";
  for (Node* st : staged) s << "    " << CT(st->meta.dt) << " w" << LN(st) << "[" << WR << "][" << WW << "];\n";
  if (staged.empty()) s << "    int unused;\n";
  s << "  };\n";
  if (ks.stream) {
    // one row of every staged input around the thread's VX cells
    s << "  static constexpr int R = " << WR << ";\n";
    s << "  template <int VX> struct Span {\n";
    for (Node* st : staged) s << "    " << CT(st->meta.dt) << " s" << LN(st) << "[VX + " << 2 * span_pad[st] << "];\n";
    s << "  };\n";
    s << "  template <int VX, bool NW> __device__ static __forceinline__ void span_row(const Args& a, const mx::Blk& k, int x, int y, const mx::RowSrc& rs, Span<VX>& sp) {\n";
    for (Node* st : staged)
      s << "    mx::ld_span<" << CT(st->meta.dt) << ", VX, " << span_pad[st] << ", H0, NW>(a.i" << LN(st) << ", k, x, y, rs, sp.s" << LN(st) << ");\n";
    s << "  }\n";
    s << "  template <int VX> __device__ static __forceinline__ void span_zero(Span<VX>& sp) {\n";
    for (Node* st : staged)
      s << "#pragma unroll\n    for (int c = 0; c < VX + " << 2 * span_pad[st] << "; c++) sp.s" << LN(st) << "[c] = 0;\n";
    s << "  }\n";
    s << "  template <int VX> __device__ static __forceinline__ void span_win(const Span<VX>* rows, Win& W) {\n";
    for (Node* st : staged)
      s << "#pragma unroll\n    for (int r = 0; r < R; r++)\n#pragma unroll\n      for (int c = PAD - H0; c < PAD + VX + H0; c++) W.w"
        << LN(st) << "[r][c] = rows[r].s" << LN(st) << "[c - PAD + " << span_pad[st] << "];\n";
    s << "  }\n";
  }
  // the thread's VX cells of every input read per cell
  s << "  struct Cells {\n";
  for (Node* n : all)
    if (cell_inputs.count(n) && n->meta.nd == D2) s << "    " << CT(n->meta.dt) << " c" << LN(n) << "[" << VX << "];\n";
  s << "    int unused;\n  };\n";
  // one cell's outputs
  s << "  struct Out {\n";
  for (Node* o : d2_outs) s << "    " << CT(o->meta.dt) << " o" << LN(o) << ";\n";
  for (size_t j = 0; j < ks.zouts.size(); j++)
    if (ks.zouts[j].node->kind == Kind::ZonalStats)
      s << "    " << CT(ks.zouts[j].dt) << " zv" << j << ";\n    int zz" << j << ";\n";
  s << "    int unused;\n  };\n";
  // accumulators
  s << "  struct Acc {\n";
  for (size_t j = 0; j < ks.zouts.size(); j++) {
    const ZonalOut& z = ks.zouts[j];
    if (z.node->kind == Kind::ZonalReduc) s << "    " << CT(z.acc_dt) << " r" << j << ";\n";
    else s << "    mx::ZoneLds<" << CT(z.dt) << "> z" << j << ";\n    mx::ZoneRun<" << CT(z.dt) << "> q" << j << ";\n";
  }
  s << "    int unused;\n  };\n";

  // sin and cos of one value in one group share one sincos call (one range
  // reduction): paired when both are emitted in the same places (halo loop
  // and/or output cell), the first by id computes both
  std::map<const Node*, const Node*> sc_pair;
  for (Node* n : g.nodes) {
    if (n->kind != Kind::Unary || n->op != SIN || !is_float_dt(n->meta.dt)) continue;
    for (Node* m : g.nodes)
      if (m->kind == Kind::Unary && m->op == COS && m->prev[0] == n->prev[0] && m->meta.dt == n->meta.dt &&
          !sc_pair.count(m) && halo.count(n) == halo.count(m) && centre.count(n) == centre.count(m)) {
        sc_pair[n] = m;
        sc_pair[m] = n;
        break;
      }
  }
  auto emit_node = [&](std::ostream& o, Node* n, bool in_pre) {
    const std::string T = CT(n->meta.dt);
    switch (n->kind) {
      case Kind::Constant:
        o << "    const " << T << " " << V(n) << " = " << lit(quantize_literal(n->cnst)) << ";\n"; break;
      case Kind::Index:
        o << "    const " << T << " " << V(n) << " = (" << T << ")("
          << (n->op == D1 ? "k.bx * k.BS0 + x" : "k.by * k.BS1 + y") << ");\n"; break;
      case Kind::Cast:
        o << "    const " << T << " " << V(n) << " = (" << T << ")" << V(n->prev[0]) << ";\n"; break;
      case Kind::Rand:  // Skeleton.cpp:183-206: counter {seed, 0}; key: the cell's linear index
        o << "    const " << T << " " << V(n) << " = mx::rand_cell<" << T << ">((mx::u32)" << V(n->prev[0])
          << ", (mx::u32)((mx::i64)(k.by * k.BS1 + y) * k.DS0 + (k.bx * k.BS0 + x)));\n"; break;
      case Kind::Unary: {
        auto pr = sc_pair.find(n);
        if (pr != sc_pair.end()) {
          const Node* sn = n->op == SIN ? n : pr->second;
          const Node* cn = n->op == COS ? n : pr->second;
          if (n->id < pr->second->id) {
            o << "    " << T << " " << V(sn) << "_s, " << V(cn) << "_c;\n";
            o << "    " << (T == std::string("float") ? (fast ? "mx::sincos_f32(" : "sincosf(") : "sincos(") << V(n->prev[0]) << ", &" << V(sn) << "_s, &"
              << V(cn) << "_c);\n";
          }
          o << "    const " << T << " " << V(n) << " = " << V(n) << (n->op == SIN ? "_s" : "_c") << ";\n";
          break;
        }
        o << "    const " << T << " " << V(n) << " = " << unary_stmt(n, V(n->prev[0]), n->prev[0]->meta.dt, fast) << ";\n"; break;
      }
      case Kind::Binary:
        o << "    const " << T << " " << V(n) << " = " << binary_stmt(n, V(n->prev[0]), V(n->prev[1]), fast) << ";\n"; break;
      case Kind::Conditional:
        o << "    const " << T << " " << V(n) << " = (" << V(n->prev[0]) << ") ? " << V(n->prev[1]) << " : "
          << V(n->prev[2]) << ";\n"; break;
      case Kind::Neighbor:
        o << "    const " << T << " " << V(n) << " = W.w" << LN(n->prev[0]) << "[H1 + (" << n->coord.arr[1]
          << ")][PAD + v + (" << n->coord.arr[0] << ")];\n"; break;
      case Kind::Convolution: {
        // acc = 0; for y (outer), x (inner): acc += S[y][x] * W[y][x]  (FocalSkeleton.cpp:300-336)
        o << "    " << T << " " << V(n) << " = 0;\n";
        const int h0 = n->halo(0), h1 = n->halo(1);
        const bool marg = is_float_dt(n->mask.dtype());  // weights from P::Args (packed %f-quantised)
        for (int i1 = -h1; i1 <= h1; i1++)
          for (int i0 = -h0; i0 <= h0; i0++) {
            const Variant wq = quantize_literal(n->mask.at(i0 + h0, i1 + h1));
            std::string w = marg ? "a.m" + LN(n) + "[" + std::to_string((i1 + h1) * n->mask.ds.arr[0] + (i0 + h0)) + "]"
                                 : lit(wq);
            const std::string tap = "W.w" + LN(n->prev[0]) + "[H1 + (" + std::to_string(i1) + ")][PAD + v + (" +
                                    std::to_string(i0) + ")]";
            if (!marg && is_float_dt(n->meta.dt) && variant_as<double>(wq) == 0.0) {
              // a zero weight: x * 0 is exact (+-0, or NaN for inf / NaN taps), so one
              // fused multiply-add gives the bits of the multiply and the add
              o << "    " << V(n) << " = " << (n->meta.dt == F32 ? "__builtin_fmaf((float)" : "__builtin_fma((double)") << tap
                << ", " << (n->meta.dt == F32 ? "0.0f" : "0.0") << ", " << V(n) << ");\n";
            } else {
              o << "    " << V(n) << " += " << tap << " * " << w << ";\n";
            }
          }
        break;
      }
      case Kind::FocalFunc: {
        Variant neu = reduction_neutral(n->op, n->meta.dt);
        o << "    " << T << " " << V(n) << " = " << lit(neu) << ";\n";
        const int h0 = n->halo(0), h1 = n->halo(1);
        for (int i1 = -h1; i1 <= h1; i1++)
          for (int i0 = -h0; i0 <= h0; i0++) {
            std::string sv = "W.w" + LN(n->prev[0]) + "[H1 + (" + std::to_string(i1) + ")][PAD + v + (" +
                             std::to_string(i0) + ")]";
            o << "    " << reduce_stmt(n->op, n->meta.dt, V(n), sv) << "\n";
          }
        break;
      }
      case Kind::SpreadNeighbor: {
        // the neighbours (y outer, x inner, centre skipped) whose D8 code points at
        // the cell: code[i1 + 1][i0 + 1] = {{2,4,8},{1,0,16},{128,64,32}}, the
        // reference's dircod (skeleton/util.cpp:628) and flowAccuGather (front/Raster.cpp:586-601)
        static const int code[3][3] = {{2, 4, 8}, {1, 0, 16}, {128, 64, 32}};
        o << "    " << T << " " << V(n) << " = " << lit(reduction_neutral(n->op, n->meta.dt)) << ";\n";
        for (int i1 = -1; i1 <= 1; i1++)
          for (int i0 = -1; i0 <= 1; i0++) {
            if (i0 == 0 && i1 == 0) continue;
            const std::string at = "[H1 + (" + std::to_string(i1) + ")][PAD + v + (" + std::to_string(i0) + ")]";
            o << "    if (W.w" << LN(n->prev[1]) << at << " == " << code[i1 + 1][i0 + 1] << ") { "
              << reduce_stmt(n->op, n->meta.dt, V(n), "W.w" + LN(n->prev[0]) + at) << " }\n";
          }
        break;
      }
      case Kind::Stats: case Kind::Barrier:  // the input's values (ZonalSkeleton.cpp:290-292, Skeleton.cpp:328-330)
        o << "    const " << T << " " << V(n) << " = " << V(n->prev[0]) << ";\n"; break;
      case Kind::LhsAccess:  // Skeleton.cpp:290-297: the one cell takes the right-hand value
        o << "    const " << T << " " << V(n) << " = (k.bx * k.BS0 + x == " << n->coord.arr[0] << " && k.by * k.BS1 + y == "
          << n->coord.arr[1] << ") ? (" << T << ")" << V(n->prev[1]) << " : " << V(n->prev[0]) << ";\n"; break;
      case Kind::FocalPercent: {
        // FocalSkeleton.cpp:375-403: count over the node's window (y outer, x inner) of
        // centre == cell (AGE) or centre > cell (ILE), in the node's type, divided by the
        // reference's nbh_size (skeleton/util.cpp:321-327), "(H0*2+1)*H1*2+1" over the
        // kernel's halo constants H0, H1 (the group's halo)
        const std::string c = "W.w" + LN(n->prev[0]) + "[H1][PAD + v]";
        o << "    " << T << " " << V(n) << " = 0;\n";
        const int h0 = n->halo(0), h1 = n->halo(1);
        for (int i1 = -h1; i1 <= h1; i1++)
          for (int i0 = -h0; i0 <= h0; i0++) {
            std::string sv = "W.w" + LN(n->prev[0]) + "[H1 + (" + std::to_string(i1) + ")][PAD + v + (" +
                             std::to_string(i0) + ")]";
            o << "    " << V(n) << " = (" << T << ")(" << V(n) << " + (" << c << (n->op == 1 ? " == " : " > ") << sv << "));\n";
          }
        o << "    " << V(n) << " = (" << T << ")(" << V(n) << " / ((H0 * 2 + 1) * H1 * 2 + 1));\n";
        break;
      }
      case Kind::Diversity: {
        // Skeleton.cpp:259-288 with the helpers of skeleton/util.cpp:439-481
        const std::string Ti = CT(n->prev[0]->meta.dt), L = LN(n), N = std::to_string(n->prev.size());
        // the lists start zeroed: the slots past dn are never read, and left undefined they let
        // the optimiser fold the unrolled loops around indeterminate values
        o << "    " << Ti << " de" << L << "[" << N << "] = {}; int dc" << L << "[" << N << "] = {}; int dn" << L << " = 0;\n";
        for (Node* p : n->prev)
          o << "    mx::div_add<" << Ti << ", " << N << ">(" << V(p) << ", de" << L << ", dc" << L << ", dn" << L << ");\n";
        const char* f = n->op == 1 ? nullptr : n->op == 2 ? "mx::div_major" : n->op == 3 ? "mx::div_minor" : "mx::div_mean";
        if (!f) o << "    const " << T << " " << V(n) << " = (" << T << ")dn" << L << ";\n";
        else o << "    const " << T << " " << V(n) << " = " << f << "<" << Ti << ", " << N << ">(de" << L << ", dc" << L << ", dn" << L << ");\n";
        break;
      }
      case Kind::ZonalReduc: case Kind::ZonalStats: break;  // accumulated below
      default: throw std::runtime_error(std::string("codegen: unexpected node kind ") + kind_name(n->kind));
    }
    (void)in_pre;
  };

  // pre_vec() / put_row(): values of the staged nodes at the VX halo cells
  // x0v..x0v+VX-1 of row y, stored to LDS as vectors; pre_vec reads the
  // inputs with ldv's border rule, put_row takes them from a Raw record that
  // fetch_row loaded earlier through a known row source (so the loads of the
  // next tile can be in flight while the current one is computed)
  std::vector<Node*> raw_in;
  for (Node* n : all)
    if (halo.count(n) && in_set.count(n) && n->meta.nd == D2) raw_in.push_back(n);
  s << "  template <int VX> struct Raw {\n";
  for (Node* n : raw_in) s << "    mx::Packed<" << CT(n->meta.dt) << ", VX> b" << LN(n) << ";\n";
  s << "    int unused;\n  };\n";
  s << "  template <int VX> __device__ static __forceinline__ void fetch_row(const Args& a, const mx::RowSrc& rs, int x0v, Raw<VX>& r) {\n";
  for (Node* n : raw_in)
    s << "    mx::ldpack<" << CT(n->meta.dt) << ", VX>(a.i" << LN(n) << ", rs, x0v, r.b" << LN(n) << ");\n";
  s << "    (void)a; (void)rs; (void)x0v; (void)r;\n  }\n";
  // both return, for a byte-lane group, the OR of the staged values (else 0)
  auto pre_body = [&](const char* fname, bool put) {
    s << "  template <int VX" << (put ? "" : ", bool NW") << "> __device__ static __forceinline__ mx::u32 " << fname
      << "(const Args& a, const mx::Blk& k, " << (put ? "const Raw<VX>& r, " : "") << "int x0v, int y, Smem& sm, int si) {\n";
    for (Node* n : raw_in) {
      s << "    " << CT(n->meta.dt) << " b" << LN(n) << "[VX];\n";
      if (put) s << "    r.b" << LN(n) << ".unpack(b" << LN(n) << ");\n";
      else s << "    mx::ldv<" << CT(n->meta.dt) << ", VX, H0, NW>(a.i" << LN(n) << ", k, x0v, y, b" << LN(n) << ");\n";
    }
    for (Node* st : staged) s << "    " << CT(st->meta.dt) << " t" << LN(st) << "[VX];\n";
    s << "#pragma unroll\n    for (int v = 0; v < VX; v++) {\n      const int x = x0v + v;\n";
    for (Node* n : all) {
      if (!halo.count(n)) continue;
      if (in_set.count(n)) {
        if (n->meta.nd == D0) s << "    const " << CT(n->meta.dt) << " " << V(n) << " = a.s" << LN(n) << ";\n";
        else s << "    const " << CT(n->meta.dt) << " " << V(n) << " = b" << LN(n) << "[v];\n";
      } else emit_node(s, n, true);
    }
    for (Node* st : staged) s << "      t" << LN(st) << "[v] = " << V(st) << ";\n";
    s << "      (void)x;\n    }\n";
    for (Node* st : staged) s << "    mx::stvec<" << CT(st->meta.dt) << ", VX>(&sm.sh" << LN(st) << "[si], t" << LN(st) << ");\n";
    s << "    (void)a; (void)k; (void)y; (void)sm; (void)si;\n";
    if (!ks.lane_r) s << "    return 0u;\n";
    else if (put) s << "    return mx::word_or(r.b" << LN(lane_s) << ".w);\n";
    else s << "    mx::u32 m = 0u;\n#pragma unroll\n    for (int v = 0; v < VX; v++) m |= (mx::u32)t" << LN(lane_s)
           << "[v];\n    return m;\n";
    s << "  }\n";
  };
  pre_body("pre_vec", false);
  pre_body("put_row", true);
  s << "  template <int VX> __device__ static __forceinline__ void pre_zero(Smem& sm, int si) {\n";
  for (Node* st : staged)
    s << "    { " << CT(st->meta.dt) << " z[VX] = {}; mx::stvec<" << CT(st->meta.dt) << ", VX>(&sm.sh" << LN(st) << "[si], z); }\n";
  s << "    (void)sm; (void)si;\n  }\n";
  // load_cells(): the thread's VX cells of every input read per cell
  s << "  template <int VX> __device__ static __forceinline__ void load_cells(const Args& a, mx::i64 off, Cells& C) {\n";
  for (Node* n : all)
    if (cell_inputs.count(n) && n->meta.nd == D2)
      s << "    mx::ldcell<" << CT(n->meta.dt) << ", VX>(a.i" << LN(n) << ", off, C.c" << LN(n) << ");\n";
  s << "    (void)a; (void)off; (void)C;\n  }\n";
  // store(): the thread's VX output cells as vectors (nv < VX: the first nv cells)
  s << "  template <int VX> __device__ static __forceinline__ void store(const Args& a, mx::i64 off, int nv, const Out (&o)[VX]) {\n";
  for (Node* on : d2_outs) {
    const std::string T = CT(on->meta.dt), L = LN(on);
    s << "    {\n      " << T << " t[VX];\n#pragma unroll\n      for (int v = 0; v < VX; v++) t[v] = o[v].o" << L << ";\n"
      << "      if (nv == VX) mx::gstv<" << T << ", VX>(a.o" << L << " + off, t);\n"
      << "      else for (int v = 0; v < nv; v++) mx::gst(a.o" << L << " + off + v, t[v]);\n    }\n";
  }
  s << "    (void)a; (void)off; (void)nv; (void)o;\n  }\n";
  // byte-lane members: post_lanes() computes the thread's VX cells four to a dword from
  // its window rows read from LDS as dwords; store_lanes() writes them
  s << "  static constexpr unsigned LANE_R = " << ks.lane_r << "u;\n";
  if (ks.lane_r) {
    const int WD = WW / 4;
    auto X = [&](const Node* p) { return std::string(lane_rg.at(p).uni ? "U" : "L") + LN(p); };
    auto tap = [&](int dx, int dy) {
      // lane_at<dx> reads dwords c + floor(dx / 4) and, off a dword boundary, the next one
      for (int j = 0; j < VX / 4; j++) {
        const int b = PAD + 4 * j + dx;
        if (b < 0 || b + 3 >= WW) throw std::runtime_error("codegen: byte-lane tap outside the window");
      }
      return "mx::lane_at<" + std::to_string(dx) + ">(w[" + std::to_string(ks.H1 + dy) + "], c)";
    };
    s << "  __device__ static __forceinline__ bool lane_args_ok(const Args& a) {\n    (void)a;\n    return true";
    for (Node* n : all)
      if (in_set.count(n) && n->meta.nd == D0)
        s << " && (long long)a.s" << LN(n) << " >= 0 && (long long)a.s" << LN(n) << " <= 127";
    s << ";\n  }\n";
    s << "  template <int VX> __device__ static __forceinline__ void post_lanes(const Args& a, const Smem& sm, int base, mx::u32 (&o)[VX / 4]) {\n";
    s << "    static_assert(VX == " << VX << ", \"byte-lane window\");\n";
    s << "    mx::u32 w[" << WR << "][" << WD << "];\n";
    s << "#pragma unroll\n    for (int r = 0; r < " << WR << "; r++)\n#pragma unroll\n      for (int c = 0; c < " << WW / VX
      << "; c++) {\n        const mx::Vec<mx::u32, VX / 4> t = *(const mx::Vec<mx::u32, VX / 4>*)&sm.sh" << LN(lane_s)
      << "[base + r * SW + c * VX];\n#pragma unroll\n        for (int e = 0; e < VX / 4; e++) w[r][c * (VX / 4) + e] = t.v[e];\n      }\n";
    // uniform values, replicated into every lane
    for (Node* n : all) {
      if (!lane_rg.at(n).uni) continue;
      if (in_set.count(n)) s << "    const mx::u32 U" << LN(n) << " = mx::lane_rep((int)a.s" << LN(n) << ");\n";
      else if (n->kind == Kind::Constant) s << "    const mx::u32 U" << LN(n) << " = mx::lane_rep(" << lane_rg.at(n).lo << ");\n";
      else s << "    const mx::u32 U" << LN(n) << " = " << X(n->prev[0]) << ";\n";  // a cast
    }
    s << "#pragma unroll\n    for (int j = 0; j < VX / 4; j++) {\n      const int c = " << PAD / 4 << " + j;\n";
    for (Node* n : all) {
      const LaneRange& r = lane_rg.at(n);
      if (r.uni) continue;
      const std::string L = "      const mx::u32 L" + LN(n) + " = ";
      if (n == lane_s) { s << L << tap(0, 0) << ";\n"; continue; }
      switch (n->kind) {
        case Kind::Neighbor: s << L << tap(n->coord.arr[0], n->coord.arr[1]) << ";\n"; break;
        case Kind::Convolution: {
          s << L << "0u";
          const int h0 = n->halo(0), h1 = n->halo(1);
          for (int i1 = -h1; i1 <= h1; i1++)
            for (int i0 = -h0; i0 <= h0; i0++)
              if (variant_as<long long>(n->mask.at(i0 + h0, i1 + h1)) != 0) s << " + " << tap(i0, i1);
          s << ";\n";
          break;
        }
        case Kind::Cast: s << L << X(n->prev[0]) << ";\n"; break;
        case Kind::Unary: s << L << X(n->prev[0]) << (n->op == NOT ? " ^ 0x01010101u" : "") << ";\n"; break;
        case Kind::Conditional:
          s << L << "mx::lane_pick(" << X(n->prev[0]) << ", " << X(n->prev[1]) << ", " << X(n->prev[2]) << ");\n";
          break;
        case Kind::Binary: {
          const Node *pa = n->prev[0], *pb = n->prev[1];
          const std::string A = X(pa), B = X(pb);
          const LaneRange &ra = lane_rg.at(pa), &rb = lane_rg.at(pb);
          const bool a01 = !ra.uni && ra.hi <= 1, b01 = !rb.uni && rb.hi <= 1;
          switch (n->op) {
            case ADD: s << L << A << " + " << B << ";\n"; break;
            case SUB: s << L << A << " - " << B << ";\n"; break;
            case MUL:
              if (a01 && b01) s << L << A << " & " << B << ";\n";
              else if (a01) s << L << "mx::lane_sel(" << A << ", " << B << ");\n";
              else s << L << "mx::lane_sel(" << B << ", " << A << ");\n";
              break;
            case EQ: s << L << "mx::lane_eq(" << A << ", " << B << ");\n"; break;
            case NE: s << L << "mx::lane_ne(" << A << ", " << B << ");\n"; break;
            case LT: s << L << "mx::lane_lt(" << A << ", " << B << ");\n"; break;
            case GE: s << L << "mx::lane_ge(" << A << ", " << B << ");\n"; break;
            case GT: s << L << "mx::lane_lt(" << B << ", " << A << ");\n"; break;
            case LE: s << L << "mx::lane_ge(" << B << ", " << A << ");\n"; break;
            case AND: case bAND: s << L << A << " & " << B << ";\n"; break;
            case OR: case bOR: s << L << A << " | " << B << ";\n"; break;
            case bXOR: s << L << A << " ^ " << B << ";\n"; break;
            default: throw std::runtime_error("codegen: byte-lane operator");
          }
          break;
        }
        default:
          if (in_set.count(n)) continue;  // the staged input is read through taps
          throw std::runtime_error("codegen: byte-lane node kind");
      }
    }
    s << "      o[j] = L" << LN(lane_out) << ";\n    }\n    (void)a;\n  }\n";
    s << "  template <int VX> __device__ static __forceinline__ void store_lanes(const Args& a, mx::i64 off, int nv, const mx::u32 (&o)[VX / 4]) {\n";
    s << "    unsigned char* p = (unsigned char*)(a.o" << LN(lane_out) << " + off);\n";
    s << "    if (nv == VX) mx::gstw<VX / 4>(p, o);\n";
    s << "    else\n      for (int v = 0; v < nv; v++) p[v] = (unsigned char)(o[v >> 2] >> (8 * (v & 3)));\n  }\n";
  }
  // load_win(): the thread's window rows from LDS with aligned vector reads
  s << "  template <int VX> __device__ static __forceinline__ void load_win(const Smem& sm, int base, Win& W) {\n";
  for (Node* st : staged) {
    const std::string T = CT(st->meta.dt), L = LN(st);
    s << "#pragma unroll\n    for (int r = 0; r < " << WR << "; r++)\n#pragma unroll\n      for (int c = 0; c < " << WW / VX
      << "; c++) {\n        const mx::Vec<" << T << ", VX> t = *(const mx::Vec<" << T << ", VX>*)&sm.sh" << L
      << "[base + r * SW + c * VX];\n#pragma unroll\n        for (int e = 0; e < VX; e++) W.w" << L
      << "[r][c * VX + e] = t.v[e];\n      }\n";
  }
  s << "    (void)sm; (void)base; (void)W;\n  }\n";

  // post(): one output cell
  s << "  __device__ static __forceinline__ void post(const Args& a, const mx::Blk& k, int x, int y, bool ok, const Win& W, const Cells& C, int v, Acc& acc, Out& out) {\n";
  for (Node* n : all) {
    if (halo.count(n)) {
      if (!centre.count(n)) continue;
      if (contains(staged, n)) s << "    const " << CT(n->meta.dt) << " " << V(n) << " = W.w" << LN(n) << "[H1][PAD + v];\n";
      else if (in_set.count(n)) {
        if (n->meta.nd == D0) s << "    const " << CT(n->meta.dt) << " " << V(n) << " = a.s" << LN(n) << ";\n";
        else s << "    const " << CT(n->meta.dt) << " " << V(n) << " = mx::ldl(a.i" << LN(n) << ", k, x, y);\n";
      } else emit_node(s, n, false);
      continue;
    }
    if (in_set.count(n)) {
      if (!cell_inputs.count(n)) continue;
      if (n->meta.nd == D0) s << "    const " << CT(n->meta.dt) << " " << V(n) << " = a.s" << LN(n) << ";\n";
      else s << "    const " << CT(n->meta.dt) << " " << V(n) << " = C.c" << LN(n) << "[v];\n";
      continue;
    }
    emit_node(s, n, false);
  }
  for (Node* o : d2_outs) s << "    out.o" << LN(o) << " = " << V(o) << ";\n";
  for (size_t j = 0; j < ks.zouts.size(); j++) {
    const ZonalOut& z = ks.zouts[j];
    Node* src = z.node->prev[0];
    if (z.node->kind == Kind::ZonalReduc)
      s << "    if (ok) acc.r" << j << " = " << red_op(z.red) << "::f(acc.r" << j << ", (" << CT(z.acc_dt) << ")" << V(src) << ");\n";
    else  // accumulated for the thread's VX cells of the row at once (zone_row)
      s << "    out.zv" << j << " = (" << CT(z.dt) << ")" << V(src) << ";\n    out.zz" << j << " = (int)" << V(z.node->prev[1]) << ";\n";
  }
  s << "    (void)a; (void)k; (void)x; (void)y; (void)ok; (void)W; (void)C; (void)v; (void)acc; (void)out;\n  }\n";
  // zone_row(): the per-zone statistics of the thread's VX cells of one row, in cell order
  s << "  template <int VX> __device__ static __forceinline__ void zone_row(Acc& acc, const Out (&o)[VX], unsigned okm) {\n";
  for (size_t j = 0; j < ks.zouts.size(); j++) {
    const ZonalOut& z = ks.zouts[j];
    if (z.node->kind != Kind::ZonalStats) continue;
    s << "    {\n      " << CT(z.dt) << " zv[VX];\n      int zz[VX];\n#pragma unroll\n      for (int v = 0; v < VX; v++) { zv[v] = o[v].zv"
      << j << "; zz[v] = o[v].zz" << j << "; }\n      mx::zone_steps<" << CT(z.dt) << ", " << z.nzones << ", VX>(acc.z" << j
      << ", acc.q" << j << ", okm, zv, zz);\n    }\n";
  }
  s << "    (void)acc; (void)o; (void)okm;\n  }\n";
  // flush(): writes one chunk's partial record and restarts the accumulators
  s << "  static constexpr bool ZONAL = " << (zonal.empty() ? "false" : "true") << ";\n";
  s << "  static constexpr unsigned PART_BYTES = " << part_bytes << "u;\n";
  s << "  __device__ static __forceinline__ void flush(Acc& acc, char* my) {\n";
  if (!zonal.empty()) s << "    __shared__ mx::u64 red_scratch[4];\n";
  for (size_t j = 0; j < ks.zouts.size(); j++) {
    const ZonalOut& z = ks.zouts[j];
    if (z.node->kind == Kind::ZonalReduc) {
      std::string A = CT(z.acc_dt);
      s << "    mx::block_reduce_store<" << red_op(z.red) << ", " << A << ">(acc.r" << j << ", (" << A << "*)(my + "
        << part_off[j] << "), (" << A << "*)red_scratch);\n";
      s << "    acc.r" << j << " = " << lit(reduction_neutral(z.red, z.acc_dt)) << ";\n";
    } else {
      s << "    mx::zone_close<" << CT(z.dt) << ", " << z.nzones << ">(acc.z" << j << ", acc.q" << j << ");\n";
      s << "    mx::zone_flush<" << CT(z.dt) << ", " << z.nzones << ">(acc.z" << j << ", (mx::ZonePart<" << CT(z.dt)
        << ">*)(my + " << part_off[j] << "));\n";
      s << "    __syncthreads();\n    mx::zone_init<" << CT(z.dt) << ", " << z.nzones << ">(acc.z" << j
        << ");\n    __syncthreads();\n";
    }
  }
  s << "    (void)acc; (void)my;\n  }\n";
  s << "};\n";
  // the host packs P::Args and TaskRec byte by byte: pin the layout
  s << "static_assert(sizeof(P::Args) == " << ks.args_size << ", \"Args layout\");\n";
  for (const Param& p : ks.params)
    s << "static_assert(__builtin_offsetof(P::Args, "
      << (p.kind == Param::InD2 ? "i" : p.kind == Param::InD0 ? "s" : p.kind == Param::Mask ? "m" : "o")
      << LN(p.node) << ") == " << p.off << ", \"Args layout\");\n";
  s << "static_assert(sizeof(mx::TaskRec<P::Args>) == " << ks.task_size << " && __builtin_offsetof(mx::TaskRec<P::Args>, L) == "
    << ks.task_launch_off << ", \"TaskRec layout\");\n";
  s << "}  // namespace\n\n";

  // main kernel
  s << "extern \"C\" __global__ void __launch_bounds__(256" << (tile.W > 0 ? ", " + std::to_string(tile.W) : std::string())
    << ") " << ks.name
    << "(P::Args a0, mx::Launch L0, const mx::TaskRec<P::Args>* tasks, const int* starts, int ntasks, int nchunks, "
       "mx::u32* queue, char* zpart) {\n";
  s << "  P::Acc acc;\n";
  for (size_t j = 0; j < ks.zouts.size(); j++) {
    const ZonalOut& z = ks.zouts[j];
    if (z.node->kind == Kind::ZonalReduc) {
      s << "  acc.r" << j << " = " << lit(reduction_neutral(z.red, z.acc_dt)) << ";\n";
    } else {
      std::string M = std::string("mx::MM<") + CT(z.dt) + ">::t";
      s << "  __shared__ double zs" << j << "[4 * " << z.nzones << "];\n";
      s << "  __shared__ " << M << " zn" << j << "[" << z.nzones << "], zx" << j << "[" << z.nzones << "];\n";
      s << "  __shared__ mx::u32 zc" << j << "[" << z.nzones << "];\n";
      s << "  acc.z" << j << ".sum = zs" << j << "; acc.z" << j << ".mn = zn" << j << "; acc.z" << j << ".mx = zx" << j
        << "; acc.z" << j << ".cnt = zc" << j << ";\n";
      s << "  mx::zone_init<" << CT(z.dt) << ", " << z.nzones << ">(acc.z" << j << ");\n";
      s << "  mx::zone_run_reset(acc.q" << j << ");\n";
    }
  }
  if (!zonal.empty()) s << "  __syncthreads();\n";
  s << "  __shared__ P::Smem sm;\n  __shared__ int slot;\n";
  s << "  mx::run_chunks<P, " << tile.TX << ", " << tile.TY << ", " << tile.VX << ", "
    << (ks.form != Form::Focal ? "mx::LOCAL" : ks.stream ? "mx::STREAM" : "mx::STAGED") << ", " << (tile.PF ? "true" : "false")
    << ", (MX_NARROW != 0)>(a0, L0, tasks, starts, ntasks, nchunks, queue, zpart, acc, sm, &slot);\n";
  s << "}\n";

  // fold kernel: folds partial records in index order. Work-group (f, t) folds
  // slice f of F of task t's records [base, base+n) of `src` into record
  // t*F + f of `dst`: task t's chunks (tasks: from the table; stride == 0: one
  // task of n_single chunks) or, with stride > 0, records [t*stride, (t+1)*stride).
  if (!zonal.empty()) {
    s << "\nextern \"C\" __global__ void __launch_bounds__(256) " << ks.fin_name
      << "(const char* src, char* dst, const mx::TaskRec<P::Args>* tasks, const int* starts, int n_single, int stride, int F) {\n";
    s << "  __shared__ mx::u64 red_scratch[4];\n";
    s << "  const int t = blockIdx.y, f = blockIdx.x;\n";
    s << "  int base = 0, n = n_single;\n";
    s << "  if (tasks) { base = tasks[t].L.part0; n = starts[t + 1] - starts[t]; }\n";
    s << "  else if (stride > 0) { base = t * stride; n = stride; }\n";
    s << "  const int b0 = base + (int)((long long)n * f / F), b1 = base + (int)((long long)n * (f + 1) / F);\n";
    s << "  char* out = dst + (size_t)(t * F + f) * " << part_bytes << "u;\n";
    for (size_t j = 0; j < ks.zouts.size(); j++) {
      const ZonalOut& z = ks.zouts[j];
      if (z.node->kind == Kind::ZonalReduc) {
        std::string A = CT(z.acc_dt);
        s << "  {\n    " << A << " r = " << lit(reduction_neutral(z.red, z.acc_dt)) << ";\n";
        s << "    for (int p = b0 + threadIdx.x; p < b1; p += 256) r = " << red_op(z.red) << "::f(r, *(const " << A
          << "*)(src + (size_t)p * " << part_bytes << "u + " << part_off[j] << "));\n";
        s << "    mx::block_reduce_store<" << red_op(z.red) << ", " << A << ">(r, (" << A << "*)(out + " << part_off[j]
          << "), (" << A << "*)red_scratch);\n  }\n";
      } else {
        std::string T = CT(z.dt);
        std::string M = std::string("mx::MM<") + T + ">::t";
        s << "  for (int zi = threadIdx.x; zi < " << z.nzones << "; zi += 256) {\n";
        s << "    double sum = 0.0; " << M << " mn = mx::mm_highest<" << T << ">(), mxv = mx::mm_lowest<" << T
          << ">(); mx::u32 cnt = 0u;\n";
        s << "    for (int p = b0; p < b1; p++) {\n";
        s << "      const mx::ZonePart<" << T << ">& q = ((const mx::ZonePart<" << T << ">*)(src + (size_t)p * "
          << part_bytes << "u + " << part_off[j] << "))[zi];\n";
        s << "      sum += q.sum; mn = mx::OpMin::f(mn, q.mn); mxv = mx::OpMax::f(mxv, q.mx); cnt += q.cnt;\n    }\n";
        s << "    mx::ZonePart<" << T << "> r; r.sum = sum; r.mn = mn; r.mx = mxv; r.cnt = cnt; r.pad = 0u;\n";
        s << "    ((mx::ZonePart<" << T << ">*)(out + " << part_off[j] << "))[zi] = r;\n  }\n";
      }
    }
    s << "}\n";
  }
  std::string src = s.str();
  // named after its source and the templates it instantiates: a template change renames every kernel
  const std::string h = hex_hash(src + dev::header_source());
  auto subst = [&](std::string t) {
    for (size_t pos; (pos = t.find("KRN_NAME")) != std::string::npos;) t.replace(pos, 8, "krn_" + h);
    return t;
  };
  ks.source = subst(src);
  ks.name = "krn_" + h;
  ks.fin_name = zonal.empty() ? "" : ks.name + "_fin";
  ks.signature = ks.source;
  g_local = nullptr;
  return ks;
}

}  // namespace mapx
