#include "runtime.hpp"
#include "rasterio.hpp"
#include "dev.hpp"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <set>
#include <sstream>
#include <stdexcept>

namespace mapx {
namespace radial {  // radial.hip
int run(void* out, const void* in, int dt_bytes, int op, int BS0, int BS1, int NB0, int NB1, int DS0, int DS1, int s0,
        int s1, void* stream);
}  // namespace radial
namespace gen {  // gen.hip
void nalloc_scratch_init(int device);
}  // namespace gen
namespace spread {  // spread.hip
int run(void* out, const void* in, const void* dir, void* cnt, int dt, int BS0, int BS1, int NB0, int NB1, int DS0,
        int DS1, void* stream);
}  // namespace spread
}  // namespace mapx

namespace mapx {

static void fail(const std::string& m) { throw std::runtime_error(m); }

namespace {
struct HostTimer {
  double& acc;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  explicit HostTimer(double& a) : acc(a) {}
  ~HostTimer() { acc += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
};
}  // namespace

// ---- Store ---------------------------------------------------------------------

void* Store::alloc(size_t bytes) {
  bytes = (bytes + 255) & ~size_t(255);
  if (cap) {
    // under the cap the buffers in use and this one must fit: resident rasters move out first
    while (in_use_ + bytes > cap)
      if (bytes > cap || !reclaim || !reclaim())
        fail("store: " + std::to_string(bytes) + " bytes do not fit the HBM cap of " + std::to_string(cap) +
             " bytes (" + std::to_string(in_use_) + " in use by the evaluation)");
  }
  auto it = free_.find(bytes);
  void* p = nullptr;
  if (it != free_.end()) {
    p = it->second;
    free_.erase(it);
    cached_ -= bytes;
    size_[p] = bytes;
    in_use_ += bytes;
    if (cap && in_use_ + cached_ > cap) trim();  // the other cached buffers beyond the cap go back
    return p;
  }
  // no room on the device: cached buffers go back first, then resident rasters
  // move out (each round frees memory or ends in the error)
  for (;;) {
    if (cap && in_use_ + cached_ + bytes > cap) trim();  // leaves in_use_ + bytes <= cap
    try {
      p = dev::malloc_device(bytes);
      break;
    } catch (const std::exception&) {
      if (!free_.empty()) { trim(); continue; }
      if (reclaim && reclaim()) continue;
      throw;
    }
  }
  size_[p] = bytes;
  in_use_ += bytes;
  return p;
}

void Store::release(void* p) {
  auto it = size_.find(p);
  if (it == size_.end()) fail("store: release of an unknown buffer");
  free_.emplace(it->second, p);
  cached_ += it->second;
  in_use_ -= it->second;
  size_.erase(it);
}

// The caller has made sure no enqueued or recorded work reads p any more.
void Store::discard(void* p) {
  auto it = size_.find(p);
  if (it == size_.end()) fail("store: discard of an unknown buffer");
  dev::free_device(p);
  in_use_ -= it->second;
  size_.erase(it);
}

// A cached buffer may still be read by a launch that has been recorded for
// graph replay but not yet enqueued (its consumer released it at launch time):
// the recorded sequence is enqueued first, so the sync below covers it, and
// recording resumes afterwards.
void Store::trim() {
  if (free_.empty()) return;
  const bool rec = dev::recording();
  if (rec) dev::record_flush();
  dev::sync();
  for (auto& f : free_) dev::free_device(f.second);
  free_.clear();
  cached_ = 0;
  if (rec) dev::record_begin();
}

Store::~Store() {
  // process exit: the device context may already be gone; leave buffers to the driver
}

// A .bin or tiled TIFF raster file (the header is read here, the payload when
// the node is first needed on the device).
Node* Dag::read(const std::string& path) {
  const Meta m = raster_file_meta(path);
  auto n = std::make_unique<Node>();
  n->kind = Kind::Read;
  n->meta = m;
  n->path = path;
  return add(std::move(n));
}

// ---- host scalar evaluation with C semantics -----------------------------------------

namespace {
enum CK { CI32, CU32, CI64, CU64, CF32, CF64 };
struct CV {
  CK k;
  int64_t i = 0;
  uint64_t u = 0;
  double d = 0;
  float f = 0;
};
CV to_cv(const Variant& v) {  // integer promotion of the stored type
  CV c;
  switch (v.dtype) {
    case B8: c.k = CI32; c.i = v.u.b8; break;
    case U8: c.k = CI32; c.i = v.u.u8; break;
    case S8: c.k = CI32; c.i = v.u.s8; break;
    case U16: c.k = CI32; c.i = v.u.u16; break;
    case S16: c.k = CI32; c.i = v.u.s16; break;
    case S32: c.k = CI32; c.i = v.u.s32; break;
    case U32: c.k = CU32; c.u = v.u.u32; break;
    case S64: c.k = CI64; c.i = v.u.s64; break;
    case U64: c.k = CU64; c.u = v.u.u64; break;
    case F32: c.k = CF32; c.f = v.u.f32; break;
    case F64: c.k = CF64; c.d = v.u.f64; break;
    default: fail("scalar: value without a type");
  }
  return c;
}
int rank(CK k) { return (k == CI64 || k == CU64) ? 2 : 1; }
bool is_uns(CK k) { return k == CU32 || k == CU64; }
CK common(CK a, CK b) {  // usual arithmetic conversions
  if (a == CF64 || b == CF64) return CF64;
  if (a == CF32 || b == CF32) return CF32;
  if (a == b) return a;
  if (is_uns(a) == is_uns(b)) return rank(a) >= rank(b) ? a : b;
  CK u = is_uns(a) ? a : b, s = is_uns(a) ? b : a;
  if (rank(u) >= rank(s)) return u;
  return s;  // I64 represents every U32
}
CV conv(const CV& c, CK k) {
  CV r; r.k = k;
  auto as_i = [&]() -> int64_t {
    switch (c.k) { case CI32: case CI64: return c.i; case CU32: case CU64: return (int64_t)c.u;
      case CF32: return (int64_t)c.f; default: return (int64_t)c.d; }
  };
  auto as_u = [&]() -> uint64_t {
    switch (c.k) { case CI32: case CI64: return (uint64_t)c.i; case CU32: case CU64: return c.u;
      case CF32: return (uint64_t)c.f; default: return (uint64_t)c.d; }
  };
  auto as_d = [&]() -> double {
    switch (c.k) { case CI32: case CI64: return (double)c.i; case CU32: case CU64: return (double)c.u;
      case CF32: return (double)c.f; default: return c.d; }
  };
  switch (k) {
    case CI32: r.i = (int32_t)as_i(); break;
    case CI64: r.i = as_i(); break;
    case CU32: r.u = (uint32_t)as_u(); break;
    case CU64: r.u = as_u(); break;
    case CF32: r.f = (c.k == CF32) ? c.f : (c.k == CI32 || c.k == CI64) ? (float)c.i
                   : (c.k == CU32 || c.k == CU64) ? (float)c.u : (float)c.d; break;
    case CF64: r.d = as_d(); break;
  }
  return r;
}
Variant from_cv(const CV& c, int dt) {
  Variant v;
  switch (c.k) {
    case CI32: v.dtype = S32; v.u.s32 = (int32_t)c.i; break;
    case CI64: v.dtype = S64; v.u.s64 = c.i; break;
    case CU32: v.dtype = U32; v.u.u32 = (uint32_t)c.u; break;
    case CU64: v.dtype = U64; v.u.u64 = c.u; break;
    case CF32: v.dtype = F32; v.u.f32 = c.f; break;
    case CF64: v.dtype = F64; v.u.f64 = c.d; break;
  }
  return variant_convert(v, dt);
}
CV cv_int(int64_t x) { CV c; c.k = CI32; c.i = x; return c; }

CV binop(int op, CV a, CV b) {
  CK k = common(a.k, b.k);
  if (op == SHL || op == SHR) k = a.k;  // shifts take the promoted left type
  CV x = conv(a, k), y = conv(b, op == SHL || op == SHR ? (rank(b.k) == 2 ? CI64 : CI32) : k), r;
  r.k = k;
  auto cmp = [&](int c) { return cv_int(c); };
  if (k == CF32 || k == CF64) {
    double l = k == CF32 ? x.f : x.d, rr = k == CF32 ? y.f : y.d;
    double res = 0;
    switch (op) {
      case ADD: res = k == CF32 ? (double)(x.f + y.f) : l + rr; break;
      case SUB: res = k == CF32 ? (double)(x.f - y.f) : l - rr; break;
      case MUL: res = k == CF32 ? (double)(x.f * y.f) : l * rr; break;
      case DIV: res = k == CF32 ? (double)(x.f / y.f) : l / rr; break;
      case EQ: return cmp(l == rr); case NE: return cmp(l != rr); case LT: return cmp(l < rr);
      case GT: return cmp(l > rr); case LE: return cmp(l <= rr); case GE: return cmp(l >= rr);
      case AND: return cmp(l != 0 && rr != 0); case OR: return cmp(l != 0 || rr != 0);
      default: fail("scalar: operator not defined on floating types");
    }
    if (k == CF32) r.f = (float)res; else r.d = res;
    return r;
  }
  if (is_uns(k)) {
    uint64_t l = x.u, rr = y.u, res = 0;
    uint64_t mask = k == CU32 ? 0xffffffffull : ~0ull;
    switch (op) {
      case ADD: res = l + rr; break; case SUB: res = l - rr; break; case MUL: res = l * rr; break;
      case DIV: if (!rr) fail("scalar: integer division by zero"); res = l / rr; break;
      case MOD: if (!rr) fail("scalar: integer modulo by zero"); res = l % rr; break;
      case EQ: return cmp(l == rr); case NE: return cmp(l != rr); case LT: return cmp(l < rr);
      case GT: return cmp(l > rr); case LE: return cmp(l <= rr); case GE: return cmp(l >= rr);
      case AND: return cmp(l && rr); case OR: return cmp(l || rr);
      case bAND: res = l & rr; break; case bOR: res = l | rr; break; case bXOR: res = l ^ rr; break;
      case SHL: res = l << (y.i & (k == CU32 ? 31 : 63)); break; case SHR: res = l >> (y.i & (k == CU32 ? 31 : 63)); break;
      default: fail("scalar: bad operator");
    }
    r.u = res & mask;
    return r;
  }
  int64_t l = x.i, rr = y.i, res = 0;
  switch (op) {
    case ADD: res = (int64_t)((uint64_t)l + (uint64_t)rr); break;
    case SUB: res = (int64_t)((uint64_t)l - (uint64_t)rr); break;
    case MUL: res = (int64_t)((uint64_t)l * (uint64_t)rr); break;
    case DIV: if (!rr) fail("scalar: integer division by zero"); res = l / rr; break;
    case MOD: if (!rr) fail("scalar: integer modulo by zero"); res = l % rr; break;
    case EQ: return cmp(l == rr); case NE: return cmp(l != rr); case LT: return cmp(l < rr);
    case GT: return cmp(l > rr); case LE: return cmp(l <= rr); case GE: return cmp(l >= rr);
    case AND: return cmp(l && rr); case OR: return cmp(l || rr);
    case bAND: res = l & rr; break; case bOR: res = l | rr; break; case bXOR: res = l ^ rr; break;
    case SHL: res = (int64_t)((uint64_t)l << (rr & 63)); break; case SHR: res = l >> (rr & 63); break;
    default: fail("scalar: bad operator");
  }
  if (k == CI32) res = (int32_t)res;
  r.i = res;
  return r;
}

double fn1(int op, double x) {
  switch (op) {
    case SIN: return std::sin(x); case COS: return std::cos(x); case TAN: return std::tan(x);
    case ASIN: return std::asin(x); case ACOS: return std::acos(x); case ATAN: return std::atan(x);
    case SINH: return std::sinh(x); case COSH: return std::cosh(x); case TANH: return std::tanh(x);
    case ASINH: return std::asinh(x); case ACOSH: return std::acosh(x); case ATANH: return std::atanh(x);
    case EXP: return std::exp(x); case EXP2: return std::exp2(x); case EXP10: return std::pow(10.0, x);
    case LOG: return std::log(x); case LOG2: return std::log2(x); case LOG10: return std::log10(x);
    case SQRT: return std::sqrt(x); case CBRT: return std::cbrt(x); case ABS: return std::fabs(x);
    case CEIL: return std::ceil(x); case FLOOR: return std::floor(x); case TRUNC: return std::trunc(x);
    case ROUND: return std::round(x);
  }
  fail("scalar: bad function");
  return 0;
}
float fn1f(int op, float x) {
  switch (op) {
    case SIN: return std::sin(x); case COS: return std::cos(x); case TAN: return std::tan(x);
    case ASIN: return std::asin(x); case ACOS: return std::acos(x); case ATAN: return std::atan(x);
    case SINH: return std::sinh(x); case COSH: return std::cosh(x); case TANH: return std::tanh(x);
    case ASINH: return std::asinh(x); case ACOSH: return std::acosh(x); case ATANH: return std::atanh(x);
    case EXP: return std::exp(x); case EXP2: return std::exp2(x); case EXP10: return (float)std::pow(10.0, (double)x);
    case LOG: return std::log(x); case LOG2: return std::log2(x); case LOG10: return std::log10(x);
    case SQRT: return std::sqrt(x); case CBRT: return std::cbrt(x); case ABS: return std::fabs(x);
    case CEIL: return std::ceil(x); case FLOOR: return std::floor(x); case TRUNC: return std::trunc(x);
    case ROUND: return std::round(x);
  }
  fail("scalar: bad function");
  return 0;
}
}  // namespace

Variant eval_scalar_node(const Node* n) {
  auto pv = [&](int i) -> const Variant& {
    const Node* p = n->prev[i];
    if (!p->evaluated) fail("scalar: operand not evaluated");
    return p->value;
  };
  const int dt = n->meta.dt;
  switch (n->kind) {
    case Kind::Constant: return n->cnst;
    case Kind::Cast: return variant_convert(pv(0), dt);
    case Kind::Scalar: return pv(0);
    case Kind::Conditional: {
      CV c = to_cv(pv(0));
      bool t = (c.k == CF32) ? c.f != 0 : (c.k == CF64) ? c.d != 0 : (is_uns(c.k) ? c.u != 0 : c.i != 0);
      return variant_convert(t ? pv(1) : pv(2), dt);
    }
    case Kind::Unary: {
      const int op = n->op;
      if (unary_is_operator(op)) {
        CV a = to_cv(pv(0)), r = a;
        switch (op) {
          case POS: break;
          case NEG:
            if (a.k == CF32) r.f = -a.f; else if (a.k == CF64) r.d = -a.d;
            else if (is_uns(a.k)) r.u = (0 - a.u) & (a.k == CU32 ? 0xffffffffull : ~0ull);
            else r.i = a.k == CI32 ? (int32_t)(0 - (uint64_t)a.i) : (int64_t)(0 - (uint64_t)a.i);
            break;
          case NOT: {
            bool z = a.k == CF32 ? a.f == 0 : a.k == CF64 ? a.d == 0 : is_uns(a.k) ? a.u == 0 : a.i == 0;
            r = cv_int(z ? 1 : 0);
            break;
          }
          case bNOT:
            if (is_uns(a.k)) r.u = ~a.u & (a.k == CU32 ? 0xffffffffull : ~0ull);
            else r.i = a.k == CI32 ? (int32_t)~a.i : ~a.i;
            break;
        }
        return from_cv(r, dt);
      }
      if (dt == F32) { Variant v = variant_zero(F32); v.u.f32 = fn1f(op, variant_as<float>(pv(0))); return v; }
      if (dt == F64) { Variant v = variant_zero(F64); v.u.f64 = fn1(op, variant_as<double>(pv(0))); return v; }
      if (op == ABS) { CV a = to_cv(pv(0)); if (!is_uns(a.k) && a.i < 0) a.i = -a.i; return from_cv(a, dt); }
      Variant d = variant_zero(F64); d.u.f64 = fn1(op, variant_as<double>(pv(0)));
      return variant_convert(d, dt);
    }
    case Kind::Binary: {
      const int op = n->op;
      if (binary_is_operator(op)) {
        if (op == MOD && (is_float_dt(pv(0).dtype) || is_float_dt(pv(1).dtype))) {
          int t = promote(pv(0).dtype, pv(1).dtype);
          Variant v = variant_zero(t);
          if (t == F32) v.u.f32 = std::fmod(variant_as<float>(pv(0)), variant_as<float>(pv(1)));
          else v.u.f64 = std::fmod(variant_as<double>(pv(0)), variant_as<double>(pv(1)));
          return variant_convert(v, dt);
        }
        return from_cv(binop(op, to_cv(pv(0)), to_cv(pv(1))), dt);
      }
      Variant l = variant_convert(pv(0), dt), r = variant_convert(pv(1), dt);
      if (op == MAX2 || op == MIN2) {
        if (dt == F32) { Variant v = l; v.u.f32 = op == MAX2 ? std::fmax(l.u.f32, r.u.f32) : std::fmin(l.u.f32, r.u.f32); return v; }
        if (dt == F64) { Variant v = l; v.u.f64 = op == MAX2 ? std::fmax(l.u.f64, r.u.f64) : std::fmin(l.u.f64, r.u.f64); return v; }
        CV a = to_cv(l), b = to_cv(r);
        CV gt = binop(GT, a, b);
        bool lg = gt.i != 0;
        return (op == MAX2) == lg ? l : r;
      }
      if (dt == F32) {
        float a = l.u.f32, b = r.u.f32, res = 0;
        switch (op) {
          case ATAN2: res = std::atan2(a, b); break; case POW: res = std::pow(a, b); break;
          case HYPOT: res = std::hypot(a, b); break; case FMOD: res = std::fmod(a, b); break;
        }
        Variant v = variant_zero(F32); v.u.f32 = res; return v;
      }
      double a = variant_as<double>(pv(0)), b = variant_as<double>(pv(1)), res = 0;
      switch (op) {
        case ATAN2: res = std::atan2(a, b); break; case POW: res = std::pow(a, b); break;
        case HYPOT: res = std::hypot(a, b); break; case FMOD: res = std::fmod(a, b); break;
      }
      Variant d = variant_zero(F64); d.u.f64 = res;
      return variant_convert(d, dt);
    }
    default: fail(std::string("scalar: cannot evaluate a ") + kind_name(n->kind) + " node on the host");
  }
  return Variant{};
}

// ---- Runtime ---------------------------------------------------------------------

Runtime& Runtime::get() {
  static Runtime* r = new Runtime();  // never destroyed: device teardown order at exit is the driver's
  return *r;
}

static void read_env(Options& o) {
  if (const char* e = std::getenv("MAPX_FORM")) {
    std::string s(e);
    if (s == "plain") o.form = FORM_PLAIN;
    else if (s == "batched") o.form = FORM_BATCHED;
  }
  if (const char* e = std::getenv("MAPX_GRAPH")) o.graph = std::atoi(e) != 0;
  if (const char* e = std::getenv("MAPX_FUSE_FOCAL_ZONAL")) o.focal_zonal = std::atoi(e) != 0;
  if (const char* e = std::getenv("MAPX_HBM_CAP")) o.hbm_cap = (size_t)std::strtoull(e, nullptr, 10);
}

void Runtime::setupDevices(const std::string& plat, unsigned dev_type, const std::string& dev_name) {
  (void)plat; (void)dev_name;
  if (dev_type != DEV_GPU && dev_type != DEV_DEF && dev_type != DEV_ACC)
    fail("setupDevices: this backend runs on the GPU only (DEV_GPU)");
  int device = 0;
  if (const char* e = std::getenv("MAPX_DEVICE")) device = std::atoi(e);
  dev::init(device);
  gen::nalloc_scratch_init(device);  // the device address of the allocation-free generator's scratch
  read_env(opt);
  store.cap = opt.hbm_cap;
  store.reclaim = [this] { return spill_one(); };
  if (const char* p = std::getenv("MAPX_TUNE_CACHE")) tune_load(p);
}

size_t Runtime::rasterBytes(const Meta& m) const {
  return (size_t)m.nb(0) * m.nb(1) * m.bs.arr[0] * m.bs.arr[1] * sizeof_dt(m.dt);
}

void* Runtime::residentPtr(Node* n) {
  auto it = resident_.find(n);
  if (it == resident_.end()) {
    if (!spilled_.count(n)) fail("raster of node " + std::to_string(n->id) + " is not resident");
    unspill(n);
    it = resident_.find(n);
  }
  return it->second;
}

void Runtime::forget_spilled(Node* n) {
  auto sp = spilled_.find(n);
  if (sp == spilled_.end()) return;
  dev::free_host(sp->second.host);
  spilled_.erase(sp);
}

void Runtime::drop(Node* n) {
  if (spilled_.count(n)) {
    if (!spilled_[n].generated || n->ref <= 0) forget_spilled(n);
    return;
  }
  auto it = resident_.find(n);
  if (it == resident_.end()) return;
  if (pins_.count(n)) return;  // an admitted request still reads it
  // a generated raster has no other source: it lives as long as its node
  if (n->host && n->host->dev == it->second && n->ref > 0) return;
  store.release(it->second);
  resident_.erase(it);
}

static int ext_refs(const Node* n) { return n->ref - (int)n->next.size(); }

void Runtime::forget(Node* n) {
  forget_spilled(n);
  auto it = resident_.find(n);
  if (it == resident_.end()) return;
  store.release(it->second);
  resident_.erase(it);
}

// Host-side stream operations (copies, synchronisation) must not overtake
// launches that are still being recorded for graph replay: recording is
// flushed before them and resumed after.
struct PauseRecording {
  bool rec;
  PauseRecording() : rec(dev::recording()) { if (rec) dev::record_flush(); }
  ~PauseRecording() { if (rec) dev::record_begin(); }
};

// ---- resident cache -------------------------------------------------------------

void Runtime::cacheBudget(size_t bytes) {
  std::lock_guard<std::recursive_mutex> lk(mtx_);
  cache_budget_ = bytes;
  while (cache_budget_ && cache_bytes_ > cache_budget_ && !cache_lru_.empty()) cache_evict(cache_lru_.back());
}

Node* Runtime::cacheGet(long long key) {
  std::lock_guard<std::recursive_mutex> lk(mtx_);
  auto it = cache_.find(key);
  if (it == cache_.end()) { cache_misses_++; return nullptr; }
  cache_hits_++;
  cache_lru_.splice(cache_lru_.begin(), cache_lru_, it->second.pos);
  return it->second.node;
}

void Runtime::cachePut(long long key, Node* n) {
  std::lock_guard<std::recursive_mutex> lk(mtx_);
  if (!n) fail("cachePut: null node");
  if (cache_.count(key)) cache_evict(key);
  CacheEntry e;
  e.node = n;
  e.bytes = n->meta.nd == D2 ? rasterBytes(n->meta) : 0;
  n->ref++;  // the cache's reference
  cache_lru_.push_front(key);
  e.pos = cache_lru_.begin();
  cache_.emplace(key, e);
  cache_bytes_ += e.bytes;
  // the new entry stays even when it alone exceeds the budget
  while (cache_budget_ && cache_bytes_ > cache_budget_ && cache_lru_.size() > 1) cache_evict(cache_lru_.back());
}

void Runtime::cacheDrop(long long key) {
  std::lock_guard<std::recursive_mutex> lk(mtx_);
  if (cache_.count(key)) cache_evict(key);
}

// Releases the cache's reference; the raster leaves HBM now if nothing else
// holds the node (a request still reading it keeps it pinned until it retires).
void Runtime::cache_evict(long long key) {
  auto it = cache_.find(key);
  if (it == cache_.end()) return;
  Node* n = it->second.node;
  cache_bytes_ -= it->second.bytes;
  cache_lru_.erase(it->second.pos);
  cache_.erase(it);
  cache_evictions_++;
  if (n->ref > 0) n->ref--;
  if (n->ref <= 0) drop(n);
}

std::string Runtime::cacheStats() const {
  std::ostringstream j;
  j << "{\"entries\": " << cache_.size() << ", \"bytes\": " << cache_bytes_ << ", \"budget\": " << cache_budget_
    << ", \"hits\": " << cache_hits_ << ", \"misses\": " << cache_misses_ << ", \"evictions\": " << cache_evictions_ << "}";
  return j.str();
}

void Runtime::release_unreferenced() {
  std::vector<Node*> gone;
  for (auto& r : resident_) if (r.first->ref <= 0) gone.push_back(r.first);
  for (auto& r : spilled_) if (r.first->ref <= 0) gone.push_back(r.first);
  for (Node* n : gone) drop(n);
}

// ---- spilling: resident rasters out of HBM and back ---------------------------------

// Called by the store when an allocation finds no room. The victim is the
// raster whose last evaluation is the oldest among those that no evaluation or
// admitted request in flight lists (Runtime::LiveTick) and no request pins. A
// raster that can be made again from its source (a file, a host array) or that
// nothing references any more leaves HBM; any other (a computed or generated
// raster still referenced) is copied to pinned host memory first. Enqueued and
// recorded work is drained before the buffer goes back to the device.
bool Runtime::spill_one() {
  const unsigned long long floor = live_ticks_.empty() ? ~0ull : *live_ticks_.begin();
  Node* v = nullptr;
  for (auto& r : resident_) {
    Node* n = r.first;
    if (n->use >= floor || pins_.count(n)) continue;
    if (!v || n->use < v->use || (n->use == v->use && n->id < v->id)) v = n;
  }
  if (!v) return false;
  void* d = resident_[v];
  const size_t bytes = rasterBytes(v->meta);
  const bool generated = v->host && v->host->dev == d;
  const bool remade = !generated && v->kind == Kind::Read;  // a file or a host array
  PauseRecording pause;
  if (remade || v->ref <= 0) {
    dev::sync();
    cnt.dropped_for_space++;
  } else {
    void* h = dev::malloc_host(bytes);
    try {
      dev::d2h(h, d, bytes);
      dev::sync();
    } catch (...) {
      dev::free_host(h);
      throw;
    }
    spilled_[v] = Spilled{h, generated};
    if (generated) v->host->dev = nullptr;
    cnt.spills++;
    cnt.bytes_spilled += (long)bytes;
  }
  resident_.erase(v);
  store.discard(d);
  return true;
}

unsigned long long Runtime::tick_begin(const std::vector<Node*>& list) {
  const unsigned long long t = ++use_clock_;
  live_ticks_.insert(t);
  for (Node* n : list) n->use = t;
  return t;
}

void Runtime::tick_end(unsigned long long t) {
  auto it = live_ticks_.find(t);
  if (it != live_ticks_.end()) live_ticks_.erase(it);
}

void Runtime::unspill(Node* n) {
  auto sp = spilled_.find(n);
  if (sp == spilled_.end()) fail("unspill: raster of node " + std::to_string(n->id) + " is not spilled");
  const size_t bytes = rasterBytes(n->meta);
  void* d = store.alloc(bytes);  // may move other rasters out; this one is not resident
  sp = spilled_.find(n);
  {
    PauseRecording pause;
    dev::h2d(d, sp->second.host, bytes);
    dev::sync();
  }
  if (sp->second.generated) n->host->dev = d;
  dev::free_host(sp->second.host);
  spilled_.erase(sp);
  resident_[n] = d;
  cnt.unspills++;
  cnt.bytes_h2d += (long)bytes;
}

// Staging chunk: files and host rasters move through two pinned buffers of
// this size, so reading chunk k overlaps the copy of chunk k-1.
static const size_t kStage = (size_t)32 << 20;

void* Runtime::staging(int i) {
  if (!stage_buf_[i]) stage_buf_[i] = dev::malloc_host(kStage);
  if (stage_ev_[i]) {
    dev::event_sync(stage_ev_[i]);
    dev::event_destroy(stage_ev_[i]);
    stage_ev_[i] = nullptr;
  }
  return stage_buf_[i];
}

void Runtime::staged(int i) { stage_ev_[i] = dev::event_record(); }

void Runtime::ensure_input(Node* n) {
  if (resident_.count(n)) return;
  if (spilled_.count(n)) { unspill(n); return; }
  if (n->kind != Kind::Read) fail("input raster of node " + std::to_string(n->id) + " is not available");
  const size_t bytes = rasterBytes(n->meta);
  if (n->host && n->host->dev) fail("generated raster of node " + std::to_string(n->id) + " was released");
  PauseRecording pause;
  std::unique_ptr<RasterFile> file;
  if (!n->host) {
    file = std::make_unique<RasterFile>(n->path);
    const Meta& m = file->meta();
    if (m.dt != n->meta.dt || !dim2_eq(m.ds, n->meta.ds) || !dim2_eq(m.bs, n->meta.bs) || file->payload_bytes() != bytes)
      fail("read: " + n->path + " changed since it was opened");
  }
  char* d = (char*)store.alloc(bytes);
  try {
    // chunk k is read (or copied) into buffer k%2 while chunk k-1 is in flight to HBM
    for (size_t off = 0, k = 0; off < bytes; off += kStage, k++) {
      const size_t len = std::min(kStage, bytes - off);
      void* s = staging((int)(k & 1));
      if (n->host) std::memcpy(s, n->host->bytes.data() + off, len);
      else file->read(off, len, s);
      dev::h2d(d + off, s, len);
      staged((int)(k & 1));
    }
  } catch (...) {
    dev::sync();
    store.release(d);
    throw;
  }
  cnt.bytes_h2d += (long)bytes;
  resident_[n] = d;
}

void Runtime::fetch(Node* n, void* host, size_t bytes) {
  std::lock_guard<std::recursive_mutex> lk(mtx_);
  if (bytes != rasterBytes(n->meta)) fail("fetch: byte count differs from the raster size");
  auto sp = spilled_.find(n);
  if (sp != spilled_.end()) {  // moved out of HBM: the host copy is the raster
    std::memcpy(host, sp->second.host, bytes);
    return;
  }
  void* d = residentPtr(n);
  PauseRecording pause;
  dev::d2h(host, d, bytes);
  dev::sync();
  cnt.bytes_d2h += (long)bytes;
}

// Streams a resident raster to a .bin or BigTIFF file: chunk k+1 is copied
// from HBM while chunk k is written.
void Runtime::write_file(Node* w) {
  Node* p = w->prev[0];
  const size_t bytes = rasterBytes(p->meta);
  PauseRecording pause;
  RasterWriter out(w->path, p->meta);
  const char* src = (const char*)residentPtr(p);
  const size_t nchunks = (bytes + kStage - 1) / kStage;
  auto issue = [&](size_t k) {
    const size_t off = k * kStage;
    void* s = staging((int)(k & 1));
    dev::d2h(s, src + off, std::min(kStage, bytes - off));
    staged((int)(k & 1));
  };
  if (nchunks) issue(0);
  for (size_t k = 0; k < nchunks; k++) {
    if (k + 1 < nchunks) issue(k + 1);
    dev::event_sync(stage_ev_[k & 1]);
    out.append(stage_buf_[k & 1], std::min(kStage, bytes - k * kStage));
  }
  out.finish();
  cnt.bytes_d2h += (long)bytes;
}

std::vector<Node*> Runtime::cut_list(const std::vector<Node*>& targets) {
  const unsigned st = ++visit_stamp_;
  for (Node* t : targets) t->tstamp = st;  // targets are listed, never cut
  std::vector<Node*> stack(targets.begin(), targets.end()), out;
  while (!stack.empty()) {
    Node* n = stack.back(); stack.pop_back();
    if (n->vstamp == st) continue;
    n->vstamp = st;
    out.push_back(n);
    if (n->tstamp != st) {
      bool have = (n->meta.nd == D2 && (resident_.count(n) || spilled_.count(n))) ||
                  (n->meta.nd == D0 && n->evaluated && n->kind != Kind::Constant);
      if (have && n->kind != Kind::Read) { n->cut = true; continue; }
    }
    for (Node* p : n->prev) stack.push_back(p);
  }
  std::sort(out.begin(), out.end(), [](Node* a, Node* b) {
    if (a->id != b->id) return a->id < b->id;
    return (int)a->isOutput() < (int)b->isOutput();
  });
  return out;
}

// `targets` sorted by address (Eval::tsorted): a group's nodes are looked up
// in it, so the cost does not grow with the number of targets.
std::vector<Node*> Runtime::d2_outputs(const Group& g, const std::vector<Node*>& targets) {
  std::vector<Node*> v;
  for (Node* n : g.outs) {
    if (n->kind == Kind::Write) v.push_back(n->prev[0]);
    else if (n->kind == Kind::Scalar) continue;
    else if (n->meta.nd == D2) v.push_back(n);
  }
  for (Node* t : g.nodes)
    if (t->meta.nd == D2 && std::binary_search(targets.begin(), targets.end(), t)) v.push_back(t);
  std::sort(v.begin(), v.end(), [](Node* a, Node* b) { return a->id < b->id; });
  v.erase(std::unique(v.begin(), v.end()), v.end());
  return v;
}

// Tile shapes are tuned per computation: the group's inputs and nodes (not
// which of its nodes are kept: every generation of an iterated script shares
// the key with its last one; nor floating mask weights, which kernels take as
// arguments) and the block size.
// The inputs are listed in the generated kernel's order (D2 before D0, each by
// id), not in the fuser's: the fuser's order depends on node ids, which differ
// between the first generation of an iterated script (its input was created
// before the constants) and the later ones.
std::string Runtime::tune_key(const Group& g) const {
  const Meta m = g.meta();
  std::vector<Node*> ins(g.ins.begin(), g.ins.end());
  std::sort(ins.begin(), ins.end(), [](Node* a, Node* b) {
    if ((a->meta.nd == D0) != (b->meta.nd == D0)) return a->meta.nd != D0;
    return a->id < b->id;
  });
  std::ostringstream s;
  for (Node* n : ins) s << n->meta.nd << name_dt(n->meta.dt) << ";";
  s << "|";
  for (Node* n : g.nodes) s << n->signature(false) << ";";
  s << "#" << m.bs.arr[0] << "x" << m.bs.arr[1];
  return s.str();
}

TileShape Runtime::tile_for(const Group& g) {
  TileShape t = default_tile(g);
  if (opt.tile.TX > 0) t = opt.tile;
  else {
    auto it = tuned_.find(tune_key(g));
    if (it != tuned_.end()) t = it->second;
  }
  // a tile never exceeds its block: halve the width (keeping 256 threads by
  // growing the height) or the height while it does
  const Meta m = g.meta();
  if (m.nd == D2) {
    // vectors never straddle blocks: VX divides the block width
    while (m.bs.arr[0] % t.VX) { t.VX /= 2; t.TX /= 2; }
    while (t.TX > m.bs.arr[0] && t.TX / t.VX > 16) { t.TX /= 2; t.TY *= 2; }
    while (t.TX > m.bs.arr[0] && t.VX > 1) { t.TX /= 2; t.VX /= 2; }
    while (t.TY > m.bs.arr[1] && t.TY > 256 / (t.TX / t.VX)) t.TY /= 2;
  }
  return t;
}

// Geometry precondition of the focal template: a halo never exceeds a block,
// so a halo lookup reaches at most the neighbour blocks (and, past an edge
// block narrower than the halo, the border rule applied again from there).
static void check_focal_geometry(const Meta& m, int H0, int H1) {
  if (H0 > m.bs.arr[0] || H1 > m.bs.arr[1]) fail("focal halo larger than the block size");
}

// Access: one cell of a resident raster (block-tiled), copied back once the
// stream has produced it.
Variant Runtime::access_value(Node* n) {
  Node* p = n->prev[0];
  const Meta& m = p->meta;
  const int x = n->coord.arr[0], y = n->coord.arr[1];
  const int bx = x / m.bs.arr[0], by = y / m.bs.arr[1];
  const size_t e = ((size_t)(by * m.nb(0) + bx) * m.bs.arr[1] + (size_t)(y - by * m.bs.arr[1])) * m.bs.arr[0] +
                   (size_t)(x - bx * m.bs.arr[0]);
  const size_t sz = sizeof_dt(m.dt);
  Variant v = variant_zero(m.dt);
  PauseRecording pause;
  dev::d2h(&v.u, (const char*)residentPtr(p) + e * sz, sz);
  dev::sync();
  cnt.bytes_d2h += (long)sz;
  return v;
}

void Runtime::scalar_task(const Group& g) {
  std::vector<Node*> v(g.nodes);
  v.insert(v.end(), g.outs.begin(), g.outs.end());
  std::sort(v.begin(), v.end(), [](Node* a, Node* b) {
    if (a->id != b->id) return a->id < b->id;
    return (int)a->isOutput() < (int)b->isOutput();
  });
  for (Node* n : v) {
    if (n->meta.nd != D0) continue;
    if (n->kind == Kind::Constant) continue;
    n->value = n->kind == Kind::Access ? access_value(n) : eval_scalar_node(n);
    n->evaluated = true;
  }
  cnt.scalar_tasks++;
}

namespace {
// mx::Launch of the device templates
struct HostLaunch { int BS0, BS1, NB0, NB1, DS0, DS1, blk0, nblk, TB0, TB1, CH, CPB, part0; };
static_assert(sizeof(HostLaunch) == 52, "mx::Launch layout");

uint32_t morton2(uint32_t x, uint32_t y) {
  auto spread = [](uint32_t v) {
    v &= 0xffff;
    v = (v | (v << 8)) & 0x00ff00ff; v = (v | (v << 4)) & 0x0f0f0f0f;
    v = (v | (v << 2)) & 0x33333333; v = (v | (v << 1)) & 0x55555555;
    return v;
  };
  return spread(x) | (spread(y) << 1);
}

// Cells per chunk of a zonal group: one partial record per chunk.
constexpr int kZonalChunkCells = 65536;
// Tiles per chunk of other groups: consecutive tiles of a block run on one
// work-group (halo rows shared with the previous tile are still in L2), at
// most kChunkTiles, fewer when the launch would otherwise leave work-groups
// idle (at least two chunks per work-group of the grid).
constexpr int kChunkTiles = 16;
int adaptive_chunk(long tiles_total, int grid_max) {
  const long per = tiles_total / (2L * std::max(1, grid_max));
  return (int)std::max(1L, std::min<long>(kChunkTiles, per));
}
}  // namespace

// One fused group over its raster: the unit a launch runs, alone or in a batch.
struct Runtime::Task {
  const Group* g = nullptr;
  KernelEntry* ke = nullptr;
  std::vector<char> args;  // P::Args
  HostLaunch L{};
  int tiles = 0;           // tiles per block
  int nchunks = 0;

  // chunk split: zonal groups keep CH fixed (it decides their partial records)
  void chunking(int chunk_tiles) {
    if (ke->spec.zouts.empty())
      L.CH = std::max(1, std::min(tiles, chunk_tiles > 0 ? chunk_tiles : kChunkTiles));
    L.CPB = (tiles + L.CH - 1) / L.CH;
    nchunks = L.nblk * L.CPB;
  }
};

// D2 nodes a group's kernel writes to HBM (not yet resident), and whether it reduces.
std::vector<Node*> Runtime::kernel_outputs(const Group& g, const std::vector<Node*>& targets, bool& has_zonal) {
  std::vector<Node*> outs = d2_outputs(g, targets), kouts;
  for (Node* n : outs)
    if (std::find(g.nodes.begin(), g.nodes.end(), n) != g.nodes.end() && !resident_.count(n) && !spilled_.count(n))
      kouts.push_back(n);
  has_zonal = false;
  for (Node* n : g.nodes) if (n->kind == Kind::ZonalReduc || n->kind == Kind::ZonalStats) has_zonal = true;
  return kouts;
}

Node* Runtime::radial_node(const Group& g) {
  for (Node* n : g.nodes)
    if (n->kind == Kind::RadialScan) return n;
  return nullptr;
}

// A radial group is its RadialScan node alone (the fuser keeps it so): the
// scan of a resident raster, launched band by band (radial.hip) in stream
// order with the other groups of its level.
void Runtime::radial_task(const Group& g, Eval& e) {
  Node* n = radial_node(g);
  if (g.nodes.size() != 1) fail("radial scan fused with other nodes");
  bool zonal = false;
  if (kernel_outputs(g, e.tsorted, zonal).empty()) return;
  for (Node* in : g.ins)
    if (in->meta.nd == D2) ensure_input(in);
  const Meta& m = n->meta;
  void* out = store.alloc(rasterBytes(m));
  resident_[n] = out;
  const int op = n->op == SUM ? 0 : n->op == PROD ? 1 : n->op == MAX ? 2 : 3;
  PauseRecording pause;
  cnt.launches += radial::run(out, residentPtr(n->prev[0]), sizeof_dt(m.dt), op, m.bs.arr[0], m.bs.arr[1], m.nb(0),
                              m.nb(1), m.ds.arr[0], m.ds.arr[1], n->coord.arr[0], n->coord.arr[1], dev::stream());
}

// The targets and the nodes evaluated with them: the MAX / MIN reductions of
// every Stats node they depend on (the reference computes them in the Stats
// task, runtime/task/StatsTask.cpp:95-160).
// ---- constant-raster prediction ----------------------------------------------------
// The reference's Predictor (runtime/visitor/Predictor.cpp:22-79, called per
// block job from Worker.cpp:84-90) skips kernels whose output blocks are
// constant. Here, over whole rasters at planning time: a D2 node holds one
// value in every cell when it is a Constant, an operator, cast, conditional
// or min/max of such values, or a convolution, neighbour or window reduction
// of a constant raster (every tap holds the value, whatever the border rule
// reads). Only statements whose host evaluation gives the device's bits are
// folded (IEEE operators, integer arithmetic in C types, exact rounding
// functions; no transcendentals, no float-to-integer casts out of range, no
// division by zero or out-of-range shifts). A zonal reduction of a constant
// raster is then known without a kernel: SUM = the value times the valid
// cell count (integer types wrap; an F32 value times at most 2^29 cells is
// exact in the FP64 accumulator, so the device's sum is that product), MAX
// and MIN the value (the neutral for NaN, which fmax/fmin skip), rAND / rOR
// the value folded with the neutral, integer PROD the value to the power of
// the count. Quirk B8 (the reference returns the value itself for SUM) is
// not reproduced.
namespace {
bool foldable(const Node* n, const std::vector<Variant>& pv) {
  switch (n->kind) {
    case Kind::Cast:
      if (is_float_dt(pv[0].dtype) && !is_float_dt(n->meta.dt)) {
        const double d = variant_as<double>(pv[0]);
        if (!(d > -2147483648.0 && d < 2147483648.0)) return false;  // C leaves the rest undefined
      }
      return true;
    case Kind::Conditional: case Kind::Stats: case Kind::Barrier:
      return true;
    case Kind::Unary:
      return unary_is_operator(n->op) || n->op == ABS || n->op == CEIL || n->op == FLOOR || n->op == TRUNC ||
             n->op == ROUND;
    case Kind::Binary: {
      const int op = n->op;
      if (!(binary_is_operator(op) || op == MAX2 || op == MIN2 || op == FMOD)) return false;
      if ((op == DIV || op == MOD) && !is_float_dt(pv[1].dtype) && variant_as<double>(pv[1]) == 0.0) return false;
      if ((op == SHL || op == SHR) && (is_float_dt(pv[1].dtype) || variant_as<double>(pv[1]) < 0 ||
                                       variant_as<double>(pv[1]) >= 8.0 * sizeof_dt(promote(pv[0].dtype, S32))))
        return false;
      return true;
    }
    default:
      return false;
  }
}

// One statement on known operand values, with the C semantics of the kernels.
Variant fold(const Node* n, const std::vector<Variant>& pv) {
  if (n->kind == Kind::Stats || n->kind == Kind::Barrier) return pv[0];
  std::vector<Node> ops(pv.size());
  Node t;
  t.kind = n->kind;
  t.op = n->op;
  t.meta = n->meta;
  for (size_t i = 0; i < pv.size(); i++) {
    ops[i].kind = Kind::Constant;
    ops[i].meta.dt = pv[i].dtype;
    ops[i].value = pv[i];
    ops[i].evaluated = true;
    t.prev.push_back(&ops[i]);
  }
  return eval_scalar_node(&t);
}

// acc = OP(acc, v) for `count` taps in the node's C type (the window statements).
Variant window_fold(const Node* n, const Variant& v) {
  const int dt = n->meta.dt;
  if (n->kind == Kind::Neighbor) return v;
  if (n->kind == Kind::FocalFunc && (n->op == MAX || n->op == MIN)) {
    if (is_float_dt(dt) && std::isnan(variant_as<double>(v))) return reduction_neutral(n->op, dt);
    return variant_convert(v, dt);
  }
  Variant acc = n->kind == Kind::Convolution ? variant_convert(Variant{{.s32 = 0}, S32}, dt) : reduction_neutral(n->op, dt);
  for (const Variant& w0 : n->mask.v) {
    if (n->kind == Kind::Convolution) {
      const Variant w = quantize_literal(w0);
      acc = from_cv(binop(ADD, to_cv(acc), binop(MUL, to_cv(v), to_cv(w))), dt);
    } else {
      const int op = n->op == SUM ? ADD : n->op == PROD ? MUL : n->op == rAND ? bAND : bOR;
      acc = from_cv(binop(op, to_cv(acc), to_cv(variant_convert(v, dt))), dt);
    }
  }
  return acc;
}
}  // namespace

int Runtime::predict(const std::vector<Node*>& targets) {
  if (!opt.predict) return 0;
  bool any = false;  // a zonal reduction still to compute: the only node predicted
  for (Node* n : reach(targets))
    if (n->kind == Kind::ZonalReduc && !n->evaluated) { any = true; break; }
  if (!any) return 0;
  std::unordered_map<const Node*, Variant> cval;
  int done = 0;
  for (Node* n : list_upwards(targets)) {
    try {
      if (n->meta.nd == D0) {
        if (n->kind == Kind::Constant) cval[n] = n->cnst;
        else if (n->evaluated) cval[n] = n->value;
        else if (n->kind == Kind::ZonalReduc && cval.count(n->prev[0])) {
          const Variant v = cval[n->prev[0]];
          const int dt = n->meta.dt;
          const Meta& m = n->prev[0]->meta;
          const uint64_t cells = (uint64_t)m.ds.arr[0] * m.ds.arr[1];
          Variant r;
          bool ok = true;
          if (dt == B8 || v.dtype != dt) ok = false;
          else if (n->op == MAX || n->op == MIN) {
            r = (is_float_dt(dt) && std::isnan(variant_as<double>(v))) ? reduction_neutral(n->op, dt) : v;
          } else if (n->op == SUM) {
            if (dt == F32 && cells <= (1ull << 29)) {
              Variant d = variant_zero(F64);
              d.u.f64 = (double)v.u.f32 * (double)cells;
              r = variant_convert(d, F32);
            } else if (!is_float_dt(dt)) {
              Variant c = variant_zero(U64);
              c.u.u64 = cells;
              r = from_cv(binop(MUL, to_cv(variant_convert(v, U64)), to_cv(c)), dt);
            } else ok = false;
          } else if (n->op == PROD && !is_float_dt(dt)) {
            uint64_t base = variant_convert(v, U64).u.u64, e = cells, acc = 1;
            while (e) { if (e & 1) acc *= base; base *= base; e >>= 1; }
            Variant a = variant_zero(U64);
            a.u.u64 = acc;
            r = variant_convert(a, dt);
          } else if ((n->op == rAND || n->op == rOR) && !is_float_dt(dt)) {
            r = from_cv(binop(n->op == rAND ? bAND : bOR, to_cv(v), to_cv(reduction_neutral(n->op, dt))), dt);
          } else ok = false;
          if (ok) {
            n->value = r;
            n->evaluated = true;
            cval[n] = r;
            done++;
          }
        }
        continue;
      }
      if (n->meta.nd != D2 || resident_.count(n)) continue;
      if (n->kind == Kind::Constant) { cval[n] = variant_convert(n->cnst, n->meta.dt); continue; }
      if (n->kind == Kind::Neighbor || n->kind == Kind::Convolution || n->kind == Kind::FocalFunc) {
        auto it = cval.find(n->prev[0]);
        if (it != cval.end()) cval[n] = window_fold(n, it->second);
        continue;
      }
      std::vector<Variant> pv;
      bool all = true;
      for (Node* p : n->prev) {
        auto it = cval.find(p);
        if (it == cval.end()) { all = false; break; }
        pv.push_back(it->second);
      }
      if (all && !n->prev.empty() && foldable(n, pv)) cval[n] = fold(n, pv);
    } catch (const std::exception&) {
      // not folded: the kernel computes it
    }
  }
  cnt.predicted += done;
  return done;
}

const std::vector<Node*>& Runtime::reach(const std::vector<Node*>& roots) {
  const unsigned st = ++visit_stamp_;
  reach_.clear();
  reach_stack_.assign(roots.begin(), roots.end());
  while (!reach_stack_.empty()) {
    Node* n = reach_stack_.back();
    reach_stack_.pop_back();
    if (n->vstamp == st) continue;
    n->vstamp = st;
    reach_.push_back(n);
    for (Node* p : n->prev) reach_stack_.push_back(p);
  }
  return reach_;
}

std::vector<Node*> Runtime::with_aux(const std::vector<Node*>& targets) {
  std::vector<Node*> owners;
  for (Node* n : reach(targets))
    if (!n->aux.empty()) owners.push_back(n);
  std::vector<Node*> out = targets;
  if (owners.empty()) return out;
  std::sort(owners.begin(), owners.end(), [](Node* a, Node* b) { return a->id < b->id; });
  std::unordered_set<Node*> have(targets.begin(), targets.end());
  for (Node* n : owners)
    for (Node* a : n->aux)
      if (have.insert(a).second) out.push_back(a);
  return out;
}

// Groups evaluated on the host: D0 chains, and cell accesses (a D0 value read
// back from a resident raster, their input).
bool Runtime::host_group(const Group& g) {
  if (g.numdim() == D0) return true;
  for (Node* n : g.nodes)
    if (n->kind == Kind::Access) return true;
  return false;
}

Node* Runtime::spread_node(const Group& g) {
  for (Node* n : g.nodes)
    if (n->kind == Kind::SpreadScan) return n;
  return nullptr;
}

// A spreading group is its SpreadScan node alone: the accumulation of a
// resident raster along the resident direction raster (spread.hip: tiles in
// LDS that also count and list the walk sources, then the walks), with a count
// word per cell and the source list taken from the store for the launches.
void Runtime::spread_task(const Group& g, Eval& e) {
  Node* n = spread_node(g);
  if (g.nodes.size() != 1) fail("spread scan fused with other nodes");
  bool zonal = false;
  if (kernel_outputs(g, e.tsorted, zonal).empty()) return;
  for (Node* in : g.ins)
    if (in->meta.nd == D2) ensure_input(in);
  const Meta& m = n->meta;
  void* out = store.alloc(rasterBytes(m));
  resident_[n] = out;
  const size_t cells = (size_t)m.bs.arr[0] * m.bs.arr[1] * m.nb(0) * m.nb(1);
  void* counts = store.alloc(cells * 8 + 8);  // a status word per cell, then the list of walk sources
  PauseRecording pause;
  cnt.launches += spread::run(out, residentPtr(n->prev[0]), residentPtr(n->prev[1]), counts, m.dt, m.bs.arr[0],
                              m.bs.arr[1], m.nb(0), m.nb(1), m.ds.arr[0], m.ds.arr[1], dev::stream());
  store.release(counts);  // reused in stream order, after the walks
}

bool Runtime::prepare_task(const Group& g, Eval& e, Task& t) {
  bool has_zonal = false;
  std::vector<Node*> kouts;
  KernelEntry* kep = nullptr;
  auto pre = e.kern.find(&g);
  if (pre != e.kern.end()) {
    kouts = pre->second.first;
    kep = pre->second.second;
    has_zonal = kep && !kep->spec.zouts.empty();
  } else {
    kouts = kernel_outputs(g, e.tsorted, has_zonal);
  }
  if (kouts.empty() && !has_zonal) {  // every output is held already: only inputs asked for are loaded
    for (Node* n : g.ins)
      if (n->meta.nd == D2 && n->kind == Kind::Read && e.tset.count(n)) ensure_input(n);
    return false;
  }
  for (Node* n : g.ins) {
    if (n->meta.nd == D2) ensure_input(n);
    else if (!n->evaluated) fail("D0 input of node " + std::to_string(n->id) + " is not evaluated");
  }

  const Meta gm = g.meta();
  for (Node* n : kouts) {
    if (!dim2_eq(n->meta.ds, gm.ds) || !dim2_eq(n->meta.bs, gm.bs))
      fail("group outputs differ in raster geometry");
    void* d = store.alloc(rasterBytes(n->meta));  // may move idle rasters out of HBM
    resident_[n] = d;
  }
  KernelEntry& ke = kep ? *kep : kernel_for(e.plan.get(), g, kouts);
  const KernelSpec& ks = ke.spec;
  if (ks.form == Form::Focal) check_focal_geometry(gm, ks.H0, ks.H1);
  t.g = &g;
  t.ke = &ke;
  t.args.assign(ks.args_size, 0);
  for (const Param& p : ks.params) {
    if (p.kind == Param::InD2 || p.kind == Param::OutD2) {
      // In<T>: the raster's base pointer (inputs are materialised rasters)
      void* d = resident_.at(p.node);
      std::memcpy(t.args.data() + p.off, &d, sizeof d);
    } else if (p.kind == Param::Mask) {
      // weights in mask order (row-major, y outer), with the %f quantisation of literals
      const Mask& mk = p.node->mask;
      const size_t es = (size_t)sizeof_dt(p.dt);
      for (int i = 0; i < (int)mk.v.size(); i++) {
        Variant w = variant_convert(quantize_literal(mk.v[i]), p.dt);
        std::memcpy(t.args.data() + p.off + i * es, &w.u, es);
      }
    } else {
      Variant v = variant_convert(p.node->value, p.dt);
      std::memcpy(t.args.data() + p.off, &v.u, (size_t)sizeof_dt(p.dt));
    }
  }
  const TileShape tile = ks.tile;
  const int BS0 = gm.bs.arr[0], BS1 = gm.bs.arr[1], NB0 = gm.nb(0), NB1 = gm.nb(1);
  const int TB0 = (BS0 + tile.TX - 1) / tile.TX, TB1 = (BS1 + tile.TY - 1) / tile.TY;
  t.tiles = TB0 * TB1;
  t.L = HostLaunch{BS0, BS1, NB0, NB1, gm.ds.arr[0], gm.ds.arr[1], 0, NB0 * NB1, TB0, TB1, 1, 1, 0};
  if (!ks.zouts.empty()) t.L.CH = std::max(1, std::min(t.tiles, kZonalChunkCells / (tile.TX * tile.TY)));
  return true;
}

// Device memory outside the store (task tables, the chunk queue): when the
// device is full, the store's cached buffers are returned, then resident
// rasters move out of HBM, as for the store's own allocations.
static void* alloc_outside_store(Store& st, size_t bytes) {
  for (;;) {
    try {
      return dev::malloc_device(bytes);
    } catch (const std::exception&) {
      if (st.cached()) { st.trim(); continue; }
      if (st.reclaim && st.reclaim()) continue;
      throw;
    }
  }
}

void* Runtime::device_table(const std::string& bytes) {
  auto it = tables_.find(bytes);
  if (it != tables_.end()) return it->second;
  if (table_bytes_ + bytes.size() > ((size_t)256 << 20)) {
    // tables are immutable while anything may read them: drop them all at once
    PauseRecording pause;
    dev::sync();
    dev::graph_clear();
    for (auto& e : tables_) dev::free_device(e.second);
    tables_.clear();
    table_bytes_ = 0;
  }
  void* d = alloc_outside_store(store, std::max<size_t>(bytes.size(), 16));
  auto ins = tables_.emplace(bytes, d).first;
  dev::h2d(d, ins->first.data(), bytes.size());  // the key keeps the source bytes alive
  table_bytes_ += bytes.size();
  return d;
}

void* Runtime::queue() {
  if (!queue_) {
    queue_ = alloc_outside_store(store, 256);
    static const unsigned zero[64] = {0};
    PauseRecording pause;
    dev::h2d(queue_, zero, sizeof zero);
    dev::sync();
  }
  return queue_;
}

void* Runtime::pinned_alloc(size_t bytes) {
  auto it = pinned_free_.lower_bound(bytes);
  if (it != pinned_free_.end() && it->first <= bytes * 2) {
    void* p = it->second;
    pinned_size_[p] = it->first;
    pinned_free_.erase(it);
    return p;
  }
  void* p = dev::malloc_host(bytes);
  pinned_size_[p] = bytes;
  return p;
}

void Runtime::pinned_release(void* p) {
  auto it = pinned_size_.find(p);
  if (it == pinned_size_.end()) return;
  pinned_free_.emplace(it->second, p);
}

void Runtime::distribute(Readback& rb) {
  for (size_t t = 0; t < rb.outs.size(); t++) {
    const unsigned char* rec = rb.host + t * rb.PB;
    for (const ZonalOut& z : rb.outs[t]) {
      Node* node = z.node;
      if (node->kind == Kind::ZonalReduc) {
        Variant a = variant_zero(z.acc_dt);
        std::memcpy(&a.u, rec + z.result_off, sizeof_dt(z.acc_dt));
        node->value = variant_convert(a, node->meta.dt);
      } else {
        node->table.assign(rec + z.result_off, rec + z.result_off + (size_t)z.nzones * 24);
      }
      node->evaluated = true;
    }
  }
  pinned_release(rb.host);
  rb.host = nullptr;
}

// Folds the partial records of the tasks of one launch (or one task) in index
// order and reads the zonal results back: at once, or (async != null) as a
// pending copy whose values are handed out once the stream has passed it.
void Runtime::fold_results(const std::vector<Task*>& batch, void* parts, void* tasks_dev, void* starts_dev, int maxn,
                           std::vector<Readback>* async) {
  const KernelEntry& ke = *batch[0]->ke;
  const KernelSpec& ks = ke.spec;
  const int T = (int)batch.size();
  const size_t PB = ks.part_bytes;
  dev::Function fin;
  fin.fn = ke.fin;
  fin.name = ks.fin_name;
  const int F = std::max(1, std::min(256, (maxn + 63) / 64));
  void* res = store.alloc(std::max<size_t>(1, (size_t)T * PB));
  void* mid = F > 1 ? store.alloc((size_t)T * F * PB) : nullptr;
  int n_single = T == 1 ? batch[0]->nchunks : 0;
  {
    void* dst = F > 1 ? mid : res;
    int stride = 0;
    void* a[7] = {&parts, &dst, &tasks_dev, &starts_dev, &n_single, &stride, (void*)&F};
    size_t z[7] = {8, 8, 8, 8, 4, 4, 4};
    dev::launch(fin, (unsigned)F, 256, a, z, 7, (unsigned)T);
    cnt.launches++;
  }
  if (F > 1) {
    void* nul = nullptr;
    int zero = 0, one = 1, stride = F;
    void* a[7] = {&mid, &res, &nul, &nul, &zero, &stride, &one};
    size_t z[7] = {8, 8, 8, 8, 4, 4, 4};
    dev::launch(fin, 1, 256, a, z, 7, (unsigned)T);
    cnt.launches++;
  }
  Readback rb;
  rb.PB = PB;
  for (Task* t : batch) rb.outs.push_back(t->ke->spec.zouts);  // each task's own nodes
  rb.host = (unsigned char*)pinned_alloc((size_t)T * PB);
  cnt.bytes_d2h += (long)((size_t)T * PB);
  if (async) {
    // the copy follows the fold launches, which are enqueued first when recorded
    if (dev::recording()) {
      PauseRecording pause;
      dev::d2h(rb.host, res, (size_t)T * PB);
    } else {
      dev::d2h(rb.host, res, (size_t)T * PB);
    }
    async->push_back(std::move(rb));
  } else {
    PauseRecording pause;
    dev::d2h(rb.host, res, (size_t)T * PB);
    dev::sync();
    distribute(rb);
  }
  if (mid) store.release(mid);
  store.release(res);
}

// Launches a batch of tasks of one kernel: per block (plain form), one task by
// value, or many tasks from a device task table with a device chunk queue.
void Runtime::run_batch(const std::vector<Task*>& batch, std::vector<Readback>* async) {
  HostTimer ht_launch(cnt.ms_launch);
  const KernelEntry& ke = *batch[0]->ke;
  const KernelSpec& ks = ke.spec;
  const bool zonal = !ks.zouts.empty();
  dev::Function fn;
  fn.fn = ke.fn;
  fn.name = ks.name;
  // a raster whose last block column or row is narrower than the halo needs the
  // stepwise border rule: the batch runs the module compiled for it (it reads
  // the other rasters of the batch alike)
  if (ks.H0 > 0 || ks.H1 > 0) {
    bool narrow = false;
    for (const Task* t : batch) {
      const HostLaunch& L = t->L;
      narrow |= L.DS0 - (L.NB0 - 1) * L.BS0 < ks.H0 || L.DS1 - (L.NB1 - 1) * L.BS1 < ks.H1;
    }
    if (narrow) {
      HostTimer ht(cnt.ms_codegen);
      fn.fn = dev::function("#define MX_NARROW 1\n" + ks.source, ks.name).fn;
      cnt.narrow_launches++;
    }
  }
  // persistent grid: every work-group resident at once (no tail wave), a
  // multiple of 8 for the XCD-contiguous split
  // work-groups per CU: the occupancy query's answer, or the option's value
  // (more than fit at once simply queue behind the resident ones)
  const int per_cu = opt.wg_per_cu > 0 ? opt.wg_per_cu : ke.occ;
  const int gmax = std::max(8, (dev::num_cus() * per_cu) & ~7);
  auto grid_for = [&](long work) { return (unsigned)((std::min<long>(work, gmax) + 7) & ~7L); };
  void* nul = nullptr;
  auto launch = [&](std::vector<char>& args, HostLaunch& L, void* tasks, void* starts, int ntasks, int nchunks,
                    void* queue, void* parts, unsigned grid) {
    void* a[8] = {args.data(), &L, &tasks, &starts, &ntasks, &nchunks, &queue, &parts};
    size_t z[8] = {args.size(), sizeof L, 8, 8, 4, 4, 8, 8};
    dev::launch(fn, grid, 256, a, z, 8);
    cnt.launches++;
  };

  if (opt.form == FORM_PLAIN) {
    for (Task* t : batch) {
      t->chunking(opt.chunk_tiles > 0 ? opt.chunk_tiles : 1);  // one block per launch: every tile its own chunk
      void* parts = zonal ? store.alloc((size_t)t->nchunks * ks.part_bytes) : nullptr;
      std::vector<std::pair<uint32_t, int>> order;
      for (int b = 0; b < t->L.nblk; b++) order.push_back({morton2(b % t->L.NB0, b / t->L.NB0), b});
      std::sort(order.begin(), order.end());
      for (auto& o : order) {
        HostLaunch L = t->L;
        L.blk0 = o.second;
        L.nblk = 1;
        L.part0 = o.second * L.CPB;
        launch(t->args, L, nul, nul, 1, L.CPB, nul, parts, grid_for(L.CPB));
      }
      if (zonal) {
        fold_results({t}, parts, nullptr, nullptr, t->nchunks, async);
        store.release(parts);
      }
    }
    return;
  }
  if (batch.size() == 1) {
    Task* t = batch[0];
    const bool q = opt.queue == 1;
    t->chunking(opt.chunk_tiles > 0 ? opt.chunk_tiles : adaptive_chunk((long)t->tiles * t->L.nblk, gmax));
    void* parts = zonal ? store.alloc((size_t)t->nchunks * ks.part_bytes) : nullptr;
    launch(t->args, t->L, nul, nul, 1, t->nchunks, q ? queue() : nul, parts, grid_for(t->nchunks));
    if (zonal) {
      fold_results(batch, parts, nullptr, nullptr, t->nchunks, async);
      store.release(parts);
    }
    return;
  }
  // task table: TaskRec per task, then starts[ntasks + 1]
  const bool q = opt.queue != 2;
  const int T = (int)batch.size();
  std::vector<int> starts(T + 1, 0);
  int maxn = 0;
  long tiles_total = 0;
  for (Task* t : batch) tiles_total += (long)t->tiles * t->L.nblk;
  const int ch = opt.chunk_tiles > 0 ? opt.chunk_tiles : adaptive_chunk(tiles_total, gmax);
  for (int i = 0; i < T; i++) {
    batch[i]->chunking(ch);
    starts[i + 1] = starts[i] + batch[i]->nchunks;
    maxn = std::max(maxn, batch[i]->nchunks);
  }
  const int total = starts[T];
  std::string tab((size_t)T * ks.task_size + (size_t)(T + 1) * sizeof(int), '\0');
  for (int i = 0; i < T; i++) {
    HostLaunch L = batch[i]->L;
    L.part0 = starts[i];
    char* rec = &tab[(size_t)i * ks.task_size];
    std::memcpy(rec, batch[i]->args.data(), ks.args_size);
    std::memcpy(rec + ks.task_launch_off, &L, sizeof L);
  }
  std::memcpy(&tab[(size_t)T * ks.task_size], starts.data(), starts.size() * sizeof(int));
  char* dtab = (char*)device_table(tab);
  void* tasks_dev = dtab;
  void* starts_dev = dtab + (size_t)T * ks.task_size;
  void* parts = zonal ? store.alloc((size_t)total * ks.part_bytes) : nullptr;
  HostLaunch L0{};
  std::vector<char> a0(ks.args_size, 0);
  launch(a0, L0, tasks_dev, starts_dev, T, total, q ? queue() : nul, parts, grid_for(total));
  if (zonal) {
    fold_results(batch, parts, tasks_dev, starts_dev, maxn, async);
    store.release(parts);
  }
}

// Runs one level of each evaluation: host scalar groups at once, kernels of
// all evaluations batched by kernel.
void Runtime::run_levels(const std::vector<std::pair<Eval*, size_t>>& work, std::vector<Readback>* async) {
  size_t n = 0;
  for (auto& w : work) n += w.first->levels[w.second].size();
  std::vector<Task> tasks(n);
  // the groups in order; unrecorded, a long list goes out in sub-batches of
  // kSubBatch groups, each launched once prepared
  std::vector<std::pair<Eval*, const Group*>> groups;
  groups.reserve(n);
  for (auto& w : work)
    for (const Group* gp : w.first->levels[w.second]) groups.push_back({w.first, gp});
  const bool whole = dev::recording() || opt.form == FORM_PLAIN;
  size_t i = 0;
  for (size_t g0 = 0, g1 = 0; g0 < groups.size(); g0 = g1) {
    // a short first sub-batch starts the device early, the rest go kSubBatch at a time
    const size_t sub = whole ? groups.size() : (g0 == 0 ? (size_t)kFirstSubBatch : (size_t)kSubBatch);
    g1 = std::min(groups.size(), g0 + sub);
    std::vector<Task*> order;
    {
      HostTimer ht_prep(cnt.ms_prepare);
      for (size_t k = g0; k < g1; k++) {
        const Group& g = *groups[k].second;
        Eval& e = *groups[k].first;
        cnt.groups++;
        Task& t = tasks[i++];
        if (host_group(g)) { scalar_task(g); continue; }
        if (radial_node(g)) { radial_task(g, e); continue; }
        if (spread_node(g)) { spread_task(g, e); continue; }
        if (prepare_task(g, e, t)) order.push_back(&t);
      }
    }
    // tasks of one kernel form one batch (in order of first appearance)
    std::vector<std::vector<Task*>> batches;
    std::unordered_map<std::string, size_t> at;
    for (Task* t : order) {
      const bool one = opt.form == FORM_PLAIN || !opt.batch_tasks;
      auto it = one ? at.end() : at.find(t->ke->spec.name);
      if (it == at.end()) {
        if (!one) at[t->ke->spec.name] = batches.size();
        batches.push_back({t});
      } else {
        batches[it->second].push_back(t);
      }
    }
    for (auto& b : batches) run_batch(b, async);
  }
  for (auto& w : work)
    for (const Group* g : w.first->levels[w.second])
      for (Node* n : g->outs)
        if (n->kind == Kind::Write) write_file(n);
}

std::string Runtime::plan_key(const std::vector<Node*>& list, const std::vector<Node*>& targets) const {
  std::string k;
  k.reserve(list.size() * 9 + targets.size() * 8 + 16);
  auto put = [&](int v) { k.append((const char*)&v, sizeof v); };
  put(opt.code_fusion); put(opt.focal_zonal); put(opt.tile.TX); put(opt.tile.TY); put(opt.tile.VX); put(opt.tile.PF); put(opt.tile.ST); put(opt.tile.W); put(opt.libm);
  for (Node* n : list) { put(n->id); k.push_back(n->cut ? 'c' : (n->isOutput() ? 'o' : 'n')); }
  k.push_back('|');
  for (Node* t : targets) put(t->id);
  return k;
}

Runtime::KernelEntry& Runtime::kernel_for(Plan* plan, const Group& g, const std::vector<Node*>& kouts) {
  auto key = std::make_pair(&g, kouts);
  auto it = plan->kernels.find(key);
  if (it != plan->kernels.end()) return it->second;
  HostTimer ht(cnt.ms_codegen);
  const TileShape tile = tile_for(g);
  const std::string skey = kernel_shape_key(g, kouts, tile, opt.libm);
  const std::vector<Node*> order = kernel_order(g);
  auto sh = shapes_.find(skey);
  if (sh != shapes_.end()) {
    KernelEntry e = sh->second.e;
    for (size_t i = 0; i < e.spec.params.size(); i++) e.spec.params[i].node = order.at(sh->second.param_node[i]);
    for (size_t i = 0; i < e.spec.d2_outs.size(); i++) e.spec.d2_outs[i] = order.at(sh->second.d2_out[i]);
    for (size_t i = 0; i < e.spec.zouts.size(); i++) e.spec.zouts[i].node = order.at(sh->second.zout[i]);
    cnt.shape_hits++;
    return plan->kernels.emplace(key, std::move(e)).first->second;
  }
  KernelEntry e;
  e.spec = generate_kernel(g, kouts, tile, opt.libm);
  kernel_group_[e.spec.name] = tune_key(g);
  kernel_src_[e.spec.name] = e.spec.source;
  dev::Function f = dev::function(e.spec.source, e.spec.name);
  e.fn = f.fn;
  e.occ = dev::occupancy(f, 256);
  kernel_occ_[e.spec.name] = e.occ;
  if (!e.spec.fin_name.empty()) e.fin = dev::function(e.spec.source, e.spec.fin_name).fn;
  cnt.compiles++;
  {
    std::unordered_map<const Node*, int> local;
    for (size_t i = 0; i < order.size(); i++) local[order[i]] = (int)i;
    ShapeEntry s;
    bool ok = true;
    auto idx = [&](const Node* n) { auto it = local.find(n); if (it == local.end()) { ok = false; return -1; } return it->second; };
    for (auto& p : e.spec.params) s.param_node.push_back(idx(p.node));
    for (Node* n : e.spec.d2_outs) s.d2_out.push_back(idx(n));
    for (auto& z : e.spec.zouts) s.zout.push_back(idx(z.node));
    if (ok) {
      if (shapes_.size() >= 4096) shapes_.clear();  // bounded
      s.e = e;
      shapes_.emplace(skey, std::move(s));
    }
  }
  return plan->kernels.emplace(key, std::move(e)).first->second;
}

// Plans an evaluation: the sub-DAG cut at resident rasters, its fusion
// groups (cached by evaluation list), their levels and raster uses. Leaves
// the cut marks set (the caller clears them with uncut()).
void Runtime::eval_plan(Eval& e) {
  HostTimer ht_plan(cnt.ms_plan);
  e.list = cut_list(e.targets);
  FuseOptions fo;
  fo.code_fusion = opt.code_fusion;
  fo.focal_zonal = opt.focal_zonal;
  const std::string pkey = plan_key(e.list, e.targets);
  auto pit = plans_.find(pkey);
  if (pit == plans_.end()) {
    if (plans_.size() >= 64) plans_.clear();  // bounded: plans of old scripts are dropped (live evaluations keep theirs)
    auto pl = std::make_shared<Plan>();
    {
      HostTimer ht(cnt.ms_fuse);
      pl->groups = Fuser(fo).fuse(e.list);
    }
    pit = plans_.emplace(pkey, std::move(pl)).first;
  }
  e.plan = pit->second;
  const auto& groups = e.plan->groups;
  // consumers of each D2 raster in this evaluation, for early release
  for (auto& g : groups)
    for (Node* n : g->ins) if (n->meta.nd == D2) e.uses[n]++;
  e.tset.insert(e.targets.begin(), e.targets.end());
  for (Node* t : e.targets) if (t->kind == Kind::Write) e.tset.insert(t->prev[0]);
  e.tsorted = e.targets;
  std::sort(e.tsorted.begin(), e.tsorted.end());
  // levels: a group runs one level after the last group it depends on; the
  // groups of one level are independent, and those sharing a kernel are batched
  std::unordered_map<const Group*, int> lvl;
  for (auto& g : groups) {
    int l = 0;
    for (Group* p : g->prev) {
      auto it = lvl.find(p);
      if (it != lvl.end()) l = std::max(l, it->second + 1);
    }
    lvl[g.get()] = l;
    if ((int)e.levels.size() <= l) e.levels.resize(l + 1);
    e.levels[l].push_back(g.get());
  }
}

void Runtime::eval_uncut(Eval& e) {
  for (Node* n : e.list) n->cut = false;
}

// Inputs of a level's groups that nothing later in the evaluation reads go back to the store.
void Runtime::eval_release(Eval& e, size_t level) {
  for (const Group* g : e.levels[level])
    for (Node* n : g->ins) {
      if (n->meta.nd != D2) continue;
      if (--e.uses[n] == 0 && ext_refs(n) <= 0 && !e.tset.count(n) && n->kind != Kind::Read) drop(n);
    }
}

// Intermediates nobody holds and nothing consumes any more.
void Runtime::eval_finish(Eval& e) {
  for (auto& g : e.plan->groups)
    for (Node* n : g->outs)
      if (n->meta.nd == D2 && n->kind != Kind::Write && ext_refs(n) <= 0 && !e.tset.count(n) && e.uses[n] == 0) drop(n);
}

void Runtime::evaluate(const std::vector<Node*>& targets) {
  std::lock_guard<std::recursive_mutex> lk(mtx_);
  if (!dev::initialized()) fail("evaluate: devices not set up (call setupDevices first)");
  if (targets.empty()) fail("evaluate: nothing to evaluate");
  if (serving()) fail("evaluate: a server is open (submit requests to it, or close it first)");
  {
    HostTimer ht(cnt.ms_collect);
    release_unreferenced();
    dag.collect_grown([this](Node* n) { forget(n); });
  }
  cnt.evals++;
  dev::Timer t;
  t.start();

  Eval e;
  {
    HostTimer ht(cnt.ms_predict);
    e.targets = with_aux(targets);
    predict(e.targets);
  }
  {
    // known values and rasters already held (in HBM or spilled) are not evaluated again
    std::vector<Node*> left;
    for (Node* n : e.targets) {
      if (n->meta.nd == D0 && n->evaluated && n->kind != Kind::Constant) continue;
      if (n->meta.nd == D2 && n->kind != Kind::Read && (resident_.count(n) || spilled_.count(n))) continue;
      left.push_back(n);
    }
    if (left.empty()) {  // every target is a known value or a held raster
      t.stop();
      cnt.ms_last_eval = t.ms();
      return;
    }
    e.targets = left;
  }
  struct Uncut {
    Runtime* r; Eval& e;
    ~Uncut() { r->eval_uncut(e); }
  } uncut{this, e};
  eval_plan(e);
  // the evaluation's rasters stay in HBM while it runs (spill_one passes them over)
  struct Live {
    Runtime* r; unsigned long long t;
    ~Live() { r->tick_end(t); }
  } live{this, tick_begin(e.list)};

  // zonal results come back as pending copies, handed out before a level that
  // reads one of them and at the end (the host does not wait for each fold)
  std::vector<Readback> pend;
  auto complete = [&] {
    if (pend.empty()) return;
    PauseRecording pause;
    dev::sync();
    for (Readback& rb : pend) distribute(rb);
    pend.clear();
  };
  if (opt.graph) dev::record_begin();
  try {
    for (size_t l = 0; l < e.levels.size(); l++) {
      bool reads_pending = false;
      for (const Group* g : e.levels[l])
        for (Node* n : g->ins)
          if (n->meta.nd == D0 && !n->evaluated) reads_pending = true;
      if (reads_pending) complete();
      // a level of many tasks runs unrecorded, in sub-batches launched as they
      // are prepared (the device starts while the host prepares the rest);
      // chains of small levels (the generations of Life) accumulate into one graph
      const bool big = e.levels[l].size() >= (size_t)kBigLevel;
      const bool rec = dev::recording();
      if (big && rec) dev::record_flush();
      run_levels({{&e, l}}, &pend);
      if (big && rec) dev::record_begin();
      eval_release(e, l);
    }
    if (dev::recording()) dev::record_flush();
    complete();
  } catch (...) {
    if (dev::recording()) { try { dev::record_flush(); } catch (...) {} }
    for (Readback& rb : pend) pinned_release(rb.host);  // copies into them are ordered before any reuse
    // rasters this evaluation was computing may be partial: none is kept
    for (Node* n : e.list)
      if (!n->cut && n->kind != Kind::Read && n->meta.nd == D2) forget(n);
    throw;
  }
  eval_finish(e);
  t.stop();
  cnt.ms_last_eval = t.ms();
}

std::string Runtime::plan(const std::vector<Node*>& targets_in) {
  std::lock_guard<std::recursive_mutex> lk(mtx_);
  const std::vector<Node*> targets = with_aux(targets_in);  // as an evaluation would run them
  std::vector<Node*> tsorted = targets;
  std::sort(tsorted.begin(), tsorted.end());
  std::vector<Node*> list = cut_list(targets);
  struct Uncut {
    std::vector<Node*>& l;
    ~Uncut() { for (Node* n : l) n->cut = false; }
  } uncut{list};
  FuseOptions fo;
  fo.code_fusion = opt.code_fusion;
  fo.focal_zonal = opt.focal_zonal;
  std::vector<std::unique_ptr<Group>> groups;
  {
    HostTimer ht(cnt.ms_fuse);
    groups = Fuser(fo).fuse(list);
  }
  auto esc = [](const std::string& s) {
    std::string o;
    for (char c : s) {
      if (c == '"' || c == '\\') { o += '\\'; o += c; }
      else if (c == '\n') o += "\\n";
      else if (c == '\t') o += "\\t";
      else o += c;
    }
    return o;
  };
  std::ostringstream j;
  j << "{\"groups\": [";
  bool fg = true;
  for (auto& g : groups) {
    if (!fg) j << ", ";
    fg = false;
    j << "{\"id\": " << g->id << ", \"pattern\": " << g->pat << ", \"numdim\": " << g->numdim();
    auto ids = [&](const char* k, const std::vector<Node*>& v) {
      j << ", \"" << k << "\": [";
      for (size_t i = 0; i < v.size(); i++) j << (i ? ", " : "") << v[i]->id;
      j << "]";
    };
    ids("ins", g->ins);
    ids("nodes", g->nodes);
    ids("outs", g->outs);
    j << ", \"kinds\": [";
    for (size_t i = 0; i < g->nodes.size(); i++) j << (i ? ", " : "") << "\"" << kind_name(g->nodes[i]->kind) << "\"";
    j << "]";
    if (Node* rn = radial_node(*g)) {  // precompiled band kernel (radial.hip), no generated source
      j << ", \"radial\": {\"reduction\": " << rn->op << ", \"start\": [" << rn->coord.arr[0] << ", " << rn->coord.arr[1]
        << "]}}";
      continue;
    }
    if (Node* sn = spread_node(*g)) {  // precompiled walk kernels (spread.hip)
      j << ", \"spread\": {\"reduction\": " << sn->op << "}}";
      continue;
    }
    if (!host_group(*g)) {
      std::vector<Node*> outs = d2_outputs(*g, tsorted), kouts;
      for (Node* n : outs)
        if (std::find(g->nodes.begin(), g->nodes.end(), n) != g->nodes.end()) kouts.push_back(n);
      bool has_zonal = false;
      for (Node* n : g->nodes) if (n->kind == Kind::ZonalReduc || n->kind == Kind::ZonalStats) has_zonal = true;
      if (!kouts.empty() || has_zonal) {
        KernelSpec ks;
        try {
          HostTimer ht(cnt.ms_codegen);
          ks = generate_kernel(*g, kouts, tile_for(*g), opt.libm);
        } catch (const std::exception& e) {
          // the grouping is still reported; evaluation would fail with this error
          j << ", \"error\": \"" << esc(e.what()) << "\"}";
          continue;
        }
        j << ", \"kernel\": \"" << ks.name << "\", \"form\": " << (int)ks.form << ", \"halo\": [" << ks.H0 << ", " << ks.H1
          << "], \"stream\": " << (ks.stream ? 1 : 0) << ", \"lane_r\": " << ks.lane_r << ", \"tile\": [" << ks.tile.TX << ", " << ks.tile.TY << ", " << ks.tile.VX << ", " << ks.tile.PF
          << "], \"tune_key\": \"" << esc(tune_key(*g)) << "\", \"shape_key\": \""
          << esc(kernel_shape_key(*g, kouts, tile_for(*g), opt.libm)) << "\", \"source\": \"" << esc(ks.source) << "\"";
        // floating mask weights as the kernel receives them (quantised like printed literals)
        j << ", \"mask_args\": [";
        bool fm = true;
        for (const Param& p : ks.params) {
          if (p.kind != Param::Mask) continue;
          j << (fm ? "[" : ", [");
          fm = false;
          for (size_t i = 0; i < p.node->mask.v.size(); i++) {
            char b[40];
            std::snprintf(b, sizeof b, "%.17g", variant_as<double>(variant_convert(quantize_literal(p.node->mask.v[i]), p.dt)));
            j << (i ? ", " : "") << b;
          }
          j << "]";
        }
        j << "]";
      }
    }
    j << "}";
  }
  j << "]}";
  return j.str();
}

std::string Runtime::autotune(const std::vector<Node*>& targets, int reps) {
  std::lock_guard<std::recursive_mutex> lk(mtx_);
  if (reps < 1) reps = 1;
  // {TX, TY, VX, PF, ST}: register-window (ST), LDS-staged, LDS-staged with prefetch (PF)
  static const TileShape cands[] = {
      {256, 16, 4, 0, 1}, {256, 32, 4, 0, 1}, {512, 16, 8, 0, 1}, {256, 32, 8, 0, 1}, {128, 32, 4, 0, 1},
      {64, 32, 1, 0, 1},
      {256, 16, 4, 0, 0}, {256, 32, 4, 0, 0}, {128, 32, 4, 0, 0}, {512, 16, 4, 0, 0}, {512, 16, 8, 0, 0},
      {256, 32, 8, 0, 0}, {512, 32, 8, 0, 0}, {256, 64, 8, 0, 0}, {1024, 16, 8, 0, 0}, {256, 16, 2, 0, 0},
      {64, 32, 1, 0, 0},
      {256, 32, 4, 1, 0}, {512, 16, 4, 1, 0}, {512, 32, 8, 1, 0}, {1024, 16, 8, 1, 0}, {256, 64, 8, 1, 0},
      {256, 16, 2, 1, 0}};
  const TileShape saved = opt.tile;
  const bool saved_graph = opt.graph, saved_prof = dev::profile_on();
  opt.graph = false;
  // Each candidate is timed by its kernel's time in one evaluation (the sum of
  // its launches), the least over the repetitions (the least disturbed). Pass 1
  // times every shape; pass 2 times again, side by side, the three fastest
  // shapes of each computation and the fastest with its registers capped for 6
  // or 8 wavefronts per SIMD (more work-groups per CU; a kernel that would spill
  // loses), and the choice comes from pass 2 alone.
  std::map<std::string, std::vector<std::pair<double, TileShape>>> times1;  // per tune key, pass 1
  std::map<std::string, std::pair<double, TileShape>> best;                  // per tune key, pass 2
  std::ostringstream rep;
  rep << "{\"candidates\": [";
  bool first = true;
  std::vector<TileShape> pass(std::begin(cands), std::end(cands));
  try {
    for (int stage = 0; stage < 2; stage++) {
      if (stage == 1) {
        pass.clear();
        auto add = [&](const TileShape& c) {
          if (std::find(pass.begin(), pass.end(), c) == pass.end()) pass.push_back(c);
        };
        for (auto& kv : times1) {
          auto v = kv.second;
          std::stable_sort(v.begin(), v.end(), [](const std::pair<double, TileShape>& a,
                                                  const std::pair<double, TileShape>& b) { return a.first < b.first; });
          for (size_t i = 0; i < v.size() && i < 3; i++) add(v[i].second);
          for (int w : {6, 8}) {
            TileShape c = v[0].second;
            c.W = w;
            add(c);
          }
        }
      }
      for (const TileShape& c : pass) {
        opt.tile = c;
        for (Node* t : targets) drop(t);
        evaluate(targets);  // compile and warm
        dev::profile_set(true);
        dev::profile_take();
        // per kernel: its time in one evaluation (all its launches), the least over the repetitions
        std::map<std::string, std::pair<double, double>> per;  // name -> (least evaluation total, sum)
        std::map<std::string, long> launches;
        for (int r = 0; r < reps; r++) {
          for (Node* t : targets) drop(t);
          evaluate(targets);
          for (auto& k : dev::profile_take()) {
            auto it = per.find(k.name);
            if (it == per.end()) per[k.name] = {k.total_ms, k.total_ms};
            else { it->second.first = std::min(it->second.first, k.total_ms); it->second.second += k.total_ms; }
            launches[k.name] += k.count;
          }
        }
        dev::profile_set(false);
        for (auto& kv : per) {
          struct { std::string name; long count; double total_ms; } k{kv.first, launches[kv.first], kv.second.second};
          auto kg = kernel_group_.find(k.name);
          if (kg == kernel_group_.end()) continue;
          const double t = kv.second.first;
          if (stage == 0) {
            times1[kg->second].push_back({t, c});
          } else {
            auto b = best.find(kg->second);
            if (b == best.end() || t < b->second.first) best[kg->second] = {t, c};
          }
          rep << (first ? "" : ", ") << "{\"pass\": " << stage + 1 << ", \"tile\": [" << c.TX << ", " << c.TY << ", " << c.VX
              << ", " << c.PF << ", " << c.ST << ", " << c.W << "], \"kernel\": \"" << k.name << "\", \"eval_ms\": " << t
              << ", \"avg_ms\": " << k.total_ms / std::max(1L, k.count) << "}";
          first = false;
        }
      }
    }
  } catch (...) {
    opt.tile = saved;
    opt.graph = saved_graph;
    dev::profile_set(saved_prof);
    throw;
  }
  opt.tile = saved;
  opt.graph = saved_graph;
  dev::profile_set(saved_prof);
  rep << "], \"chosen\": [";
  first = true;
  for (auto& b : best) {
    tuned_[b.first] = b.second.second;
    rep << (first ? "" : ", ") << "{\"eval_ms\": " << b.second.first << ", \"tile\": [" << b.second.second.TX << ", "
        << b.second.second.TY << ", " << b.second.second.VX << ", " << b.second.second.PF << ", " << b.second.second.ST << ", "
        << b.second.second.W << "]}";
    first = false;
  }
  rep << "]}";
  for (Node* t : targets) drop(t);
  // cached plans hold kernels generated with the previous tile shapes
  plans_.clear();
  if (const char* p = std::getenv("MAPX_TUNE_CACHE")) tune_save(p);
  return rep.str();
}

// The tuning cache file: one line per entry, "TX TY VX PF ST W <tune key>".
void Runtime::tune_load(const std::string& path) {
  std::ifstream f(path);
  if (!f) return;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream l(line);
    TileShape t;
    std::string key;
    if (!(l >> t.TX >> t.TY >> t.VX >> t.PF >> t.ST >> t.W)) continue;
    std::getline(l, key);
    if (!key.empty() && key[0] == ' ') key.erase(0, 1);
    if (!key.empty() && t.TX > 0 && t.VX > 0 && t.TY > 0) tuned_[key] = t;
  }
}

void Runtime::tune_save(const std::string& path) const {
  std::ofstream f(path, std::ios::trunc);
  for (auto& e : tuned_)
    f << e.second.TX << " " << e.second.TY << " " << e.second.VX << " " << e.second.PF << " " << e.second.ST << " "
      << e.second.W << " " << e.first << "\n";
}

}  // namespace mapx
