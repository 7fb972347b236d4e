// extern "C" ma_* entry points (front/bindings.cpp:16-275 of the reference):
// node factories over the DAG, evaluation through the runtime, and the
// extension entry points declared in include/mapx.h.
#include "dev.hpp"
#include "runtime.hpp"
#include "rasterio.hpp"

namespace mapx { struct Node; }
#define MAPX_NODE_TYPE mapx::Node
#include "../../include/mapx.h"

#include <cstring>
#include <sstream>
#include <string>

namespace mapx {
namespace gen {
size_t dem_scratch_bytes(int ng);
size_t nalloc_scratch_bytes();
void* nalloc_scratch(int device);
size_t voronoi_scratch_bytes(int np);
void dem(void* out, int BS0, int BS1, int NB0, int NB1, int DS0, int DS1, const double* params, int ng,
         uint64_t seed, void* scratch, void* stream);
void voronoi(void* out, int dt_bytes, bool is_signed, int BS0, int BS1, int NB0, int NB1, int DS0, int DS1,
             const int* px, const int* py, int np, void* scratch, void* stream);
}  // namespace gen
namespace mathcheck {
void run(const std::string& name, uint64_t lo, uint64_t hi, void* scratch, void* stream, double* max_ulp,
         unsigned long long* n_over, unsigned long long* n, unsigned* worst_in);
}  // namespace mathcheck
}  // namespace mapx

using namespace mapx;

static_assert(sizeof(ma_variant_t) == sizeof(Variant), "variant layout");
static_assert(sizeof(ma_dim2_t) == sizeof(Dim2), "dim2 layout");

namespace {
thread_local std::string g_err;
thread_local std::string g_json;

Variant V(const ma_variant_t& v) { Variant r; std::memcpy(&r, &v, sizeof r); return r; }
ma_variant_t MV(const Variant& v) { ma_variant_t r; std::memcpy(&r, &v, sizeof r); return r; }
Dim2 D(const ma_dim2_t& d) { Dim2 r; std::memcpy(&r, &d, sizeof r); return r; }
ma_dim2_t MD(const Dim2& d) { ma_dim2_t r; std::memcpy(&r, &d, sizeof r); return r; }

template <class F> auto guard(F f, decltype(f()) bad) -> decltype(f()) {
  try {
    g_err.clear();
    return f();
  } catch (const std::exception& e) {
    g_err = e.what();
  } catch (...) {
    g_err = "unknown error";
  }
  return bad;
}
Runtime& R() { return Runtime::get(); }
Node* need(Node* n) {
  if (!n) throw std::runtime_error("null node argument");
  return n;
}
}  // namespace

extern "C" {

const char* ma_lastError(void) { return g_err.c_str(); }

void ma_setupDevices(const char* plat_name, ma_devicetype_t dev, const char* dev_name) {
  guard([&] {
    R().setupDevices(plat_name ? plat_name : "", (unsigned)dev, dev_name ? dev_name : "");
    return 0;
  }, 0);
}

void ma_increaseRef(ma_node_t* node) { if (node) node->ref++; }
void ma_decreaseRef(ma_node_t* node) { if (node && node->ref > 0) node->ref--; }

void ma_eval(ma_node_t** vec, int num) {
  guard([&] {
    if (num <= 0 || !vec) throw std::runtime_error("eval: no nodes");
    std::vector<Node*> v;
    for (int i = 0; i < num; i++) v.push_back(need(vec[i]));
    R().evaluate(v);
    return 0;
  }, 0);
}

ma_variant_t ma_value(ma_node_t* node) {
  Variant bad = variant_zero(NONE_DATATYPE);
  return MV(guard([&] {
    need(node);
    if (node->meta.nd != D0) throw std::runtime_error("value: node is not D0");
    if (!node->evaluated) throw std::runtime_error("value: node has not been evaluated");
    return node->value;
  }, bad));
}

int ma_nodeid(ma_node_t* node) { return node ? node->id : -1; }
ma_datatype_t ma_datatype(ma_node_t* node) { return node ? node->meta.dt : 0; }
ma_numdim_t ma_numdim(ma_node_t* node) { return node ? node->meta.nd : 0; }
ma_memorder_t ma_memorder(ma_node_t* node) { return node ? node->meta.mo : 0; }
ma_dim2_t ma_datasize(ma_node_t* node) { return MD(node ? node->meta.ds : dim2_none()); }
ma_dim2_t ma_blocksize(ma_node_t* node) { return MD(node ? node->meta.bs : dim2_none()); }

ma_node_t* ma_read(const char* file_path) {
  return guard([&] { return R().dag.read(file_path ? file_path : ""); }, (Node*)nullptr);
}

int ma_write(ma_node_t* node, const char* file_path) {
  return guard([&] {
    Node* w = R().dag.write(need(node), file_path ? file_path : "");
    w->ref++;
    try {
      R().evaluate({w});
    } catch (...) {
      w->ref--;
      throw;
    }
    w->ref--;
    return 0;
  }, -1);
}

ma_node_t* ma_constant(ma_variant_t var, ma_dim2_t ds, ma_datatype_t dt, ma_memorder_t mo, ma_dim2_t bs) {
  return guard([&] { return R().dag.constant(V(var), D(ds), dt, mo, D(bs)); }, (Node*)nullptr);
}
ma_node_t* ma_cast(ma_node_t* node, ma_datatype_t type) {
  return guard([&] { return R().dag.cast(need(node), type); }, (Node*)nullptr);
}
ma_node_t* ma_rand(ma_node_t* seed, ma_datatype_t dt, ma_memorder_t mo) {
  return guard([&] { return R().dag.rand(need(seed), dt, mo); }, (Node*)nullptr);
}
ma_node_t* ma_index(ma_node_t* node, ma_numdim_t dim) {
  return guard([&] { return R().dag.index(need(node), dim); }, (Node*)nullptr);
}
ma_node_t* ma_conditional(ma_node_t* cond, ma_node_t* lhs, ma_node_t* rhs) {
  return guard([&] { return R().dag.conditional(need(cond), need(lhs), need(rhs)); }, (Node*)nullptr);
}
ma_node_t* ma_unary(ma_node_t* node, ma_unary_t type) {
  return guard([&] { return R().dag.unary(need(node), type); }, (Node*)nullptr);
}
ma_node_t* ma_binary(ma_node_t* lhs, ma_node_t* rhs, ma_binary_t type) {
  return guard([&] { return R().dag.binary(need(lhs), need(rhs), type); }, (Node*)nullptr);
}
ma_node_t* ma_neighbor(ma_node_t* node, ma_dim2_t coord) {
  return guard([&] { return R().dag.neighbor(need(node), D(coord)); }, (Node*)nullptr);
}

static Mask make_mask(ma_variant_t* v, int s, ma_dim2_t ds) {
  if (!v || s <= 0) throw std::runtime_error("mask: no values");
  Mask m;
  m.ds = D(ds);
  for (int i = 0; i < s; i++) m.v.push_back(V(v[i]));
  return m;
}
ma_node_t* ma_convolution(ma_node_t* node, ma_variant_t* mask_v, int mask_s, ma_dim2_t mask_ds) {
  return guard([&] { return R().dag.convolution(need(node), make_mask(mask_v, mask_s, mask_ds)); }, (Node*)nullptr);
}
ma_node_t* ma_focalFunc(ma_node_t* node, ma_variant_t* mask_v, int mask_s, ma_dim2_t mask_ds, ma_reduction_t type) {
  return guard([&] { return R().dag.focalFunc(need(node), make_mask(mask_v, mask_s, mask_ds), type); }, (Node*)nullptr);
}
ma_node_t* ma_focalPercent(ma_node_t* node, ma_variant_t* mask_v, int mask_s, ma_dim2_t mask_ds, ma_percent_t type) {
  return guard([&] { return R().dag.focalPercent(need(node), make_mask(mask_v, mask_s, mask_ds), type); },
               (Node*)nullptr);
}
ma_node_t* ma_diversity(ma_node_t** vec, int size, ma_diversity_t type) {
  return guard([&] {
    if (!vec || size < 0) throw std::runtime_error("diversity: bad node list");
    std::vector<Node*> ps;
    for (int i = 0; i < size; i++) ps.push_back(need(vec[i]));
    return R().dag.diversity(ps, type);
  }, (Node*)nullptr);
}
ma_node_t* ma_radialScan(ma_node_t* node, ma_reduction_t type, ma_dim2_t coord) {
  return guard([&] { return R().dag.radialScan(need(node), type, D(coord)); }, (Node*)nullptr);
}

ma_node_t* ma_zonalReduc(ma_node_t* node, ma_reduction_t type) {
  return guard([&] { return R().dag.zonalReduc(need(node), type); }, (Node*)nullptr);
}
ma_node_t* ma_boundedNbh(ma_node_t* node, ma_node_t* cx, ma_node_t* cy) {
  return guard([&]() -> Node* {
    need(node); need(cx); need(cy);
    throw std::runtime_error("boundedNbh: not supported (the reference generates no kernel statement for it)");
  }, (Node*)nullptr);
}

// Symbolic-loop digestion (front/bindings.cpp:244-275): a `while` over a D0
// raster runs as ordinary Python here (Raster.__bool__), so none of these
// builds a Loop node.
static const char* kNoSymLoop =
    "symbolic loops: not supported; a while over a D0 raster runs as ordinary Python (its condition is evaluated "
    "on the device each iteration)";
void ma_loopStart(void) { guard([]() -> int { throw std::runtime_error(kNoSymLoop); }, 0); }
void ma_loopCond(ma_node_t*) { guard([]() -> int { throw std::runtime_error(kNoSymLoop); }, 0); }
void ma_loopBody(void) { guard([]() -> int { throw std::runtime_error(kNoSymLoop); }, 0); }
void ma_loopAgain(void) { guard([]() -> int { throw std::runtime_error(kNoSymLoop); }, 0); }
ma_node_t* ma_loopAssemble(void) { return guard([]() -> Node* { throw std::runtime_error(kNoSymLoop); }, (Node*)nullptr); }
void ma_loopEnd(void) { guard([]() -> int { throw std::runtime_error(kNoSymLoop); }, 0); }
void ma_loopAgainTail(ma_node_t*, ma_node_t*** agains, ma_node_t*** tails, int* num) {
  if (agains) *agains = nullptr;
  if (tails) *tails = nullptr;
  if (num) *num = 0;
  guard([]() -> int { throw std::runtime_error(kNoSymLoop); }, 0);
}

ma_node_t* ma_spreadNeighbor(ma_node_t* node, ma_node_t* dir, ma_reduction_t type) {
  return guard([&] { return R().dag.spreadNeighbor(need(node), need(dir), type); }, (Node*)nullptr);
}
ma_node_t* ma_spreadScan(ma_node_t* node, ma_node_t* dir, ma_reduction_t type) {
  return guard([&] { return R().dag.spreadScan(need(node), need(dir), type); }, (Node*)nullptr);
}
ma_node_t* ma_stats(ma_node_t* node) {
  return guard([&] { return R().dag.stats(need(node)); }, (Node*)nullptr);
}
ma_node_t* ma_barrier(ma_node_t* node) {
  return guard([&] { return R().dag.barrier(need(node)); }, (Node*)nullptr);
}
ma_node_t* ma_access(ma_node_t* node, ma_dim2_t coord) {
  return guard([&] { return R().dag.access(need(node), D(coord)); }, (Node*)nullptr);
}
ma_node_t* ma_lhsAccess(ma_node_t* lhs, ma_node_t* rhs, ma_dim2_t coord) {
  return guard([&] { return R().dag.lhsAccess(need(lhs), need(rhs), D(coord)); }, (Node*)nullptr);
}

// ---- extensions

int ma_statsRange(ma_node_t* node, ma_variant_t* max, ma_variant_t* min) {
  return guard([&] {
    need(node);
    if (node->kind != Kind::Stats || node->aux.size() != 2) throw std::runtime_error("statsRange: not a Stats node");
    for (Node* a : node->aux)
      if (!a->evaluated) throw std::runtime_error("statsRange: the Stats node has not been evaluated");
    if (max) *max = MV(node->aux[0]->value);
    if (min) *min = MV(node->aux[1]->value);
    return 0;
  }, -1);
}

ma_node_t* ma_zonalStats(ma_node_t* data, ma_node_t* zone, int nzones) {
  return guard([&] { return R().dag.zonalStats(need(data), need(zone), nzones); }, (Node*)nullptr);
}

int ma_zonalTable(ma_node_t* node, double* sum, double* mn, double* mx, long long* count, int nzones) {
  return guard([&] {
    need(node);
    if (node->kind != Kind::ZonalStats) throw std::runtime_error("zonalTable: not a zonal-statistics node");
    if (!node->evaluated) throw std::runtime_error("zonalTable: node has not been evaluated");
    if (nzones != node->nzones) throw std::runtime_error("zonalTable: wrong number of zones");
    const int dt = node->prev[0]->meta.dt;
    for (int z = 0; z < nzones; z++) {
      const unsigned char* r = node->table.data() + (size_t)z * 24;
      double s; std::memcpy(&s, r, 8);
      uint32_t bmn, bmx, c;
      std::memcpy(&bmn, r + 8, 4); std::memcpy(&bmx, r + 12, 4); std::memcpy(&c, r + 16, 4);
      auto conv = [&](uint32_t b) -> double {
        if (dt == F32) { float f; std::memcpy(&f, &b, 4); return f; }
        if (dt == S32 || dt == S8 || dt == S16) return (double)(int32_t)b;
        return (double)b;
      };
      if (sum) sum[z] = s;
      if (mn) mn[z] = conv(bmn);
      if (mx) mx[z] = conv(bmx);
      if (count) count[z] = c;
    }
    return 0;
  }, -1);
}

ma_node_t* ma_input(const void* host, size_t nbytes, ma_dim2_t ds, ma_datatype_t dt, ma_dim2_t bs) {
  return guard([&] {
    static uint64_t uid = 1;
    auto r = std::make_shared<HostRaster>();
    r->meta.dt = dt; r->meta.nd = D2; r->meta.mo = ROW + BLK;
    r->meta.ds = D(ds); r->meta.bs = D(bs);
    if (dt <= NONE_DATATYPE || dt >= N_DATATYPE) throw std::runtime_error("input: bad data type");
    if (r->meta.ds.n != 2 || r->meta.bs.n != 2) throw std::runtime_error("input: sizes need 2 dimensions");
    if (r->meta.ds.arr[0] <= 0 || r->meta.ds.arr[1] <= 0 || r->meta.bs.arr[0] <= 0 || r->meta.bs.arr[1] <= 0)
      throw std::runtime_error("input: sizes must be positive");
    const size_t need_bytes = R().rasterBytes(r->meta);
    if (nbytes != need_bytes) throw std::runtime_error("input: byte count differs from the block-tiled size");
    if (!host) throw std::runtime_error("input: null data");
    r->bytes.assign((const unsigned char*)host, (const unsigned char*)host + nbytes);
    r->nbytes = nbytes;
    r->uid = uid++;
    return R().dag.input(r);
  }, (Node*)nullptr);
}

int ma_fetch(ma_node_t* node, void* host, size_t nbytes) {
  return guard([&] { R().fetch(need(node), host, nbytes); return 0; }, -1);
}

int ma_fileLoad(const char* file_path, void* host, size_t nbytes) {
  return guard([&] {
    RasterFile f(file_path ? file_path : "");
    if (!host || nbytes != f.payload_bytes())
      throw std::runtime_error("fileLoad: the buffer must hold the whole block-tiled payload (" +
                               std::to_string(f.payload_bytes()) + " bytes)");
    f.read(0, nbytes, host);
    return 0;
  }, -1);
}

int ma_fileStore(const char* file_path, const void* host, size_t nbytes, ma_dim2_t ds, ma_datatype_t dt, ma_dim2_t bs) {
  return guard([&] {
    Meta m;
    m.dt = dt; m.nd = D2; m.mo = ROW + BLK; m.ds = D(ds); m.bs = D(bs);
    if (dt <= NONE_DATATYPE || dt >= N_DATATYPE || m.ds.arr[0] <= 0 || m.ds.arr[1] <= 0 || m.bs.arr[0] <= 0 ||
        m.bs.arr[1] <= 0)
      throw std::runtime_error("fileStore: bad raster description");
    RasterWriter w(file_path ? file_path : "", m);
    if (!host || nbytes != w.payload_bytes())
      throw std::runtime_error("fileStore: the buffer must hold the whole block-tiled payload (" +
                               std::to_string(w.payload_bytes()) + " bytes)");
    w.append(host, nbytes);
    w.finish();
    return 0;
  }, -1);
}

int ma_release(ma_node_t* node) {
  return guard([&] { R().drop(need(node)); return 0; }, -1);
}

int ma_setOption(const char* key, long long value) {
  return guard([&] {
    std::string k(key ? key : "");
    Options& o = R().opt;
    if (k == "form") {
      if (value != FORM_PLAIN && value != FORM_BATCHED) throw std::runtime_error("setOption: form must be 0 or 1");
      o.form = (int)value;
    } else if (k == "graph") o.graph = value != 0;
    else if (k == "focal_zonal") o.focal_zonal = value != 0;
    else if (k == "code_fusion") o.code_fusion = value != 0;
    else if (k == "profile") dev::profile_set(value != 0);
    else if (k == "tile_tx") o.tile.TX = (int)value;
    else if (k == "tile_ty") o.tile.TY = (int)value;
    else if (k == "tile_vx") o.tile.VX = (int)value;
    else if (k == "tile_pf") o.tile.PF = value != 0;
    else if (k == "tile_st") o.tile.ST = value != 0;
    else if (k == "tile_w") {
      if (value < 0 || value > 8) throw std::runtime_error("setOption: tile_w (wavefronts per SIMD) in [0, 8]");
      o.tile.W = (int)value;
    }
    else if (k == "libm") o.libm = value != 0;
    else if (k == "predict") o.predict = value != 0;
    else if (k == "hbm_cap") {
      if (value < 0) throw std::runtime_error("setOption: hbm_cap is a byte count (0: no cap)");
      o.hbm_cap = (size_t)value;
      R().store.cap = o.hbm_cap;
    } else if (k == "wg_per_cu") {
      if (value < 0 || value > 64) throw std::runtime_error("setOption: wg_per_cu in [0, 64]");
      o.wg_per_cu = (int)value;
    } else if (k == "batch_tasks") o.batch_tasks = value != 0;
    else if (k == "chunk_tiles") {
      if (value < 0 || value > 4096) throw std::runtime_error("setOption: chunk_tiles in [0, 4096]");
      o.chunk_tiles = (int)value;
    }
    else if (k == "queue") {
      if (value < 0 || value > 2) throw std::runtime_error("setOption: queue in {0 batched only, 1 always, 2 never}");
      o.queue = (int)value;
    } else throw std::runtime_error("setOption: unknown option " + k);
    return 0;
  }, -1);
}

long long ma_getOption(const char* key) {
  std::string k(key ? key : "");
  Options& o = R().opt;
  if (k == "form") return o.form;
  if (k == "graph") return o.graph;
  if (k == "focal_zonal") return o.focal_zonal;
  if (k == "code_fusion") return o.code_fusion;
  if (k == "profile") return dev::profile_on();
  if (k == "tile_tx") return o.tile.TX;
  if (k == "tile_ty") return o.tile.TY;
  if (k == "tile_vx") return o.tile.VX;
  if (k == "tile_pf") return o.tile.PF;
  if (k == "tile_st") return o.tile.ST;
  if (k == "tile_w") return o.tile.W;
  if (k == "libm") return o.libm;
  if (k == "predict") return o.predict;
  if (k == "hbm_cap") return (long long)o.hbm_cap;
  if (k == "wg_per_cu") return o.wg_per_cu;
  if (k == "batch_tasks") return o.batch_tasks;
  if (k == "queue") return o.queue;
  if (k == "chunk_tiles") return o.chunk_tiles;
  return -1;
}

const char* ma_plan(ma_node_t** vec, int num) {
  return guard([&] {
    if (num <= 0 || !vec) throw std::runtime_error("plan: no nodes");
    std::vector<Node*> v;
    for (int i = 0; i < num; i++) v.push_back(need(vec[i]));
    g_json = R().plan(v);
    return g_json.c_str();
  }, (const char*)nullptr);
}

int ma_compilePlan(ma_node_t** vec, int num) {
  return guard([&] {
    if (num <= 0 || !vec) throw std::runtime_error("compilePlan: no nodes");
    std::vector<Node*> v;
    for (int i = 0; i < num; i++) v.push_back(need(vec[i]));
    std::string j = R().plan(v);
    // every "source" string of the plan is compiled
    size_t pos = 0;
    int n = 0;
    while ((pos = j.find("\"source\": \"", pos)) != std::string::npos) {
      pos += 11;
      std::string src;
      for (; pos < j.size() && j[pos] != '"'; pos++) {
        if (j[pos] == '\\' && pos + 1 < j.size()) {
          char c = j[++pos];
          src += c == 'n' ? '\n' : c == 't' ? '\t' : c;
        } else src += j[pos];
      }
      dev::compile(src);
      n++;
    }
    return n;
  }, -1);
}

const char* ma_runStats(void) {
  Runtime& r = R();
  std::ostringstream s;
  s << "{\"evals\": " << r.cnt.evals << ", \"groups\": " << r.cnt.groups << ", \"launches\": " << r.cnt.launches
    << ", \"compiles\": " << r.cnt.compiles << ", \"shape_hits\": " << r.cnt.shape_hits
    << ", \"predicted\": " << r.cnt.predicted << ", \"narrow_launches\": " << r.cnt.narrow_launches
    << ", \"spills\": " << r.cnt.spills
    << ", \"unspills\": " << r.cnt.unspills << ", \"dropped_for_space\": " << r.cnt.dropped_for_space
    << ", \"bytes_spilled\": " << r.cnt.bytes_spilled
    << ", \"scalar_tasks\": " << r.cnt.scalar_tasks << ", \"bytes_h2d\": " << r.cnt.bytes_h2d
    << ", \"bytes_d2h\": " << r.cnt.bytes_d2h << ", \"store_in_use\": " << r.store.in_use()
    << ", \"store_cached\": " << r.store.cached() << ", \"graphs\": " << dev::graph_cache_size()
    << ", \"ms_last_eval\": " << r.cnt.ms_last_eval << ", \"ms_fuse\": " << r.cnt.ms_fuse
    << ", \"ms_codegen\": " << r.cnt.ms_codegen << ", \"ms_launch\": " << r.cnt.ms_launch
    << ", \"ms_plan\": " << r.cnt.ms_plan << ", \"ms_prepare\": " << r.cnt.ms_prepare
    << ", \"ms_collect\": " << r.cnt.ms_collect << ", \"ms_predict\": " << r.cnt.ms_predict << ", \"occupancy\": {";
  bool first = true;
  for (auto& k : r.kernel_occ_) { s << (first ? "" : ", ") << "\"" << k.first << "\": " << k.second; first = false; }
  s << "}, \"cus\": " << (dev::initialized() ? dev::num_cus() : 0) << "}";
  g_json = s.str();
  return g_json.c_str();
}

double ma_lastEvalMs(void) { return R().cnt.ms_last_eval; }

const char* ma_kernelTimes(void) {
  return guard([&] {
    std::ostringstream s;
    s << "{";
    bool first = true;
    for (auto& k : dev::profile_take()) {
      s << (first ? "" : ", ") << "\"" << k.name << "\": {\"count\": " << k.count << ", \"total_ms\": " << k.total_ms
        << ", \"min_ms\": " << k.min_ms << ", \"max_ms\": " << k.max_ms << "}";
      first = false;
    }
    s << "}";
    g_json = s.str();
    return g_json.c_str();
  }, (const char*)nullptr);
}

int ma_serveOpen(long long hbm_budget, int max_inflight) {
  return guard([&] {
    if (hbm_budget < 0) throw std::runtime_error("serveOpen: negative budget");
    R().serveOpen((size_t)hbm_budget, max_inflight);
    return 0;
  }, -1);
}

long long ma_serveSubmit(ma_node_t** vec, int num) {
  return guard([&] {
    if (num <= 0 || !vec) throw std::runtime_error("serveSubmit: no nodes");
    std::vector<Node*> v;
    for (int i = 0; i < num; i++) v.push_back(need(vec[i]));
    return R().serveSubmit(v);
  }, (long long)-1);
}

int ma_serveStep(int wait_us) {
  return guard([&] { return R().serveStep(wait_us); }, -1);
}

int ma_serveTake(long long* ids, double* t_submit, double* t_admit, double* t_done, int cap) {
  return guard([&] {
    if (cap < 0) throw std::runtime_error("serveTake: negative capacity");
    auto d = R().serveTake((size_t)cap);
    for (size_t i = 0; i < d.size(); i++) {
      if (ids) ids[i] = d[i].id;
      if (t_submit) t_submit[i] = d[i].t_submit;
      if (t_admit) t_admit[i] = d[i].t_admit;
      if (t_done) t_done[i] = d[i].t_done;
    }
    return (int)d.size();
  }, -1);
}

const char* ma_serveStats(void) {
  return guard([&] {
    g_json = R().serveStats();
    return g_json.c_str();
  }, (const char*)nullptr);
}

int ma_serveClose(void) {
  return guard([&] { R().serveClose(); return 0; }, -1);
}

const char* ma_kernelSource(const char* name) {
  return guard([&] {
    g_json = R().kernelSource(name ? name : "");
    if (g_json.empty()) throw std::runtime_error("kernelSource: no kernel of that name was built");
    return g_json.c_str();
  }, (const char*)nullptr);
}

const char* ma_autotune(ma_node_t** vec, int num, int reps) {
  return guard([&] {
    if (num <= 0 || !vec) throw std::runtime_error("autotune: no nodes");
    std::vector<Node*> v;
    for (int i = 0; i < num; i++) v.push_back(need(vec[i]));
    g_json = R().autotune(v, reps);
    return g_json.c_str();
  }, (const char*)nullptr);
}

int ma_tuneFile(const char* path, int save) {
  return guard([&] {
    if (!path) throw std::runtime_error("tuneFile: null path");
    if (save) R().tune_save(path);
    else R().tune_load(path);
    return (int)R().tuned_count();
  }, -1);
}

int ma_fusionRules(int top_or_left, int bot_or_right) {
  return guard([&] {
    const bool fz = R().opt.focal_zonal;
    return (int)can_pipe_fuse(top_or_left, bot_or_right, fz) | ((int)can_flat_fuse(top_or_left, bot_or_right, fz) << 1);
  }, -1);
}

ma_node_t* ma_genDem(ma_dim2_t ds, ma_dim2_t bs, const double* gauss, int ng, unsigned long long noise_seed) {
  return guard([&] {
    Runtime& r = R();
    Meta m;
    m.dt = F32; m.nd = D2; m.mo = ROW + BLK; m.ds = D(ds); m.bs = D(bs);
    auto h = std::make_shared<HostRaster>();
    h->meta = m;
    h->nbytes = r.rasterBytes(m);
    static uint64_t uid = 1ull << 40;
    h->uid = uid++;
    if (ng < 0 || ng > 256) throw std::runtime_error("genDem: at most 256 Gaussians");
    h->dev = r.store.alloc(h->nbytes);
    void* scratch = nullptr;
    try {
      scratch = r.store.alloc(gen::dem_scratch_bytes(ng));
      gen::dem(h->dev, m.bs.arr[0], m.bs.arr[1], m.nb(0), m.nb(1), m.ds.arr[0], m.ds.arr[1], gauss, ng, noise_seed,
               scratch, dev::stream());
    } catch (...) {
      if (scratch) r.store.release(scratch);
      r.store.release(h->dev);
      throw;
    }
    r.store.release(scratch);
    // the runtime owns this buffer: it is released with the node
    h->dev_owned = true;
    Node* n = r.dag.input(h);
    r.adopt(n, h->dev);
    return n;
  }, (Node*)nullptr);
}

int ma_genDemNoAlloc(void* out, size_t out_bytes, ma_dim2_t ds, ma_dim2_t bs, const double* gauss, int ng,
                     unsigned long long noise_seed) {
  return guard([&] {
    if (!dev::initialized()) throw std::runtime_error("genDemNoAlloc: no device (call ma_setupDevices)");
    Meta m;
    m.dt = F32; m.nd = D2; m.mo = ROW + BLK; m.ds = D(ds); m.bs = D(bs);
    if (m.ds.arr[0] <= 0 || m.ds.arr[1] <= 0 || m.bs.arr[0] <= 0 || m.bs.arr[1] <= 0)
      throw std::runtime_error("genDemNoAlloc: bad raster geometry");
    const size_t max_ng = gen::nalloc_scratch_bytes() / gen::dem_scratch_bytes(1);
    if (ng < 0 || (size_t)ng > max_ng)
      throw std::runtime_error("genDemNoAlloc: at most " + std::to_string(max_ng) + " Gaussians (its scratch holds " +
                               std::to_string(gen::nalloc_scratch_bytes()) + " bytes)");
    if (ng > 0 && !gauss) throw std::runtime_error("genDemNoAlloc: null Gaussian table");
    const size_t need = (size_t)m.nb(0) * m.nb(1) * m.bs.arr[0] * m.bs.arr[1] * sizeof(float);
    if (!out || out_bytes < need)
      throw std::runtime_error("genDemNoAlloc: the output needs " + std::to_string(need) + " bytes of device memory");
    gen::dem(out, m.bs.arr[0], m.bs.arr[1], m.nb(0), m.nb(1), m.ds.arr[0], m.ds.arr[1], gauss, ng, noise_seed,
             gen::nalloc_scratch(dev::device_id()), dev::stream());
    return 0;
  }, -1);
}

ma_node_t* ma_genZones(ma_dim2_t ds, ma_dim2_t bs, ma_datatype_t dt, const int* px, const int* py, int np) {
  return guard([&] {
    Runtime& r = R();
    if (!(dt == U8 || dt == U16 || dt == S32 || dt == U32)) throw std::runtime_error("genZones: type must be U8, U16, S32 or U32");
    Meta m;
    m.dt = dt; m.nd = D2; m.mo = ROW + BLK; m.ds = D(ds); m.bs = D(bs);
    auto h = std::make_shared<HostRaster>();
    h->meta = m;
    h->nbytes = r.rasterBytes(m);
    static uint64_t uid = 1ull << 41;
    h->uid = uid++;
    if (np <= 0 || np > 1024) throw std::runtime_error("genZones: 1..1024 points");
    h->dev = r.store.alloc(h->nbytes);
    void* scratch = nullptr;
    try {
      scratch = r.store.alloc(gen::voronoi_scratch_bytes(np));
      gen::voronoi(h->dev, sizeof_dt(dt), dt == S32, m.bs.arr[0], m.bs.arr[1], m.nb(0), m.nb(1), m.ds.arr[0],
                   m.ds.arr[1], px, py, np, scratch, dev::stream());
    } catch (...) {
      if (scratch) r.store.release(scratch);
      r.store.release(h->dev);
      throw;
    }
    r.store.release(scratch);
    h->dev_owned = true;
    Node* n = r.dag.input(h);
    r.adopt(n, h->dev);
    return n;
  }, (Node*)nullptr);
}

int ma_cacheBudget(long long bytes) {
  return guard([&] {
    if (bytes < 0) throw std::runtime_error("cacheBudget: negative budget");
    R().cacheBudget((size_t)bytes);
    return 0;
  }, -1);
}
ma_node_t* ma_cacheGet(long long key) {
  return guard([&] { return R().cacheGet(key); }, (Node*)nullptr);
}
int ma_cachePut(long long key, ma_node_t* node) {
  return guard([&] { R().cachePut(key, need(node)); return 0; }, -1);
}
int ma_cacheDrop(long long key) {
  return guard([&] { R().cacheDrop(key); return 0; }, -1);
}
const char* ma_cacheStats(void) {
  return guard([&] { g_json = R().cacheStats(); return g_json.c_str(); }, (const char*)nullptr);
}

int ma_mathCheck(const char* fn, unsigned long long lo, unsigned long long hi, double* max_ulp,
                 unsigned long long* n_over, unsigned long long* n, unsigned* worst_in) {
  return guard([&] {
    if (!dev::initialized()) throw std::runtime_error("mathCheck: no device (call ma_setupDevices)");
    if (!fn || !max_ulp || !n_over || !n || !worst_in) throw std::runtime_error("mathCheck: null argument");
    Runtime& r = R();
    void* scratch = r.store.alloc(64);
    try {
      mathcheck::run(fn, lo, hi, scratch, dev::stream(), max_ulp, n_over, n, worst_in);
    } catch (...) {
      r.store.release(scratch);
      throw;
    }
    r.store.release(scratch);
    return 0;
  }, -1);
}

}  // extern "C"
