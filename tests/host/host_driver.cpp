// Drives the runtime's host code through the C ABI (include/mapx.h) on the
// device double (dev_double.cpp), built with AddressSanitizer and
// UndefinedBehaviorSanitizer: node lifetimes (references held and dropped in
// every order), evaluation with intermediates released early, batched
// evaluation of many rasters, graph recording, the resident cache under a
// budget, spilling under an HBM cap (and the error past it), serving with
// continuous batching over cached rasters, and DAG collection as nodes come
// and go. No value is checked (the double computes nothing); the run must
// end without a sanitizer report, a double check, or a failed call that
// should have succeeded. Prints "host driver: ok" at the end.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/mapx.h"

namespace {

enum { F32 = 1, U8 = 4, S32 = 10, D0 = 2, D2 = 6, ROWBLK = 3, DEV_GPU = 4 };
enum { ADD = 1, SUB = 2, MUL = 3, DIV = 4, GT = 9 };
enum { COS = 7, ATAN = 11, SQRT = 24 };
enum { SUM = 1, MAX = 6, MIN = 7 };

[[noreturn]] void fail(const char* what) {
  std::fprintf(stderr, "host driver: %s failed: %s\n", what, ma_lastError());
  std::exit(2);
}

// A held node, as a Python Raster holds one.
struct R {
  ma_node_t* n = nullptr;
  R() {}
  explicit R(ma_node_t* p, const char* what = "factory") : n(p) {
    if (!n) fail(what);
    ma_increaseRef(n);
  }
  R(const R& o) : n(o.n) { if (n) ma_increaseRef(n); }
  R& operator=(const R& o) {
    if (o.n) ma_increaseRef(o.n);
    if (n) ma_decreaseRef(n);
    n = o.n;
    return *this;
  }
  ~R() { if (n) ma_decreaseRef(n); }
};

ma_dim2_t dim(int a, int b) { ma_dim2_t d; d.arr[0] = a; d.arr[1] = b; d.n = 2; return d; }
ma_dim2_t none() { ma_dim2_t d; d.arr[0] = 0; d.arr[1] = 0; d.n = 0; return d; }
ma_variant_t f32(float v) { ma_variant_t x; std::memset(&x, 0, sizeof x); x.u.f32 = v; x.dtype = F32; return x; }
ma_variant_t s32(int v) { ma_variant_t x; std::memset(&x, 0, sizeof x); x.u.s32 = v; x.dtype = S32; return x; }

R k(float v) { return R(ma_constant(f32(v), none(), F32, ROWBLK, none()), "constant"); }
R bin(const R& a, const R& b, int op) { return R(ma_binary(a.n, b.n, op), "binary"); }
R un(const R& a, int op) { return R(ma_unary(a.n, op), "unary"); }
R conv(const R& a, const int (&w)[9]) {
  ma_variant_t v[9];
  for (int i = 0; i < 9; i++) v[i] = s32(w[i]);
  return R(ma_convolution(a.n, v, 9, dim(3, 3)), "convolution");
}

R input(int w, int h, int bw, int bh, std::mt19937& rng) {
  const int nb0 = (w + bw - 1) / bw, nb1 = (h + bh - 1) / bh;
  std::vector<float> a((size_t)nb0 * nb1 * bw * bh);
  for (float& x : a) x = (float)(rng() % 1000);
  return R(ma_input(a.data(), a.size() * 4, dim(w, h), F32, dim(bw, bh)), "input");
}

R slope(const R& d) {
  static const int h[9] = {-1, 0, 1, -2, 0, 2, -1, 0, 1}, v[9] = {-1, -2, -1, 0, 0, 0, 1, 2, 1};
  R x = bin(conv(d, h), k(8.0f), DIV), y = bin(conv(d, v), k(8.0f), DIV);
  return un(un(bin(bin(x, x, MUL), bin(y, y, MUL), ADD), SQRT), ATAN);
}
R shade(const R& d) { return bin(un(slope(d), COS), k(0.7f), MUL); }

void eval(std::vector<R> v) {
  std::vector<ma_node_t*> p;
  for (R& r : v) p.push_back(r.n);
  ma_eval(p.data(), (int)p.size());
  if (std::strlen(ma_lastError())) fail("eval");
}
void fetch(const R& r) {
  const ma_dim2_t ds = ma_datasize(r.n), bs = ma_blocksize(r.n);
  const size_t n = (size_t)((ds.arr[0] + bs.arr[0] - 1) / bs.arr[0]) * ((ds.arr[1] + bs.arr[1] - 1) / bs.arr[1]) *
                   bs.arr[0] * bs.arr[1] * (ma_datatype(r.n) == U8 ? 1 : 4);
  std::vector<char> h(n);
  if (ma_fetch(r.n, h.data(), n) != 0) fail("fetch");
}

// a raster whose last block column and row are narrower than a 5x5 window's halo
void scenario_narrow(std::mt19937& rng) {
  R d = input(129, 65, 128, 64, rng);
  ma_variant_t w[25];
  for (int i = 0; i < 25; i++) w[i] = s32(1 + i % 3);
  R c = R(ma_convolution(d.n, w, 25, dim(5, 5)), "convolution");
  eval({c});
  fetch(c);
  if (!std::strstr(ma_runStats(), "\"narrow_launches\": 0,") == false) fail("the narrow module (not used)");
}

void scenario_lifetimes(std::mt19937& rng) {
  for (int rep = 0; rep < 3; rep++) {
    R d = input(300, 200, 128, 64, rng);
    eval({d});                  // an input alone: loaded, held
    fetch(d);
    R s = slope(d);
    R hs = shade(d);
    eval({hs});                 // s is held but not a target: computed inside, not kept
    eval({s, hs});              // hs held: the cut takes it; s computed
    fetch(s);
    fetch(hs);
    R z = R(ma_cast(bin(s, k(1.0f), GT).n, S32), "cast");
    R zs = R(ma_zonalStats(d.n, z.n, 2), "zonalStats");
    R zr = R(ma_zonalReduc(hs.n, SUM), "zonalReduc");
    eval({zs, zr});
    double sum[2], mn[2], mx[2];
    long long cnt[2];
    if (ma_zonalTable(zs.n, sum, mn, mx, cnt, 2) != 0) fail("zonalTable");
    ma_value(zr.n);
    s = R();                    // drop in an order different from creation
    d = R();
    fetch(hs);
  }
}

void scenario_batched(std::mt19937& rng, bool graph) {
  ma_setOption("graph", graph);
  std::vector<R> outs, dems;
  for (int i = 0; i < 640; i++) {  // past one sub-batch of prepared groups
    const int w = 64 + (int)(rng() % 300), h = 64 + (int)(rng() % 200);
    R d = input(w, h, 256, 256, rng);
    R z = R(ma_cast(bin(d, k(500.0f), GT).n, S32), "cast");
    dems.push_back(d);
    outs.push_back(R(ma_zonalStats(shade(d).n, z.n, 2), "zonalStats"));
  }
  for (int rep = 0; rep < 3; rep++) {
    for (R& o : outs) ma_release(o.n);
    eval(outs);
  }
  outs.resize(outs.size() / 2);  // half dropped while the rest stay
  eval(outs);
  ma_setOption("graph", 0);
}

// a chain of small levels, recorded and replayed from graphs
void scenario_chain(std::mt19937& rng) {
  ma_setOption("graph", 1);
  R d = input(400, 300, 128, 128, rng);
  static const int ring[9] = {1, 1, 1, 1, 0, 1, 1, 1, 1};
  R s = d;
  for (int g = 0; g < 20; g++) s = bin(conv(s, ring), k(0.125f), MUL);
  R z = R(ma_zonalReduc(s.n, MAX), "zonalReduc");
  for (int rep = 0; rep < 3; rep++) {
    ma_release(s.n);
    eval({s, z});
    fetch(s);
  }
  ma_setOption("graph", 0);
}

void scenario_cache_and_spill(std::mt19937& rng) {
  ma_cacheBudget(6LL << 20);
  for (int i = 0; i < 24; i++) {
    const long long key = i % 9;
    ma_node_t* c = ma_cacheGet(key);
    if (!c) {
      R d = input(512, 384, 128, 128, rng);
      R s = slope(d);
      eval({s});
      if (ma_cachePut(key, s.n) != 0) fail("cachePut");
    } else {
      R s(c, "cacheGet");
      R hs = bin(s, k(2.0f), MUL);
      eval({hs});
      fetch(hs);
    }
  }
  // spilling: held rasters past a small cap leave and come back
  std::vector<R> held;
  for (int i = 0; i < 8; i++) {
    R d = input(512, 512, 128, 128, rng);
    held.push_back(slope(d));
    eval({held.back()});
  }
  if (ma_setOption("hbm_cap", 5LL << 20) != 0) fail("setOption hbm_cap");
  for (int rep = 0; rep < 2; rep++)
    for (R& h : held) {
      R t = bin(h, k(1.5f), MUL);
      eval({t});
      fetch(t);
      fetch(h);
    }
  // one evaluation larger than the cap fails and leaves nothing half-computed
  {
    R big = input(2048, 2048, 256, 256, rng);
    R t = bin(shade(big), slope(bin(big, k(2.0f), MUL)), ADD);
    std::vector<ma_node_t*> p{t.n};
    ma_eval(p.data(), 1);
    if (std::strstr(ma_lastError(), "HBM cap") == nullptr) fail("the evaluation past the cap (expected an error)");
  }
  ma_setOption("hbm_cap", 0);
  for (R& h : held) fetch(h);
  ma_cacheBudget(0);
  for (long long key = 0; key < 9; key++) ma_cacheDrop(key);
}

void scenario_serving(std::mt19937& rng) {
  std::vector<R> dems;
  for (int i = 0; i < 12; i++) dems.push_back(input(200 + 37 * i, 150 + 23 * i, 128, 128, rng));
  if (ma_serveOpen(64LL << 20, 8) != 0) fail("serveOpen");
  std::vector<std::pair<long long, std::vector<R>>> live;
  long long done = 0, submitted = 0;
  for (int step = 0; step < 400; step++) {
    if (submitted < 300 && rng() % 3 != 0) {
      const R& d = dems[rng() % dems.size()];
      std::vector<R> t;
      switch (rng() % 3) {
        case 0: t.push_back(shade(d)); break;
        case 1: t.push_back(slope(d)); t.push_back(shade(d)); break;
        default: {
          R z = R(ma_cast(bin(d, k(300.0f), GT).n, S32), "cast");
          t.push_back(R(ma_zonalStats(d.n, z.n, 2), "zonalStats"));
        }
      }
      std::vector<ma_node_t*> p;
      for (R& r : t) p.push_back(r.n);
      const long long id = ma_serveSubmit(p.data(), (int)p.size());
      if (id < 0) fail("serveSubmit");
      live.push_back({id, t});
      submitted++;
    }
    if (ma_serveStep(0) < 0) fail("serveStep");
    long long ids[64];
    double a[64], b[64], c[64];
    const int n = ma_serveTake(ids, a, b, c, 64);
    if (n < 0) fail("serveTake");
    for (int i = 0; i < n; i++)
      for (size_t j = 0; j < live.size(); j++)
        if (live[j].first == ids[i]) { live.erase(live.begin() + j); done++; break; }
  }
  if (ma_serveClose() != 0) fail("serveClose");
  long long ids[512];
  double a[512], b[512], c[512];
  int n;
  while ((n = ma_serveTake(ids, a, b, c, 512)) > 0) done += n;
  if (done != submitted) {
    std::fprintf(stderr, "host driver: %lld of %lld requests completed\n", done, submitted);
    std::exit(2);
  }
}

// serving while rasters spill: held DEMs past a small cap, requests reading them
void scenario_serving_under_cap(std::mt19937& rng) {
  std::vector<R> dems;
  for (int i = 0; i < 6; i++) {
    dems.push_back(slope(input(400, 300, 128, 128, rng)));  // computed rasters: spilled, not dropped
    eval({dems.back()});
  }
  if (ma_setOption("hbm_cap", 4LL << 20) != 0) fail("setOption hbm_cap");
  if (ma_serveOpen(3LL << 20, 4) != 0) fail("serveOpen");
  long long submitted = 0, done = 0;
  std::vector<std::vector<R>> keep;
  for (int step = 0; step < 120; step++) {
    if (submitted < 60) {
      std::vector<R> t{shade(dems[rng() % dems.size()])};
      std::vector<ma_node_t*> p{t[0].n};
      if (ma_serveSubmit(p.data(), 1) < 0) fail("serveSubmit");
      keep.push_back(t);
      submitted++;
    }
    if (ma_serveStep(0) < 0) fail("serveStep");
    long long ids[64];
    double a[64], b[64], c[64];
    const int n = ma_serveTake(ids, a, b, c, 64);
    if (n < 0) fail("serveTake");
    done += n;
  }
  if (ma_serveClose() != 0) fail("serveClose");
  long long ids[64];
  double a[64], b[64], c[64];
  int n;
  while ((n = ma_serveTake(ids, a, b, c, 64)) > 0) done += n;
  if (done != submitted) fail("requests under the cap (not all completed)");
  ma_setOption("hbm_cap", 0);
  for (auto& t : keep) fetch(t[0]);
}

void scenario_churn(std::mt19937& rng) {
  R d = input(128, 128, 64, 64, rng);
  for (int i = 0; i < 3000; i++) {
    R t = bin(bin(d, k((float)i), ADD), k(2.0f), MUL);  // new constants: new nodes each time
    if (i % 500 == 0) eval({t});
  }
  eval({shade(d)});
}

}  // namespace

int main() {
  ma_setupDevices("", DEV_GPU, "");
  std::mt19937 rng(7);
  scenario_lifetimes(rng);
  scenario_narrow(rng);
  scenario_batched(rng, false);
  scenario_batched(rng, true);
  scenario_chain(rng);
  ma_setOption("form", 0);  // plain form: one launch per block
  scenario_lifetimes(rng);
  ma_setOption("form", 1);
  scenario_cache_and_spill(rng);
  scenario_serving(rng);
  scenario_serving_under_cap(rng);
  scenario_churn(rng);
  std::printf("host driver: ok %s\n", ma_runStats());
  return 0;
}
