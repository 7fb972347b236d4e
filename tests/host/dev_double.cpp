// Host double of the device layer (map_amd/csrc/dev.hpp) and of the
// ahead-of-time device entry points (gen.hip, radial.hip, spread.hip,
// mathcheck.hip), so that the runtime's host code (DAG, fusion, code
// generation, planning, the store, spilling, the resident cache, serving)
// builds and runs on a machine without a GPU, under AddressSanitizer and
// UndefinedBehaviorSanitizer. Nothing is computed: "device memory" is host
// memory (zero-filled), launches and copies are bookkeeping, events complete
// at once, and kernels are never compiled. What the double checks by itself:
// - every device buffer freed is a live allocation of this layer;
// - no device buffer goes back while launches recorded for graph replay are
//   still pending (they would read it when the graph runs);
// - no read-back or wait while recorded launches are pending (it would not
//   see or cover them);
// - every copy stays inside one live allocation.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "dev.hpp"
#include "mapx_dev_hiph.inc"

namespace mapx {
namespace dev {
namespace {

struct State {
  bool init = false;
  int device = 0;
  bool recording = false;
  long pending = 0;                       // launches recorded since record_begin
  std::map<char*, size_t> dev_allocs;      // "device" buffers by base address
  std::map<char*, size_t> host_allocs;     // pinned host buffers
  bool profile = false;
  std::map<std::string, KernelTime> times;
  long graphs = 0;
};
State& S() {
  static State* s = new State();
  return *s;
}
[[noreturn]] void die(const std::string& m) {
  std::fprintf(stderr, "dev double: %s\n", m.c_str());
  std::fflush(stderr);
  std::abort();
}
// the allocation holding [p, p + n), or die
void check_range(const std::map<char*, size_t>& m, const void* p, size_t n, const char* what) {
  char* c = (char*)p;
  auto it = m.upper_bound(c);
  if (it == m.begin()) die(std::string(what) + ": address outside every allocation");
  --it;
  if (c + n > it->first + it->second) die(std::string(what) + ": range past the end of its allocation");
}

}  // namespace

bool available() { return true; }
void init(int device) {
  State& s = S();
  if (s.init && device != s.device) throw std::runtime_error("dev::init: the runtime already uses another device");
  s.init = true;
  s.device = device;
}
bool initialized() { return S().init; }
int device_id() { return S().device; }
void* stream() {
  if (!S().init) throw std::runtime_error("device not set up (call setupDevices first)");
  return (void*)&S();
}

void* malloc_device(size_t bytes) {
  const char* lim = std::getenv("MAPX_DOUBLE_HBM");  // a device this small refuses larger totals
  size_t held = 0;
  for (auto& a : S().dev_allocs) held += a.second;
  if (lim && held + bytes > (size_t)std::strtoull(lim, nullptr, 10)) throw std::runtime_error("hipMalloc: out of memory");
  char* p = (char*)std::calloc(bytes ? bytes : 1, 1);
  if (!p) throw std::runtime_error("hipMalloc: out of memory");
  S().dev_allocs[p] = bytes;
  return p;
}
void free_device(void* p) {
  if (!p) return;
  State& s = S();
  auto it = s.dev_allocs.find((char*)p);
  if (it == s.dev_allocs.end()) die("free_device of an address that is not a live device allocation");
  if (s.recording && s.pending > 0) die("free_device while recorded launches are pending (they would read it)");
  s.dev_allocs.erase(it);
  std::free(p);
}
void* malloc_host(size_t bytes) {
  char* p = (char*)std::calloc(bytes ? bytes : 1, 1);
  if (!p) throw std::runtime_error("hipHostMalloc: out of memory");
  S().host_allocs[p] = bytes;
  return p;
}
void free_host(void* p) {
  if (!p) return;
  auto it = S().host_allocs.find((char*)p);
  if (it == S().host_allocs.end()) die("free_host of an address that is not a live pinned allocation");
  S().host_allocs.erase(it);
  std::free(p);
}
void h2d(void* d, const void* h, size_t n) {
  check_range(S().dev_allocs, d, n, "h2d destination");
  std::memcpy(d, h, n);
}
void d2h(void* h, const void* d, size_t n) {
  if (S().recording && S().pending > 0) die("d2h while recorded launches are pending (it would read before they run)");
  check_range(S().dev_allocs, d, n, "d2h source");
  std::memcpy(h, d, n);
}
void d2d(void* d, const void* s, size_t n) {
  check_range(S().dev_allocs, d, n, "d2d destination");
  check_range(S().dev_allocs, s, n, "d2d source");
  std::memmove(d, s, n);
}
void sync() {
  if (S().recording && S().pending > 0) die("sync while recorded launches are pending (the wait would not cover them)");
}
size_t mem_free() { return (size_t)64 << 30; }
size_t mem_total() { return (size_t)288 << 30; }

std::vector<char> compile(const std::string& source, std::string* log) {
  if (log) log->clear();
  return std::vector<char>(source.begin(), source.begin() + std::min<size_t>(source.size(), 16));
}
Function function(const std::string& source, const std::string& name) {
  if (source.find(name) == std::string::npos) throw std::runtime_error("function: " + name + " not in its source");
  static std::map<std::string, int> ids;
  Function f;
  f.fn = (void*)(uintptr_t)(0x1000 + 16 * (ids.emplace(name, (int)ids.size()).first->second));
  f.name = name;
  return f;
}

void launch(const Function& f, unsigned grid, unsigned block, void** params, const size_t* sizes, int nparams,
            unsigned grid_y) {
  State& s = S();
  if (!s.init) die("launch before init");
  if (!f.fn || grid == 0 || block == 0 || grid_y == 0) die("launch of " + f.name + " with an empty grid");
  for (int i = 0; i < nparams; i++)
    if (!params[i] && sizes[i]) die("launch of " + f.name + ": null argument pointer");
  if (s.recording) s.pending++;
  if (s.profile) {
    KernelTime& k = s.times[f.name];
    k.name = f.name;
    k.count++;
  }
}

void record_begin() {
  S().recording = true;
  S().pending = 0;
}
void record_flush() {
  if (S().pending) S().graphs++;
  S().recording = false;
  S().pending = 0;
}
bool recording() { return S().recording; }
size_t graph_cache_size() { return (size_t)S().graphs; }
void graph_clear() { S().graphs = 0; }

void* event_record() { return new int(1); }
bool event_done(void* e) { return e != nullptr; }
void event_destroy(void* e) { delete (int*)e; }
void event_sync(void*) {}

Timer::Timer() {}
Timer::~Timer() {}
void Timer::start() {}
void Timer::stop() {}
float Timer::ms() { return 0.0f; }

void profile_set(bool on) { S().profile = on; }
bool profile_on() { return S().profile; }
std::vector<KernelTime> profile_take() {
  std::vector<KernelTime> v;
  for (auto& k : S().times) v.push_back(k.second);
  S().times.clear();
  return v;
}

int num_cus() { return 256; }
int occupancy(const Function&, int) { return 2; }

const std::string& header_source() {
  static const std::string h(kDevHeader);
  return h;
}
std::string cache_dir() { return "/tmp"; }

}  // namespace dev

// ahead-of-time device entry points: bounds-checked, nothing computed
namespace gen {
size_t dem_scratch_bytes(int ng) { return (size_t)(ng > 0 ? ng : 1) * 40; }
size_t voronoi_scratch_bytes(int np) { return (size_t)(np > 0 ? np : 1) * 8; }
size_t nalloc_scratch_bytes() { return 256 * 40; }
void* nalloc_scratch(int) {
  static void* p = dev::malloc_device(nalloc_scratch_bytes());
  return p;
}
void nalloc_scratch_init(int) {}
void dem(void* out, int BS0, int BS1, int NB0, int NB1, int, int, const double*, int ng, uint64_t, void* scratch,
         void*) {
  dev::h2d(out, std::vector<char>((size_t)BS0 * BS1 * NB0 * NB1 * 4).data(), (size_t)BS0 * BS1 * NB0 * NB1 * 4);
  (void)ng; (void)scratch;
}
void voronoi(void* out, int dt_bytes, bool, int BS0, int BS1, int NB0, int NB1, int, int, const int*, const int*, int,
             void*, void*) {
  const size_t n = (size_t)BS0 * BS1 * NB0 * NB1 * dt_bytes;
  dev::h2d(out, std::vector<char>(n).data(), n);
}
}  // namespace gen
namespace radial {
int run(void* out, const void* in, int dt_bytes, int, int BS0, int BS1, int NB0, int NB1, int, int, int, int, void*) {
  const size_t n = (size_t)BS0 * BS1 * NB0 * NB1 * dt_bytes;
  dev::d2d(out, in, n);
  return 1;
}
}  // namespace radial
namespace spread {
int run(void* out, const void* in, const void* dir, void* cnt, int dt, int BS0, int BS1, int NB0, int NB1, int, int,
        void*) {
  (void)dir; (void)dt;
  const size_t cells = (size_t)BS0 * BS1 * NB0 * NB1;
  std::vector<char> z(cells * 8 + 8);
  dev::h2d(cnt, z.data(), z.size());
  (void)out; (void)in;
  return 1;
}
}  // namespace spread
namespace mathcheck {
void run(const std::string&, uint64_t, uint64_t, void*, void*, double* max_ulp, unsigned long long* n_over,
         unsigned long long* n, unsigned* worst_in) {
  *max_ulp = 0; *n_over = 0; *n = 0; *worst_in = 0;
}
}  // namespace mathcheck
}  // namespace mapx
