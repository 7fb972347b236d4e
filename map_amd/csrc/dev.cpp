#include "dev.hpp"
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

#include "mapx_dev_hiph.inc"  // static const char kDevHeader[]: generated from device/mapx_dev.hiph

namespace mapx {
namespace dev {

namespace {

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::ostringstream s;
    s << what << ": " << hipGetErrorName(e) << " (" << hipGetErrorString(e) << ")";
    throw std::runtime_error(s.str());
  }
}

struct State {
  bool init = false;
  int device = -1;
  hipStream_t stream = nullptr;
  int cus = 0;
  std::mutex mtx;
  std::unordered_map<std::string, hipModule_t> modules;  // by source hash
  // launch recording
  bool recording = false;
  struct Rec {
    hipFunction_t f;
    unsigned grid, grid_y, block;
    std::vector<std::vector<char>> args;
  };
  std::vector<Rec> pending;
  std::unordered_map<std::string, hipGraphExec_t> graphs;  // by launch-sequence key
  // per-launch timing
  bool profile = false;
  struct Ev {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Ev> evs;
  std::vector<hipEvent_t> pool;
};
State& S() {
  static State s;
  return s;
}

std::string fnv(const std::string& s) {
  uint64_t h = 1469598103934665603ull;
  for (unsigned char c : s) { h ^= c; h *= 1099511628211ull; }
  char buf[24];
  std::snprintf(buf, sizeof buf, "%016llx", (unsigned long long)h);
  return buf;
}

const char* kOpts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                       "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
const int kNumOpts = sizeof(kOpts) / sizeof(kOpts[0]);

}  // namespace

const std::string& header_source() {
  static const std::string h(kDevHeader);
  return h;
}

std::string cache_dir() {
  const char* e = std::getenv("MAPX_KCACHE");
  std::string d;
  if (e && *e) d = e;
  else {
    const char* t = std::getenv("TMPDIR");
    d = std::string(t && *t ? t : "/tmp") + "/mapx-kcache-" + std::to_string((long)getuid());
  }
  mkdir(d.c_str(), 0700);
  return d;
}

bool available() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) { (void)hipGetLastError(); return false; }
  return n > 0;
}

void init(int device) {
  State& s = S();
  std::lock_guard<std::mutex> g(s.mtx);
  if (s.init) {
    if (device != s.device) throw std::runtime_error("dev::init: the runtime already uses another device");
    return;
  }
  int n = 0;
  check(hipGetDeviceCount(&n), "hipGetDeviceCount");
  if (device < 0 || device >= n) throw std::runtime_error("dev::init: no such device");
  check(hipSetDevice(device), "hipSetDevice");
  check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate");
  hipDeviceProp_t p;
  check(hipGetDeviceProperties(&p, device), "hipGetDeviceProperties");
  s.cus = p.multiProcessorCount;
  s.device = device;
  s.init = true;
}

bool initialized() { return S().init; }
int device_id() { return S().device; }
void* stream() {
  if (!S().init) throw std::runtime_error("device not set up (call setupDevices first)");
  return S().stream;
}
int num_cus() { return S().cus; }

int occupancy(const Function& f, int block) {
  int n = 0;
  check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&n, (hipFunction_t)f.fn, block, 0),
        "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
  return n > 0 ? n : 1;
}

void* malloc_device(size_t bytes) {
  void* p = nullptr;
  const hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
  if (e != hipSuccess) {
    // a failed allocation leaves its error as the thread's last error; the caller may
    // recover (free cached buffers and retry), so a later launch check must not see it
    (void)hipGetLastError();
    check(e, "hipMalloc");
  }
  return p;
}
void free_device(void* p) { if (p) check(hipFree(p), "hipFree"); }
void* malloc_host(size_t bytes) {
  void* p = nullptr;
  check(hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault), "hipHostMalloc");
  return p;
}
void free_host(void* p) { if (p) check(hipHostFree(p), "hipHostFree"); }
void h2d(void* d, const void* h, size_t n) {
  check(hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, (hipStream_t)stream()), "hipMemcpyAsync H2D");
}
void d2h(void* h, const void* d, size_t n) {
  check(hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, (hipStream_t)stream()), "hipMemcpyAsync D2H");
}
void d2d(void* d, const void* s, size_t n) {
  check(hipMemcpyAsync(d, s, n, hipMemcpyDeviceToDevice, (hipStream_t)stream()), "hipMemcpyAsync D2D");
}
void sync() { check(hipStreamSynchronize((hipStream_t)stream()), "hipStreamSynchronize"); }
size_t mem_free() { size_t f = 0, t = 0; check(hipMemGetInfo(&f, &t), "hipMemGetInfo"); return f; }
size_t mem_total() { size_t f = 0, t = 0; check(hipMemGetInfo(&f, &t), "hipMemGetInfo"); return t; }

std::vector<char> compile(const std::string& source, std::string* log) {
  const std::string key = fnv(source + "|" + header_source());
  const std::string path = cache_dir() + "/" + key + ".co";
  {
    std::ifstream f(path, std::ios::binary);
    if (f) {
      std::vector<char> co((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
      if (!co.empty()) return co;
    }
  }
  hiprtcProgram prog;
  const char* hdr = header_source().c_str();
  const char* hname = "mapx_dev.hiph";
  if (hiprtcCreateProgram(&prog, source.c_str(), "mapx_group.hip", 1, &hdr, &hname) != HIPRTC_SUCCESS)
    throw std::runtime_error("hiprtcCreateProgram failed");
  hiprtcResult r = hiprtcCompileProgram(prog, kNumOpts, kOpts);
  size_t ls = 0;
  hiprtcGetProgramLogSize(prog, &ls);
  std::string lg(ls, '\0');
  if (ls) hiprtcGetProgramLog(prog, &lg[0]);
  if (log) *log = lg;
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    throw std::runtime_error("hiprtc compilation failed: " + std::string(hiprtcGetErrorString(r)) + "\n" + lg);
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  std::vector<char> co(cs);
  hiprtcGetCode(prog, co.data());
  hiprtcDestroyProgram(&prog);
  // atomic publish into the disk cache
  std::string tmp = path + ".tmp" + std::to_string((long)getpid());
  {
    std::ofstream f(tmp, std::ios::binary);
    f.write(co.data(), (std::streamsize)co.size());
  }
  std::rename(tmp.c_str(), path.c_str());
  return co;
}

Function function(const std::string& source, const std::string& name) {
  State& s = S();
  if (!s.init) throw std::runtime_error("device not set up (call setupDevices first)");
  const std::string key = fnv(source);
  hipModule_t mod = nullptr;
  {
    std::lock_guard<std::mutex> g(s.mtx);
    auto it = s.modules.find(key);
    if (it != s.modules.end()) mod = it->second;
  }
  if (!mod) {
    std::vector<char> co = compile(source);
    check(hipModuleLoadData(&mod, co.data()), "hipModuleLoadData");
    std::lock_guard<std::mutex> g(s.mtx);
    s.modules[key] = mod;
  }
  hipFunction_t f = nullptr;
  check(hipModuleGetFunction(&f, mod, name.c_str()), "hipModuleGetFunction");
  Function r;
  r.fn = (void*)f;
  r.name = name;
  return r;
}

void launch(const Function& f, unsigned grid, unsigned block, void** params, const size_t* sizes, int nparams,
            unsigned grid_y) {
  State& s = S();
  if (grid == 0 || grid_y == 0) return;
  if (s.recording) {
    State::Rec r;
    r.f = (hipFunction_t)f.fn;
    r.grid = grid;
    r.grid_y = grid_y;
    r.block = block;
    for (int i = 0; i < nparams; i++)
      r.args.emplace_back((const char*)params[i], (const char*)params[i] + sizes[i]);
    s.pending.push_back(std::move(r));
    return;
  }
  State::Ev ev;
  if (s.profile) {
    for (hipEvent_t* e : {&ev.a, &ev.b}) {
      if (!s.pool.empty()) { *e = s.pool.back(); s.pool.pop_back(); }
      else check(hipEventCreate(e), "hipEventCreate");
    }
    ev.name = f.name;
    check(hipEventRecord(ev.a, s.stream), "hipEventRecord");
  }
  check(hipModuleLaunchKernel((hipFunction_t)f.fn, grid, grid_y, 1, block, 1, 1, 0, s.stream, params, nullptr),
        "hipModuleLaunchKernel");
  if (s.profile) {
    check(hipEventRecord(ev.b, s.stream), "hipEventRecord");
    s.evs.push_back(ev);
  }
}

void profile_set(bool on) { S().profile = on; }
bool profile_on() { return S().profile; }

std::vector<KernelTime> profile_take() {
  State& s = S();
  std::vector<KernelTime> out;
  std::unordered_map<std::string, size_t> at;
  for (auto& e : s.evs) {
    check(hipEventSynchronize(e.b), "hipEventSynchronize");
    float ms = 0;
    check(hipEventElapsedTime(&ms, e.a, e.b), "hipEventElapsedTime");
    auto it = at.find(e.name);
    if (it == at.end()) {
      at[e.name] = out.size();
      KernelTime k;
      k.name = e.name;
      k.min_ms = k.max_ms = ms;
      out.push_back(k);
      it = at.find(e.name);
    }
    KernelTime& k = out[it->second];
    k.count++;
    k.total_ms += ms;
    if (ms < k.min_ms) k.min_ms = ms;
    if (ms > k.max_ms) k.max_ms = ms;
    s.pool.push_back(e.a);
    s.pool.push_back(e.b);
  }
  s.evs.clear();
  return out;
}

void record_begin() {
  State& s = S();
  if (s.recording) throw std::runtime_error("launch recording already active");
  s.recording = true;
  s.pending.clear();
}

bool recording() { return S().recording; }
size_t graph_cache_size() { return S().graphs.size(); }

void graph_clear() {
  State& s = S();
  for (auto& g : s.graphs) (void)hipGraphExecDestroy(g.second);
  s.graphs.clear();
}

void record_flush() {
  State& s = S();
  if (!s.recording) return;
  s.recording = false;
  if (s.pending.empty()) return;
  std::string key;
  key.reserve(s.pending.size() * 64);
  for (auto& r : s.pending) {
    key.append((const char*)&r.f, sizeof r.f);
    key.append((const char*)&r.grid, sizeof r.grid);
    key.append((const char*)&r.grid_y, sizeof r.grid_y);
    key.append((const char*)&r.block, sizeof r.block);
    for (auto& a : r.args) { key.append(a.data(), a.size()); key.push_back('\x1f'); }
    key.push_back('\x1e');
  }
  auto it = s.graphs.find(key);
  hipGraphExec_t exec = nullptr;
  if (it != s.graphs.end()) exec = it->second;
  else {
    check(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    for (auto& r : s.pending) {
      std::vector<void*> ptrs;
      for (auto& a : r.args) ptrs.push_back((void*)a.data());
      hipError_t e = hipModuleLaunchKernel(r.f, r.grid, r.grid_y, 1, r.block, 1, 1, 0, s.stream, ptrs.data(), nullptr);
      if (e != hipSuccess) {
        hipGraph_t g;
        (void)hipStreamEndCapture(s.stream, &g);
        check(e, "hipModuleLaunchKernel (capture)");
      }
    }
    hipGraph_t graph = nullptr;
    check(hipStreamEndCapture(s.stream, &graph), "hipStreamEndCapture");
    check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    check(hipGraphDestroy(graph), "hipGraphDestroy");
    if (s.graphs.size() > 64) {  // bounded cache
      for (auto& g : s.graphs) (void)hipGraphExecDestroy(g.second);
      s.graphs.clear();
    }
    s.graphs[key] = exec;
  }
  s.pending.clear();
  check(hipGraphLaunch(exec, s.stream), "hipGraphLaunch");
}

void* event_record() {
  hipEvent_t e = nullptr;
  check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
  check(hipEventRecord(e, S().stream), "hipEventRecord");
  return (void*)e;
}
bool event_done(void* e) {
  hipError_t r = hipEventQuery((hipEvent_t)e);
  if (r == hipSuccess) return true;
  if (r == hipErrorNotReady) return false;
  check(r, "hipEventQuery");
  return false;
}
void event_destroy(void* e) { if (e) (void)hipEventDestroy((hipEvent_t)e); }
void event_sync(void* e) { check(hipEventSynchronize((hipEvent_t)e), "hipEventSynchronize"); }

Timer::Timer() {
  check(hipEventCreate((hipEvent_t*)&a), "hipEventCreate");
  check(hipEventCreate((hipEvent_t*)&b), "hipEventCreate");
}
Timer::~Timer() {
  if (a) (void)hipEventDestroy((hipEvent_t)a);
  if (b) (void)hipEventDestroy((hipEvent_t)b);
}
void Timer::start() { check(hipEventRecord((hipEvent_t)a, (hipStream_t)stream()), "hipEventRecord"); }
void Timer::stop() { check(hipEventRecord((hipEvent_t)b, (hipStream_t)stream()), "hipEventRecord"); }
float Timer::ms() {
  check(hipEventSynchronize((hipEvent_t)b), "hipEventSynchronize");
  float m = 0;
  check(hipEventElapsedTime(&m, (hipEvent_t)a, (hipEvent_t)b), "hipEventElapsedTime");
  return m;
}

}  // namespace dev
}  // namespace mapx
