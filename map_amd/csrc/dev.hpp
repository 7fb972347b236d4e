// Device layer over HIP: device selection, the compute stream, memory, hiprtc
// compilation with an on-disk code-object cache, kernel launches and
// hipGraph capture of launch sequences.
//
// Replaces cle/OclEnv and the raw OpenCL call sites of the reference:
//   Runtime.cpp:111-148   clenv.init / clCreateContext / clCreateCommandQueue
//   Cache.cpp:61,66,141   clCreateBuffer / clCreateSubBuffer
//   Cache.cpp:158-160     pinned staging buffers
//   Block.cpp:123,137     clEnqueueWriteBuffer / clEnqueueReadBuffer
//   Version.cpp:57,97-113 clCreateProgramWithSource / clBuildProgram / clCreateKernel
//   Task.cpp:344-385      clSetKernelArg / clEnqueueNDRangeKernel / clFinish
// Errors are reported with exceptions inside the library and with return
// codes plus mx_dev_last_error() across the C ABI; nothing aborts.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace mapx {
namespace dev {

struct Function {
  void* fn = nullptr;  // hipFunction_t
  std::string name;
};

bool available();                 // a HIP device is present
void init(int device);            // select the device, create the compute stream
bool initialized();
int device_id();
void* stream();                   // hipStream_t of the compute stream

void* malloc_device(size_t bytes);
void free_device(void* p);
void* malloc_host(size_t bytes);  // pinned
void free_host(void* p);
void h2d(void* dst, const void* src, size_t bytes);   // async on the compute stream
void d2h(void* dst, const void* src, size_t bytes);   // async on the compute stream
void d2d(void* dst, const void* src, size_t bytes);
void sync();                       // waits for the compute stream
size_t mem_free();
size_t mem_total();

// Compiles `source` (device code including "mapx_dev.hiph") with hiprtc for
// gfx950; returns the code object. Cached in memory and on disk.
std::vector<char> compile(const std::string& source, std::string* log = nullptr);
// Loads (once) the module of `source` and returns the named kernel.
Function function(const std::string& source, const std::string& name);

// Kernel launch on the compute stream. `params[i]` points at argument i of
// `sizes[i]` bytes (the sizes let recorded launches copy their arguments).
void launch(const Function& f, unsigned grid, unsigned block, void** params, const size_t* sizes, int nparams,
            unsigned grid_y = 1);

// Launch recording for hipGraph replay: while recording, launches are queued;
// flush() replays an identical earlier sequence from its instantiated graph,
// or captures and instantiates a new one.
void record_begin();
void record_flush();
bool recording();
size_t graph_cache_size();
void graph_clear();  // destroys the instantiated graphs (call with the stream idle)

// Events on the compute stream (completion tracking without synchronising).
void* event_record();          // records a new event on the compute stream
bool event_done(void* e);      // the stream has passed the event
void event_destroy(void* e);
void event_sync(void* e);      // waits until the stream has passed the event

// Device timing of the compute stream with HIP events.
struct Timer {
  void* a = nullptr;
  void* b = nullptr;
  Timer();
  ~Timer();
  void start();
  void stop();
  float ms();  // elapsed milliseconds (synchronises on the stop event)
};

// Per-launch kernel timing with HIP events around every direct launch on the
// compute stream (launches replayed from graphs are not timed).
struct KernelTime {
  std::string name;
  int count = 0;
  double total_ms = 0, min_ms = 0, max_ms = 0;
};
void profile_set(bool on);
bool profile_on();
std::vector<KernelTime> profile_take();  // synchronises, returns and clears the timings

// Number of compute units of the device (for grid sizing).
int num_cus();
// Work-groups of `block` threads of this kernel that fit on one CU at once.
int occupancy(const Function& f, int block);

const std::string& header_source();  // text of mapx_dev.hiph embedded at build time
std::string cache_dir();

}  // namespace dev
}  // namespace mapx
