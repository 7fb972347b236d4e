// Synthetic input rasters generated in HBM (block-tiled layout):
//   DEM:    sum of 2D Gaussians (parameters drawn by the caller from a seeded
//           PCG64 stream) plus 8-octave value-noise fBm scaled to +-20 m,
//           computed in FP64 and rounded once to FP32.
//   zones:  Voronoi cells of integer seed points, nearest by exact int64
//           squared distance, ties to the lowest point index.
// The cell functions are __host__ __device__ so that the same definition is
// documented in one place; the oracle holds an independent CPU restatement.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace mapx {
namespace gen {

// every product and sum rounds on its own, as in the host twin of these generators
// (compiled without contraction), so both give the same FP64 value up to exp()
#pragma clang fp contract(off)

struct Gauss {
  double cx, cy, inv2s2, amp, r2max;
};

__host__ __device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Lattice value in [-1, 1) of octave o at integer lattice point (i, j).
__host__ __device__ inline double lattice(uint64_t seed, int o, int64_t i, int64_t j) {
  uint64_t z = splitmix64(seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(o + 1)));
  z = splitmix64(z ^ (uint64_t)i * 0x9E3779B97F4A7C15ull);
  z = splitmix64(z ^ (uint64_t)j * 0xC2B2AE3D27D4EB4Full);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

__host__ __device__ inline double fbm(uint64_t seed, double x, double y) {
  double sum = 0.0, norm = 0.0, amp = 1.0, period = 256.0;
  for (int o = 0; o < 8; o++) {
    const double u = x / period, v = y / period;
    const double fi = floor(u), fj = floor(v);
    const double fu = u - fi, fv = v - fj;
    const double su = fu * fu * (3.0 - 2.0 * fu), sv = fv * fv * (3.0 - 2.0 * fv);
    const int64_t i = (int64_t)fi, j = (int64_t)fj;
    const double a = lattice(seed, o, i, j), b = lattice(seed, o, i + 1, j);
    const double c = lattice(seed, o, i, j + 1), d = lattice(seed, o, i + 1, j + 1);
    const double top = a + (b - a) * su, bot = c + (d - c) * su;
    sum += amp * (top + (bot - top) * sv);
    norm += amp;
    amp *= 0.5;
    period *= 0.5;
  }
  return 20.0 * sum / norm;
}

__host__ __device__ inline float dem_cell(int64_t x, int64_t y, const Gauss* g, int ng, uint64_t seed) {
  double h = 0.0;
  for (int k = 0; k < ng; k++) {
    const double dx = (double)x - g[k].cx, dy = (double)y - g[k].cy;
    const double r2 = dx * dx + dy * dy;
    if (r2 <= g[k].r2max) h += g[k].amp * exp(-r2 * g[k].inv2s2);
  }
  h += fbm(seed, (double)x, (double)y);
  return (float)h;
}

struct Geo {
  int BS0, BS1, NB0, NB1, DS0, DS1;
};

__device__ inline bool cell_of(const Geo& G, uint64_t e, int64_t& gx, int64_t& gy) {
  const uint64_t bsz = (uint64_t)G.BS0 * G.BS1;
  const uint64_t b = e / bsz, r = e % bsz;
  const int bx = (int)(b % G.NB0), by = (int)(b / G.NB0);
  gx = (int64_t)bx * G.BS0 + (int64_t)(r % G.BS0);
  gy = (int64_t)by * G.BS1 + (int64_t)(r / G.BS0);
  return gx < G.DS0 && gy < G.DS1;
}

__global__ void __launch_bounds__(256) k_dem(float* out, Geo G, const Gauss* g, int ng, uint64_t seed, uint64_t n) {
  __shared__ Gauss sg[256];
  for (int i = threadIdx.x; i < ng; i += blockDim.x) sg[i] = g[i];
  __syncthreads();
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
    int64_t x, y;
    out[e] = cell_of(G, e, x, y) ? dem_cell(x, y, sg, ng, seed) : 0.0f;
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_voronoi(T* out, Geo G, const int* px, const int* py, int np, uint64_t n) {
  __shared__ int sx[1024], sy[1024];
  for (int i = threadIdx.x; i < np; i += blockDim.x) { sx[i] = px[i]; sy[i] = py[i]; }
  __syncthreads();
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
    int64_t x, y;
    if (!cell_of(G, e, x, y)) { out[e] = 0; continue; }
    int best = 0;
    int64_t bd = INT64_MAX;
    for (int k = 0; k < np; k++) {
      const int64_t dx = x - sx[k], dy = y - sy[k];
      const int64_t d = dx * dx + dy * dy;
      if (d < bd) { bd = d; best = k; }
    }
    out[e] = (T)best;
  }
}

static void check(hipError_t e, const char* w) {
  if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}

size_t dem_scratch_bytes(int ng) { return sizeof(Gauss) * (size_t)(ng > 0 ? ng : 1); }
size_t voronoi_scratch_bytes(int np) { return 2 * sizeof(int) * (size_t)(np > 0 ? np : 1); }

// params: ng rows of (cx, cy, sigma, amplitude). scratch: a device buffer of
// dem_scratch_bytes(ng) (the Gaussians' table), free to reuse once this returns.
void dem(void* out, int BS0, int BS1, int NB0, int NB1, int DS0, int DS1, const double* params, int ng,
         uint64_t seed, void* scratch, void* stream) {
  if (ng < 0 || ng > 256) throw std::runtime_error("gen dem: at most 256 Gaussians");
  Gauss h[256];  // on the stack: this path allocates nothing
  for (int k = 0; k < ng; k++) {
    const double s = params[4 * k + 2];
    h[k].cx = params[4 * k + 0];
    h[k].cy = params[4 * k + 1];
    h[k].inv2s2 = 1.0 / (2.0 * s * s);
    h[k].amp = params[4 * k + 3];
    h[k].r2max = (12.0 * s) * (12.0 * s);  // exp(-72) * 3000 m is far below FP32 resolution
  }
  Gauss* d = (Gauss*)scratch;
  check(hipMemcpyAsync(d, h, dem_scratch_bytes(ng), hipMemcpyHostToDevice, (hipStream_t)stream), "H2D");
  Geo G{BS0, BS1, NB0, NB1, DS0, DS1};
  const uint64_t n = (uint64_t)BS0 * BS1 * NB0 * NB1;
  const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_dem, dim3(grid), dim3(256), 0, (hipStream_t)stream, (float*)out, G, d, ng, seed, n);
  const hipError_t le = hipGetLastError();
  // the table's host copy must outlive the copy; the scratch is reused after return
  check(hipStreamSynchronize((hipStream_t)stream), "sync");
  check(le, "k_dem launch");
}

// scratch: a device buffer of voronoi_scratch_bytes(np) (the points), free to reuse once this returns.
void voronoi(void* out, int dt_bytes, bool is_signed, int BS0, int BS1, int NB0, int NB1, int DS0, int DS1,
             const int* px, const int* py, int np, void* scratch, void* stream) {
  if (np <= 0 || np > 1024) throw std::runtime_error("gen zones: 1..1024 points");
  if (!(dt_bytes == 1 || dt_bytes == 2 || dt_bytes == 4)) throw std::runtime_error("gen zones: zone type must be 1, 2 or 4 bytes");
  int* dx = (int*)scratch;
  int* dy = dx + np;
  hipStream_t s = (hipStream_t)stream;
  check(hipMemcpyAsync(dx, px, sizeof(int) * np, hipMemcpyHostToDevice, s), "H2D");
  check(hipMemcpyAsync(dy, py, sizeof(int) * np, hipMemcpyHostToDevice, s), "H2D");
  Geo G{BS0, BS1, NB0, NB1, DS0, DS1};
  const uint64_t n = (uint64_t)BS0 * BS1 * NB0 * NB1;
  const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 65536);
  if (dt_bytes == 1) hipLaunchKernelGGL(k_voronoi<uint8_t>, dim3(grid), dim3(256), 0, s, (uint8_t*)out, G, dx, dy, np, n);
  else if (dt_bytes == 2) hipLaunchKernelGGL(k_voronoi<uint16_t>, dim3(grid), dim3(256), 0, s, (uint16_t*)out, G, dx, dy, np, n);
  else if (is_signed) hipLaunchKernelGGL(k_voronoi<int32_t>, dim3(grid), dim3(256), 0, s, (int32_t*)out, G, dx, dy, np, n);
  else hipLaunchKernelGGL(k_voronoi<uint32_t>, dim3(grid), dim3(256), 0, s, (uint32_t*)out, G, dx, dy, np, n);
  const hipError_t le = hipGetLastError();
  check(hipStreamSynchronize(s), "sync");
  check(le, "k_voronoi launch");
}

// ---- scratch of the allocation-free DEM entry point --------------------------------
// ma_genDemNoAlloc keeps the Gaussians' table (at most 256 rows) in this
// file-scope device array instead of memory the library allocates. Each GPU
// holds its own copy until the process ends; every call on a GPU overwrites
// it, so calls that overlap on one GPU race. Host code never passes the
// array's name to a kernel: its device address comes from
// hipGetSymbolAddress, once per device, with that device current.
static __device__ __align__(256) unsigned char scratch[256 * sizeof(Gauss)];
static std::mutex g_scratch_mtx;
static void* g_scratch[64] = {};

size_t nalloc_scratch_bytes() { return sizeof(scratch); }

void nalloc_scratch_init(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mtx);
  if (device < 0 || device >= 64) throw std::runtime_error("gen: device index out of range for the scratch table");
  if (g_scratch[device]) return;
  void* p = nullptr;
  check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  g_scratch[device] = p;
}

void* nalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mtx);
  if (device < 0 || device >= 64 || !g_scratch[device]) throw std::runtime_error("gen: scratch not set up on this device");
  return g_scratch[device];
}

}  // namespace gen
}  // namespace mapx
