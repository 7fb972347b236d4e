// Spreading scan (the reference's SpreadScan with SUM: runtime/dag/SpreadScan.cpp,
// runtime/skeleton/SpreadingSkeleton.cpp:288-388, runtime/task/SpreadingTask.cpp).
// Each cell c holds a D8 direction code dir[c] (1 E, 2 SE, 4 S, 8 SW, 16 W,
// 32 NW, 64 N, 128 NE; y grows south); its downstream cell is c + off(dir[c])
// when the code is one of the eight and the cell lies inside the raster. The
// total that flows through c is
//   T(c) = in(c) + sum of T(u) over the upstream neighbours u of c
// (the neighbours whose downstream cell is c, in the order y outer, x inner),
// in the raster's C type, and the result is out(c) = T(c) when c has a
// downstream cell: the amount that leaves c. A cell without one keeps the
// neutral value 0, as in the reference, whose iterations move the amounts
// downstream until nothing changes and add to each cell what leaves it.
//
// The reference iterates over the whole raster until it is stable (one
// step per cell of the longest flow path). Here every cell is visited once:
//   k_init counts each cell's upstream neighbours (and marks the cells with
//   none, the sources) and clears the output;
//   k_walk starts one walk per source. A walk computes T(c) from the finished
//   upstream totals, stores it, and decrements the count of c's downstream
//   cell with an acquire-release atomic; the walk that takes the count to zero
//   has seen every other upstream total of that cell (each was stored before
//   its own decrement) and continues there; the others end.
// Each cell is computed by exactly one walk, all walks end (each step takes a
// counter that only falls), and the sums follow the fixed neighbour order, so
// the result does not depend on the schedule. Cells on a direction cycle, or
// downstream of one, never become ready and keep 0 (the reference's iteration
// does not end for them).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#pragma clang fp contract(off)

namespace mapx {
namespace spread {

constexpr int NT = 256;
constexpr unsigned SRC = 0x100u;  // count word of a source cell (no upstream neighbour); never decremented

struct Geo {
  int BS0, BS1, NB0, NB1, DS0, DS1;
};

// block-tiled element index of cell (x, y)
__device__ __forceinline__ uint64_t at(const Geo& g, int x, int y) {
  const int bx = x / g.BS0, by = y / g.BS1;
  return ((uint64_t)(by * g.NB0 + bx) * g.BS1 + (uint64_t)(y - by * g.BS1)) * g.BS0 + (uint64_t)(x - bx * g.BS0);
}

// code of the neighbour at (i0, i1) from a cell when it flows into that cell
// (the reference's dircod table {{2,4,8},{1,0,16},{128,64,32}}, skeleton/util.cpp:628)
__device__ __forceinline__ unsigned inflow_code(int i0, int i1) {
  const unsigned t[9] = {2, 4, 8, 1, 0, 16, 128, 64, 32};
  return t[(i1 + 1) * 3 + (i0 + 1)];
}

// downstream offset of a code; false for anything but the eight codes
__device__ __forceinline__ bool outflow(unsigned code, int& dx, int& dy) {
  switch (code) {
    case 1: dx = 1; dy = 0; return true;
    case 2: dx = 1; dy = 1; return true;
    case 4: dx = 0; dy = 1; return true;
    case 8: dx = -1; dy = 1; return true;
    case 16: dx = -1; dy = 0; return true;
    case 32: dx = -1; dy = -1; return true;
    case 64: dx = 0; dy = -1; return true;
    case 128: dx = 1; dy = -1; return true;
    default: return false;
  }
}

// cell (x, y) of element e of the block-tiled raster; false for block padding
__device__ __forceinline__ bool cell_of(const Geo& g, uint64_t e, int& x, int& y) {
  const uint64_t bsz = (uint64_t)g.BS0 * g.BS1;
  const int b = (int)(e / bsz);
  const int r = (int)(e - (uint64_t)b * bsz);
  const int by = b / g.NB0, bx = b - by * g.NB0;
  x = bx * g.BS0 + r % g.BS0;
  y = by * g.BS1 + r / g.BS0;
  return x < g.DS0 && y < g.DS1;
}

template <class T>
__global__ void __launch_bounds__(NT) k_init(const uint8_t* __restrict__ dir, unsigned* __restrict__ cnt,
                                             T* __restrict__ out, Geo g, uint64_t n) {
  const uint64_t e = (uint64_t)blockIdx.x * NT + threadIdx.x;
  if (e >= n) return;
  int x, y;
  unsigned c = 0;
  if (cell_of(g, e, x, y)) {
    for (int i1 = -1; i1 <= 1; i1++)
      for (int i0 = -1; i0 <= 1; i0++) {
        const int u0 = x + i0, u1 = y + i1;
        if ((i0 | i1) == 0 || u0 < 0 || u1 < 0 || u0 >= g.DS0 || u1 >= g.DS1) continue;
        c += dir[at(g, u0, u1)] == inflow_code(i0, i1);
      }
    if (c == 0) c = SRC;
  }
  cnt[e] = c;
  out[e] = T(0);
}

// Totals move between walks through the device's coherence point: agent-scope
// loads and stores (they bypass the XCD-local L2 caches), ordered by waiting for the
// store to complete before the count is decremented. A release / acquire fence pair
// would do the same by writing back and invalidating a whole L2 at every step.
template <class T> __device__ __forceinline__ T ld_total(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ void st_total(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
__global__ void __launch_bounds__(NT) k_walk(const T* __restrict__ in, const uint8_t* __restrict__ dir,
                                             unsigned* cnt, T* out, Geo g, uint64_t n) {
  const uint64_t e0 = (uint64_t)blockIdx.x * NT + threadIdx.x;
  if (e0 >= n) return;
  if (cnt[e0] != SRC) return;  // only sources start; their word is never written
  int x, y;
  cell_of(g, e0, x, y);
  uint64_t e = e0;
  for (;;) {
    // one round of loads for the cell and its eight neighbours (clamped into the raster;
    // a neighbour outside it never matches), then one round for the upstream totals
    uint64_t en[9];
    bool inr[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const int u0 = x + k % 3 - 1, u1 = y + k / 3 - 1;
      inr[k] = u0 >= 0 && u1 >= 0 && u0 < g.DS0 && u1 < g.DS1;
      en[k] = inr[k] ? at(g, u0, u1) : e;
    }
    unsigned dn[9];
#pragma unroll
    for (int k = 0; k < 9; k++) dn[k] = dir[en[k]];
    const T own = in[e];
    bool up[9];
#pragma unroll
    for (int k = 0; k < 9; k++) up[k] = k != 4 && inr[k] && dn[k] == inflow_code(k % 3 - 1, k / 3 - 1);
    T tv[9];
#pragma unroll
    for (int k = 0; k < 9; k++) tv[k] = ld_total(&out[up[k] ? en[k] : e]);
    T acc = own;
#pragma unroll
    for (int k = 0; k < 9; k++)
      if (up[k]) acc = (T)(acc + tv[k]);
    int dx, dy;
    if (!outflow(dn[4], dx, dy)) return;  // no downstream code: the neutral value stays
    x += dx;
    y += dy;
    if (x < 0 || y < 0 || x >= g.DS0 || y >= g.DS1) return;  // flows off the raster
    st_total(&out[e], acc);
    const uint64_t ed = at(g, x, y);
    // the total has reached the coherence point before the count falls, so the walk
    // that takes the count to zero reads it (its loads are issued after the atomic returns)
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    const unsigned before = __hip_atomic_fetch_add(&cnt[ed], 0xffffffffu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (before != 1u) return;
    e = ed;
  }
}

static void check(hipError_t e, const char* w) {
  if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}

template <class T>
static void launch(void* out, const void* in, const void* dir, void* cnt, const Geo& g, uint64_t n, hipStream_t s) {
  const unsigned grid = (unsigned)((n + NT - 1) / NT);
  hipLaunchKernelGGL((k_init<T>), dim3(grid), dim3(NT), 0, s, (const uint8_t*)dir, (unsigned*)cnt, (T*)out, g, n);
  check(hipGetLastError(), "spread init launch");
  hipLaunchKernelGGL((k_walk<T>), dim3(grid), dim3(NT), 0, s, (const T*)in, (const uint8_t*)dir, (unsigned*)cnt,
                     (T*)out, g, n);
  check(hipGetLastError(), "spread walk launch");
}

// dt: the DataTypeEnum of the raster.
// cnt: 4 bytes per element of the block-tiled raster. Returns the launch count.
int run(void* out, const void* in, const void* dir, void* cnt, int dt, int BS0, int BS1, int NB0, int NB1, int DS0,
        int DS1, void* stream) {
  if (BS0 <= 0 || BS1 <= 0 || NB0 <= 0 || NB1 <= 0 || DS0 <= 0 || DS1 <= 0 || DS0 > NB0 * BS0 || DS1 > NB1 * BS1)
    throw std::runtime_error("spread: bad raster geometry");
  const Geo g{BS0, BS1, NB0, NB1, DS0, DS1};
  const uint64_t n = (uint64_t)BS0 * BS1 * NB0 * NB1;
  if ((n + NT - 1) / NT > 0x7fffffffull) throw std::runtime_error("spread: raster too large for one launch");
  hipStream_t s = (hipStream_t)stream;
  switch (dt) {  // DataTypeEnum: F32 1, F64 2, U8 4, U16 5, U32 6, U64 7, S8 8, S16 9, S32 10, S64 11
    case 1: launch<float>(out, in, dir, cnt, g, n, s); break;
    case 2: launch<double>(out, in, dir, cnt, g, n, s); break;
    case 4: launch<uint8_t>(out, in, dir, cnt, g, n, s); break;
    case 5: launch<uint16_t>(out, in, dir, cnt, g, n, s); break;
    case 6: launch<uint32_t>(out, in, dir, cnt, g, n, s); break;
    case 7: launch<uint64_t>(out, in, dir, cnt, g, n, s); break;
    case 8: launch<int8_t>(out, in, dir, cnt, g, n, s); break;
    case 9: launch<int16_t>(out, in, dir, cnt, g, n, s); break;
    case 10: launch<int32_t>(out, in, dir, cnt, g, n, s); break;
    case 11: launch<int64_t>(out, in, dir, cnt, g, n, s); break;
    default: throw std::runtime_error("spread: unsupported data type");
  }
  return 2;
}

}  // namespace spread
}  // namespace mapx
