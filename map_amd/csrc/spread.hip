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
// step per cell of the longest flow path). Here every cell is computed once,
// in three passes:
//   k_tile: a work-group per 64x64 tile computes, in LDS, every cell whose
//     upstream cells all lie in the tile ("closed" cells: their totals need
//     nothing from outside), in rounds of cells whose upstream is done, and
//     marks the others "open" (a cell with an upstream neighbour outside the
//     tile, or downstream of an open cell, or on a direction cycle);
//   k_count: every open cell counts its open upstream neighbours and lists the
//     sources (open cells with none);
//   k_walk: a persistent grid walks from the listed sources. A walk computes T(c) from the finished
//     upstream totals, stores it, and decrements the count of c's downstream
//     cell; the walk that takes the count to zero has seen every other
//     upstream total of that cell (each was stored before its own decrement)
//     and continues there; the others end.
// Each cell is computed exactly once, every walk ends (each step takes a
// counter that only falls), and every sum follows the fixed neighbour order,
// so the result does not depend on the schedule. Cells on a direction cycle
// never have all their upstream totals and keep 0 (the reference's iteration
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

constexpr unsigned CLOSED = 0x200u;  // status word of a cell computed by k_tile (or block padding)
constexpr int TS = 64;               // tile side of k_tile
constexpr int CPT = TS * TS / NT;    // cells per thread in k_tile

template <class T>
struct TileLds {
  T tot[TS * TS];              // totals of the tile's cells (input values until computed)
  uint8_t dir[TS + 2][TS + 2];  // directions with a one-cell ring (0 outside the raster)
  unsigned left[TS * TS];      // in-tile upstream neighbours not yet computed
  uint8_t open[TS * TS];       // needs a total from outside the tile
  uint8_t done[TS * TS];
  int any[2];                  // a round computed something (two slots: one is reset while the other is read)
};

template <class T>
__global__ void __launch_bounds__(NT) k_tile(const T* __restrict__ in, const uint8_t* __restrict__ dir,
                                             unsigned* __restrict__ cnt, T* __restrict__ out, unsigned* queue, Geo g,
                                             int ntx) {
  __shared__ TileLds<T> L;
  const int tx0 = (int)(blockIdx.x % (unsigned)ntx) * TS, ty0 = (int)(blockIdx.x / (unsigned)ntx) * TS;
  const int t = threadIdx.x;
  if (blockIdx.x == 0 && t == 0) queue[0] = queue[1] = 0;  // sources listed, sources taken
  for (int i = t; i < (TS + 2) * (TS + 2); i += NT) {
    const int lx = i % (TS + 2) - 1, ly = i / (TS + 2) - 1;
    const int x = tx0 + lx, y = ty0 + ly;
    L.dir[ly + 1][lx + 1] = (x >= 0 && y >= 0 && x < g.DS0 && y < g.DS1) ? dir[at(g, x, y)] : 0;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < CPT; j++) {
    const int i = t + j * NT, lx = i % TS, ly = i / TS;
    const int x = tx0 + lx, y = ty0 + ly;
    const bool inr = x < g.DS0 && y < g.DS1;
    unsigned kin = 0, ext = 0;
    for (int i1 = -1; i1 <= 1; i1++)
      for (int i0 = -1; i0 <= 1; i0++) {
        if ((i0 | i1) == 0 || L.dir[ly + 1 + i1][lx + 1 + i0] != inflow_code(i0, i1)) continue;
        const bool inside = lx + i0 >= 0 && lx + i0 < TS && ly + i1 >= 0 && ly + i1 < TS;
        kin += inside;
        ext |= !inside;
      }
    L.left[i] = kin;
    L.open[i] = (uint8_t)ext;
    L.done[i] = !inr;
    L.tot[i] = inr ? in[at(g, x, y)] : T(0);
  }
  if (t == 0) L.any[0] = L.any[1] = 0;
  __syncthreads();
  // rounds: the cells whose in-tile upstream cells are all done; a cell with an open
  // upstream cell is open, any other gets in + its upstream totals in the fixed order
  for (int round = 0;; round++) {
    unsigned ready = 0;
#pragma unroll
    for (int j = 0; j < CPT; j++) {
      const int i = t + j * NT;
      if (!L.done[i] && L.left[i] == 0) ready |= 1u << j;
    }
    __syncthreads();  // the round's set is fixed before any of it changes
    if (ready) {
#pragma unroll
      for (int j = 0; j < CPT; j++) {
        if (!(ready >> j & 1u)) continue;
        const int i = t + j * NT, lx = i % TS, ly = i / TS;
        const bool op = L.open[i];
        if (!op) {
          T acc = L.tot[i];
          for (int i1 = -1; i1 <= 1; i1++)
            for (int i0 = -1; i0 <= 1; i0++)
              if ((i0 | i1) != 0 && L.dir[ly + 1 + i1][lx + 1 + i0] == inflow_code(i0, i1))
                acc = (T)(acc + L.tot[(ly + i1) * TS + lx + i0]);  // inside the tile: the cell is closed
          L.tot[i] = acc;
        }
        L.done[i] = 1;
        int dx, dy;
        if (outflow(L.dir[ly + 1][lx + 1], dx, dy) && lx + dx >= 0 && lx + dx < TS && ly + dy >= 0 && ly + dy < TS) {
          const int d = (ly + dy) * TS + lx + dx;
          if (op) L.open[d] = 1;
          __hip_atomic_fetch_add(&L.left[d], 0xffffffffu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
      L.any[round & 1] = 1;
    }
    __syncthreads();
    const bool more = L.any[round & 1];
    if (t == 0) L.any[(round + 1) & 1] = 0;  // last read before the previous round's second barrier
    if (!more) break;
  }
  // closed cells are final; open (and cycle) cells start at 0 for the walks
#pragma unroll
  for (int j = 0; j < CPT; j++) {
    const int i = t + j * NT, lx = i % TS, ly = i / TS;
    const int x = tx0 + lx, y = ty0 + ly;
    if (x >= g.DS0 || y >= g.DS1) continue;
    const bool closed = L.done[i] && !L.open[i];
    int dx, dy;
    const bool down = outflow(L.dir[ly + 1][lx + 1], dx, dy) && x + dx >= 0 && x + dx < g.DS0 && y + dy >= 0 &&
                      y + dy < g.DS1;
    const uint64_t e = at(g, x, y);
    out[e] = closed && down ? L.tot[i] : T(0);
    cnt[e] = closed ? CLOSED : 0u;
  }
}

// Upstream neighbours of (x, y) that are open: their word is not CLOSED (it is 0,
// open and not yet counted, or a count, never CLOSED).
__device__ __forceinline__ unsigned open_upstream(const uint8_t* dir, unsigned* cnt, const Geo& g, int x, int y) {
  unsigned c = 0;
  for (int i1 = -1; i1 <= 1; i1++)
    for (int i0 = -1; i0 <= 1; i0++) {
      const int u0 = x + i0, u1 = y + i1;
      if ((i0 | i1) == 0 || u0 < 0 || u1 < 0 || u0 >= g.DS0 || u1 >= g.DS1) continue;
      const uint64_t eu = at(g, u0, u1);
      c += dir[eu] == inflow_code(i0, i1) && __hip_atomic_load(&cnt[eu], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != CLOSED;
    }
  return c;
}

// Open cells count their open upstream neighbours (a word that is not CLOSED) and the
// sources among them (none) are listed in queue[2..]; block padding becomes CLOSED.
__global__ void __launch_bounds__(NT) k_count(const uint8_t* __restrict__ dir, unsigned* cnt, unsigned* queue, Geo g,
                                              uint64_t n) {
  const uint64_t e = (uint64_t)blockIdx.x * NT + threadIdx.x;
  bool src = false;
  if (e < n) {
    int x, y;
    if (!cell_of(g, e, x, y)) {
      cnt[e] = CLOSED;
    } else if (cnt[e] != CLOSED) {
      const unsigned c = open_upstream(dir, cnt, g, x, y);
      __hip_atomic_store(&cnt[e], c == 0 ? SRC : c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      src = c == 0;
    }
  }
  // one atomic per wavefront reserves the slots of its sources
  const unsigned long long m = __ballot(src);
  if (m == 0) return;
  const int lane = __lane_id();
  unsigned base = 0;
  if (lane == __ffsll((long long)m) - 1) base = atomicAdd(&queue[0], (unsigned)__popcll(m));
  base = __shfl(base, __ffsll((long long)m) - 1, 64);
  if (src) queue[2 + base + __popcll(m & ((1ull << lane) - 1))] = (unsigned)e;
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

// A walk carries the total of the cell it leaves in a register. Entering a cell whose
// count is 1 (every other upstream cell is final, or has already handed over), it goes
// on without an atomic; at a confluence it stores its total, waits for the store,
// then decrements the count, and only the walk that takes it to zero goes on.
template <class T>
__device__ void walk(const T* __restrict__ in, const uint8_t* __restrict__ dir, unsigned* cnt, T* out, const Geo& g,
                     uint64_t e) {
  int x, y;
  cell_of(g, e, x, y);
  uint64_t from = ~0ull;  // the cell the walk came from and its total
  T carried = T(0);
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
    for (int k = 0; k < 9; k++) tv[k] = ld_total(&out[up[k] && en[k] != from ? en[k] : e]);
    T acc = own;
#pragma unroll
    for (int k = 0; k < 9; k++)
      if (up[k]) acc = (T)(acc + (en[k] == from ? carried : tv[k]));
    int dx, dy;
    if (!outflow(dn[4], dx, dy)) return;  // no downstream code: the neutral value stays
    x += dx;
    y += dy;
    if (x < 0 || y < 0 || x >= g.DS0 || y >= g.DS1) return;  // flows off the raster
    st_total(&out[e], acc);
    const uint64_t ed = at(g, x, y);
    if (__hip_atomic_load(&cnt[ed], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 1u) {
      // the total has reached the coherence point before the count falls, so the walk
      // that takes the count to zero reads it (its loads are issued after the atomic returns)
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      const unsigned before =
          __hip_atomic_fetch_add(&cnt[ed], 0xffffffffu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (before != 1u) return;
    }
    from = e;
    carried = acc;
    e = ed;
  }
}

// A persistent grid: each thread takes the next listed source when its walk ends,
// so long walks do not hold back the start of the others.
template <class T>
__global__ void __launch_bounds__(NT) k_walk(const T* __restrict__ in, const uint8_t* __restrict__ dir,
                                             unsigned* cnt, T* out, unsigned* queue, Geo g) {
  const unsigned total = queue[0];
  for (;;) {
    const unsigned i = atomicAdd(&queue[1], 1u);
    if (i >= total) return;
    walk<T>(in, dir, cnt, out, g, queue[2 + i]);
  }
}

static void check(hipError_t e, const char* w) {
  if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}

template <class T>
static void launch(void* out, const void* in, const void* dir, void* cnt, const Geo& g, uint64_t n, hipStream_t s) {
  unsigned* queue = (unsigned*)cnt + n;  // 2 counters, then up to n listed sources
  const int ntx = (g.DS0 + TS - 1) / TS, nty = (g.DS1 + TS - 1) / TS;
  hipLaunchKernelGGL((k_tile<T>), dim3((unsigned)(ntx * nty)), dim3(NT), 0, s, (const T*)in, (const uint8_t*)dir,
                     (unsigned*)cnt, (T*)out, queue, g, ntx);
  check(hipGetLastError(), "spread tile launch");
  const unsigned grid = (unsigned)((n + NT - 1) / NT);
  hipLaunchKernelGGL(k_count, dim3(grid), dim3(NT), 0, s, (const uint8_t*)dir, (unsigned*)cnt, queue, g, n);
  check(hipGetLastError(), "spread count launch");
  int dev = 0, cus = 0, per_cu = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
  check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_walk<T>), NT, 0),
        "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  const unsigned wgrid = (unsigned)((cus > 0 ? cus : 1) * (per_cu > 0 ? per_cu : 1));
  hipLaunchKernelGGL((k_walk<T>), dim3(wgrid), dim3(NT), 0, s, (const T*)in, (const uint8_t*)dir, (unsigned*)cnt,
                     (T*)out, queue, g);
  check(hipGetLastError(), "spread walk launch");
}

// dt: the DataTypeEnum of the raster.
// cnt: 8 bytes per element of the block-tiled raster, plus 8. Returns the launch count.
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
  return 3;
}

}  // namespace spread
}  // namespace mapx
