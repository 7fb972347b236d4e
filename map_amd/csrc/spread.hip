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
// in two passes:
//   k_tile: a work-group per 32x32 tile computes, in LDS, every cell whose
//     upstream cells all lie in the tile ("closed" cells: their totals need
//     nothing from outside), by walks from the tile's sources with the
//     hand-off described below, and marks the others "open" (a cell with an
//     upstream neighbour outside the tile, or downstream of an open cell, or
//     on a direction cycle); it assembles each open cell's count of open
//     upstream neighbours with atomics and lists the open cells whose count
//     is zero (the walk sources);
//   k_walk: a persistent grid walks from the listed sources. A walk computes
//     T(c) from the finished upstream totals, stores it, and decrements the
//     count of c's downstream cell; the walk that takes the count to zero has
//     seen every other upstream total of that cell (each was stored before
//     its own decrement) and continues there; the others end.
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

struct Geo {
  int BS0, BS1, NB0, NB1, DS0, DS1;
};

// The block-tiled element index of cell (x, y) splits into a row term and a column
// term: ((by*NB0 + bx)*BS1 + ly)*BS0 + lx = rowt(y) + colt(x), with by, ly = y / BS1,
// y % BS1 and bx, lx = x / BS0, x % BS0.
__device__ __forceinline__ uint64_t rowt(const Geo& g, int y) {
  const int by = y / g.BS1;
  return ((uint64_t)by * g.NB0 * g.BS1 + (uint64_t)(y - by * g.BS1)) * g.BS0;
}
__device__ __forceinline__ uint64_t colt(const Geo& g, int x) {
  const int bx = x / g.BS0;
  return (uint64_t)bx * g.BS1 * g.BS0 + (uint64_t)(x - bx * g.BS0);
}
__device__ __forceinline__ uint64_t at(const Geo& g, int x, int y) { return rowt(g, y) + colt(g, x); }

// Element indices of the 3x3 neighbourhood of an inside cell (x, y), k = (i1+1)*3 + (i0+1);
// a neighbour outside the raster gets the cell's own index and inr = false. One
// division pair per axis; the neighbouring rows and columns follow by carries.
__device__ __forceinline__ void hood(const Geo& g, int x, int y, uint64_t (&en)[9], bool (&inr)[9]) {
  const int by = y / g.BS1, ly = y - by * g.BS1, bx = x / g.BS0, lx = x - bx * g.BS0;
  const uint64_t blk = (uint64_t)g.BS1 * g.BS0, row = (uint64_t)g.NB0 * blk;
  uint64_t r[3], c[3];
  r[1] = (uint64_t)by * row + (uint64_t)ly * g.BS0;
  r[0] = ly > 0 ? r[1] - g.BS0 : r[1] - row + (uint64_t)(g.BS1 - 1) * g.BS0;
  r[2] = ly + 1 < g.BS1 ? r[1] + g.BS0 : r[1] + row - (uint64_t)(g.BS1 - 1) * g.BS0;
  c[1] = (uint64_t)bx * blk + lx;
  c[0] = lx > 0 ? c[1] - 1 : c[1] - blk + (g.BS0 - 1);
  c[2] = lx + 1 < g.BS0 ? c[1] + 1 : c[1] + blk - (g.BS0 - 1);
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const int u0 = x + k % 3 - 1, u1 = y + k / 3 - 1;
    inr[k] = u0 >= 0 && u1 >= 0 && u0 < g.DS0 && u1 < g.DS1;
    en[k] = inr[k] ? r[k / 3] + c[k % 3] : r[1] + c[1];
  }
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

// cell (x, y) of element e of the block-tiled raster (e < 2^32); false for block padding
__device__ __forceinline__ bool cell_of(const Geo& g, uint64_t e, int& x, int& y) {
  const unsigned bsz = (unsigned)g.BS0 * (unsigned)g.BS1;
  const unsigned b = (unsigned)e / bsz, r = (unsigned)e - b * bsz;
  const unsigned by = b / (unsigned)g.NB0, bx = b - by * (unsigned)g.NB0;
  const unsigned ry = r / (unsigned)g.BS0;
  x = (int)(bx * g.BS0 + (r - ry * g.BS0));
  y = (int)(by * g.BS1 + ry);
  return x < g.DS0 && y < g.DS1;
}

constexpr int TS = 32;               // tile side of k_tile (46 KB of LDS at 64 left more CUs idle than busy)
constexpr int CPT = TS * TS / NT;    // cells per thread in k_tile

// Cell j of thread t in a tile: row ly = t/TS + (NT/TS)j, column lx = (t - ly) mod TS. Every
// row and every column of the tile meets TS different threads, so the sources along a tile
// edge (where flow enters) start their walks in parallel whatever the flow direction.
__device__ __forceinline__ int tile_cell(int t, int j) {
  const int ly = t / TS + (NT / TS) * j;
  return ly * TS + ((t % TS - ly) & (TS - 1));
}

template <class T>
struct TileLds {
  T tot[TS * TS];              // totals of the tile's cells (input values until computed)
  uint8_t dir[TS + 2][TS + 2];  // directions with a one-cell ring (0 outside the raster)
  unsigned left[TS * TS];      // in-tile upstream neighbours not yet computed
  uint8_t open[TS * TS];       // needs a total from outside the tile
  uint8_t done[TS * TS];
  uint64_t rt[TS + 2], ct[TS + 2];  // row and column terms of the tile and its ring
  unsigned src[TS * TS];       // walk sources found by this work-group
  unsigned nsrc, base;
};

template <class T>
__global__ void __launch_bounds__(NT) k_tile(const T* __restrict__ in, const uint8_t* __restrict__ dir,
                                             unsigned* __restrict__ cnt, T* __restrict__ out, unsigned* queue, Geo g,
                                             int ntx) {
  __shared__ TileLds<T> L;
  const int tx0 = (int)(blockIdx.x % (unsigned)ntx) * TS, ty0 = (int)(blockIdx.x / (unsigned)ntx) * TS;
  const int t = threadIdx.x;
  if (t < TS + 2) L.rt[t] = ty0 + t - 1 >= 0 && ty0 + t - 1 < g.DS1 ? rowt(g, ty0 + t - 1) : 0;
  else if (t < 2 * (TS + 2)) L.ct[t - (TS + 2)] = tx0 + t - (TS + 2) - 1 >= 0 && tx0 + t - (TS + 2) - 1 < g.DS0 ? colt(g, tx0 + t - (TS + 2) - 1) : 0;
  __syncthreads();
  {
    constexpr int RING = (TS + 2) * (TS + 2), PER = (RING + NT - 1) / NT;
    uint8_t v[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {  // every load in flight before the first LDS write
      const int i = t + j * NT, lx = i % (TS + 2), ly = i / (TS + 2);
      const int x = tx0 + lx - 1, y = ty0 + ly - 1;
      v[j] = (i < RING && x >= 0 && y >= 0 && x < g.DS0 && y < g.DS1) ? dir[L.rt[ly] + L.ct[lx]] : 0;
    }
#pragma unroll
    for (int j = 0; j < PER; j++) {
      const int i = t + j * NT;
      if (i < RING) L.dir[i / (TS + 2)][i % (TS + 2)] = v[j];
    }
  }
  __syncthreads();
  T own[CPT];
#pragma unroll
  for (int j = 0; j < CPT; j++) {  // every load in flight before the first is used
    const int i = tile_cell(t, j), lx = i % TS, ly = i / TS;
    own[j] = tx0 + lx < g.DS0 && ty0 + ly < g.DS1 ? in[L.rt[ly + 1] + L.ct[lx + 1]] : T(0);
  }
#pragma unroll
  for (int j = 0; j < CPT; j++) {
    const int i = tile_cell(t, j), lx = i % TS, ly = i / TS;
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
    L.tot[i] = own[j];
  }
  __syncthreads();
  // the tile's sources (no upstream cell inside the tile), fixed before any count falls
  unsigned srcs = 0;
#pragma unroll
  for (int j = 0; j < CPT; j++) {
    const int i = tile_cell(t, j);
    if (!L.done[i] && L.left[i] == 0) srcs |= 1u << j;
  }
  __syncthreads();
  // walks in LDS from each source, with the hand-off of the global walks: a cell with
  // an open upstream cell is open, any other gets in + its upstream totals in the
  // fixed order; the walk that takes the downstream count to zero goes on
#pragma unroll 1
  for (int j = 0; j < CPT; j++) {
    if (!(srcs >> j & 1u)) continue;
    int i = tile_cell(t, j);
    for (;;) {
      const int lx = i % TS, ly = i / TS;
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
      if (!outflow(L.dir[ly + 1][lx + 1], dx, dy) || lx + dx < 0 || lx + dx >= TS || ly + dy < 0 || ly + dy >= TS) break;
      const int d = (ly + dy) * TS + lx + dx;
      if (op) L.open[d] = 1;
      // release / acquire at work-group scope: the total and the open mark are in LDS
      // before the count falls, and the walk that takes it to zero reads them after
      if (__hip_atomic_fetch_add(&L.left[d], 0xffffffffu, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) != 1u) break;
      i = d;
    }
  }
  __syncthreads();
  // closed cells are final; open (and cycle) cells start at 0 for the walks. The count of
  // an open cell (its open upstream neighbours) is assembled in `cnt` (cleared before) from
  // two sides, in any order: its own tile adds its open in-tile upstream cells plus all its
  // upstream cells outside the tile, and the tile of each such cell that is closed takes one
  // back. Every step before the last leaves a non-zero value (negative before the addition,
  // above the final count after it), so the atomic that leaves zero is the last one, and it
  // lists the cell as a source of the walks.
  if (t == 0) L.nsrc = 0;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < CPT; j++) {
    const int i = tile_cell(t, j), lx = i % TS, ly = i / TS;
    const int x = tx0 + lx, y = ty0 + ly;
    if (x >= g.DS0 || y >= g.DS1) continue;
    const bool closed = L.done[i] && !L.open[i];
    int dx, dy;
    const bool down = outflow(L.dir[ly + 1][lx + 1], dx, dy) && x + dx >= 0 && x + dx < g.DS0 && y + dy >= 0 &&
                      y + dy < g.DS1;
    const uint64_t e = L.rt[ly + 1] + L.ct[lx + 1];
    out[e] = closed && down ? L.tot[i] : T(0);
    if (closed) {
      if (down && (lx + dx < 0 || lx + dx >= TS || ly + dy < 0 || ly + dy >= TS)) {
        const uint64_t ed = L.rt[ly + 1 + dy] + L.ct[lx + 1 + dx];
        if (atomicAdd(&cnt[ed], 0xffffffffu) == 1u) L.src[atomicAdd(&L.nsrc, 1u)] = (unsigned)ed;
      }
    } else {
      unsigned k = 0;
      for (int i1 = -1; i1 <= 1; i1++)
        for (int i0 = -1; i0 <= 1; i0++) {
          if ((i0 | i1) == 0 || L.dir[ly + 1 + i1][lx + 1 + i0] != inflow_code(i0, i1)) continue;
          const int u = (ly + i1) * TS + lx + i0;
          const bool inside = lx + i0 >= 0 && lx + i0 < TS && ly + i1 >= 0 && ly + i1 < TS;
          k += !inside || !L.done[u] || L.open[u];
        }
      if (atomicAdd(&cnt[e], k) + k == 0u) L.src[atomicAdd(&L.nsrc, 1u)] = (unsigned)e;
    }
  }
  __syncthreads();
  if (t == 0) L.base = L.nsrc ? atomicAdd(&queue[0], L.nsrc) : 0u;
  __syncthreads();
  for (unsigned q = t; q < L.nsrc; q += NT) queue[2 + L.base + q] = L.src[q];
}

// Clears the count words and the two list counters before the tile pass.
__global__ void __launch_bounds__(NT) k_clear(unsigned* cnt, unsigned* queue, uint64_t n) {
  const uint64_t e = (uint64_t)blockIdx.x * NT + threadIdx.x;
  if (e < n) cnt[e] = 0u;
  if (e < 2) queue[e] = 0u;  // sources listed, sources taken
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
  // the cell's 3x3 directions and input value (clamped into the raster; a neighbour
  // outside it never matches); each step loads the next cell's while its own upstream
  // totals and the next count are in flight, one round trip per step
  uint64_t en[9];
  bool inr[9];
  hood(g, x, y, en, inr);
  unsigned dn[9];
#pragma unroll
  for (int k = 0; k < 9; k++) dn[k] = dir[en[k]];
  T own = in[e];
  for (;;) {
    bool up[9];
#pragma unroll
    for (int k = 0; k < 9; k++) up[k] = k != 4 && inr[k] && dn[k] == inflow_code(k % 3 - 1, k / 3 - 1);
    int dx = 0, dy = 0;
    const bool code = outflow(dn[4], dx, dy);
    const bool down = code && x + dx >= 0 && y + dy >= 0 && x + dx < g.DS0 && y + dy < g.DS1;
    uint64_t ed = en[4];
#pragma unroll
    for (int k = 0; k < 9; k++)  // a select per slot: a computed index would put en[] in scratch memory
      if (k == (dy + 1) * 3 + dx + 1) ed = en[k];
    // next cell's loads (the cell itself when there is none), then this cell's totals
    const int x2 = down ? x + dx : x, y2 = down ? y + dy : y;
    uint64_t en2[9];
    bool inr2[9];
    hood(g, x2, y2, en2, inr2);
    unsigned dn2[9];
#pragma unroll
    for (int k = 0; k < 9; k++) dn2[k] = dir[en2[k]];
    const T own2 = in[ed];
    const unsigned left = __hip_atomic_load(&cnt[ed], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    T tv[9];
#pragma unroll
    for (int k = 0; k < 9; k++) tv[k] = ld_total(&out[up[k] && en[k] != from ? en[k] : e]);
    T acc = own;
#pragma unroll
    for (int k = 0; k < 9; k++)
      if (up[k]) acc = (T)(acc + (en[k] == from ? carried : tv[k]));
    if (!down) return;  // no downstream cell inside the raster: the neutral value stays
    st_total(&out[e], acc);
    if (left != 1u) {
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
    x = x2;
    y = y2;
#pragma unroll
    for (int k = 0; k < 9; k++) {
      en[k] = en2[k];
      inr[k] = inr2[k];
      dn[k] = dn2[k];
    }
    own = own2;
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
  const unsigned grid = (unsigned)((n + NT - 1) / NT);
  hipLaunchKernelGGL(k_clear, dim3(grid), dim3(NT), 0, s, (unsigned*)cnt, queue, n);
  check(hipGetLastError(), "spread clear launch");
  const int ntx = (g.DS0 + TS - 1) / TS, nty = (g.DS1 + TS - 1) / TS;
  hipLaunchKernelGGL((k_tile<T>), dim3((unsigned)(ntx * nty)), dim3(NT), 0, s, (const T*)in, (const uint8_t*)dir,
                     (unsigned*)cnt, (T*)out, queue, g, ntx);
  check(hipGetLastError(), "spread tile launch");
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
  // cell indices, walk sources and queue entries are 32-bit element indices
  if (n >= (1ull << 32)) throw std::runtime_error("spread: rasters of 2^32 or more cells are not supported");
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
