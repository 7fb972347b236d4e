// Radial scan (the reference's RadialScan / RadiatingSkeleton,
// runtime/skeleton/RadiatingSkeleton.cpp:47-325, runtime/skeleton/util.cpp:543-616,
// runtime/task/RadiatingTask.cpp:133-271): for an observer cell s, every cell
// p gets R(p) = OP(base(p), in(p)) (in(p) skipped when not finite), where the
// plane around s is cut into 8 octants (first direction F, second S; the
// cells with d0 = (p-s).F >= d1 = (p-s).S >= 0) and
//   base(s) = neutral; base = R(p-F) on the F axis (d1 = 0); base = R(p-F-S)
//   on the diagonal (d1 = d0); else the ray's crossing of column d0-1,
//   a*w + b*(1-w), a = R(p-F), b = R(p-F-S), w = d1/d0*(d0-1) - (d1-1).
// Cells on an axis or a diagonal belong to two octants and get the same
// value from both.
//
// Column d0 of an octant depends only on column d0-1, so a launch advances
// every octant by one band of K columns. A work-group owns a strip of HS rows
// (one row per thread) and also runs the K rows below it (a trapezoid: row r
// at column c needs rows r and r-1 at column c-1, so after K columns only the
// HS owned rows are exact); strips never wait on each other. Everything off
// the recurrence comes first (the band's input loads, all in flight together,
// and the interpolation weights); each column then exchanges the previous
// column's values through LDS (one barrier) with no global memory operation
// in the chain (a store there would make the next barrier wait for it); the
// results are stored after the band to `out` (block-tiled, like every raster),
// from which the next band reads its starting column.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

// every product and sum rounds on its own, as the generated kernels do (-ffp-contract=off)
#pragma clang fp contract(off)

namespace mapx {
namespace radial {

constexpr int NT = 256;  // threads (rows) per work-group

struct Oct {
  int fx, fy, sx, sy;  // unit vectors of the first and second directions
  int L0, L1;          // last column (along F) and last row (along S) inside the raster
};

struct Params {
  int BS0, BS1, NB0, NB1;
  int s0, s1;          // observer cell
  int c0;              // first column of the band
  Oct oct[8];
  int first[9];        // first work-group of each octant in this launch; first[8] = total
};

// Block-tiled position of one cell, stepped along F one column at a time.
struct Pos {
  int bx, ix, by, iy;
  __device__ void init(const Params& P, int x, int y) {
    bx = x / P.BS0; ix = x - bx * P.BS0;
    by = y / P.BS1; iy = y - by * P.BS1;
  }
  __device__ void step(const Params& P, int fx, int fy) {
    ix += fx;
    if (ix == P.BS0) { ix = 0; bx++; } else if (ix < 0) { ix = P.BS0 - 1; bx--; }
    iy += fy;
    if (iy == P.BS1) { iy = 0; by++; } else if (iy < 0) { iy = P.BS1 - 1; by--; }
  }
  __device__ int64_t off(const Params& P) const {
    return (((int64_t)by * P.NB0 + bx) * P.BS1 + iy) * P.BS0 + ix;
  }
};

enum Op { SUM = 0, PROD = 1, MAX = 2, MIN = 3 };

template <class T, int OP> __device__ __forceinline__ T neutral() {
  if (OP == SUM) return (T)0;
  if (OP == PROD) return (T)1;
  if (OP == MAX) return -__builtin_huge_val();
  return __builtin_huge_val();
}
template <int OP> __device__ __forceinline__ float combine(float b, float x) {
  if (OP == SUM) return b + x;
  if (OP == PROD) return b * x;
  if (OP == MAX) return fmaxf(b, x);
  return fminf(b, x);
}
template <int OP> __device__ __forceinline__ double combine(double b, double x) {
  if (OP == SUM) return b + x;
  if (OP == PROD) return b * x;
  if (OP == MAX) return fmax(b, x);
  return fmin(b, x);
}

template <class T, int OP, int K>
__global__ void __launch_bounds__(NT) k_radial(const T* __restrict__ in, T* __restrict__ out, Params P) {
  constexpr int HS = NT - K;  // rows a work-group owns
  __shared__ T xch[2][NT];
  const int g = blockIdx.x;
  int o = 0;
  while (o < 7 && g >= P.first[o + 1]) o++;
  const Oct O = P.oct[o];
  const int j = g - P.first[o];
  const int t = threadIdx.x;
  const int r = j * HS - K + t;                  // this thread's row (d1)
  const int c0 = P.c0;
  const int cend = min(c0 + K, O.L0 + 1);        // columns [c0, cend)
  const bool row_ok = r >= 0 && r <= O.L1;
  const bool owner = t >= K;
  // Row clamped into the octant's rows, so every address below is a raster
  // cell: all loads are issued unconditionally (no branch between them), so
  // they are in flight together; values of cells outside the octant are unused.
  const int rc = r < 0 ? 0 : (r > O.L1 ? O.L1 : r);
  const int x0 = P.s0 + c0 * O.fx + rc * O.sx, y0 = P.s1 + c0 * O.fy + rc * O.sy;  // cell (c0, rc)

  // the previous column's value of this row
  T prev = (T)0;
  if (c0 > 0) {  // uniform
    Pos q;
    q.init(P, x0 - O.fx, y0 - O.fy);
    prev = out[q.off(P)];
  }
  // the band's inputs and each column's interpolation weights (w, 1 - w), all
  // before the per-column chain
  T val[K], w[K], v1[K];
  Pos pos;
  pos.init(P, x0, y0);
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int c = c0 + k;
    val[k] = in[pos.off(P)];
    if (c + 1 < cend) pos.step(P, O.fx, O.fy);  // uniform: the last column repeats
    w[k] = (T)0;
    // linear_interpolation's weight (runtime/skeleton/util.cpp:564; its float literals are exact here)
    if (r > 0 && r < c) w[k] = (T)r / (T)c * (T)(c - 1) - (T)(r - 1);
    v1[k] = (T)1 - w[k];
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int c = c0 + k;
    if (c >= cend) break;  // uniform over the work-group
    xch[k & 1][t] = prev;
    __syncthreads();
    const T nb = t > 0 ? xch[k & 1][t - 1] : prev;  // row r-1 at column c-1 (the halo's first row is never owned)
    if (row_ok && r <= c) {
      const T mix = prev * w[k] + nb * v1[k];  // a*w + b*(1-w)
      T base = r == c ? nb : mix;
      base = r == 0 ? prev : base;
      base = c == 0 ? neutral<T, OP>() : base;
      const T x = val[k];
      prev = __builtin_isfinite(x) ? combine<OP>(base, x) : base;
      val[k] = prev;
    }
  }
  // the owned rows' results of the band
  if (row_ok && owner) {
    pos.init(P, x0, y0);
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int c = c0 + k;
      if (c >= cend) break;
      if (r <= c) out[pos.off(P)] = val[k];
      if (c + 1 < cend) pos.step(P, O.fx, O.fy);
    }
  }
}

static void check(hipError_t e, const char* w) {
  if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}

// Octants in the reference's order (RadialCase, runtime/Direction.hpp:68):
// NORTH_EAST, EAST_NORTH, EAST_SOUTH, SOUTH_EAST, SOUTH_WEST, WEST_SOUTH,
// WEST_NORTH, NORTH_WEST; x grows east, y grows south.
static const int kDirs[8][4] = {{0, -1, 1, 0}, {1, 0, 0, -1}, {1, 0, 0, 1}, {0, 1, 1, 0},
                                {0, 1, -1, 0}, {-1, 0, 0, 1}, {-1, 0, 0, -1}, {0, -1, -1, 0}};

static int reach(int s, int ds, int u0, int u1) {  // cells from s to the edge along (u0, u1) on its axis
  const int u = u0 ? u0 : u1;
  return u > 0 ? ds - 1 - s : s;
}

// dt_bytes 4: float, 8: double; op: 0 sum, 1 prod, 2 max, 3 min. Returns the launch count.
int run(void* out, const void* in, int dt_bytes, int op, int BS0, int BS1, int NB0, int NB1, int DS0, int DS1, int s0,
        int s1, void* stream) {
  if (s0 < 0 || s1 < 0 || s0 >= DS0 || s1 >= DS1) throw std::runtime_error("radial: start cell outside the raster");
  if (op < 0 || op > 3) throw std::runtime_error("radial: bad reduction");
  Params P{};
  P.BS0 = BS0; P.BS1 = BS1; P.NB0 = NB0; P.NB1 = NB1; P.s0 = s0; P.s1 = s1;
  int maxL0 = 0;
  for (int o = 0; o < 8; o++) {
    Oct& O = P.oct[o];
    O.fx = kDirs[o][0]; O.fy = kDirs[o][1]; O.sx = kDirs[o][2]; O.sy = kDirs[o][3];
    O.L0 = O.fx ? reach(s0, DS0, O.fx, 0) : reach(s1, DS1, 0, O.fy);
    O.L1 = O.sx ? reach(s0, DS0, O.sx, 0) : reach(s1, DS1, 0, O.sy);
    maxL0 = O.L0 > maxL0 ? O.L0 : maxL0;
  }
  const int K = dt_bytes == 8 ? 32 : 64, HS = NT - K;
  hipStream_t s = (hipStream_t)stream;
  int launches = 0;
  for (int c0 = 0; c0 <= maxL0; c0 += K) {
    P.c0 = c0;
    int total = 0;
    for (int o = 0; o < 8; o++) {
      P.first[o] = total;
      const Oct& O = P.oct[o];
      if (c0 > O.L0) continue;
      const int rmax = (c0 + K - 1 < O.L1 ? c0 + K - 1 : O.L1);
      total += (rmax + 1 + HS - 1) / HS;
    }
    P.first[8] = total;
    if (!total) continue;
#define MAPX_RADIAL(T, OPV, KV) hipLaunchKernelGGL((k_radial<T, OPV, KV>), dim3(total), dim3(NT), 0, s, (const T*)in, (T*)out, P)
    if (dt_bytes == 4) {
      if (op == SUM) MAPX_RADIAL(float, SUM, 64);
      else if (op == PROD) MAPX_RADIAL(float, PROD, 64);
      else if (op == MAX) MAPX_RADIAL(float, MAX, 64);
      else MAPX_RADIAL(float, MIN, 64);
    } else if (dt_bytes == 8) {
      if (op == SUM) MAPX_RADIAL(double, SUM, 32);
      else if (op == PROD) MAPX_RADIAL(double, PROD, 32);
      else if (op == MAX) MAPX_RADIAL(double, MAX, 32);
      else MAPX_RADIAL(double, MIN, 32);
    } else {
      throw std::runtime_error("radial: F32 or F64 rasters only");
    }
#undef MAPX_RADIAL
    check(hipGetLastError(), "k_radial launch");
    launches++;
  }
  return launches;
}

}  // namespace radial
}  // namespace mapx
