/*
 * CPU oracle for the map-algebra block-task path (built with -ffp-contract=off,
 * so every float product and sum is rounded separately, as in C). TEST INFRASTRUCTURE ONLY:
 * used by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py as the checker; never by the product path.
 *
 * Restates, on whole rasters in row-major order (rows = y):
 *  - the block-relative focal border rule of load_F
 *    (runtime/skeleton/util.cpp:492-506): a cell outside the central block is
 *    read from the neighbour block that holds it; when that neighbour lies
 *    outside the raster both coordinates are mirrored inside the central block
 *    (invert), also for diagonal lookups whose row or column neighbour exists.
 *    Edge blocks of sizes that are not block multiples mirror at their valid
 *    extent (the backend's masking of padding; the reference mirrors at BS).
 *    A lookup that lands outside its new block's valid extent (an edge block
 *    narrower than the halo) applies the rule again from that block, until it
 *    is inside one: load_F followed across the blocks it needs.
 *  - the focal correlation of FocalSkeleton.cpp:300-336: acc = 0, then for
 *    y = -h..h (outer), x = -h..h (inner): acc += S[y][x] * W[y][x], in the
 *    C types of the operands (integer: exact then wrapped to the node type;
 *    float: each product and sum rounded to FP32 in that order).
 *  - the Game of Life rule of python/life.py:24-36 with the conv node in S32
 *    and the update (nbh==3) + (nbh==2)*state in U8.
 *  - per-zone statistics (sum in FP64, min, max, count).
 *  - the synthetic input generators (an independent restatement of the
 *    definitions used by the backend's generator kernels).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline int or_invert(int i, int d) { return i < 0 ? -i - 1 : (i >= d ? d - (i - d + 1) : i); }
static inline int or_modulo(int i, int d) { return i < 0 ? i + d : (i >= d ? i - d : i); }

/* Global source cell of the lookup at block-local (lx, ly) of block (bx, by). */
void or_src(int ds0, int ds1, int bs0, int bs1, int bx, int by, int lx, int ly, int *gx, int *gy) {
  const int nb0 = (ds0 + bs0 - 1) / bs0, nb1 = (ds1 + bs1 - 1) / bs1;
  int x = lx, y = ly;
  /* each step lands in range, moves to a neighbour block, or reflects the coordinate
     (two reflections shift it by twice the extent towards the block): it ends */
  for (int step = 0; step < 4096; step++) {
    const int e0 = ds0 - bx * bs0 < bs0 ? ds0 - bx * bs0 : bs0;
    const int e1 = ds1 - by * bs1 < bs1 ? ds1 - by * bs1 : bs1;
    const int i0 = x < 0 ? 0 : (x >= e0 ? 2 : 1);
    const int i1 = y < 0 ? 0 : (y >= e1 ? 2 : 1);
    if (i0 == 1 && i1 == 1) break;
    const int nbx = bx + i0 - 1, nby = by + i1 - 1;
    if (nbx < 0 || nby < 0 || nbx >= nb0 || nby >= nb1) {
      x = or_invert(x, e0);
      y = or_invert(y, e1);
    } else {
      x = or_modulo(x, bs0);
      y = or_modulo(y, bs1);
      bx = nbx;
      by = nby;
    }
  }
  *gx = bx * bs0 + x;
  *gy = by * bs1 + y;
}

/* Index map of every tap: idx[(cell * nt) + t] = source cell (row-major), for
 * the taps t of a (2*h0+1) x (2*h1+1) window in y-outer, x-inner order. */
void or_tap_index(int ds0, int ds1, int bs0, int bs1, int h0, int h1, int64_t *idx) {
  const int nt = (2 * h0 + 1) * (2 * h1 + 1);
#pragma omp parallel for schedule(static)
  for (int gy = 0; gy < ds1; gy++) {
    for (int gx = 0; gx < ds0; gx++) {
      const int bx = gx / bs0, by = gy / bs1, x = gx % bs0, y = gy % bs1;
      int64_t *o = idx + ((int64_t)gy * ds0 + gx) * nt;
      int t = 0;
      for (int dy = -h1; dy <= h1; dy++)
        for (int dx = -h0; dx <= h0; dx++) {
          int sx, sy;
          or_src(ds0, ds1, bs0, bs1, bx, by, x + dx, y + dy, &sx, &sy);
          o[t++] = (int64_t)sy * ds0 + sx;
        }
    }
  }
}

/* FP32 correlation: out = sum over taps in order, each product and each
 * addition rounded to float (weights already converted to float). */
void or_conv_f32(const float *in, const int64_t *idx, int64_t n, int nt, const float *w, float *out) {
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < n; c++) {
    float acc = 0.0f;
    const int64_t *ix = idx + c * nt;
    for (int t = 0; t < nt; t++) {
      const float p = in[ix[t]] * w[t];
      acc = acc + p;
    }
    out[c] = acc;
  }
}

/* FP64 correlation for the FP64 oracle. */
void or_conv_f64(const double *in, const int64_t *idx, int64_t n, int nt, const double *w, double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < n; c++) {
    double acc = 0.0;
    const int64_t *ix = idx + c * nt;
    for (int t = 0; t < nt; t++) acc = acc + in[ix[t]] * w[t];
    out[c] = acc;
  }
}

/* Integer correlation: exact in int64 (the caller wraps to the node type,
 * which equals C's wrapping arithmetic modulo 2^k). */
void or_conv_i64(const int64_t *in, const int64_t *idx, int64_t n, int nt, const int64_t *w, int64_t *out) {
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < n; c++) {
    uint64_t acc = 0;
    const int64_t *ix = idx + c * nt;
    for (int t = 0; t < nt; t++) acc += (uint64_t)in[ix[t]] * (uint64_t)w[t];
    out[c] = (int64_t)acc;
  }
}

/* Focal min/max/sum over the window with the same border rule (FocalFunc,
 * weights ignored as in FocalSkeleton.cpp:338-372): op 1 SUM, 6 MAX, 7 MIN. */
void or_focal_f64(const double *in, const int64_t *idx, int64_t n, int nt, int op, double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < n; c++) {
    const int64_t *ix = idx + c * nt;
    double acc = op == 1 ? 0.0 : (op == 6 ? -INFINITY : INFINITY);
    for (int t = 0; t < nt; t++) {
      double v = in[ix[t]];
      if (op == 1) acc += v;
      else if (op == 6) acc = v > acc ? v : acc;
      else acc = v < acc ? v : acc;
    }
    out[c] = acc;
  }
}

/* `gens` generations of Life on a U8 raster (rows = y), ping-ponging between
 * a and b; the result ends in a when gens is even, else in b. */
void or_life(uint8_t *a, uint8_t *b, int ds0, int ds1, int bs0, int bs1, int gens) {
  for (int g = 0; g < gens; g++) {
    const uint8_t *s = (g & 1) ? b : a;
    uint8_t *d = (g & 1) ? a : b;
#pragma omp parallel for schedule(static)
    for (int gy = 0; gy < ds1; gy++) {
      const int by = gy / bs1, y = gy % bs1;
      const int e1 = ds1 - by * bs1 < bs1 ? ds1 - by * bs1 : bs1;
      for (int gx = 0; gx < ds0; gx++) {
        const int bx = gx / bs0, x = gx % bs0;
        const int e0 = ds0 - bx * bs0 < bs0 ? ds0 - bx * bs0 : bs0;
        int32_t nbh = 0;
        if (x >= 1 && y >= 1 && x + 1 < e0 && y + 1 < e1) {
          const uint8_t *r0 = s + (int64_t)(gy - 1) * ds0 + gx, *r1 = r0 + ds0, *r2 = r1 + ds0;
          nbh = r0[-1] + r0[0] + r0[1] + r1[-1] + r1[1] + r2[-1] + r2[0] + r2[1];
        } else {
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
              if (!dx && !dy) continue;
              int sx, sy;
              or_src(ds0, ds1, bs0, bs1, bx, by, x + dx, y + dy, &sx, &sy);
              nbh += s[(int64_t)sy * ds0 + sx];
            }
        }
        const uint8_t st = s[(int64_t)gy * ds0 + gx];
        const uint8_t eq3 = (uint8_t)(nbh == 3), eq2 = (uint8_t)(nbh == 2);
        d[(int64_t)gy * ds0 + gx] = (uint8_t)(eq3 + (uint8_t)(eq2 * st));
      }
    }
  }
}

/* Per-zone statistics of FP32 data (zones outside [0, nz) ignored). */
void or_zstats_f32(const float *v, const int32_t *z, int64_t n, int nz, double *sum, float *mn, float *mx,
                   int64_t *cnt) {
  for (int k = 0; k < nz; k++) { sum[k] = 0.0; mn[k] = INFINITY; mx[k] = -INFINITY; cnt[k] = 0; }
  for (int64_t i = 0; i < n; i++) {
    const int k = z[i];
    if (k < 0 || k >= nz) continue;
    sum[k] += (double)v[i];
    if (v[i] < mn[k]) mn[k] = v[i];
    if (v[i] > mx[k]) mx[k] = v[i];
    cnt[k]++;
  }
}

/* ---- input generators (independent restatement of the backend's definitions) ---- */

static inline uint64_t or_splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline double or_lattice(uint64_t seed, int o, int64_t i, int64_t j) {
  uint64_t z = or_splitmix64(seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(o + 1)));
  z = or_splitmix64(z ^ (uint64_t)i * 0x9E3779B97F4A7C15ull);
  z = or_splitmix64(z ^ (uint64_t)j * 0xC2B2AE3D27D4EB4Full);
  return (double)(z >> 11) / 9007199254740992.0 * 2.0 - 1.0;
}

/* DEM value (FP64, rounded once to FP32) at global cells [x0,x0+w) x [y0,y0+h),
 * Gaussians g[k] = (cx, cy, sigma, amplitude), truncated beyond 12 sigma. */
void or_dem(float *out, int64_t x0, int64_t y0, int w, int h, const double *g, int ng, uint64_t seed) {
#pragma omp parallel for schedule(static)
  for (int r = 0; r < h; r++)
    for (int c = 0; c < w; c++) {
      const double x = (double)(x0 + c), y = (double)(y0 + r);
      double v = 0.0;
      for (int k = 0; k < ng; k++) {
        const double s = g[4 * k + 2], dx = x - g[4 * k], dy = y - g[4 * k + 1];
        const double r2 = dx * dx + dy * dy;
        if (r2 <= (12.0 * s) * (12.0 * s)) v += g[4 * k + 3] * exp(-r2 * (1.0 / (2.0 * s * s)));
      }
      double sum = 0.0, norm = 0.0, amp = 1.0, period = 256.0;
      for (int o = 0; o < 8; o++) {
        const double u = x / period, q = y / period, fi = floor(u), fj = floor(q);
        const double fu = u - fi, fv = q - fj;
        const double su = fu * fu * (3.0 - 2.0 * fu), sv = fv * fv * (3.0 - 2.0 * fv);
        const int64_t i = (int64_t)fi, j = (int64_t)fj;
        const double a = or_lattice(seed, o, i, j), b = or_lattice(seed, o, i + 1, j);
        const double cc = or_lattice(seed, o, i, j + 1), d = or_lattice(seed, o, i + 1, j + 1);
        const double top = a + (b - a) * su, bot = cc + (d - cc) * su;
        sum += amp * (top + (bot - top) * sv);
        norm += amp;
        amp *= 0.5;
        period *= 0.5;
      }
      out[(int64_t)r * w + c] = (float)(v + 20.0 * sum / norm);
    }
}

/* Voronoi zone index (nearest point by int64 squared distance, ties -> lowest index). */
void or_voronoi(int32_t *out, int64_t x0, int64_t y0, int w, int h, const int32_t *px, const int32_t *py, int np) {
#pragma omp parallel for schedule(static)
  for (int r = 0; r < h; r++)
    for (int c = 0; c < w; c++) {
      const int64_t x = x0 + c, y = y0 + r;
      int best = 0;
      int64_t bd = INT64_MAX;
      for (int k = 0; k < np; k++) {
        const int64_t dx = x - px[k], dy = y - py[k], d = dx * dx + dy * dy;
        if (d < bd) { bd = d; best = k; }
      }
      out[(int64_t)r * w + c] = best;
    }
}
