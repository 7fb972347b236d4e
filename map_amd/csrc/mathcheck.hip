// Accuracy check of the FP32 fast paths of the generated kernels
// (mx::sqrt_f32, mx::sincos_f32 / sin_f32 / cos_f32, mx::atan2_f32 in
// device/mapx_dev.hiph) against FP64 on the device: every FP32 bit pattern of a
// range (atan2: each pattern x paired with a y drawn from it), the error in
// ulps of the FP32 value nearest the FP64 result. Non-finite references must
// be matched exactly (NaN by NaN).
#include <hip/hip_runtime.h>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "device/mapx_dev.hiph"

namespace mapx {
namespace mathcheck {

enum Fn : int { SQRT = 0, SIN = 1, COS = 2, SINCOS = 3, ATAN2 = 4, ATAN2_NEAR = 5 };

struct Result {
  unsigned long long max_ulp_bits;  // the largest error (a non-negative double, ordered as its bits)
  unsigned long long n_over;        // results more than 4 ulp away
  unsigned long long n;             // results checked
  unsigned long long worst_in;      // input bits of the largest error
};

__device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// ulps between an FP32 result and the FP64 reference (spacing of the FP32
// value nearest the reference); infinite for a NaN / infinity mismatch
__device__ inline double ulps(float got, double ref) {
  if (ref != ref) return got != got ? 0.0 : __builtin_inf();
  const float fr = (float)ref;
  if (__builtin_isinf(fr) || __builtin_isinf(ref)) return (double)got == ref || got == fr ? 0.0 : __builtin_inf();
  if (got != got) return __builtin_inf();
  const float a = __builtin_fabsf(fr);
  const int e = a >= 0x1p-126f ? ((int)((__float_as_uint(a) >> 23) & 0xff) - 127) : -126;
  const double sp = __builtin_ldexp(1.0, e - 23);
  return __builtin_fabs((double)got - ref) / sp;
}

__global__ void __launch_bounds__(256) k_check(int fn, uint64_t lo, uint64_t hi, Result* out) {
  double worst = 0.0;
  unsigned long long over = 0, n = 0, win = 0;
  for (uint64_t i = lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t b = (uint32_t)i;
    const float x = __uint_as_float(b);
    double e0 = 0.0, e1 = 0.0;
    if (fn == SQRT) {
      e0 = ulps(mx::sqrt_f32(x), __builtin_sqrt((double)x));
    } else if (fn == SIN) {
      e0 = ulps(mx::sin_f32(x), sin((double)x));
    } else if (fn == COS) {
      e0 = ulps(mx::cos_f32(x), cos((double)x));
    } else if (fn == SINCOS) {
      float s, c;
      mx::sincos_f32(x, &s, &c);
      e0 = ulps(s, sin((double)x));
      e1 = ulps(c, cos((double)x));
    } else {
      // y: random bits (ATAN2) or x's sign and exponent with random low bits
      // (ATAN2_NEAR, |y| / |x| near 1: the octant boundaries)
      const uint64_t h = mix64(i);
      const uint32_t yb = fn == ATAN2 ? (uint32_t)h : ((b & 0xff800000u) ^ ((uint32_t)h & 0x807fffffu));
      const float y = __uint_as_float(yb);
      e0 = ulps(mx::atan2_f32(y, x), atan2((double)y, (double)x));
      const float y2 = __uint_as_float((uint32_t)(h >> 32));
      e1 = ulps(mx::atan2_f32(x, y2), atan2((double)x, (double)y2));
    }
    const double e = e0 > e1 ? e0 : e1;
    if (e > worst) { worst = e; win = b; }
    over += (e0 > 4.0) + (e1 > 4.0);
    n += (fn == SINCOS || fn >= ATAN2) ? 2 : 1;
  }
  __shared__ double sw[256];
  __shared__ unsigned long long so[256], sn[256], si[256];
  sw[threadIdx.x] = worst; so[threadIdx.x] = over; sn[threadIdx.x] = n; si[threadIdx.x] = win;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      if (sw[threadIdx.x + s] > sw[threadIdx.x]) { sw[threadIdx.x] = sw[threadIdx.x + s]; si[threadIdx.x] = si[threadIdx.x + s]; }
      so[threadIdx.x] += so[threadIdx.x + s];
      sn[threadIdx.x] += sn[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const unsigned long long wb = (unsigned long long)__double_as_longlong(sw[0]);
    const unsigned long long prev = atomicMax(&out->max_ulp_bits, wb);
    if (wb > prev) atomicExch(&out->worst_in, si[0]);
    atomicAdd(&out->n_over, so[0]);
    atomicAdd(&out->n, sn[0]);
  }
}

static void check(hipError_t e, const char* w) {
  if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}

// Checks bit patterns [lo, hi) (hi <= 2^32) of function `name`; `scratch`: a
// device buffer of at least sizeof(Result) bytes.
void run(const std::string& name, uint64_t lo, uint64_t hi, void* scratch, void* stream, double* max_ulp,
         unsigned long long* n_over, unsigned long long* n, unsigned* worst_in) {
  int fn;
  if (name == "sqrt") fn = SQRT;
  else if (name == "sin") fn = SIN;
  else if (name == "cos") fn = COS;
  else if (name == "sincos") fn = SINCOS;
  else if (name == "atan2") fn = ATAN2;
  else if (name == "atan2_near") fn = ATAN2_NEAR;
  else throw std::runtime_error("mathCheck: unknown function " + name);
  if (lo >= hi || hi > (1ull << 32)) throw std::runtime_error("mathCheck: bad range");
  hipStream_t s = (hipStream_t)stream;
  Result h{0, 0, 0, 0};
  check(hipMemcpyAsync(scratch, &h, sizeof h, hipMemcpyHostToDevice, s), "H2D");
  hipLaunchKernelGGL(k_check, dim3(4096), dim3(256), 0, s, fn, lo, hi, (Result*)scratch);
  check(hipGetLastError(), "mathCheck launch");
  check(hipMemcpyAsync(&h, scratch, sizeof h, hipMemcpyDeviceToHost, s), "D2H");
  check(hipStreamSynchronize(s), "sync");
  *max_ulp = __builtin_bit_cast(double, h.max_ulp_bits);
  *n_over = h.n_over;
  *n = h.n;
  *worst_in = (unsigned)h.worst_in;
}

}  // namespace mathcheck
}  // namespace mapx
