// arith_check.hip -- device checks of the in-range square root and division
// of npx_device.hpp against the compiler's correctly rounded sqrtf and '/'.
//
// Built by novaphy_amd/build.py into tests/hip/libnpx_arith_check.so with
// the library's own flags (correctly rounded divide / sqrt, no contraction);
// tests/test_gpu_arith.py calls it through ctypes.
//
//   npxt_sqrt_check   every float bit pattern in [lo, hi]: npx_sqrt_in(x)
//                     against sqrtf(x);
//   npxt_div_check    n seeded operand pairs: npx_div_in(a, b) against a / b.
//                     mode 0: random a, b in the range npx_div_in states;
//                     mode 1: spiky_grad's (-cs) diff^2 and |r| for h drawn
//                             log-uniformly from [2^-6, 2^6] and r^2 over the
//                             support, a share of them at its outer edge;
//                     mode 2: the s_corr ratio's poly6(r^2) and
//                             poly6((dq h)^2) for h as in mode 1, dq in
//                             (0, 1) and r^2 in [0, h^2).
//                     Modes 1 and 2 also count pairs outside the stated
//                     range, which must not occur for h in [2^-6, 2^6].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../novaphy_amd/csrc/npx_device.hpp"

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ float unit(uint64_t r) {  // [0, 1)
  return (float)(r >> 40) * 0x1.0p-24f;
}

__device__ __forceinline__ bool in_range(float a, float b) {
  const float aa = fabsf(a), ab = fabsf(b);
  if (!(ab >= 0x1.0p-60f && ab <= 0x1.0p60f)) return false;
  if (a == 0.0f) return __float_as_uint(a) == 0u;  // +0 only
  if (!(aa >= 0x1.0p-100f && aa <= 0x1.0p60f)) return false;
  const double q = (double)aa / (double)ab;
  return q >= 0x1.0p-120 && q <= 0x1.0p90;
}

__global__ void k_sqrt_check(uint32_t lo, uint64_t count, unsigned long long* bad, unsigned* first) {
  unsigned long long nb = 0;
  unsigned f = 0xffffffffu;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const uint32_t u = lo + (uint32_t)i;
    const float x = __uint_as_float(u);
    if (__float_as_uint(npx_sqrt_in(x)) != __float_as_uint(sqrtf(x))) {
      ++nb;
      f = u < f ? u : f;
    }
  }
  if (nb) {
    atomicAdd(bad, nb);
    atomicMin(first, f);
  }
}

__device__ void h_consts(uint64_t r, float* h, float* h2, float* c6, float* cs) {
  const float e = -6.0f + 12.0f * unit(r);  // log2 h in [-6, 6)
  const float hh = exp2f(e);
  const float hh2 = hh * hh;
  const float h9 = hh2 * hh2 * hh2 * hh2 * hh;
  const float h6 = hh * hh * hh * hh * hh * hh;
  const float pi = 3.14159265358979323846f;
  *h = hh;
  *h2 = hh2;
  *c6 = 315.0f / ((64.0f * pi) * h9);
  *cs = 45.0f / (pi * h6);
}

__global__ void k_div_check(int mode, uint64_t n, uint64_t seed, unsigned long long* out) {
  unsigned long long bad = 0, outside = 0, tested = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t r0 = mix64(seed ^ (4 * i)), r1 = mix64(seed ^ (4 * i + 1));
    const uint64_t r2 = mix64(seed ^ (4 * i + 2)), r3 = mix64(seed ^ (4 * i + 3));
    float a, b;
    if (mode == 0) {
      const int ea = -100 + (int)(r0 % 161);  // 2^-100 .. 2^60
      const int eb = -60 + (int)(r1 % 121);   // 2^-60 .. 2^60
      a = ldexpf(1.0f + unit(r2), ea) * ((r3 & 1) ? -1.0f : 1.0f);
      b = ldexpf(1.0f + unit(r3 >> 1), eb) * ((r3 & 2) ? -1.0f : 1.0f);
      if ((r0 >> 32) % 64 == 0) a = 0.0f;
      if (!in_range(a, b)) continue;  // ratio outside the range: not a claim
    } else {
      float h, h2, c6, cs;
      h_consts(r0, &h, &h2, &c6, &cs);
      float r_sq;
      if ((r1 & 7) == 0) {  // the outer edge of the support, a few ulps inside
        r_sq = __uint_as_float(__float_as_uint(h2) - 1u - (uint32_t)((r1 >> 3) % 64));
      } else {
        r_sq = h2 * unit(r1 >> 3);
      }
      if (mode == 1) {
        if ((r2 & 15) == 0) r_sq = 1e-14f * (1.0f + 4.0f * unit(r2 >> 4));  // inner edge
        const float r_len = sqrtf(r_sq);
        if (r_len <= 1e-7f || r_len >= h) continue;  // outside the support: no division
        const float diff = h - r_len;
        a = (-cs) * diff * diff;
        b = r_len;
      } else {
        const float dq = unit(r2);
        const float dqh = dq * h;
        const float x = dqh * dqh;
        if (x >= h2) continue;
        const float dd = h2 - x;
        const float wdq = c6 * dd * dd * dd;
        if (!(wdq > 1e-10f)) continue;  // s_corr off
        const float d = h2 - r_sq;
        a = (r_sq >= h2) ? 0.0f : c6 * d * d * d;
        b = wdq;
      }
      if (!in_range(a, b)) ++outside;
    }
    ++tested;
    if (__float_as_uint(npx_div_in(a, b)) != __float_as_uint(a / b)) ++bad;
  }
  atomicAdd(&out[0], bad);
  atomicAdd(&out[1], outside);
  atomicAdd(&out[2], tested);
}

int check(hipError_t e) { return e == hipSuccess ? 0 : -(int)e; }

}  // namespace

extern "C" {

// Mismatches of npx_sqrt_in against sqrtf over the bit patterns lo..hi; *first
// receives the smallest mismatching pattern (0xffffffff if none).
int npxt_sqrt_check(unsigned lo, unsigned hi, unsigned long long* mismatches, unsigned* first) {
  if (hi < lo) return -1;
  unsigned long long* d_bad = nullptr;
  unsigned* d_first = nullptr;
  int rc = check(hipMalloc(&d_bad, sizeof(*d_bad)));
  if (!rc) rc = check(hipMalloc(&d_first, sizeof(*d_first)));
  const unsigned init_first = 0xffffffffu;
  if (!rc) rc = check(hipMemset(d_bad, 0, sizeof(*d_bad)));
  if (!rc) rc = check(hipMemcpy(d_first, &init_first, sizeof(init_first), hipMemcpyHostToDevice));
  if (!rc) {
    const uint64_t count = (uint64_t)hi - lo + 1;
    hipLaunchKernelGGL(k_sqrt_check, dim3(4096), dim3(256), 0, 0, lo, count, d_bad, d_first);
    rc = check(hipGetLastError());
  }
  if (!rc) rc = check(hipDeviceSynchronize());
  if (!rc) rc = check(hipMemcpy(mismatches, d_bad, sizeof(*d_bad), hipMemcpyDeviceToHost));
  if (!rc) rc = check(hipMemcpy(first, d_first, sizeof(*first), hipMemcpyDeviceToHost));
  if (d_bad) (void)hipFree(d_bad);
  if (d_first) (void)hipFree(d_first);
  return rc;
}

// out[0] mismatches of npx_div_in against '/', out[1] pairs outside the
// stated operand range (modes 1 and 2), out[2] pairs compared.
int npxt_div_check(int mode, unsigned long long n, unsigned long long seed, unsigned long long* out) {
  if (mode < 0 || mode > 2) return -1;
  unsigned long long* d_out = nullptr;
  int rc = check(hipMalloc(&d_out, 3 * sizeof(*d_out)));
  if (!rc) rc = check(hipMemset(d_out, 0, 3 * sizeof(*d_out)));
  if (!rc) {
    hipLaunchKernelGGL(k_div_check, dim3(4096), dim3(256), 0, 0, mode, (uint64_t)n, (uint64_t)seed, d_out);
    rc = check(hipGetLastError());
  }
  if (!rc) rc = check(hipDeviceSynchronize());
  if (!rc) rc = check(hipMemcpy(out, d_out, 3 * sizeof(*d_out), hipMemcpyDeviceToHost));
  if (d_out) (void)hipFree(d_out);
  return rc;
}

}  // extern "C"
