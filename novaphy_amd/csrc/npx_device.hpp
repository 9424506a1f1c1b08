// npx_device.hpp -- per-scene constants and the SPH arithmetic of the PBF path.
//
// Every function here reproduces the reference's float operation order
// (include/novaphy/fluid/sph_kernel.h, src/fluid/pbf_solver.cpp) when built
// with -ffp-contract=off and correctly rounded sqrt / divide, so the GPU step
// is bitwise equal to the FP32 CPU restatement on the same input.  The FAST
// variants (NPX_FLAG_FAST_MATH) trade that for FMA / rsqrt arithmetic.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define NPX_BLOCK 256      // scene segment alignment and block size of the sort kernels
#define NPX_MAX_BLOCK 256  // largest block size of the neighbour passes (tunable, <= NPX_BLOCK)
#define NPX_SLOT_PAD 64     // elements of padding after the x* buffers (reads past a slot range are masked)
#define NPX_WINDOW_MARGIN 4

// Constants of one scene, computed on the host with the reference's float
// operations (each value is what the reference computes per call).
struct __align__(16) SceneC {
  float h, h2, inv_h, c6;          // c6 = 315 / (64*PI*h^9)        sph_kernel.h:29
  float cs, rho0, inv_rho0, mass;  // cs = 45 / (PI*h^6)            sph_kernel.h:44
  float m_inv_rho0, eps, xsph_c, vort_eps;
  float corr_k, wdq, dt, inv_dt;   // wdq = poly6((dq*h)^2)         pbf_solver.cpp:188-189
  float max_speed, max_speed_sq, gdt0, gdt1;
  float gdt2, clo0, clo1, clo2;    // clamp box: domain_lower + r .. domain_upper - r
  float chi0, chi1, chi2, cut2;     // cut2: beyond it poly6 and spiky_grad both vanish
  int corr_n, iters, bounds, xsph_on;
  int wlo0, wlo1, wlo2, wx;        // cell window origin (global cell coords) and dims
  int wy, wz, cls_base, ncell;     // class ids of this scene: [cls_base, cls_base + ncell)
  int seg_off, n, seg_blocks, vort_on;
  int nbnd, fdiv, pad_i2, pad_i3;  // nbnd: boundary particles of the scene (batched coupling);
                                   // fdiv: h in [2^-6, 2^6], so npx_div_in serves the per-pair divisions
  float sp_lo, sp_hi;               // spiky support on r^2: (sp_lo, sp_hi), i.e. 1e-7 < |r| < h
  float skin_cut2, skin_allow;      // skin radius^2; largest summed displacement a skin list allows
};

struct __align__(16) GridWin {  // window of an auxiliary grid (boundary / committed fluid)
  float inv_h;
  int wlo0, wlo1, wlo2;
  int wx, wy, wz, ncell;
};

// ------------------------------------------------------------ arithmetic
__device__ __forceinline__ float sqn3(float x, float y, float z) {
  return x * x + (y * y + z * z);  // Eigen 3-vector squaredNorm order
}

// SPHKernels::poly6 (sph_kernel.h:24-30)
__device__ __forceinline__ float poly6(float r_sq, float h2, float c6) {
  if (r_sq >= h2) return 0.0f;
  float d = h2 - r_sq;
  return c6 * d * d * d;
}

// SPHKernels::spiky_grad (sph_kernel.h:39-46): returns the scalar that
// multiplies r, or flags out-of-support with *in = false (zero vector).
__device__ __forceinline__ float spiky_coeff(float r_sq, float h, float cs, bool* in) {
  float r_len = sqrtf(r_sq);
  if (r_len <= 1e-7f || r_len >= h) {
    *in = false;
    return 0.0f;
  }
  *in = true;
  float diff = h - r_len;
  return (-cs) * diff * diff / r_len;
}

// Correctly rounded square root of a finite x >= 2^-96.  This is the
// sequence the compiler emits for a correctly rounded sqrtf (the hardware
// estimate, within one ulp, corrected by the residuals of its two
// neighbours) without the steps that only act on other inputs: the 2^32
// pre-scaling of x < 2^-96 and the pass-through of zeros and infinities.
// Same bits as sqrtf on that range (checked exhaustively on the device by
// tests/hip/arith_check.hip).
__device__ __forceinline__ float npx_sqrt_in(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sd = __uint_as_float(__float_as_uint(s) - 1u);
  const float su = __uint_as_float(__float_as_uint(s) + 1u);
  const float rd = __builtin_fmaf(-sd, s, x);
  const float ru = __builtin_fmaf(-su, s, x);
  const float t = (rd <= 0.0f) ? sd : s;
  return (ru > 0.0f) ? su : t;
}

// Correctly rounded a / b for operands where the compiler's correctly
// rounded division sequence scales nothing and fixes nothing up: b normal
// with 2^-60 <= |b| <= 2^60, a = 0 or 2^-100 <= |a| <= 2^60, and
// 2^-120 <= |a / b| <= 2^90.  There v_div_scale returns its input with VCC
// clear, v_div_fmas is a plain fma and v_div_fixup returns the refined
// quotient, so the Newton-Raphson refinement below is that sequence step for
// step.  Same bits as a / b on that range (checked on the device by
// tests/hip/arith_check.hip).  For h in [2^-6, 2^6] every per-pair division
// of the neighbour passes is in range: spiky_grad's (-cs) diff^2 / |r| has
// |r| in (1e-7, h), diff > h 2^-25 and cs = 45 / (PI h^6), and the s_corr
// ratio poly6(r^2) / poly6((dq h)^2) has a numerator above 2^-93 or zero and a
// denominator in (1e-10, 2^19).
__device__ __forceinline__ float npx_div_in(float a, float b) {
  float r = __builtin_amdgcn_rcpf(b);
  const float e = __builtin_fmaf(-b, r, 1.0f);
  r = __builtin_fmaf(e, r, r);
  float q = a * r;
  const float e1 = __builtin_fmaf(-b, q, a);
  q = __builtin_fmaf(e1, r, q);
  const float e2 = __builtin_fmaf(-b, q, a);
  return __builtin_fmaf(e2, r, q);
}

// a / b, through npx_div_in when the scene's constants keep the operands in
// its range (fd, uniform over a block).
__device__ __forceinline__ float npx_div(float a, float b, bool fd) {
  if (fd) return npx_div_in(a, b);
  return a / b;
}

// spiky_coeff with the support test done on r^2 against the exact bounds
// sp_lo = max{x : sqrtf(x) <= 1e-7} and sp_hi = min{x : sqrtf(x) >= h}, so the
// square root is taken only inside the support (where r^2 > 1e-14 > 2^-96);
// same results bit for bit.
__device__ __forceinline__ float spiky_coeff_sq(float r_sq, float sp_lo, float sp_hi, float h, float cs, bool fd,
                                                bool* in) {
  if (r_sq <= sp_lo || r_sq >= sp_hi) {
    *in = false;
    return 0.0f;
  }
  *in = true;
  const float r_len = npx_sqrt_in(r_sq);
  const float diff = h - r_len;
  return npx_div((-cs) * diff * diff, r_len, fd);
}

// Fast variant: one rsqrt, FMA-friendly; support tested on r_sq
// (|r| <= 1e-7  <=>  r_sq <= 1e-14, |r| >= h  <=>  r_sq >= h^2).
__device__ __forceinline__ float spiky_coeff_fast(float r_sq, float h, float h2, float cs, bool* in) {
  if (r_sq <= 1e-14f || r_sq >= h2) {
    *in = false;
    return 0.0f;
  }
  *in = true;
  float rinv = __frsqrt_rn(r_sq);
  float diff = h - r_sq * rinv;
  return (-cs) * diff * diff * rinv;
}

__device__ __forceinline__ uint64_t teschner_key(int cx, int cy, int cz) {
  return ((uint64_t)(int64_t)cx * 73856093ull) ^ ((uint64_t)(int64_t)cy * 19349663ull) ^
         ((uint64_t)(int64_t)cz * 83492791ull);
}

__device__ __forceinline__ int cell_coord(float p, float inv_h) {
  return (int)floorf(p * inv_h);  // SpatialHashGrid::world_to_cell, neighbor_search.cpp:99-105
}

// Visits the candidate list of a particle whose cell has local window index
// l, in the reference's order: dz outer, dy, dx inner, each offset's bucket in
// ascending particle index, a bucket visited again when two offsets share a
// Teschner key (query_neighbors, neighbor_search.cpp:28-49).  clsmap maps a
// window cell to its class (cells with equal keys share one class); cstart is
// the exclusive prefix of class populations, minus `shift` giving the slot.
template <class F>
__device__ __forceinline__ void for_each_candidate(int l, int wx, int wxy, const int* __restrict__ clsmap,
                                                   const int* __restrict__ cstart, int shift, F&& f) {
#pragma unroll 1
  for (int dz = -1; dz <= 1; ++dz) {
#pragma unroll 1
    for (int dy = -1; dy <= 1; ++dy) {
      const int base = l + dy * wx + dz * wxy;
      const int c0 = clsmap[base - 1];
      const int c1 = clsmap[base];
      const int c2 = clsmap[base + 1];
      if (c1 == c0 + 1 && c2 == c0 + 2) {
        const int b = cstart[c0] - shift, e = cstart[c0 + 3] - shift;
        for (int j = b; j < e; ++j) f(j);
      } else {
        const int cc[3] = {c0, c1, c2};
#pragma unroll 1
        for (int k = 0; k < 3; ++k) {
          const int b = cstart[cc[k]] - shift, e = cstart[cc[k] + 1] - shift;
          for (int j = b; j < e; ++j) f(j);
        }
      }
    }
  }
}
