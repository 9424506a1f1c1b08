// npx_repadd.hpp -- k sequential float additions of one value, exactly.
//
// npx_rep_add(a, t, k) returns the value of
//     for (i = 0; i < k; ++i) a = a + t;
// in IEEE binary32 with round-to-nearest-even, bit for bit, in O(number of
// binades crossed) instead of O(k).  The neighbour passes use it where a run
// of candidates with identical state contributes the same term k times in a
// row (particles stacked on one point).
//
// Why it is exact.  Take a > 0 (the negative case is its mirror image, since
// rounding to nearest-even is symmetric) in the binade [2^E, 2^(E+1)) with
// unit u = 2^(E-23), so a = M u with an integer significand M in
// [2^23, 2^24).  When the exact sum a + t lies in [2^23 u, (2^24 - 1) u], the
// rounded sum is (M + round(t/u)) u, where a tie (frac(t/u) = 1/2) goes to the
// even result.  Without a tie the increment is the same integer for every M.
// With a tie the result is even, so after one tie step M is even, every later
// increment is even, and it is again the same for every M.  Hence after two
// explicit steps whose exact sums stay in that range, the increment D of the
// second one repeats for every further step whose exact sum stays in range,
// and the count of such steps follows from integer arithmetic.  Steps across
// a binade edge, at zero, subnormals, inf and NaN are done one by one.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NPX_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define NPX_HD inline
#endif

NPX_HD uint32_t npx_f2u(float x) {
  uint32_t u;
  __builtin_memcpy(&u, &x, 4);
  return u;
}

NPX_HD float npx_u2f(uint32_t u) {
  float x;
  __builtin_memcpy(&x, &u, 4);
  return x;
}

// Largest n >= 0 with m + (n - 1) d staying within bound (steps of size
// |d| >= 1 from m towards the bound, all values below 2^25): n = k when the
// whole remaining run fits, so the division is only taken at a binade edge.
NPX_HD int npx_steps_within(int span, int d, int k) {  // span = |bound - m| >= 0, d > 0
  if ((long long)(k - 1) * d <= span) return k;
  return span / d + 1;
}

NPX_HD float npx_rep_add(float a, float t, int k) {
  for (;;) {
    if (k <= 3) {
      for (; k > 0; --k) a = a + t;
      return a;
    }
    const float a1 = a + t;
    const float a2 = a1 + t;
    const float a3 = a2 + t;
    k -= 3;
    a = a3;
    if (a3 == a2) return a3;  // a fixed point of x -> x + t (inf included)
    const uint32_t b1 = npx_f2u(a1), b2 = npx_f2u(a2), b3 = npx_f2u(a3);
    const uint32_t se = b1 >> 23;  // sign and exponent
    const uint32_t e = se & 0xffu;
    if ((b2 >> 23) != se || (b3 >> 23) != se || e == 0u || e == 255u) continue;
    // t / u in the positive frame; exact: a2 != a3 gives |t| >= u / 2, and
    // the range test keeps it far from overflow
    const float tu = ldexpf((b1 >> 31) ? -t : t, 150 - (int)e);
    if (!(fabsf(tu) < 4194304.0f)) continue;
    const int tf = (int)floorf(tu), tc = (int)ceilf(tu);
    const int m1 = (int)((b1 & 0x7fffffu) | 0x800000u);
    const int m2 = (int)((b2 & 0x7fffffu) | 0x800000u);
    const int m3 = (int)((b3 & 0x7fffffu) | 0x800000u);
    // exact sums m1 + tu and m2 + tu within [2^23, 2^24 - 1]
    if (m1 + tf < 8388608 || m1 + tc > 16777215 || m2 + tf < 8388608 || m2 + tc > 16777215) continue;
    const int d = m3 - m2;
    int n;  // further steps from m3 whose exact sums stay in [2^23, 2^24 - 1]
    if (d > 0) {
      const int lim = 16777215 - tc;  // m + tu <= 2^24 - 1
      n = (m3 > lim) ? 0 : npx_steps_within(lim - m3, d, k);
    } else {
      const int lo = 8388608 - tf;  // m + tu >= 2^23
      n = (m3 < lo) ? 0 : npx_steps_within(m3 - lo, -d, k);
    }
    if (n > 0) {
      const int m = m3 + n * d;
      a = npx_u2f((se << 23) | (uint32_t)(m - 0x800000));
      k -= n;
    }
  }
}
