// Checks npx_rep_add(a, t, k) against k sequential float additions, bit for
// bit, over seeded random cases: mixed magnitudes and signs, exact ties
// (t = (n + 1/2) ulp), sequences that cross zero or binade edges, zeros,
// subnormals, inf and NaN.  Prints the first mismatches; exit status 1 if any.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <initializer_list>

#include "../../novaphy_amd/csrc/npx_repadd.hpp"

static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint64_t rnd() {
  st ^= st << 13;
  st ^= st >> 7;
  st ^= st << 17;
  return st;
}
static double uni() { return (rnd() >> 11) * (1.0 / 9007199254740992.0); }
static float logu(double lo, double hi) { return (float)exp(log(lo) + uni() * (log(hi) - log(lo))); }
static float sgn(float x) { return (rnd() & 1) ? -x : x; }

static long fails = 0, cases = 0;
static void check(float a, float t, int k) {
  float ref = a;
  for (int i = 0; i < k; ++i) ref = ref + t;
  const float got = npx_rep_add(a, t, k);
  ++cases;
  const bool same = (isnan(ref) && isnan(got)) || npx_f2u(ref) == npx_f2u(got);
  if (!same) {
    if (fails < 20)
      printf("MISMATCH a=%a t=%a k=%d ref=%a got=%a\n", (double)a, (double)t, k, (double)ref, (double)got);
    ++fails;
  }
}

int main(int argc, char** argv) {
  const int n = argc > 1 ? atoi(argv[1]) : 100000;
  for (int i = 0; i < n; ++i) {
    const int k = 1 + (int)(rnd() % 2000);
    // mixed magnitudes, t relative to a
    float a = sgn(logu(1e-30, 1e30));
    check(a, a * sgn(logu(1e-12, 10.0)), k);
    // accumulators starting at +0 (the neighbour sums)
    check(0.0f, sgn(logu(1e-20, 1e20)), k);
    // exact ties: a = M u, t = (m + 1/2) u
    const int e = 100 + (int)(rnd() % 50);
    const float u = ldexpf(1.0f, e - 150);
    const uint32_t M = 0x800000u | (uint32_t)(rnd() & 0x7fffffu);
    a = npx_u2f(((uint32_t)e << 23) | (M - 0x800000u));
    const float half = (float)(rnd() % 64) + 0.5f;
    check(sgn(a), sgn(half * u), k);
    check(a, -half * u, k);
    // crossing zero
    check(sgn(logu(1e-3, 1e3)), -logu(1e-3, 1.0) , k);
    // near a binade edge
    const float edge = ldexpf(1.0f, (int)(rnd() % 40) - 20);
    check(sgn(edge) * (1.0f - logu(1e-7, 1e-3)), sgn(edge) * logu(1e-9, 1e-3) * edge, k);
    // subnormal range
    check(sgn(logu(1e-44, 1e-38)), sgn(logu(1e-45, 1e-40)), 1 + (int)(rnd() % 200));
  }
  const float sp[] = {0.0f, -0.0f, 1.0f, -1.0f, INFINITY, -INFINITY, NAN, 1e-45f, -1e-45f, 3.4e38f, -3.4e38f};
  for (float x : sp)
    for (float y : sp)
      for (int k : {1, 2, 3, 4, 7, 100, 1000}) check(x, y, k);
  // long runs
  for (int i = 0; i < 40; ++i) {
    const float a = sgn(logu(1e-6, 1e6));
    check(a, a * sgn(logu(1e-9, 1e-2)), 200000 + (int)(rnd() % 800000));
    check(0.0f, sgn(logu(1e-6, 1.0)), 1000000);
  }
  printf("%ld cases, %ld mismatches\n", cases, fails);
  return fails ? 1 : 0;
}
