/* CPU check of the correctly rounded square root the HIP kernels build from
 * the hardware's approximate one (sqrt_rn in alluvion_amd/csrc/alv_device.hpp).
 * The hardware result is only known to lie within one unit in the last
 * place, so the check runs the sequence from each of the three floats nearest
 * the exact root (the correctly rounded one and its two neighbours) and
 * requires the IEEE result (sqrtf, correctly rounded) from all three.  Built
 * and loaded by tests/test_round_check.py. */
#include <math.h>
#include <stdint.h>
#include <string.h>

static uint32_t bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
static float from_bits(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

/* the sequence of sqrt_rn, from the hardware estimate s */
static float sqrt_from(float x, float s) {
  const float s_dn = from_bits(bits(s) - 1u);
  const float s_up = from_bits(bits(s) + 1u);
  const float r_dn = fmaf(-s_dn, s, x);
  const float r_up = fmaf(-s_up, s, x);
  float t = r_dn <= 0.0f ? s_dn : s;
  return r_up > 0.0f ? s_up : t;
}

/* every float x in [lo, hi] (positive) and zero; returns mismatches, the
 * first one's x in *bad */
long long check_sqrt(float lo, float hi, float* bad) {
  long long miss = 0;
  for (uint32_t u = bits(lo); u <= bits(hi) + 1u; ++u) {
    const float x = u > bits(hi) ? 0.0f : from_bits(u);
    const float want = sqrtf(x);
    for (int k = -1; k <= 1; ++k) {
      if (want == 0.0f && k < 0) continue;
      const float s = from_bits(bits(want) + (uint32_t)k);
      if (bits(sqrt_from(x, s)) != bits(want)) {
        if (!miss) *bad = x;
        ++miss;
      }
    }
  }
  return miss;
}

/* the sequence of rcp_rn, from the hardware estimate r0 */
static float rcp_from(float y, float r0) {
  const float e = fmaf(-y, r0, 1.0f);
  return fmaf(e, r0, r0);
}

/* Every significand y in [1, 2) and every estimate r0 within one ulp of 1/y
 * (|r0 y - 1| < ulp(r0) y, exact in double): counts the (y, r0) whose Newton
 * step misses RN(1/y) = 1.0f / y, the last one's y in *bad and r0 in
 * *bad_r0, and the estimates tried in *tried.  The reciprocal's significand
 * depends on y's significand only, so one binade covers every normal y whose
 * reciprocal is normal. */
long long check_rcp(float* bad, float* bad_r0, long long* tried) {
  long long miss = 0, n = 0;
  for (uint32_t m = 0; m < (1u << 23); ++m) {
    const float y = from_bits(0x3f800000u | m);
    const float want = 1.0f / y;
    for (int k = -2; k <= 2; ++k) {
      const float r0 = from_bits(bits(want) + (uint32_t)k);
      const double ulp = ldexp(1.0, ilogb(r0) - 23);
      if (fabs((double)r0 * (double)y - 1.0) >= ulp * (double)y) continue;
      ++n;
      if (bits(rcp_from(y, r0)) != bits(want)) {
        *bad = y;
        *bad_r0 = r0;
        ++miss;
      }
    }
  }
  *tried = n;
  return miss;
}

/* The gradient's outer branch (1 - q)(q - 1) / r, q = RN(r / h), for every
 * float r in (h/2, h], by the known-divisor sequence with r's correctly
 * rounded reciprocal (div_rn); returns mismatches against the IEEE division,
 * the first one's r in *bad. */
long long check_div_gradient(float h, float* bad) {
  long long miss = 0;
  for (uint32_t u = bits(0.5f * h) + 1u; u <= bits(h); ++u) {
    const float r = from_bits(u);
    const float q = r / h;
    const float x = (1.0f - q) * (q - 1.0f);
    const float want = x / r;
    const float rr = 1.0f / r;
    const float q0 = x * rr;
    const float rem = fmaf(-q0, r, x);
    if (bits(fmaf(rem, rr, q0)) != bits(want)) {
      if (!miss) *bad = r;
      ++miss;
    }
  }
  return miss;
}
