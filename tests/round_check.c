/* CPU checks of the correctly rounded square root and division the HIP
 * kernels build from the hardware's approximate ones (sqrt_rn and div_rn in
 * alluvion_amd/csrc/alv_device.hpp).  The hardware results are only known to
 * lie within one unit in the last place, so every check runs the sequences
 * from each of the three floats nearest the exact value (the correctly
 * rounded one and its two neighbours) and requires the IEEE result
 * (sqrtf / SSE division, correctly rounded) from all three.  Built and
 * loaded by tests/test_round_check.py. */
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

/* the sequence of div_rn, from the hardware reciprocal estimate y0 */
static float div_from(float a, float b, float y0) {
  const float e = fmaf(-b, y0, 1.0f);
  const float y = fmaf(e, y0, y0);
  const float q0 = a * y;
  const float r0 = fmaf(-b, q0, a);
  const float q1 = fmaf(r0, y, q0);
  const float r1 = fmaf(-b, q1, a);
  return fmaf(r1, y, q1);
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

/* The kernel's use: a = (1 - q)(q - 1) over b = r, q = RN(r / h), for every
 * float r with 1/2 < q <= 1 (the branch of the cubic spline gradient that
 * divides by r). */
long long check_div_gradient(float h, float* bad) {
  long long miss = 0;
  for (uint32_t u = bits(0.5f * h) - 8u; u <= bits(h) + 8u; ++u) {
    const float b = from_bits(u);
    const float q = b / h;
    if (!(q > 0.5f && q <= 1.0f)) continue;
    const float a = (1.0f - q) * (q - 1.0f);
    const float want = a / b;
    const float y = 1.0f / b;
    for (int k = -1; k <= 1; ++k) {
      const float y0 = from_bits(bits(y) + (uint32_t)k);
      if (bits(div_from(a, b, y0)) != bits(want)) {
        if (!miss) *bad = b;
        ++miss;
      }
    }
  }
  return miss;
}

/* n pseudo-random (a, b): b with exponent in [-20, 20), a/b in [2^-40, 2^40]
 * in magnitude, both signs (xorshift64 from seed) */
long long check_div_random(uint64_t seed, long long n, float* bad_a, float* bad_b) {
  long long miss = 0;
  for (long long i = 0; i < n; ++i) {
    seed ^= seed << 13;
    seed ^= seed >> 7;
    seed ^= seed << 17;
    const float b = ldexpf(1.0f + (float)(seed & 0x7fffff) / 8388608.0f,
                           (int)((seed >> 23) % 40) - 20) * ((seed >> 63) ? -1.0f : 1.0f);
    const float a = b * ldexpf(1.0f + (float)((seed >> 29) & 0x7fffff) / 8388608.0f,
                               (int)((seed >> 52) % 80) - 40) * (((seed >> 62) & 1) ? -1.0f : 1.0f);
    const float want = a / b;
    const float y = 1.0f / b;
    for (int k = -1; k <= 1; ++k) {
      const float y0 = from_bits(bits(y) + (uint32_t)k);
      if (bits(div_from(a, b, y0)) != bits(want)) {
        if (!miss) {
          *bad_a = a;
          *bad_b = b;
        }
        ++miss;
      }
    }
  }
  return miss;
}
