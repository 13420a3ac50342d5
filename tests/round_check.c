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
