/* Checks the known-divisor division used by the HIP kernels (div_known in
 * alluvion_amd/csrc/alv_device.hpp): q = RN(x rd), r = fma(-q, d, x),
 * RN(q + r rd) with rd = RN(1/d), against the IEEE single-precision quotient
 * x / d (SSE division, correctly rounded).  Built and loaded by
 * tests/test_div_known.py. */
#include <math.h>
#include <stdint.h>
#include <string.h>

static float div_known(float x, float d, float rd) {
  const float q = x * rd;
  const float r = fmaf(-q, d, x);
  return fmaf(r, rd, q);
}

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

/* Every float x with lo <= x <= hi (both positive), and -x: returns the
 * number of mismatches; the first one's x goes to *bad. */
long long check_range(float d, float lo, float hi, float* bad) {
  const float rd = 1.0f / d;
  long long miss = 0;
  for (uint32_t u = bits(lo); u <= bits(hi); ++u) {
    const float x = from_bits(u);
    for (int s = 0; s < 2; ++s) {
      const float xs = s ? -x : x;
      const float want = xs / d;
      const float got = div_known(xs, d, rd);
      if (bits(want) != bits(got)) {
        if (!miss) *bad = xs;
        ++miss;
      }
    }
  }
  return miss;
}

/* n pseudo-random (d, x) pairs: d with a random significand in
 * [2^-12, 2^4), x in [2^-10 d, 2^6 d]; xorshift64 from seed. */
long long check_random(uint64_t seed, long long n, float* bad_d, float* bad_x) {
  long long miss = 0;
  for (long long k = 0; k < n; ++k) {
    seed ^= seed << 13;
    seed ^= seed >> 7;
    seed ^= seed << 17;
    const float d = ldexpf(1.0f + (float)(seed & 0x7fffff) / 8388608.0f,
                           (int)((seed >> 23) % 16) - 12);
    const float x = d * ldexpf(1.0f + (float)((seed >> 31) & 0x7fffff) / 8388608.0f,
                               (int)((seed >> 54) % 16) - 10);
    const float rd = 1.0f / d;
    if (bits(x / d) != bits(div_known(x, d, rd))) {
      if (!miss) {
        *bad_d = d;
        *bad_x = x;
      }
      ++miss;
    }
  }
  return miss;
}
