// Host build of the engine's Ziggurat-tail log1p (mcell_amd/csrc/device/rng.h)
// for tests/test_rng_tail.py: it must equal the C library's log1p on every
// input the tail gives it, -u with u = k * 2^-32.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../mcell_amd/csrc/device/rng.h"

extern "C" {

// Inputs k0, k0 + stride, ... (n of them, modulo 2^32); returns how many
// differ from log1p bit for bit, and the first such k in *first.
uint64_t log1p_tail_mismatches(uint64_t k0, uint64_t stride, uint64_t n, uint64_t* first) {
  uint64_t bad = 0;
  *first = ~0ull;
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t k = (uint32_t)(k0 + i * stride);
    const double x = -(2.3283064365386962890625e-10 * (double)k);
    const double a = log1p(x), b = mcg::log1p_neg(x);
    uint64_t ua, ub;
    memcpy(&ua, &a, 8);
    memcpy(&ub, &b, 8);
    if (ua != ub) {
      if (bad == 0) *first = k;
      ++bad;
    }
  }
  return bad;
}
}
