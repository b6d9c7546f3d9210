// Counter-based random draws for the diffuse-react kernels.
//
// Every molecule owns an independent uint32 stream per (iteration, pass):
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11) with key = 64-bit seed and counter = {block, molecule id,
// iteration, (replica << 8) | pass}.  Draws are consumed x0, x1, x2, x3 of
// block 0, then block 1, ...  A molecule's draws are therefore fixed by its
// id and the iteration, whatever order the GPU runs it in, and a re-run of a
// molecule replays exactly the same numbers.
//
// Gaussians use the Marsaglia-Tsang Ziggurat in exactly the form of MCell's
// rng_gauss (src/rng.c:173-218) with MCell's strip tables, so the same uint32
// draws give the same Gaussian as the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ziggurat_tables.h"

namespace mcg {

struct PhiloxStream {
  uint32_t k0, k1;
  uint32_t c0, c1, c2, c3;
  uint32_t b0, b1, b2, b3;
  uint32_t idx;

  __device__ __forceinline__ void init(uint64_t seed, uint32_t id, uint32_t iteration, uint32_t replica,
                                       uint32_t pass) {
    k0 = (uint32_t)seed;
    k1 = (uint32_t)(seed >> 32);
    c0 = 0;
    c1 = id;
    c2 = iteration;
    c3 = (replica << 8) | (pass & 0xFFu);
    idx = 4;
  }

  __device__ __forceinline__ void refill() {
    uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3;
    uint32_t ka = k0, kb = k1;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
      if (r > 0) {
        ka += 0x9E3779B9u;
        kb += 0xBB67AE85u;
      }
      uint32_t lo0 = 0xD2511F53u * x0;
      uint32_t hi0 = __umulhi(0xD2511F53u, x0);
      uint32_t lo1 = 0xCD9E8D57u * x2;
      uint32_t hi1 = __umulhi(0xCD9E8D57u, x2);
      uint32_t n0 = hi1 ^ x1 ^ ka;
      uint32_t n2 = hi0 ^ x3 ^ kb;
      x0 = n0;
      x1 = lo1;
      x2 = n2;
      x3 = lo0;
    }
    b0 = x0;
    b1 = x1;
    b2 = x2;
    b3 = x3;
    ++c0;
    idx = 0;
  }

  __device__ __forceinline__ uint32_t u32() {
    if (idx == 4) refill();
    uint32_t v = (idx == 0) ? b0 : (idx == 1) ? b1 : (idx == 2) ? b2 : b3;
    ++idx;
    return v;
  }

  // rng_dbl: DBL32 * uint32 (src/isaac64.h:33,71-75)
  __device__ __forceinline__ double dbl() { return 2.3283064365386962890625e-10 * (double)u32(); }
};

// src/rng.c:173-218.  The table reads go through the read-only constant
// tables; the fast path (about 98.8% of draws) is one uint32, one multiply and
// one compare.
__device__ __forceinline__ double gauss(PhiloxStream& r, const ZigguratTables& T) {
  const double SCALE_FACTOR = 3.442619855899;
  const double RECIP_SCALE_FACTOR = 1.0 / 3.442619855899;
  double x, y = 0.0;
  double sign;
  for (;;) {
    uint32_t bits = r.u32();
    sign = (bits & 0x80u) ? -1.0 : 1.0;
    uint32_t region = bits & 0x7Fu;
    uint32_t pos_within_region = bits & 0xFFFFFF00u;
    x = (double)pos_within_region * T.w[region];
    if (pos_within_region < T.k[region]) break;
    if (region != 0) {
      double yB = T.y[region];
      double yR = T.y[region - 1] - yB;
      y = yB + yR * r.dbl();
    } else {
      x = SCALE_FACTOR - log1p(-r.dbl()) * RECIP_SCALE_FACTOR;
      y = exp(-SCALE_FACTOR * (x - 0.5 * SCALE_FACTOR)) * r.dbl();
    }
    if (!(y >= exp(-0.5 * x * x))) break;
  }
  return sign * x;
}

}  // namespace mcg
