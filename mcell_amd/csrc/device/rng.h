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
      // one 32 x 32 -> 64 multiply per product (v_mad_u64_u32) instead of
      // separate low and high halves
      const uint64_t p0 = (uint64_t)0xD2511F53u * x0;
      const uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
      uint32_t lo0 = (uint32_t)p0;
      uint32_t hi0 = (uint32_t)(p0 >> 32);
      uint32_t lo1 = (uint32_t)p1;
      uint32_t hi1 = (uint32_t)(p1 >> 32);
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

// log1p(x) for x in (-1, 0] with the host C library's rounding.  The
// Ziggurat tail (src/rng.c:205) feeds log1p straight into the Gaussian, and
// the device library's log1p differs from the host's in the last bit for
// about one input in a thousand, so the tail is evaluated here in IEEE basic
// operations only: the fdlibm method (1 + x = 2^k (1 + f) with
// sqrt(2)/2 <= 1 + f < sqrt(2), s = f / (2 + f),
// log(1 + f) = f - (f^2/2 - s (f^2/2 + R(s^2))) with R the degree-7 minimax
// polynomial of fdlibm's Lp1..Lp7, k ln 2 in a hi + lo split, and the
// rounding error of 1 + x carried as c / u), with the polynomial summed in
// the split order of glibc's dbl-64 log1p (R1 + z^2 R2 + z^4 R3 + z^6 R4).
// Equal to glibc's log1p(-u) for every u = k * 2^-32 (all 2^32 k checked).
__host__ __device__ __forceinline__ double log1p_neg(double x) {
  constexpr double ln2_hi = 6.93147180369123816490e-01;  // 0x3FE62E42FEE00000
  constexpr double ln2_lo = 1.90821492927058770002e-10;  // 0x3DEA39EF35793C76
  constexpr double Lp1 = 6.666666666666735130e-01, Lp2 = 3.999999999940941908e-01,
                   Lp3 = 2.857142874366239149e-01, Lp4 = 2.222219843214978396e-01,
                   Lp5 = 1.818357216161805012e-01, Lp6 = 1.531383769920937332e-01,
                   Lp7 = 1.479819860511658591e-01;
  const uint64_t xb = __builtin_bit_cast(uint64_t, x);
  const int32_t hx = (int32_t)(uint32_t)(xb >> 32);
  const int32_t ax = hx & 0x7fffffff;
  int k = 1;
  double f = 0.0, c = 0.0;
  uint32_t hu = 0;
  if (ax >= 0x3ff00000) return -__builtin_inf();  // x = -1
  if (ax < 0x3e200000) {                          // |x| < 2^-29
    if (ax < 0x3c900000) return x;
    return x - x * x * 0.5;
  }
  // (the sign bit makes hx negative: hx <= 0xbfd2bec3 as a signed word is
  // -0.2929 <= x < 0)
  if (hx > (int32_t)0xbfd2bec3) {  // x < -0.2929: reduce
    const double u0 = 1.0 + x;
    uint64_t ub = __builtin_bit_cast(uint64_t, u0);
    hu = (uint32_t)(ub >> 32);
    k = (int)(hu >> 20) - 1023;
    c = (k > 0) ? 1.0 - (u0 - x) : x - (u0 - 1.0);
    c /= u0;
    hu &= 0x000fffffu;
    double u;
    if (hu < 0x6a09eu) {
      u = __builtin_bit_cast(double, ((uint64_t)(hu | 0x3ff00000u) << 32) | (ub & 0xffffffffull));
    } else {
      k += 1;
      u = __builtin_bit_cast(double, ((uint64_t)(hu | 0x3fe00000u) << 32) | (ub & 0xffffffffull));
      hu = (0x00100000u - hu) >> 2;
    }
    f = u - 1.0;
  } else {  // -0.2929 <= x < 0: no reduction
    k = 0;
    f = x;
    hu = 1;
  }
  const double hfsq = 0.5 * f * f;
  if (hu == 0) {  // |f| < 2^-20
    if (f == 0.0) {
      if (k == 0) return 0.0;
      c += k * ln2_lo;
      return k * ln2_hi + c;
    }
    const double R = hfsq * (1.0 - 0.66666666666666666 * f);
    if (k == 0) return f - R;
    return k * ln2_hi - ((R - (k * ln2_lo + c)) - f);
  }
  const double s = f / (2.0 + f);
  const double z = s * s;
  const double R1 = z * Lp1, z2 = z * z;
  const double R2 = Lp2 + z * Lp3, z4 = z2 * z2;
  const double R3 = Lp4 + z * Lp5, z6 = z4 * z2;
  const double R4 = Lp6 + z * Lp7;
  const double R = R1 + z2 * R2 + z4 * R3 + z6 * R4;
  if (k == 0) return f - (hfsq - s * (hfsq + R));
  return k * ln2_hi - ((hfsq - (s * (hfsq + R) + (k * ln2_lo + c))) - f);
}

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
      x = SCALE_FACTOR - log1p_neg(-r.dbl()) * RECIP_SCALE_FACTOR;
      y = exp(-SCALE_FACTOR * (x - 0.5 * SCALE_FACTOR)) * r.dbl();
    }
    if (!(y >= exp(-0.5 * x * x))) break;
  }
  return sign * x;
}

}  // namespace mcg
