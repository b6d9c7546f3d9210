// The common case of a step's Brownian displacement, for the light passes.
//
// rng_gauss (src/rng.c:173-218) takes one uint32 per Gaussian when the draw
// falls inside its Ziggurat strip (pos_within_region < KTAB[region], about
// 98.8 % of draws); only the other draws read more of the stream (wedge and
// tail tests, with exp and log1p).  A step's three Gaussians are therefore
// the first three uint32 of the molecule's stream -- one Philox block --
// whenever all three fall inside their strips (about 96.4 % of steps).  The
// light passes compute exactly that case and hand every other step to the
// full step, which draws the same stream through rng_gauss itself: the
// light passes carry no wedge/tail code, no exp or log, and no second
// Philox block, and a wave never waits on one lane's rejection loop.
#pragma once
#include "rng.h"

namespace mcg {

// The strip tables of the fast path, staged into LDS per workgroup: the
// acceptance bounds, the widths as the reference holds them (double) and
// their FP32 roundings for the FP32 form.
struct FastZiggurat {
  uint32_t k[128];
  double w[128];
  float w32[128];
};

__device__ __forceinline__ void stage_fast_ziggurat(FastZiggurat& F, const ZigguratTables& C) {
  for (int r = threadIdx.x; r < 128; r += blockDim.x) {
    F.k[r] = C.k[r];
    F.w[r] = C.w[r];
    F.w32[r] = (float)C.w[r];
  }
  __syncthreads();
}

// Block 0 of a molecule's stream (PhiloxStream::init + the first refill):
// its first four uint32 draws.
__device__ __forceinline__ void philox_block0(uint64_t seed, uint32_t id, uint32_t iteration, uint32_t replica,
                                              uint32_t pass, uint32_t& r0, uint32_t& r1, uint32_t& r2,
                                              uint32_t& r3) {
  PhiloxStream s;
  s.init(seed, id, iteration, replica, pass);
  s.refill();
  r0 = s.b0;
  r1 = s.b1;
  r2 = s.b2;
  r3 = s.b3;
}

// rng_gauss of one draw when it falls inside its strip (src/rng.c:185-190):
// true and the Gaussian sign * (pos_within_region * WTAB[region]) in FP64,
// bit for bit the reference's value.
__device__ __forceinline__ bool gauss_fast(uint32_t bits, const FastZiggurat& T, double& g) {
  const uint32_t region = bits & 0x7Fu;
  const uint32_t pos = bits & 0xFFFFFF00u;
  const double x = (double)pos * T.w[region];
  g = (bits & 0x80u) ? -x : x;
  return pos < T.k[region];
}

// The same in FP32: pos_within_region has 24 significant bits and is exact
// in FP32; the width is FP32-rounded, so the result is within one FP32 ulp
// of the FP64 Gaussian rounded to FP32.
__device__ __forceinline__ bool gauss_fast32(uint32_t bits, const FastZiggurat& T, float& g) {
  const uint32_t region = bits & 0x7Fu;
  const uint32_t pos = bits & 0xFFFFFF00u;
  const float x = (float)pos * T.w32[region];
  g = (bits & 0x80u) ? -x : x;
  return pos < T.k[region];
}

// The FP64 displacement of a full-length step whose three Gaussians are on
// the fast path: (scale * g) * sqrt(1/2) per axis, as draw_displacement
// computes it.  False when any draw leaves its strip.
__device__ __forceinline__ bool fast_displacement(uint32_t r0, uint32_t r1, uint32_t r2, const FastZiggurat& T,
                                                  double scale, double& dx, double& dy, double& dz) {
  double g0, g1, g2;
  const bool f0 = gauss_fast(r0, T, g0), f1 = gauss_fast(r1, T, g1), f2 = gauss_fast(r2, T, g2);
  dx = scale * g0 * 0.70710678118654752440;
  dy = scale * g1 * 0.70710678118654752440;
  dz = scale * g2 * 0.70710678118654752440;
  return f0 && f1 && f2;
}

// The FP32 displacement: g32 * fp32(scale * sqrt(1/2)); each component is
// within kMoveUlps FP32 ulps of the FP64 displacement rounded to FP32.
constexpr float kMoveUlps = 4.0f;
__device__ __forceinline__ bool fast_displacement32(uint32_t r0, uint32_t r1, uint32_t r2, const FastZiggurat& T,
                                                    float scale32, float& dx, float& dy, float& dz) {
  float g0, g1, g2;
  const bool f0 = gauss_fast32(r0, T, g0), f1 = gauss_fast32(r1, T, g1), f2 = gauss_fast32(r2, T, g2);
  dx = g0 * scale32;
  dy = g1 * scale32;
  dz = g2 * scale32;
  return f0 && f1 && f2;
}

}  // namespace mcg
