// Test helper: exposes the engine's FP32 wall filter (device/filter32.h,
// compiled for the host with the same IEEE single-precision arithmetic and no
// contraction) so that tests can check it against the FP64 collide_wall.
#include <math.h>
#include <stdint.h>

#include "../mcell_amd/csrc/device/filter32.h"

// move_ulps > 0: each FP32 move component is moved by up to move_ulps ulps
// (a fixed pseudo-random pattern), as a move computed in FP32 may differ from
// the FP64 one, and the filter is told so (FilterParams::mrel).
extern "C" void filter_batch_perturbed(const float* wall16, float S, float move_ulps, uint64_t n, const double* pos,
                                       const double* move, int32_t* certain_miss, float* end_dist);

extern "C" void filter_batch(const float* wall16, float S, uint64_t n, const double* pos, const double* move,
                             int32_t* certain_miss, float* end_dist) {
  filter_batch_perturbed(wall16, S, 0.0f, n, pos, move, certain_miss, end_dist);
}

static float nudge(float x, int k) {
  for (; k > 0; --k) x = nextafterf(x, 3.4e38f);
  for (; k < 0; ++k) x = nextafterf(x, -3.4e38f);
  return x;
}

extern "C" void filter_batch_perturbed(const float* wall16, float S, float move_ulps, uint64_t n, const double* pos,
                                       const double* move, int32_t* certain_miss, float* end_dist) {
  mcg::FilterWall w;
  const float* f = wall16;
  w.nx = f[0]; w.ny = f[1]; w.nz = f[2]; w.d = f[3];
  w.ux = f[4]; w.uy = f[5]; w.uz = f[6]; w.uv1u = f[7];
  w.vx = f[8]; w.vy = f[9]; w.vz = f[10]; w.uv2u = f[11];
  w.p0x = f[12]; w.p0y = f[13]; w.p0z = f[14]; w.uv2v = f[15];
  mcg::FilterParams F = mcg::make_filter_params(S, move_ulps);
  const int u = (int)move_ulps;
  for (uint64_t i = 0; i < n; ++i) {
    float de = 0;
    uint64_t h = (i + 1) * 0x9E3779B97F4A7C15ull;
    float m[3];
    for (int a = 0; a < 3; ++a) {
      h ^= h >> 29;
      h *= 0xBF58476D1CE4E5B9ull;
      const int k = u ? (int)((h >> 40) % (uint64_t)(2 * u + 1)) - u : 0;
      m[a] = nudge((float)move[3 * i + a], k);
    }
    certain_miss[i] = mcg::certain_miss32(w, F, (float)pos[3 * i], (float)pos[3 * i + 1], (float)pos[3 * i + 2],
                                          m[0], m[1], m[2], &de)
                          ? 1
                          : 0;
    end_dist[i] = de;
  }
}
