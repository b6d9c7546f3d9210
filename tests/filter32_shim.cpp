// Test helper: exposes the engine's FP32 wall filter (device/filter32.h,
// compiled for the host with the same IEEE single-precision arithmetic and no
// contraction) so that tests can check it against the FP64 collide_wall.
#include <stdint.h>

#include "../mcell_amd/csrc/device/filter32.h"

extern "C" void filter_batch(const float* wall16, float S, uint64_t n, const double* pos, const double* move,
                             int32_t* certain_miss, float* end_dist) {
  mcg::FilterWall w;
  const float* f = wall16;
  w.nx = f[0]; w.ny = f[1]; w.nz = f[2]; w.d = f[3];
  w.ux = f[4]; w.uy = f[5]; w.uz = f[6]; w.uv1u = f[7];
  w.vx = f[8]; w.vy = f[9]; w.vz = f[10]; w.uv2u = f[11];
  w.p0x = f[12]; w.p0y = f[13]; w.p0z = f[14]; w.uv2v = f[15];
  mcg::FilterParams F = mcg::make_filter_params(S);
  for (uint64_t i = 0; i < n; ++i) {
    float de = 0;
    certain_miss[i] = mcg::certain_miss32(w, F, (float)pos[3 * i], (float)pos[3 * i + 1], (float)pos[3 * i + 2],
                                          (float)move[3 * i], (float)move[3 * i + 1], (float)move[3 * i + 2], &de)
                          ? 1
                          : 0;
    end_dist[i] = de;
  }
}
