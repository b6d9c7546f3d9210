// Brownian displacement of a volume molecule (pick_vol_displacement,
// diffusion_utils.inl:113-119, and compute_vol_displacement's time clipping,
// :366-432).
#pragma once
#include "path64.h"
#include "rng.h"

namespace mcg {

// Full-length step: scale = space_step.
__device__ __forceinline__ D3 draw_displacement(PhiloxStream& rng, const ZigguratTables& T, double scale) {
  D3 d;
  d.x = scale * gauss(rng, T) * 0.70710678118654752440;
  d.y = scale * gauss(rng, T) * 0.70710678118654752440;
  d.z = scale * gauss(rng, T) * 0.70710678118654752440;
  return d;
}

// A step clipped to max_time (units of the time step): steps and t_steps as
// compute_vol_displacement sets them; scale = sqrt(steps) * space_step when
// steps != 1.  Returns r_rate_factor.
__device__ __forceinline__ double step_scale(double space_step, double time_step, double max_time, double& t_steps,
                                             double& scale) {
  const double EPS = 1e-12;
  double steps = 1.0;
  t_steps = steps * time_step;
  if (t_steps > max_time) {
    t_steps = max_time;
    steps = max_time / time_step;
  }
  if (steps < EPS) {
    steps = EPS;
    t_steps = EPS * time_step;
  }
  double r_rate_factor = 1.0;
  scale = space_step;
  if (!(steps == 1.0)) {
    const double rate_factor = sqrt(steps);
    r_rate_factor = 1.0 / rate_factor;
    scale = rate_factor * space_step;
  }
  return r_rate_factor;
}

// The displacement of such a step drawn through rng_gauss.
__device__ __forceinline__ double clipped_displacement(PhiloxStream& rng, const ZigguratTables& T, double space_step,
                                                       double time_step, double max_time, double& t_steps, D3& d) {
  double scale;
  const double r_rate_factor = step_scale(space_step, time_step, max_time, t_steps, scale);
  d = draw_displacement(rng, T, scale);
  return r_rate_factor;
}

}  // namespace mcg
