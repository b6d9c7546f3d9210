// Diffusion-time bookkeeping shared by the volume and surface phases:
// offsets of a molecule's diffusion time from the start of the iteration.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "react_step.h"

namespace mcg {

// diffuse_single_molecule's time bookkeeping (diffuse_react_event.cpp:
// 201-337): diffusion_time += t_steps, then rounded to a whole iteration when
// within SQRT_EPS of one.  Returns the offset past the next iteration start.
__device__ __forceinline__ double next_toff(uint32_t iteration, double toff, double t_steps) {
  double t = ((double)iteration + toff) + t_steps;
  const double r = rint(t);
  if (cmp_eq_d(t, r, 1e-6)) t = r;
  return t - (double)(iteration + 1);
}

// A molecule that cannot diffuse is scheduled at its unimolecular reaction
// time, or never (TIME_FOREVER: an infinite offset) without one
// (diffuse_single_molecule, diffuse_react_event.cpp:320-333).  Returns the
// offset from this iteration's start.
__device__ __forceinline__ double immobile_toff(uint32_t iteration, double utime) {
  return utime != kTimeInvalid ? utime - (double)iteration : __builtin_inf();
}

// Partition::get_molecules_ready_for_diffusion (partition.h:232-246): the
// molecule's diffusion time lies before the iteration's end,
// cmp_lt(diffusion_time, iteration + 1, EPS).
__device__ __forceinline__ bool ready_this_iteration(uint32_t iteration, double toff) {
  const double t = (double)iteration + toff, e = (double)(iteration + 1);
  return cmp_lt_d(t, e, 1e-12);
}
// The offset of an unmoved molecule's diffusion time past the next
// iteration's start (exact: both are times of one iteration range).
__device__ __forceinline__ double toff_past_iteration(uint32_t iteration, double toff) {
  return ((double)iteration + toff) - (double)(iteration + 1);
}

}  // namespace mcg
