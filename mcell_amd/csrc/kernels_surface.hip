// The surface phase of an iteration: every surface molecule of a species
// that diffuses takes its step (diffuse_surf_molecule,
// diffuse_react_event.cpp:1071-1246) after the iteration's volume phases,
// one thread per tile slot.  A molecule whose end point lies on its own tile
// moves there at once; one whose end point lies on another tile that is
// free proposes the move and claims the tile (smallest id wins); a tile
// that is taken makes it draw again (move_sm_on_same_triangle /
// move_sm_to_new_triangle, diffusion_utils.inl:453-548).  Winners move,
// losers repeat their step from the start of their stream in the next
// round, seeing the tiles the winners took (and left).  The reference moves
// molecules one at a time in id order, so a tile whose occupant has a
// smaller id and has not yet taken its step may still be vacated: a mover
// that meets one waits for the next round instead of drawing again.
// Species with time steps below one iteration step again in further passes
// of the phase (the reference's later diffuse actions of the iteration),
// each pass reading the next pass of the molecule's stream.  The order the
// waits and claims follow is the reference's: molecules whose diffusion
// time falls on the iteration's start first, then the delayed ones
// (diffuse_molecules, diffuse_react_event.cpp:66-113), each group by id.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device/schedule.h"
#include "device/surface.h"
#include "kernels_react.h"
#include "kernels_surface.h"

namespace mcg {

namespace {

inline dim3 blocks_for(uint32_t n) { return dim3((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

__global__ void k_fill_claims(unsigned long long* claim, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) claim[i] = ~0ull;
}

// A molecule's place in the phase's order: delayed << 32 | id.  In the
// first pass, delayed is read from the diffusion-time offset the molecule
// starts the iteration with (cmp_eq(delay, 0) fails); later passes keep
// the first pass's order through the tile word.
__device__ __forceinline__ bool surf_delayed(const SurfPhase& P, uint32_t tw, double toff) {
  return P.pass == 0 ? !cmp_eq_d(toff, 0.0, 1e-12) : (tw & kTileDelayed) != 0;
}
__device__ __forceinline__ unsigned long long surf_key(bool delayed, uint32_t id) {
  return ((unsigned long long)delayed << 32) | id;
}

// The time a surface molecule's step covers (diffuse_surf_molecule, :1081-
// 1116) with get_max_time's limits (:164-198): the time up to the barrier,
// shortened for a young molecule of a species with a long time step (no
// unimolecular clipping on this path).
struct SurfTiming {
  double t_steps, space_factor;
  bool mature;  // the molecule becomes mature with this step
};

__device__ __forceinline__ SurfTiming surf_timing(const SurfPhase& P, const SpeciesRec& sp, uint32_t tw, double toff,
                                                  uint32_t s) {
  SurfTiming r;
  double max_time = P.barrier - toff;
  r.mature = false;
  if (!(tw & kTileMature) && sp.time_step > 1.0) {
    const double age_dependent_max_time = 1.0 + 0.2 * (((double)P.iteration + toff) - P.birth[s]);
    if (max_time > age_dependent_max_time) max_time = age_dependent_max_time;
    r.mature = age_dependent_max_time > sp.time_step;
  }
  constexpr double EPS = 1e-12;
  r.t_steps = sp.time_step > max_time ? max_time : sp.time_step;
  double steps = 1.0;
  if (sp.time_step > max_time) {
    steps = max_time / sp.time_step;
    if (steps < EPS) {
      r.t_steps = EPS * sp.time_step;
      steps = EPS;
    }
  }
  r.space_factor = steps == 1.0 ? sp.space_step : sp.space_step * sqrt(steps);
  return r;
}

// The end of a molecule's step in slot s (diffuse_single_molecule's
// rescheduling, :287-318): diffusion_time += t_steps; still before the
// iteration's end, the molecule steps again in the next pass (its offset
// kept from this iteration's start), else its offset moves past the next
// iteration's start.
__device__ __forceinline__ void surf_step_over(const SurfPhase& P, uint32_t s, uint32_t tw, double toff,
                                               const SurfTiming& tm, bool delayed) {
  const double t_new = ((double)P.iteration + toff) + tm.t_steps;
  const bool again = cmp_lt_d(t_new, (double)P.iteration + 1.0, 1e-12);
  const uint32_t tw2 = (tw & ~kTileDelayed) | (tm.mature ? kTileMature : 0u) | (delayed ? kTileDelayed : 0u);
  if (tw2 != tw) P.tiles[s] = tw2;
  P.ttoff[s] = again ? t_new - (double)P.iteration : next_toff(P.iteration, toff, tm.t_steps);
  if (again) {
    if (P.pass + 1 >= kSurfPasses) {
      atomicOr(P.err, (uint32_t)ERR_PASSES);
      P.done[s] = P.stamp_base | kDoneFinal;
      return;
    }
    P.done[s] = P.stamp_base | (P.pass + 1);
    P.clist[atomicAdd(P.ccount, 1u)] = s;
  } else {
    P.done[s] = P.stamp_base | kDoneFinal;
  }
}

// One surface molecule's step, for the slots of `list` (every slot when
// list is null: the first round of the first pass, which also books
// molecules due in later iterations).
__global__ __launch_bounds__(256) void k_surf_move(SurfPhase P, const uint32_t* __restrict__ list,
                                                   const uint32_t* __restrict__ count, uint32_t n_all,
                                                   uint32_t* __restrict__ lose_list,
                                                   uint32_t* __restrict__ lose_count) {
  const uint32_t n = list ? *count : n_all;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = list ? list[i] : i;
  const uint32_t tw = P.tiles[s];
  if (tw == kTileEmpty) return;
  const uint32_t rep = s / P.total_tiles, lt = s - rep * P.total_tiles;
  const SpeciesRec& sp = P.species.rec[rep * P.species.stride + tile_species(tw)];
  if (!(sp.flags & SPF_SURFACE) || !(sp.space_step > 0)) {
    P.done[s] = P.stamp_base | kDoneFinal;
    return;
  }
  const double toff = P.ttoff[s];
  if (!list && !ready_this_iteration(P.iteration, toff)) {
    P.ttoff[s] = toff_past_iteration(P.iteration, toff);
    P.done[s] = P.stamp_base | kDoneFinal;
    return;
  }
  const SurfTiming tm = surf_timing(P, sp, tw, toff, s);
  const uint32_t id = P.tile_id[s];
  const bool delayed = surf_delayed(P, tw, toff);
  const unsigned long long key = surf_key(delayed, id);
  PhiloxStream rng;
  rng.init(P.seed, id, P.iteration, rep, P.pass);
  const uint32_t wi = P.tile_wall[lt];
  const uint32_t t0 = lt - P.grids[wi].tile_base;
  const double u = P.tu[s], v = P.tv[s];
  uint32_t err = 0;
  for (int attempt = 0; attempt <= kSurfRetries; ++attempt) {
    double du, dv;
    surf_displacement(rng, tm.space_factor, du, dv);
    double nu = 0, nv = 0;
    const uint32_t nw = ray_trace_surf(P.walls, P.edges, wi, u, v, du, dv, nu, nv, err);
    if (nw == kInvalid) continue;  // an ambiguous edge crossing: draw again
    const WallGrid& g = P.grids[nw];
    const uint32_t nt = uv2grid(P.walls[nw], g, nu, nv);
    if (nt == kInvalid) {
      err |= ERR_GRID;
      break;
    }
    if (nw == wi && nt == t0) {  // the same tile: the molecule moves within it
      P.tu[s] = nu;
      P.tv[s] = nv;
      surf_step_over(P, s, tw, toff, tm, delayed);
      if (err) atomicOr(P.err, err);
      return;
    }
    const uint32_t ts = rep * P.total_tiles + g.tile_base + nt;
    const uint32_t occ = P.tiles[ts];
    if (occ != kTileEmpty) {
      // an occupant ahead in the order that still steps in this pass: wait
      // for it (in the first pass: one not yet stamped in this iteration;
      // later: one listed for this pass and not yet done)
      const SpeciesRec& os = P.species.rec[rep * P.species.stride + tile_species(occ)];
      const uint32_t d = P.done[ts];
      const bool pending = P.pass == 0 ? (d & ~15u) != P.stamp_base : d == (P.stamp_base | P.pass);
      if (pending && (os.flags & SPF_SURFACE) && os.space_step > 0 &&
          surf_key(surf_delayed(P, occ, P.ttoff[ts]), P.tile_id[ts]) < key) {
        lose_list[atomicAdd(lose_count, 1u)] = s;
        if (err) atomicOr(P.err, err);
        return;
      }
      continue;  // taken: draw again
    }
    P.prop[s] = ts;
    P.pu[s] = nu;
    P.pv[s] = nv;
    atomicMin(P.claim + ts, key);
    P.plist[atomicAdd(P.hdr + kHdrSurfProps, 1u)] = s;
    if (err) atomicOr(P.err, err);
    return;
  }
  // no position found: the molecule stays where it was for this step
  surf_step_over(P, s, tw, toff, tm, delayed);
  if (err) atomicOr(P.err, err);
}

// Claims settled: the first in the order moves to the tile it proposed (the tile
// it leaves was taken when the round's proposals were made, so no other
// mover enters it in this round) and ends its step there; the others are
// listed for the next round.
__global__ __launch_bounds__(256) void k_surf_commit(SurfPhase P, uint32_t* __restrict__ lose_list,
                                                     uint32_t* __restrict__ lose_count, uint32_t max_props) {
  const uint32_t n = P.hdr[kHdrSurfProps];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n && i < max_props; i += gridDim.x * blockDim.x) {
    const uint32_t s = P.plist[i];
    const uint32_t ts = P.prop[s];
    const uint32_t id = P.tile_id[s];
    const uint32_t tw = P.tiles[s];
    const double toff = P.ttoff[s];
    const bool delayed = surf_delayed(P, tw, toff);
    if (P.claim[ts] == surf_key(delayed, id)) {
      const uint32_t rep = s / P.total_tiles;
      const SpeciesRec& sp = P.species.rec[rep * P.species.stride + tile_species(tw)];
      // the step's timing again, from the slot it leaves
      const SurfTiming tm = surf_timing(P, sp, tw, toff, s);
      P.tiles[ts] = tw;
      P.tile_id[ts] = id;
      P.tu[ts] = P.pu[s];
      P.tv[ts] = P.pv[s];
      P.birth[ts] = P.birth[s];
      P.tile_consumed[ts] = 0;  // the molecule reacts with the volume movers of later iterations
      surf_step_over(P, ts, tw, toff, tm, delayed);
      P.tiles[s] = kTileEmpty;
      P.tile_id[s] = kInvalid;
    } else {
      lose_list[atomicAdd(lose_count, 1u)] = s;
    }
  }
}

__global__ __launch_bounds__(256) void k_surf_unclaim(SurfPhase P, uint32_t max_props) {
  const uint32_t n = P.hdr[kHdrSurfProps];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n && i < max_props; i += gridDim.x * blockDim.x)
    P.claim[P.prop[P.plist[i]]] = ~0ull;
}

__global__ void k_fill_ones(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = ~0u;
}

__global__ void k_zero_word(uint32_t* p) {
  if (threadIdx.x == 0) *p = 0;
}

}  // namespace

// no stamp of any iteration
hipError_t launch_fill_done(hipStream_t s, uint32_t* done, uint32_t n) {
  if (n) hipLaunchKernelGGL(k_fill_ones, blocks_for(n), dim3(256), 0, s, done, n);
  return hipGetLastError();
}

hipError_t launch_zero_word(hipStream_t s, uint32_t* word) {
  hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(64), 0, s, word);
  return hipGetLastError();
}

hipError_t launch_fill_claims(hipStream_t s, unsigned long long* claim, uint32_t n) {
  if (n) hipLaunchKernelGGL(k_fill_claims, blocks_for(n), dim3(256), 0, s, claim, n);
  return hipGetLastError();
}

hipError_t surf_round(hipStream_t s, const SurfPhase& P, const uint32_t* list, const uint32_t* count,
                      uint32_t max_movers, uint32_t* lose_list, uint32_t* lose_count) {
  if (max_movers == 0) return hipSuccess;
  const uint32_t n_all = P.total_tiles * P.num_replicas;
  hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(64), 0, s, P.hdr + kHdrSurfProps);
  hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(64), 0, s, lose_count);
  hipLaunchKernelGGL(k_surf_move, blocks_for(max_movers), dim3(256), 0, s, P, list, count, n_all, lose_list,
                     lose_count);
  const dim3 grid = dim3(std::min<uint32_t>(blocks_for(max_movers).x, 2048u));
  hipLaunchKernelGGL(k_surf_commit, grid, dim3(256), 0, s, P, lose_list, lose_count, max_movers);
  hipLaunchKernelGGL(k_surf_unclaim, grid, dim3(256), 0, s, P, max_movers);
  return hipGetLastError();
}

}  // namespace mcg
