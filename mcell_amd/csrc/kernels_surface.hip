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
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device/schedule.h"
#include "device/surface.h"
#include "kernels_react.h"
#include "kernels_surface.h"

namespace mcg {

namespace {

inline dim3 blocks_for(uint32_t n) { return dim3((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

__global__ void k_fill_claims(uint32_t* claim, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) claim[i] = kInvalid;
}

// One surface molecule's step, for the slots of `list` (every slot when
// list is null: the round that also books molecules due in later
// iterations).
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
  const uint32_t stamp = P.iteration + 1;
  if (!(sp.flags & SPF_SURFACE) || !(sp.space_step > 0)) {
    P.done[s] = stamp;
    return;
  }
  const double toff = P.ttoff[s];
  if (!list && !ready_this_iteration(P.iteration, toff)) {
    P.ttoff[s] = toff_past_iteration(P.iteration, toff);
    P.done[s] = stamp;
    return;
  }
  // the step's time (diffuse_surf_molecule, :1081-1116); no unimolecular
  // clipping, so max_time is the time up to the barrier
  const double max_time = P.barrier - toff;
  constexpr double EPS = 1e-12;
  double t_steps = sp.time_step > max_time ? max_time : sp.time_step;
  double steps = 1.0;
  if (sp.time_step > max_time) {
    steps = max_time / sp.time_step;
    if (steps < EPS) {
      t_steps = EPS * sp.time_step;
      steps = EPS;
    }
  }
  const double space_factor = steps == 1.0 ? sp.space_step : sp.space_step * sqrt(steps);
  const double new_toff = next_toff(P.iteration, toff, t_steps);
  const uint32_t id = P.tile_id[s];
  PhiloxStream rng;
  rng.init(P.seed, id, P.iteration, rep, 0);
  const uint32_t wi = P.tile_wall[lt];
  const uint32_t t0 = lt - P.grids[wi].tile_base;
  const double u = P.tu[s], v = P.tv[s];
  uint32_t err = 0;
  for (int attempt = 0; attempt <= kSurfRetries; ++attempt) {
    double du, dv;
    surf_displacement(rng, space_factor, du, dv);
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
      P.ttoff[s] = new_toff;
      P.done[s] = stamp;
      if (err) atomicOr(P.err, err);
      return;
    }
    const uint32_t ts = rep * P.total_tiles + g.tile_base + nt;
    const uint32_t occ = P.tiles[ts];
    if (occ != kTileEmpty) {
      // an occupant with a smaller id that may still move: wait for it
      const SpeciesRec& os = P.species.rec[rep * P.species.stride + tile_species(occ)];
      if (P.tile_id[ts] < id && P.done[ts] != stamp && (os.flags & SPF_SURFACE) && os.space_step > 0) {
        lose_list[atomicAdd(lose_count, 1u)] = s;
        if (err) atomicOr(P.err, err);
        return;
      }
      continue;  // taken: draw again
    }
    P.prop[s] = ts;
    P.pu[s] = nu;
    P.pv[s] = nv;
    P.ptoff[s] = new_toff;
    atomicMin(P.claim + ts, id);
    P.plist[atomicAdd(P.hdr + kHdrSurfProps, 1u)] = s;
    if (err) atomicOr(P.err, err);
    return;
  }
  // no position found: the molecule stays where it was for this step
  P.ttoff[s] = new_toff;
  P.done[s] = stamp;
  if (err) atomicOr(P.err, err);
}

// Claims settled: the smallest id moves to the tile it proposed (the tile
// it leaves was taken when the round's proposals were made, so no other
// mover enters it in this round); the others are listed for the next round.
__global__ __launch_bounds__(256) void k_surf_commit(SurfPhase P, uint32_t* __restrict__ lose_list,
                                                     uint32_t* __restrict__ lose_count, uint32_t max_props) {
  const uint32_t n = P.hdr[kHdrSurfProps];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n && i < max_props; i += gridDim.x * blockDim.x) {
    const uint32_t s = P.plist[i];
    const uint32_t ts = P.prop[s];
    const uint32_t id = P.tile_id[s];
    if (P.claim[ts] == id) {
      P.tiles[ts] = P.tiles[s];
      P.tile_id[ts] = id;
      P.tu[ts] = P.pu[s];
      P.tv[ts] = P.pv[s];
      P.ttoff[ts] = P.ptoff[s];
      P.tile_consumed[ts] = 0;  // the molecule reacts with the volume movers of later iterations
      P.done[ts] = P.iteration + 1;
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
    P.claim[P.prop[P.plist[i]]] = kInvalid;
}

__global__ void k_zero_fill(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_zero_word(uint32_t* p) {
  if (threadIdx.x == 0) *p = 0;
}

}  // namespace

hipError_t launch_fill_done(hipStream_t s, uint32_t* done, uint32_t n) {
  if (n) hipLaunchKernelGGL(k_zero_fill, blocks_for(n), dim3(256), 0, s, done, n);
  return hipGetLastError();
}

hipError_t launch_fill_claims(hipStream_t s, uint32_t* claim, uint32_t n) {
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
