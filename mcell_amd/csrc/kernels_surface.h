// Host-side interface of the surface-molecule diffusion phase
// (kernels_surface.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device/surface.h"

namespace mcg {

// The arrays of one surface phase.  Tile slots are replica-major,
// [replica * total_tiles + tile_base + tile].
struct SurfPhase {
  const WallRec64* __restrict__ walls;
  const WallGrid* __restrict__ grids;
  const WallEdgeRec* __restrict__ edges;     // 3 per wall
  const uint32_t* __restrict__ tile_wall;    // per tile of one replica
  SpeciesTable species;
  uint32_t* tiles;                           // tile words (kTileEmpty: free)
  uint32_t* tile_id;
  uint32_t* tile_consumed;
  double* tu;                                // molecule positions in the wall's uv frame
  double* tv;
  double* ttoff;                             // diffusion-time offsets from the iteration's start
  uint32_t* claim;                           // per slot: smallest id proposing to move there
  uint32_t* prop;                            // per mover slot: the slot it proposes
  double* pu;                                // the proposed position and offset
  double* pv;
  double* ptoff;
  uint32_t* plist;                           // proposing slots of the round
  uint32_t* done;                            // per slot: iteration + 1 once its molecule's step is over
  uint32_t* hdr;
  uint32_t* err;
  uint32_t total_tiles, num_replicas;
  uint64_t seed;
  uint32_t iteration;
  double barrier;                            // time up to the next barrier from the iteration's start
};

// One round of the phase over `list` (every slot when null; else the
// movers of the previous round that lost a tile), then the claims'
// settlement: winners move, losers are listed for the next round.
hipError_t surf_round(hipStream_t s, const SurfPhase& P, const uint32_t* list, const uint32_t* count,
                      uint32_t max_movers, uint32_t* lose_list, uint32_t* lose_count);
hipError_t launch_fill_done(hipStream_t s, uint32_t* done, uint32_t n);
hipError_t launch_fill_claims(hipStream_t s, uint32_t* claim, uint32_t n);

}  // namespace mcg
