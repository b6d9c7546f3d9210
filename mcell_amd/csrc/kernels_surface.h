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
  double* birth;                             // birthdays (absolute)
  unsigned long long* claim;                 // per slot: smallest order key proposing to move there
  uint32_t* prop;                            // per mover slot: the slot it proposes
  double* pu;                                // the proposed position
  double* pv;
  uint32_t* plist;                           // proposing slots of the round
  // per slot: the iteration's stamp once its molecule's step of a pass is
  // over, stamp_base | kDoneFinal when it has no further step in the
  // iteration, stamp_base | (pass + 1) when it steps again in the next pass
  uint32_t* done;
  uint32_t* clist;                           // slots of the molecules that step again (next pass)
  uint32_t* ccount;
  uint32_t* hdr;
  uint32_t* err;
  uint32_t total_tiles, num_replicas;
  uint64_t seed;
  uint32_t iteration;
  uint32_t pass;                             // the pass (a molecule's n-th step in the iteration)
  uint32_t stamp_base;                       // ((iteration + 1) mod 2^28) << 4
  double barrier;                            // time up to the next barrier from the iteration's start
};

constexpr uint32_t kDoneFinal = 15u;
constexpr uint32_t kSurfPasses = 8;          // steps per iteration (time steps of at least 1/8)

// One round of a pass over `list` (every slot when null: the first pass's
// first round; else the molecules stepping again in this pass, or the
// movers of the previous round that lost a tile or waited), then the
// claims' settlement: winners move, losers are listed for the next round.
hipError_t surf_round(hipStream_t s, const SurfPhase& P, const uint32_t* list, const uint32_t* count,
                      uint32_t max_movers, uint32_t* lose_list, uint32_t* lose_count);
hipError_t launch_fill_done(hipStream_t s, uint32_t* done, uint32_t n);
hipError_t launch_zero_word(hipStream_t s, uint32_t* word);
hipError_t launch_fill_claims(hipStream_t s, unsigned long long* claim, uint32_t n);

}  // namespace mcg
