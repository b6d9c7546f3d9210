// Launch wrappers of the diffuse-react kernels (host-visible declarations).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "device/types.h"

namespace mcg {

// Structure-of-arrays molecule state in HBM (FP64 positions).
struct MolView64 {
  double* x;
  double* y;
  double* z;
  uint32_t* id;
  uint32_t* tag;
};

// Structure-of-arrays molecule state in HBM (FP32 positions).
struct MolView32 {
  float* x;
  float* y;
  float* z;
  uint32_t* id;
  uint32_t* tag;
};

struct WallView64 {
  const WallRec64* walls;
  const uint32_t* off;    // CSR offsets, n^3 + 1
  const uint32_t* idx;    // wall indices, ascending within a subpartition
  const WallRec64* ent;   // walls[idx[k]] for every list entry k
  const WallRec32* ent32; // its FP32 form
  const uint32_t* box;    // sub-cell box of every list entry
};

// Slow-path queue: kQueueShards counters and shards of shard_cap slots.
// Free diffusion runs three passes: light, middle (displacements off the
// Ziggurat's fast path) and full step, with one queue in front of each of
// the last two.
struct SlowQueue {
  uint32_t* slots;
  uint32_t* count;
  uint32_t shard_cap;
  uint32_t nshards;  // shards the light pass can have filled: min(kQueueShards, its workgroups)
};
constexpr uint32_t kNumQueueShards = 2048;
// Slots per shard when n movers' workgroups append round-robin to
// `shards` shards (all of them, or one half per kind of mover).
uint32_t shard_capacity(uint32_t n, uint32_t shards = kNumQueueShards);

hipError_t launch_diffuse_reflect64(hipStream_t s, const MolView64& m, uint32_t n, SpeciesTable species,
                                    const SurfaceClasses& SC,
                                    const PartitionParams& P, const WallView64& W, uint64_t seed, uint32_t iteration,
                                    SlowQueue MQ, SlowQueue Q, uint32_t* err, const HitTrace& trace,
                                    unsigned long long* ctr);

hipError_t launch_diffuse_reflect32(hipStream_t s, const MolView32& m, uint32_t n, SpeciesTable species,
                                    const SurfaceClasses& SC,
                                    const PartitionParams& P, const WallView64& W, const WallRec32* w32,
                                    float max_abs_coord, uint64_t seed, uint32_t iteration, SlowQueue MQ,
                                    SlowQueue Q, uint32_t* err, const HitTrace& trace, unsigned long long* ctr);

hipError_t launch_copy(hipStream_t s, const void* src, void* dst, size_t bytes);

}  // namespace mcg
