// HIP kernels of the diffuse-react engine (gfx950).
//
// One iteration of non-reacting volume molecules is two launches:
//   k_light*  every molecule: draw the displacement; if the straight move
//             certainly touches no wall, commit it; otherwise append the
//             molecule to the hard queue (one atomic per wave).
//   k_hard*   queued molecules: the full ray trace with reflections, which
//             re-draws the same displacement from the molecule's stream.
// The light kernel holds few registers (high occupancy for the HBM stream);
// the register-heavy trace runs only for the few molecules near walls.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device/displace.h"
#include "device/light.h"
#include "device/path32.h"
#include "device/path64.h"
#include "device/rng.h"
#include "device/types.h"
#include "kernels.h"

namespace mcg {

static __constant__ ZigguratTables c_zig = MCG_ZIGGURAT_TABLES_INIT;

// Stages the Ziggurat tables (2.5 KB) into LDS so that the per-lane strip
// lookups of a wave are LDS gathers rather than scattered cache reads.
__device__ __forceinline__ void stage_ziggurat(ZigguratTables& T) {
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&c_zig);
  uint32_t* dst = reinterpret_cast<uint32_t*>(&T);
  constexpr int words = sizeof(ZigguratTables) / 4;
  for (int k = threadIdx.x; k < words; k += blockDim.x) dst[k] = src[k];
  __syncthreads();
}

__global__ void k_reset_counters(uint32_t* c, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = 0;
}

// ------------------------------------------------------------------- FP64 --
__global__ __launch_bounds__(256) void k_light64(MolView64 m, uint32_t n, SpeciesTable species,
                                                 PartitionParams P, WallView64 W, uint64_t seed, uint32_t iteration,
                                                 SlowQueue Q, HitTrace trace) {
  __shared__ ZigguratTables T;
  stage_ziggurat(T);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool hard = false;
  if (i < n) {
    const uint32_t tag = m.tag[i];
    const SpeciesRec sp = species.of(tag);
    if (!(tag_flags(tag) & MF_DEFUNCT) && sp.space_step > 0) {
      PhiloxStream rng;
      rng.init(seed, m.id[i], iteration, tag_replica(tag), 0);
      const D3 pos = {m.x[i], m.y[i], m.z[i]};
      const D3 disp = draw_displacement(rng, T, sp.space_step);
      const D3 end = {pos.x + disp.x, pos.y + disp.y, pos.z + disp.z};
      int lo[3], hi[3];
      if (move_index_box(P, pos, end, lo, hi) && all_walls_miss64(P, W.ent, W.off, W.box, pos, disp, lo, hi)) {
        m.x[i] = end.x;
        m.y[i] = end.y;
        m.z[i] = end.z;
        if (trace.enabled) trace.count[i] = 0;
      } else {
        hard = true;
      }
    }
  }
  const uint32_t shard = blockIdx.x % kQueueShards;
  wave_append(hard, i, Q.slots + (size_t)shard * Q.shard_cap, Q.count + shard);
}

__global__ __launch_bounds__(256) void k_hard64(MolView64 m, SpeciesTable species, SurfaceClasses SC,
                                                PartitionParams P, WallView64 W, uint64_t seed, uint32_t iteration,
                                                SlowQueue Q, uint32_t* __restrict__ err_out, HitTrace trace,
                                                unsigned long long* __restrict__ ctr) {
  __shared__ ZigguratTables T;
  __shared__ uint32_t pre[kQueueShards + 1];
  __shared__ uint32_t part[256];
  // the shards read as one list, grid-stride
  const uint32_t total = queue_prefix(Q.count, Q.nshards, pre, part);
  if ((uint64_t)blockIdx.x * blockDim.x >= total) return;
  stage_ziggurat(T);
  uint32_t n_reflect = 0, n_cross = 0;
  // every thread of the block runs the same number of rounds (the counters
  // are wave-summed after the loop)
  for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
    const uint32_t q = base + threadIdx.x;
    if (q >= total) continue;
    const uint32_t i = queue_at(Q.slots, Q.shard_cap, pre, q);
    const uint32_t tag = m.tag[i];
    const SpeciesRec sp = species.of(tag);
    PhiloxStream rng;
    rng.init(seed, m.id[i], iteration, tag_replica(tag), 0);
    D3 pos = {m.x[i], m.y[i], m.z[i]};
    const D3 disp = draw_displacement(rng, T, sp.space_step);
    uint32_t err = 0, tn = 0;
    uint32_t* tw = trace.enabled ? trace.wall + (size_t)i * kTraceHits : nullptr;
    const bool alive = diffuse_reflect64(P, W.walls, W.off, W.idx, W.box, SC, tag_species(tag), pos, disp, rng, err,
                                         tw, tn, n_reflect, n_cross);
    m.x[i] = pos.x;
    m.y[i] = pos.y;
    m.z[i] = pos.z;
    if (!alive) m.tag[i] = tag | (MF_DEFUNCT << 8);
    if (trace.enabled) trace.count[i] = tn;
    if (err) atomicOr(err_out, err);
  }
  wave_count(ctr + CTR_REFLECTIONS, n_reflect);
  wave_count(ctr + CTR_CROSSINGS, n_cross);
}

// ------------------------------------------------------------------- FP32 --
__global__ __launch_bounds__(256) void k_light32(MolView32 m, uint32_t n, SpeciesTable species,
                                                 PartitionParams P, Walls32View W, FilterParams F, float band,
                                                 uint64_t seed, uint32_t iteration, SlowQueue Q, HitTrace trace,
                                                 uint32_t* __restrict__ err_out) {
  __shared__ ZigguratTables T;
  stage_ziggurat(T);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool hard = false;
  if (i < n) {
    const uint32_t tag = m.tag[i];
    const SpeciesRec sp = species.of(tag);
    if (!(tag_flags(tag) & MF_DEFUNCT) && sp.space_step > 0) {
      PhiloxStream rng;
      rng.init(seed, m.id[i], iteration, tag_replica(tag), 0);
      const D3 pos = {(double)m.x[i], (double)m.y[i], (double)m.z[i]};
      const D3 disp = draw_displacement(rng, T, sp.space_step);
      const D3 end = {pos.x + disp.x, pos.y + disp.y, pos.z + disp.z};
      int lo[3], hi[3];
      bool near_plane = false;
      if (move_index_box(P, pos, end, lo, hi) &&
          all_walls_miss32(P, W.ent32, W.off, W.box, F, band, pos, disp, lo, hi, near_plane)) {
        // an end point next to a wall plane keeps its side in FP32, as the
        // full step's would
        float fx, fy, fz;
        if (near_plane) {
          if (!round_to_fp32_same_side(P, W.w64, W.off, W.idx, W.box, end.x, end.y, end.z, &fx, &fy, &fz))
            atomicOr(err_out, ERR_ROUND32);
        } else {
          fx = (float)end.x;
          fy = (float)end.y;
          fz = (float)end.z;
        }
        m.x[i] = fx;
        m.y[i] = fy;
        m.z[i] = fz;
        if (trace.enabled) trace.count[i] = 0;
      } else {
        hard = true;
      }
    }
  }
  const uint32_t shard = blockIdx.x % kQueueShards;
  wave_append(hard, i, Q.slots + (size_t)shard * Q.shard_cap, Q.count + shard);
}

__global__ __launch_bounds__(256) void k_hard32(MolView32 m, SpeciesTable species, SurfaceClasses SC,
                                                PartitionParams P, Walls32View W,
                                                uint64_t seed, uint32_t iteration, SlowQueue Q,
                                                uint32_t* __restrict__ err_out, HitTrace trace,
                                                unsigned long long* __restrict__ ctr) {
  __shared__ ZigguratTables T;
  __shared__ uint32_t pre[kQueueShards + 1];
  __shared__ uint32_t part[256];
  // the shards read as one list, grid-stride
  const uint32_t total = queue_prefix(Q.count, Q.nshards, pre, part);
  if ((uint64_t)blockIdx.x * blockDim.x >= total) return;
  stage_ziggurat(T);
  uint32_t n_reflect = 0, n_cross = 0;
  for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
    const uint32_t q = base + threadIdx.x;
    if (q >= total) continue;
    const uint32_t i = queue_at(Q.slots, Q.shard_cap, pre, q);
    const uint32_t tag = m.tag[i];
    const SpeciesRec sp = species.of(tag);
    PhiloxStream rng;
    rng.init(seed, m.id[i], iteration, tag_replica(tag), 0);
    D3 pos = {(double)m.x[i], (double)m.y[i], (double)m.z[i]};
    const D3 disp = draw_displacement(rng, T, sp.space_step);
    uint32_t err = 0, tn = 0;
    uint32_t* tw = trace.enabled ? trace.wall + (size_t)i * kTraceHits : nullptr;
    // the FP64 wall search (a queued molecule meets few walls after the
    // entry-box pre-test; the FP32 filter in front of them cost more than
    // it saved), then the end point rounded to FP32 on its side of nearby
    // wall planes
    const bool alive = diffuse_reflect64(P, W.w64, W.off, W.idx, W.box, SC, tag_species(tag), pos, disp, rng, err,
                                         tw, tn, n_reflect, n_cross);
    if (!alive) m.tag[i] = tag | (MF_DEFUNCT << 8);
    float fx, fy, fz;
    if (!round_to_fp32_same_side(P, W.w64, W.off, W.idx, W.box, pos.x, pos.y, pos.z, &fx, &fy, &fz))
      err |= ERR_ROUND32;
    m.x[i] = fx;
    m.y[i] = fy;
    m.z[i] = fz;
    if (trace.enabled) trace.count[i] = tn;
    if (err) atomicOr(err_out, err);
  }
  wave_count(ctr + CTR_REFLECTIONS, n_reflect);
  wave_count(ctr + CTR_CROSSINGS, n_cross);
}

// Copy kernel of known size (device-to-device bandwidth reference).
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

// -------------------------------------------------------------- launchers --
static_assert(kNumQueueShards == kQueueShards, "host and device shard counts differ");
static_assert(kQueueShards % 256 == 0, "counter reset grid");

uint32_t shard_capacity(uint32_t n, uint32_t shards) {
  uint32_t blocks = (n + 255) / 256;
  return 256u * ((blocks + shards - 1) / shards);
}


hipError_t launch_diffuse_reflect64(hipStream_t s, const MolView64& m, uint32_t n, SpeciesTable species,
                                    const SurfaceClasses& SC,
                                    const PartitionParams& P, const WallView64& W, uint64_t seed, uint32_t iteration,
                                    SlowQueue Q, uint32_t* err, const HitTrace& trace, unsigned long long* ctr) {
  if (n == 0) return hipSuccess;
  Q.shard_cap = shard_capacity(n);
  hipLaunchKernelGGL(k_reset_counters, dim3(kQueueShards / 256), dim3(256), 0, s, Q.count, kQueueShards);
  hipLaunchKernelGGL(k_light64, dim3((n + 255) / 256), dim3(256), 0, s, m, n, species, P, W, seed, iteration, Q,
                     trace);
  Q.nshards = std::min<uint32_t>(kQueueShards, (n + 255) / 256);
  hipLaunchKernelGGL(k_hard64, dim3(Q.nshards), dim3(256), 0, s, m, species, SC, P, W, seed, iteration, Q, err, trace,
                     ctr);
  return hipGetLastError();
}

hipError_t launch_diffuse_reflect32(hipStream_t s, const MolView32& m, uint32_t n, SpeciesTable species,
                                    const SurfaceClasses& SC,
                                    const PartitionParams& P, const WallView64& W, const WallRec32* w32,
                                    float max_abs_coord, uint64_t seed, uint32_t iteration, SlowQueue Q,
                                    uint32_t* err, const HitTrace& trace, unsigned long long* ctr) {
  if (n == 0) return hipSuccess;
  FilterParams F = make_filter_params(max_abs_coord);
  float band = 2.0f * F.e_dd + 8.0f * F.eps * max_abs_coord;
  Walls32View W32{W.walls, reinterpret_cast<const FilterWall*>(w32), W.off, W.idx,
                  reinterpret_cast<const FilterWall*>(W.ent32), W.box};
  Q.shard_cap = shard_capacity(n);
  hipLaunchKernelGGL(k_reset_counters, dim3(kQueueShards / 256), dim3(256), 0, s, Q.count, kQueueShards);
  hipLaunchKernelGGL(k_light32, dim3((n + 255) / 256), dim3(256), 0, s, m, n, species, P, W32, F, band, seed,
                     iteration, Q, trace, err);
  Q.nshards = std::min<uint32_t>(kQueueShards, (n + 255) / 256);
  hipLaunchKernelGGL(k_hard32, dim3(Q.nshards), dim3(256), 0, s, m, species, SC, P, W32, seed, iteration, Q, err,
                     trace, ctr);
  return hipGetLastError();
}

hipError_t launch_copy(hipStream_t s, const void* src, void* dst, size_t bytes) {
  size_t n = bytes / sizeof(float4);
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, s, static_cast<const float4*>(src), static_cast<float4*>(dst),
                     n);
  return hipGetLastError();
}

}  // namespace mcg
