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
#include "device/fastdraw.h"
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

// The light passes run persistent: a grid of a few workgroups per CU walks
// the molecules tile by tile (256 per tile), so the strip tables are staged
// once per workgroup.  A tile's queued molecules go to shard tile % 2048.
//
// Queues: light passes append molecules whose displacement left the
// Ziggurat's fast path to the middle queue (shard = tile % 2048) and
// molecules a wall may stop to the first half of the full-step queue (shard
// = tile % 1024); the middle pass appends the molecules it does not clear to
// the second half (shard = 1024 + its tile % 1024).  Every shard holds the
// slots of all the tiles that map to it, so neither can overflow.
constexpr uint32_t kHalfShards = kQueueShards / 2;
__device__ __forceinline__ void append_light(bool mid, bool hard, uint32_t i, uint32_t tile, const SlowQueue& MQ,
                                             const SlowQueue& HQ) {
  const uint32_t ms = tile % kQueueShards, hs = tile % kHalfShards;
  wave_append(mid, i, MQ.slots + (size_t)ms * MQ.shard_cap, MQ.count + ms);
  wave_append(hard, i, HQ.slots + (size_t)hs * HQ.shard_cap, HQ.count + hs);
}
__device__ __forceinline__ void append_mid(bool hard, uint32_t i, uint32_t base, const SlowQueue& HQ) {
  const uint32_t hs = kHalfShards + (base / 256u) % kHalfShards;
  wave_append(hard, i, HQ.slots + (size_t)hs * HQ.shard_cap, HQ.count + hs);
}

// The move p -> p + m (FP32) certainly meets no wall and stays inside the
// partition, with its end point clear of every tested wall plane by `band`.
__device__ __forceinline__ bool clear_move32(const PartitionParams32& Q, const Walls32View& W, const FilterParams& F,
                                             float band, float px, float py, float pz, float mx, float my,
                                             float mz) {
  const float ex = px + mx, ey = py + my, ez = pz + mz;
  int lo[3], hi[3];
  if (!index_box32(Q, px, py, pz, ex, ey, ez, lo, hi)) return false;
  return walls_clear(Q, W.off, W.box, px, py, pz, ex, ey, ez, lo, hi, [&](uint32_t k) {
    float de;
    return certain_miss32(W.ent32[k], F, px, py, pz, mx, my, mz, &de) && fabs_f(de) >= band;
  });
}

// The FP64 move certainly meets no wall (the reference's collide_wall of
// every wall of the index box returns MISS: bit-exact) and stays inside the
// partition.  The index and entry boxes are FP32 and conservative.
__device__ __forceinline__ bool clear_move64(const PartitionParams32& Q, const WallView64& W, const D3& pos,
                                             const D3& disp, const D3& end) {
  const float px = (float)pos.x, py = (float)pos.y, pz = (float)pos.z;
  const float ex = (float)end.x, ey = (float)end.y, ez = (float)end.z;
  int lo[3], hi[3];
  if (!index_box32(Q, px, py, pz, ex, ey, ez, lo, hi)) return false;
  return walls_clear(Q, W.off, W.box, px, py, pz, ex, ey, ez, lo, hi,
                     [&](uint32_t k) { return wall_certain_miss64(W.ent[k], pos, disp); });
}

// ------------------------------------------------------------------- FP64 --
// Every molecule whose three Gaussians are on the Ziggurat's fast path and
// whose straight move certainly meets no wall (FP64 collide_wall of every
// wall in the move's index box, bit-exact: wall_certain_miss64) commits here;
// the others go to k_hard64, which draws the same stream through rng_gauss.
__global__ __launch_bounds__(256) void k_light64(MolView64 m, uint32_t n, SpeciesTable species,
                                                 PartitionParams32 Q, WallView64 W, uint64_t seed,
                                                 uint32_t iteration, SlowQueue MQ, SlowQueue HQ, HitTrace trace) {
  __shared__ FastZiggurat T;
  stage_fast_ziggurat(T, c_zig);
  for (uint32_t tile = blockIdx.x; tile * 256u < n; tile += gridDim.x) {
    const uint32_t i = tile * 256u + threadIdx.x;
    bool mid = false, hard = false;
    if (i < n) {
      const uint32_t tag = m.tag[i];
      const uint32_t idv = m.id[i];
      const D3 pos = {m.x[i], m.y[i], m.z[i]};
      const SpeciesRec sp = species.of(tag);
      if (!(tag_flags(tag) & MF_DEFUNCT) && sp.space_step > 0) {
        uint32_t r0, r1, r2, r3;
        philox_block0(seed, idv, iteration, tag_replica(tag), 0, r0, r1, r2, r3);
        D3 disp;
        mid = !fast_displacement(r0, r1, r2, T, sp.space_step, disp.x, disp.y, disp.z);
        bool clear = !mid;
        const D3 end = {pos.x + disp.x, pos.y + disp.y, pos.z + disp.z};
        clear = clear && clear_move64(Q, W, pos, disp, end);
        if (clear) {
          m.x[i] = end.x;
          m.y[i] = end.y;
          m.z[i] = end.z;
          if (trace.enabled) trace.count[i] = 0;
        }
        hard = !clear && !mid;
      }
    }
    append_light(mid, hard, i, tile, MQ, HQ);
  }
}

__global__ __launch_bounds__(256) void k_mid64(MolView64 m, SpeciesTable species, PartitionParams32 Q,
                                               WallView64 W, uint64_t seed, uint32_t iteration, SlowQueue MQ,
                                               SlowQueue HQ, HitTrace trace) {
  __shared__ ZigguratTables T;
  __shared__ uint32_t pre[kQueueShards + 1];
  __shared__ uint32_t part[256];
  const uint32_t total = queue_prefix(MQ.count, MQ.nshards, pre, part);
  if ((uint64_t)blockIdx.x * blockDim.x >= total) return;
  stage_ziggurat(T);
  for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
    const uint32_t q = base + threadIdx.x;
    uint32_t i = 0;
    bool hard = false;
    if (q < total) {
      i = queue_at(MQ.slots, MQ.shard_cap, pre, q);
      const uint32_t tag = m.tag[i];
      const SpeciesRec sp = species.of(tag);
      PhiloxStream rng;
      rng.init(seed, m.id[i], iteration, tag_replica(tag), 0);
      const D3 disp = draw_displacement(rng, T, sp.space_step);
      const D3 pos = {m.x[i], m.y[i], m.z[i]};
      const D3 end = {pos.x + disp.x, pos.y + disp.y, pos.z + disp.z};
      hard = !clear_move64(Q, W, pos, disp, end);
      if (!hard) {
        m.x[i] = end.x;
        m.y[i] = end.y;
        m.z[i] = end.z;
        if (trace.enabled) trace.count[i] = 0;
      }
    }
    append_mid(hard, i, base, HQ);
  }
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
// Everything in FP32: the displacement (within kMoveUlps of the FP64 one),
// the end point, the index and entry boxes, and the wall filter (filter32.h,
// widened by the displacement's error), which certifies misses of the FP64
// move as well.  An end point within `band` of a tested wall's plane takes
// the full step, which rounds the FP64 end point onto its side.
__global__ __launch_bounds__(256) void k_light32(MolView32 m, uint32_t n, SpeciesTable species,
                                                 PartitionParams32 Q, Walls32View W, FilterParams F, float band,
                                                 uint64_t seed, uint32_t iteration, SlowQueue MQ, SlowQueue HQ,
                                                 HitTrace trace) {
  __shared__ FastZiggurat T;
  stage_fast_ziggurat(T, c_zig);
  for (uint32_t tile = blockIdx.x; tile * 256u < n; tile += gridDim.x) {
    const uint32_t i = tile * 256u + threadIdx.x;
    bool mid = false, hard = false;
    if (i < n) {
      const uint32_t tag = m.tag[i];
      const uint32_t idv = m.id[i];
      const float px = m.x[i], py = m.y[i], pz = m.z[i];
      const SpeciesRec sp = species.of(tag);
      if (!(tag_flags(tag) & MF_DEFUNCT) && sp.space_step > 0) {
        uint32_t r0, r1, r2, r3;
        philox_block0(seed, idv, iteration, tag_replica(tag), 0, r0, r1, r2, r3);
        float mx, my, mz;
        mid = !fast_displacement32(r0, r1, r2, T, (float)(sp.space_step * 0.70710678118654752440), mx, my, mz);
        const bool clear = !mid && clear_move32(Q, W, F, band, px, py, pz, mx, my, mz);
        if (clear) {
          m.x[i] = px + mx;
          m.y[i] = py + my;
          m.z[i] = pz + mz;
          if (trace.enabled) trace.count[i] = 0;
        }
        hard = !clear && !mid;
      }
    }
    append_light(mid, hard, i, tile, MQ, HQ);
  }
}

// Middle pass: the molecules whose displacement left the fast path.  The
// displacement is drawn through rng_gauss from the same stream (FP64, the
// reference's), then the light checks decide as in the light passes; the
// molecules they do not clear go on to the full step.
__global__ __launch_bounds__(256) void k_mid32(MolView32 m, SpeciesTable species, PartitionParams32 Q,
                                               Walls32View W, FilterParams F, float band, uint64_t seed,
                                               uint32_t iteration, SlowQueue MQ, SlowQueue HQ, HitTrace trace) {
  __shared__ ZigguratTables T;
  __shared__ uint32_t pre[kQueueShards + 1];
  __shared__ uint32_t part[256];
  const uint32_t total = queue_prefix(MQ.count, MQ.nshards, pre, part);
  if ((uint64_t)blockIdx.x * blockDim.x >= total) return;
  stage_ziggurat(T);
  for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
    const uint32_t q = base + threadIdx.x;
    uint32_t i = 0;
    bool hard = false;
    if (q < total) {
      i = queue_at(MQ.slots, MQ.shard_cap, pre, q);
      const uint32_t tag = m.tag[i];
      const SpeciesRec sp = species.of(tag);
      PhiloxStream rng;
      rng.init(seed, m.id[i], iteration, tag_replica(tag), 0);
      const D3 disp = draw_displacement(rng, T, sp.space_step);
      const float px = m.x[i], py = m.y[i], pz = m.z[i];
      const float mx = (float)disp.x, my = (float)disp.y, mz = (float)disp.z;
      hard = !clear_move32(Q, W, F, band, px, py, pz, mx, my, mz);
      if (!hard) {
        m.x[i] = px + mx;
        m.y[i] = py + my;
        m.z[i] = pz + mz;
        if (trace.enabled) trace.count[i] = 0;
      }
    }
    append_mid(hard, i, base, HQ);
  }
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

// Persistent grid of a 256-thread kernel: as many workgroups as can be
// resident at once (CUs x the kernel's occupancy), never more than the tiles
// -- a workgroup beyond what fits would only start after others finish and
// then walk its share of the tiles alone.
template <typename Kernel>
static dim3 persistent_grid(Kernel kernel, uint32_t n) {
  static int resident = 0;
  if (resident == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess || per_cu <= 0)
      per_cu = 1;
    resident = cus * per_cu;
  }
  const uint32_t tiles = (n + 255) / 256;
  return dim3(std::max<uint32_t>(1u, std::min<uint32_t>(tiles, (uint32_t)resident)));
}

uint32_t shard_capacity(uint32_t n, uint32_t shards) {
  uint32_t blocks = (n + 255) / 256;
  return 256u * ((blocks + shards - 1) / shards);
}


hipError_t launch_diffuse_reflect64(hipStream_t s, const MolView64& m, uint32_t n, SpeciesTable species,
                                    const SurfaceClasses& SC,
                                    const PartitionParams& P, const WallView64& W, uint64_t seed, uint32_t iteration,
                                    SlowQueue MQ, SlowQueue Q, uint32_t* err, const HitTrace& trace,
                                    unsigned long long* ctr) {
  if (n == 0) return hipSuccess;
  const PartitionParams32 P32 = make_partition32(P);
  MQ.shard_cap = shard_capacity(n);
  MQ.nshards = kQueueShards;
  Q.shard_cap = shard_capacity(n, kHalfShards);
  Q.nshards = kQueueShards;
  hipLaunchKernelGGL(k_reset_counters, dim3(kQueueShards / 256), dim3(256), 0, s, MQ.count, kQueueShards);
  hipLaunchKernelGGL(k_reset_counters, dim3(kQueueShards / 256), dim3(256), 0, s, Q.count, kQueueShards);
  hipLaunchKernelGGL(k_light64, persistent_grid(k_light64, n), dim3(256), 0, s, m, n, species, P32, W, seed,
                     iteration, MQ, Q, trace);
  hipLaunchKernelGGL(k_mid64, persistent_grid(k_mid64, n), dim3(256), 0, s, m, species, P32, W, seed, iteration, MQ,
                     Q, trace);
  hipLaunchKernelGGL(k_hard64, dim3(std::min<uint32_t>(kQueueShards, (n + 255) / 256)), dim3(256), 0, s, m, species,
                     SC, P, W, seed, iteration, Q, err, trace, ctr);
  return hipGetLastError();
}

hipError_t launch_diffuse_reflect32(hipStream_t s, const MolView32& m, uint32_t n, SpeciesTable species,
                                    const SurfaceClasses& SC,
                                    const PartitionParams& P, const WallView64& W, const WallRec32* w32,
                                    float max_abs_coord, uint64_t seed, uint32_t iteration, SlowQueue MQ,
                                    SlowQueue Q, uint32_t* err, const HitTrace& trace, unsigned long long* ctr) {
  if (n == 0) return hipSuccess;
  FilterParams F = make_filter_params(max_abs_coord, kMoveUlps);
  float band = 2.0f * F.e_dd + 8.0f * F.eps * max_abs_coord;
  Walls32View W32{W.walls, reinterpret_cast<const FilterWall*>(w32), W.off, W.idx,
                  reinterpret_cast<const FilterWall*>(W.ent32), W.box};
  const PartitionParams32 P32 = make_partition32(P);
  MQ.shard_cap = shard_capacity(n);
  MQ.nshards = kQueueShards;
  Q.shard_cap = shard_capacity(n, kHalfShards);
  Q.nshards = kQueueShards;
  hipLaunchKernelGGL(k_reset_counters, dim3(kQueueShards / 256), dim3(256), 0, s, MQ.count, kQueueShards);
  hipLaunchKernelGGL(k_reset_counters, dim3(kQueueShards / 256), dim3(256), 0, s, Q.count, kQueueShards);
  hipLaunchKernelGGL(k_light32, persistent_grid(k_light32, n), dim3(256), 0, s, m, n, species, P32, W32, F, band,
                     seed, iteration, MQ, Q, trace);
  hipLaunchKernelGGL(k_mid32, persistent_grid(k_mid32, n), dim3(256), 0, s, m, species, P32, W32, F, band, seed,
                     iteration, MQ, Q, trace);
  hipLaunchKernelGGL(k_hard32, dim3(std::min<uint32_t>(kQueueShards, (n + 255) / 256)), dim3(256), 0, s, m, species,
                     SC, P, W32, seed, iteration, Q, err, trace, ctr);
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
