// The common case of a diffusion step, kept register-light: draw the
// displacement, and if the straight move cannot touch any wall, commit it.
//
// A move touches only subpartitions inside the index box spanned by its start
// and end subpartitions (the segment lies inside the bounding box of its end
// points), and every wall is listed in every subpartition it meets, so
// testing the walls of that index box (at most 2x2x2 boxes for moves shorter
// than a subpartition) decides whether any wall can be hit.  If none can, the
// reference's ray trace would find no collision and commit pos + disp
// (diffuse_react_event.cpp:774-777); otherwise the molecule is handed to the
// full ray-trace kernel, which re-draws the same displacement from the same
// counter-based stream.
#pragma once
#include "filter32.h"
#include "path64.h"

namespace mcg {

// The slow-path queue is split into kQueueShards shards; workgroup b appends
// to shard b % kQueueShards, so the per-wave atomics of different workgroups
// hit different counters (one shared counter serialises them in L2).  A
// shard holds the slots of all workgroups that map to it, so it cannot
// overflow: shard_cap = 256 * ceil(num_workgroups / kQueueShards).
constexpr uint32_t kQueueShards = 2048;

// The drains read the shards as one list: every workgroup prefix-sums the
// shard counts in LDS (pre[0..kQueueShards], part[256] scratch; all 256
// threads call this) and deals list positions out grid-stride, so that
// movers crowded into a few shards (hard movers cluster near surfaces and the
// light pass is cell-ordered) still spread over the whole GPU.  Returns the
// number of queued movers.
__device__ __forceinline__ uint32_t queue_prefix(const uint32_t* __restrict__ count, uint32_t nshards, uint32_t* pre,
                                                uint32_t* part) {
  static_assert(kQueueShards % 256 == 0, "shards per thread");
  constexpr uint32_t kPer = kQueueShards / 256;
  uint32_t local[kPer];
  uint32_t sum = 0;
  for (uint32_t k = 0; k < kPer; ++k) {
    const uint32_t sh = threadIdx.x * kPer + k;
    local[k] = sh < nshards ? count[sh] : 0u;
    sum += local[k];
  }
  part[threadIdx.x] = sum;
  __syncthreads();
  for (uint32_t off = 1; off < 256; off <<= 1) {
    const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = part[threadIdx.x] - sum;
  for (uint32_t k = 0; k < kPer; ++k) {
    pre[threadIdx.x * kPer + k] = run;
    run += local[k];
  }
  if (threadIdx.x == 255) pre[kQueueShards] = part[255];
  __syncthreads();
  return pre[kQueueShards];
}

// The queued value at list position q < total: the last shard k with
// pre[k] <= q holds it at offset q - pre[k].
__device__ __forceinline__ uint32_t queue_at(const uint32_t* __restrict__ slots, uint32_t shard_cap,
                                            const uint32_t* pre, uint32_t q) {
  uint32_t lo = 0, hi = kQueueShards;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pre[mid] <= q) lo = mid;
    else hi = mid;
  }
  return slots[(size_t)lo * shard_cap + (q - pre[lo])];
}

// Wave-aggregated append of the `want` lanes' value to a shard: one atomic
// per wave, slot order by lane.
__device__ __forceinline__ void wave_append(bool want, uint32_t value, uint32_t* __restrict__ shard_slots,
                                            uint32_t* __restrict__ shard_count) {
  const uint64_t mask = __ballot(want);
  if (mask == 0) return;
  const uint32_t lane = __lane_id();
  const uint32_t leader = (uint32_t)(__ffsll((unsigned long long)mask) - 1);
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(shard_count, (uint32_t)__popcll(mask));
  base = __shfl(base, (int)leader);
  if (want) {
    const uint64_t below = mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
    shard_slots[base + (uint32_t)__popcll(below)] = value;
  }
}

// Index box of the move; false if the end point leaves the partition.
__device__ __forceinline__ bool move_index_box(const PartitionParams& P, D3 pos, D3 end, int lo[3], int hi[3]) {
  if (!in_partition(P, end)) return false;
  int a[3], b[3];
  subpart_3d(P, pos, a[0], a[1], a[2]);
  subpart_3d(P, end, b[0], b[1], b[2]);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lo[k] = a[k] < b[k] ? a[k] : b[k];
    hi[k] = a[k] < b[k] ? b[k] : a[k];
    if (lo[k] < 0 || hi[k] >= (int)P.n) return false;
  }
  return true;
}

// FP64: every wall of the index box is an exact WALL_MISS of collide_wall
// (no hit, no re-draw).  Exact in the reference's arithmetic.
// `ent` holds the wall records in list order (ent[k] = walls[idx[k]]).
__device__ __forceinline__ bool all_walls_miss64(const PartitionParams& P, const WallRec64* __restrict__ ent,
                                                 const uint32_t* __restrict__ off, const uint32_t* __restrict__ box,
                                                 D3 pos, D3 disp, const int lo[3], const int hi[3]) {
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        const uint32_t s = subpart_of(P, x, y, z);
        const uint32_t k0 = off[s], k1 = off[s + 1];
        if (k0 == k1) continue;
        const uint32_t sb = segment_box(P, s, pos.x, pos.y, pos.z, disp.x, disp.y, disp.z);
        for (uint32_t k = next_overlap<kLightOverlapGroup>(box, k0, k1, sb); k < k1;
             k = next_overlap<kLightOverlapGroup>(box, k + 1, k1, sb))
          if (!wall_certain_miss64(ent[k], pos, disp)) return false;
      }
  return true;
}

// FP32 filter over the index box; *near_plane set when the FP32 end point is
// within `band` of some wall plane.
__device__ __forceinline__ bool all_walls_miss32(const PartitionParams& P, const FilterWall* __restrict__ ent,
                                                 const uint32_t* __restrict__ off, const uint32_t* __restrict__ box,
                                                 const FilterParams& F, float band, D3 pos, D3 disp,
                                                 const int lo[3], const int hi[3], bool& near_plane) {
  const float px = (float)pos.x, py = (float)pos.y, pz = (float)pos.z;
  const float mx = (float)disp.x, my = (float)disp.y, mz = (float)disp.z;
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        const uint32_t s = subpart_of(P, x, y, z);
        const uint32_t k0 = off[s], k1 = off[s + 1];
        if (k0 == k1) continue;
        const uint32_t sb = segment_box(P, s, pos.x, pos.y, pos.z, disp.x, disp.y, disp.z);
        for (uint32_t k = next_overlap<kLightOverlapGroup>(box, k0, k1, sb); k < k1;
             k = next_overlap<kLightOverlapGroup>(box, k + 1, k1, sb)) {
          float de;
          if (!certain_miss32(ent[k], F, px, py, pz, mx, my, mz, &de)) return false;
          if (fabs_f(de) < band) near_plane = true;
        }
      }
  return true;
}

// The partition's geometry in FP32 for the light passes' conservative
// index and entry boxes.
struct PartitionParams32 {
  float ox, oy, oz;
  float sp_len, sp_rcp;
  float box_scale;  // kSubCells / subpartition edge
  float nf;         // subpartitions per edge
  uint32_t n, n2;
};

__host__ __device__ inline PartitionParams32 make_partition32(const PartitionParams& P) {
  PartitionParams32 Q;
  Q.ox = (float)P.ox;
  Q.oy = (float)P.oy;
  Q.oz = (float)P.oz;
  Q.sp_len = (float)P.sp_len;
  Q.sp_rcp = (float)P.sp_rcp;
  Q.box_scale = (float)((double)kSubCells * P.sp_rcp);
  Q.nf = (float)P.n;
  Q.n = P.n;
  Q.n2 = P.n2;
  return Q;
}

// Subpartition index box of the segment p -> e, computed in FP32 and widened
// by kIndexMargin subpartition widths on every side: every subpartition the
// segment touches lies inside it (the FP32 rounding of the cell coordinates
// is below 1e-4 widths for partitions of up to ~10^3 subpartitions per
// edge).  False when the widened box reaches a face of the partition, so a
// true result also means the end point lies inside the partition: the move
// takes the full step otherwise (clip at the partition face, escape checks).
constexpr float kIndexMargin = 1e-3f;
__device__ __forceinline__ bool index_box32(const PartitionParams32& Q, float px, float py, float pz, float ex,
                                            float ey, float ez, int lo[3], int hi[3]) {
  const float a[3] = {px, py, pz}, b[3] = {ex, ey, ez}, o[3] = {Q.ox, Q.oy, Q.oz};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float t0 = (a[k] - o[k]) * Q.sp_rcp, t1 = (b[k] - o[k]) * Q.sp_rcp;
    const float mn = (t0 < t1 ? t0 : t1) - kIndexMargin, mx = (t0 < t1 ? t1 : t0) + kIndexMargin;
    if (!(mn >= 0.0f && mx < Q.nf)) return false;  // (a NaN fails too)
    lo[k] = (int)mn;
    hi[k] = (int)mx;
  }
  return true;
}

// Walks the wall lists of the index box: for every entry whose box overlaps
// the segment's (FP32 segment boxes), test(k) must return true (the wall
// certainly cannot be hit); false as soon as one does not.
template <typename Test>
__device__ __forceinline__ bool walls_clear(const PartitionParams32& Q, const uint32_t* __restrict__ off,
                                            const uint32_t* __restrict__ box, float px, float py, float pz,
                                            float ex, float ey, float ez, const int lo[3], const int hi[3],
                                            Test test) {
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        const uint32_t s = (uint32_t)x + (uint32_t)y * Q.n + (uint32_t)z * Q.n2;
        const uint32_t k0 = off[s], k1 = off[s + 1];
        if (k0 == k1) continue;
        const uint32_t sb =
            segment_box32(Q.ox + (float)x * Q.sp_len, Q.oy + (float)y * Q.sp_len, Q.oz + (float)z * Q.sp_len,
                          Q.box_scale, px, py, pz, ex, ey, ez);
        for (uint32_t k = next_overlap<kLightOverlapGroup>(box, k0, k1, sb); k < k1;
             k = next_overlap<kLightOverlapGroup>(box, k + 1, k1, sb))
          if (!test(k)) return false;
      }
  return true;
}

}  // namespace mcg
