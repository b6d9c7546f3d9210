// HIP kernels of the diffuse-react engine (gfx950).
#include <hip/hip_runtime.h>

#include "device/path64.h"
#include "device/rng.h"
#include "device/types.h"
#include "kernels.h"

namespace mcg {

__constant__ ZigguratTables c_zig = MCG_ZIGGURAT_TABLES_INIT;

// Stages the Ziggurat tables (3 KB) into LDS so that the per-lane strip
// lookups of a wave are LDS gathers rather than scattered cache reads.
__device__ __forceinline__ void stage_ziggurat(ZigguratTables& T) {
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&c_zig);
  uint32_t* dst = reinterpret_cast<uint32_t*>(&T);
  constexpr int words = sizeof(ZigguratTables) / 4;
  for (int k = threadIdx.x; k < words; k += blockDim.x) dst[k] = src[k];
  __syncthreads();
}

// One iteration of every live volume molecule that has no volume reactions:
// draw the displacement (pick_vol_displacement, diffusion_utils.inl:113-119),
// ray-trace it through the subpartition grid and reflect off walls.
__global__ __launch_bounds__(256) void k_diffuse_reflect64(MolView64 m, uint32_t n,
                                                           const SpeciesRec* __restrict__ species,
                                                           PartitionParams P, WallView64 W, uint64_t seed,
                                                           uint32_t iteration, uint32_t* __restrict__ err_out,
                                                           HitTrace trace) {
  __shared__ ZigguratTables T;
  stage_ziggurat(T);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t tag = m.tag[i];
  if (tag_flags(tag) & MF_DEFUNCT) return;
  const SpeciesRec sp = species[tag_species(tag)];
  if (!(sp.space_step > 0)) return;
  PhiloxStream rng;
  rng.init(seed, m.id[i], iteration, tag_replica(tag), 0);
  D3 pos = {m.x[i], m.y[i], m.z[i]};
  const double scale = sp.space_step;
  D3 disp;
  disp.x = scale * gauss(rng, T) * 0.70710678118654752440;
  disp.y = scale * gauss(rng, T) * 0.70710678118654752440;
  disp.z = scale * gauss(rng, T) * 0.70710678118654752440;
  uint32_t err = 0;
  uint32_t tn = 0;
  uint32_t* tw = trace.enabled ? trace.wall + (size_t)i * kTraceHits : nullptr;
  diffuse_reflect64(P, W.walls, W.off, W.idx, pos, disp, rng, err, tw, tn);
  m.x[i] = pos.x;
  m.y[i] = pos.y;
  m.z[i] = pos.z;
  if (trace.enabled) trace.count[i] = tn;
  if (err) atomicOr(err_out, err);
}

// Copy kernel of known size: calibrates the HBM byte counters and gives the
// achievable stream bandwidth on this device.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_diffuse_reflect64(hipStream_t s, const MolView64& m, uint32_t n, const SpeciesRec* species,
                                    const PartitionParams& P, const WallView64& W, uint64_t seed, uint32_t iteration,
                                    uint32_t* err, const HitTrace& trace) {
  if (n == 0) return hipSuccess;
  dim3 block(256), grid((n + 255) / 256);
  hipLaunchKernelGGL(k_diffuse_reflect64, grid, block, 0, s, m, n, species, P, W, seed, iteration, err, trace);
  return hipGetLastError();
}

hipError_t launch_copy(hipStream_t s, const void* src, void* dst, size_t bytes) {
  size_t n = bytes / sizeof(float4);
  if (n == 0) return hipSuccess;
  dim3 block(256), grid(2048);
  hipLaunchKernelGGL(k_copy, grid, block, 0, s, static_cast<const float4*>(src), static_cast<float4*>(dst), n);
  return hipGetLastError();
}

}  // namespace mcg
