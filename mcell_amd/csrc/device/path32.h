// FP32+fallback form: wall views and the FP32 rounding of end points.
//
// Positions live in HBM as FP32.  Inside a step the molecule's position and
// displacement are carried in FP64 registers.  The light passes test every
// wall of a move's box with the FP32 filter (filter32.h) and commit only
// moves it certifies as misses; every other move takes the full step, whose
// wall tests are the FP64 collide_wall (which also yields the FP64
// reflection point).  Time bookkeeping is FP64.  At the end of a step the
// position is rounded to FP32; if the end point lies within a few FP32 ulps
// of the plane of a nearby wall the rounded point is moved back onto the side
// the FP64 point is on.
#pragma once
#include "filter32.h"
#include "path64.h"

namespace mcg {

struct Walls32View {
  const WallRec64* __restrict__ w64;
  const FilterWall* __restrict__ w32;
  const uint32_t* __restrict__ off;
  const uint32_t* __restrict__ idx;
  const FilterWall* __restrict__ ent32;  // w32[idx[k]] for every list entry k
  const uint32_t* __restrict__ box;      // sub-cell box of every list entry k
};

// Rounds the FP64 end point to FP32 keeping it on the same side of the plane
// of every wall of its subpartition that passes next to it (entry box within
// the box margin of the point).  The plane of a wall whose triangle is farther
// away may be crossed: a later move from the rounded point meets that plane
// within rounding distance of the point, outside the triangle, and misses.
// Each pass steps the point one ulp per axis towards the FP64 side of every
// plane it is on or across; near an acute edge one step can cross another
// plane, so passes repeat until one moves nothing.  Returns false when
// kRoundPasses passes did not settle (the caller reports ERR_ROUND32).
constexpr int kRoundPasses = 8;
__device__ __forceinline__ bool round_to_fp32_same_side(const PartitionParams& P, const WallRec64* __restrict__ walls,
                                                     const uint32_t* __restrict__ off,
                                                     const uint32_t* __restrict__ idx,
                                                     const uint32_t* __restrict__ box, double px, double py,
                                                     double pz, float* ox, float* oy, float* oz) {
  float fx = (float)px, fy = (float)py, fz = (float)pz;
  bool settled = true;
  int sx, sy, sz;
  subpart_3d(P, D3{px, py, pz}, sx, sy, sz);
  if (!(sx < 0 || sy < 0 || sz < 0 || sx >= (int)P.n || sy >= (int)P.n || sz >= (int)P.n)) {
    uint32_t s = subpart_of(P, sx, sy, sz);
    const uint32_t pb = segment_box(P, s, px, py, pz, 0.0, 0.0, 0.0);
    settled = false;
    for (int pass = 0; pass < kRoundPasses && !settled; ++pass) {
      bool moved = false;
      const uint32_t e1 = off[s + 1];
      for (uint32_t k = next_overlap(box, off[s], e1, pb); k < e1; k = next_overlap(box, k + 1, e1, pb)) {
        const WallRec64& w = walls[idx[k]];
        double sd = dot3(D3{w.nx, w.ny, w.nz}, D3{px, py, pz}) - w.d;
        double sf = dot3(D3{w.nx, w.ny, w.nz}, D3{(double)fx, (double)fy, (double)fz}) - w.d;
        // strictly on the FP64 point's side; a rounded point exactly on the
        // plane counts as crossed (collide_wall lets a move leave from there)
        if (sd == 0 || (sd > 0 && sf > 0) || (sd < 0 && sf < 0)) continue;
        // step the FP32 point one ulp per axis towards the FP64 side
        float dir = sd > 0 ? 1.0f : -1.0f;
        if (w.nx * dir > 0) fx = nextafterf(fx, 3.4e38f);
        else if (w.nx * dir < 0) fx = nextafterf(fx, -3.4e38f);
        if (w.ny * dir > 0) fy = nextafterf(fy, 3.4e38f);
        else if (w.ny * dir < 0) fy = nextafterf(fy, -3.4e38f);
        if (w.nz * dir > 0) fz = nextafterf(fz, 3.4e38f);
        else if (w.nz * dir < 0) fz = nextafterf(fz, -3.4e38f);
        moved = true;
      }
      settled = !moved;
    }
  }
  *ox = fx;
  *oy = fy;
  *oz = fz;
  return settled;
}

}  // namespace mcg
