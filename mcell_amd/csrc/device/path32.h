// FP32+fallback volume-molecule step.
//
// Positions live in HBM as FP32.  Inside a step the molecule's position and
// displacement are carried in FP64 registers; every wall test goes through
// the FP32 filter (filter32.h) and only candidates are re-solved by the FP64
// collide_wall, which also yields the FP64 reflection point.  Time bookkeeping
// is FP64.  At the end of the step the position is rounded to FP32; if the
// end point lies within a few FP32 ulps of a wall plane the rounded point is
// moved back onto the side the FP64 point is on.
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

// The FP64 test of one candidate wall (inlined: an out-of-line call forces
// the caller's live registers through scratch memory).
__device__ __forceinline__ int collide_wall64_ool(const WallRec64* __restrict__ walls, uint32_t wi, double px, double py,
                                               double pz, double* mx, double* my, double* mz, double* t, double* hx,
                                               double* hy, double* hz, PhiloxStream* rng) {
  D3 pos = {px, py, pz};
  D3 move = {*mx, *my, *mz};
  D3 hit;
  double tt = 0;
  int ct = collide_wall64(walls[wi], pos, move, tt, hit, *rng);
  *mx = move.x;
  *my = move.y;
  *mz = move.z;
  *t = tt;
  *hx = hit.x;
  *hy = hit.y;
  *hz = hit.z;
  return ct;
}

// get_closest_wall_collision with the FP32 filter in front of the FP64 test.
// `end_band` reports whether the end point of the move lies within `band` of
// any tested wall plane (FP32 estimate).
__device__ __forceinline__ bool closest_wall_in_subpart32(const PartitionParams& P, const Walls32View& W,
                                                          const FilterParams& F, uint32_t subpart, uint32_t last_hit,
                                                          D3 pos, D3& move, PhiloxStream& rng, WallHit& best,
                                                          uint32_t& err, float band, bool& end_band) {
  const uint32_t k0 = W.off[subpart], k1 = W.off[subpart + 1];
  if (k0 == k1) return false;
  int redo = 0;
  bool found;
  for (;;) {
    found = false;
    double closest = 1e20;
    bool restarted = false;
    const float px = (float)pos.x, py = (float)pos.y, pz = (float)pos.z;
    const float mx = (float)move.x, my = (float)move.y, mz = (float)move.z;
    const uint32_t sb = segment_box(P, subpart, pos.x, pos.y, pos.z, move.x, move.y, move.z);
    for (uint32_t k = k0; k < k1; ++k) {
      // a wall the move cannot come near: a certain miss whose end point
      // is not next to the wall itself
      if (!boxes_overlap(W.box[k], sb)) continue;
      uint32_t wi = W.idx[k];
      float de;
      if (wi == last_hit) {
        // not hit again, but the end point may lie next to its plane (a
        // short remainder after the reflection): FP32 rounding must see it
        (void)certain_miss32(W.ent32[k], F, px, py, pz, mx, my, mz, &de);
        if (fabs_f(de) < band) end_band = true;
        continue;
      }
      if (certain_miss32(W.ent32[k], F, px, py, pz, mx, my, mz, &de)) {
        if (fabs_f(de) < band) end_band = true;
        continue;
      }
      end_band = true;
      double t, hx, hy, hz;
      int ct = collide_wall64_ool(W.w64, wi, pos.x, pos.y, pos.z, &move.x, &move.y, &move.z, &t, &hx, &hy, &hz, &rng);
      if (ct == WALL_REDO) {
        restarted = true;
        break;
      }
      if (ct != WALL_MISS) {
        D3 hit = {hx, hy, hz};
        if (subpart_index(P, hit) != subpart) continue;
        if (t < closest) {
          found = true;
          closest = t;
          best.wall = wi;
          best.type = ct;
          best.t = t;
          best.pos = hit;
        }
      }
    }
    if (!restarted) break;
    if (++redo > kMaxRedo) {
      err |= ERR_REDO;
      return false;
    }
  }
  return found;
}

// Rounds the FP64 end point to FP32 keeping it on the same side of the plane
// of every wall of its subpartition that passes next to it (entry box within
// the box margin of the point).  The plane of a wall whose triangle is farther
// away may be crossed: a later move from the rounded point meets that plane
// within rounding distance of the point, outside the triangle, and misses.
__device__ __forceinline__ void round_to_fp32_same_side(const PartitionParams& P, const WallRec64* __restrict__ walls,
                                                     const uint32_t* __restrict__ off,
                                                     const uint32_t* __restrict__ idx,
                                                     const uint32_t* __restrict__ box, double px, double py,
                                                     double pz, float* ox, float* oy, float* oz) {
  float fx = (float)px, fy = (float)py, fz = (float)pz;
  int sx, sy, sz;
  subpart_3d(P, D3{px, py, pz}, sx, sy, sz);
  if (!(sx < 0 || sy < 0 || sz < 0 || sx >= (int)P.n || sy >= (int)P.n || sz >= (int)P.n)) {
    uint32_t s = subpart_of(P, sx, sy, sz);
    const uint32_t pb = segment_box(P, s, px, py, pz, 0.0, 0.0, 0.0);
    for (int pass = 0; pass < 4; ++pass) {
      bool moved = false;
      for (uint32_t k = off[s]; k < off[s + 1]; ++k) {
        if (!boxes_overlap(box[k], pb)) continue;
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
      if (!moved) break;
    }
  }
  *ox = fx;
  *oy = fy;
  *oz = fz;
}

// diffuse_reflect64 with the FP32 filter; returns whether the end point is
// within `band` of a wall plane.
// `alive` becomes false when a surface class absorbs the molecule.
__device__ __forceinline__ bool diffuse_reflect32(const PartitionParams& P, const Walls32View& W,
                                                  const FilterParams& F, float band, const SurfaceClasses& SC,
                                                  uint32_t species, D3& pos, D3 disp, PhiloxStream& rng,
                                                  uint32_t& err, uint32_t* trace_walls, uint32_t& trace_n,
                                                  bool& alive) {
  uint32_t last_hit = kInvalid;
  int seg;
  bool end_band = false;
  alive = true;
  for (seg = 0; seg < kMaxSegments; ++seg) {
    end_band = false;
    D3 end = {pos.x + disp.x, pos.y + disp.y, pos.z + disp.z};
    D3 part_disp = in_partition(P, end) ? disp : displacement_up_to_partition_boundary(P, pos, disp);
    int sx, sy, sz;
    subpart_3d(P, pos, sx, sy, sz);
    if (sx < 0 || sy < 0 || sz < 0 || sx >= (int)P.n || sy >= (int)P.n || sz >= (int)P.n) {
      err |= ERR_ESCAPED;
      return false;
    }
    uint32_t start = subpart_of(P, sx, sy, sz);
    WallHit hit;
    bool found = closest_wall_in_subpart32(P, W, F, start, last_hit, pos, disp, rng, hit, err, band, end_band);
    if (!found) {
      // the walk only runs when the move leaves the start subpartition
      D3 dst = {pos.x + part_disp.x, pos.y + part_disp.y, pos.z + part_disp.z};
      int dx, dy, dz;
      subpart_3d(P, dst, dx, dy, dz);
      if (subpart_of(P, dx, dy, dz) != start) {
        SubpartWalk64 walk;
        walk.init(P, pos, start, part_disp);
        while (walk.next(P, err)) {
          if (closest_wall_in_subpart32(P, W, F, walk.cur, last_hit, pos, disp, rng, hit, err, band, end_band)) {
            found = true;
            break;
          }
        }
      }
    }
    if (!found) {
      pos = {pos.x + disp.x, pos.y + disp.y, pos.z + disp.z};
      break;
    }
    if (trace_walls) {
      if (trace_n < (uint32_t)kTraceHits) trace_walls[trace_n] = hit.wall;
      else err |= ERR_TRACE;
      ++trace_n;
    }
    const uint32_t act = SC.of(hit.wall, species, hit.type == WALL_FRONT);
    if (act == SC_ABSORB) {
      (void)rng.dbl();
      (void)rng.dbl();
      pos = hit.pos;
      alive = false;
      return false;
    }
    const WallRec64& w = W.w64[hit.wall];
    double t_hit = hit.t;
    pos = hit.pos;
    last_hit = hit.wall;
    double s = 1.0 - t_hit;
    if (act == SC_TRANSPARENT) {
      disp = {disp.x * s, disp.y * s, disp.z * s};
      continue;
    }
    double rf = -2.0 * dot3(disp, d3(w.nx, w.ny, w.nz));
    disp = {(disp.x + rf * w.nx) * s, (disp.y + rf * w.ny) * s, (disp.z + rf * w.nz) * s};
  }
  if (seg == kMaxSegments) err |= ERR_SEGMENTS;
  if (!in_partition(P, pos)) err |= ERR_ESCAPED;
  return end_band;
}

}  // namespace mcg
