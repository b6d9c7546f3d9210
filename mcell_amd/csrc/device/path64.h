// FP64 volume-molecule ray trace: subpartition walk, ray-triangle test with
// edge re-draw (REDO), reflection.  Every expression keeps the association
// order of the reference (glm dot = (x*x' + y*y') + z*z') and the file is
// compiled with -ffp-contract=off, so results are bit-identical to a plain
// FP64 CPU evaluation of the same formulas.
#pragma once
#include "rng.h"
#include "types.h"
#include "wallbox.h"

namespace mcg {

struct D3 {
  double x, y, z;
};
__host__ __device__ __forceinline__ D3 d3(double x, double y, double z) { return {x, y, z}; }
__host__ __device__ __forceinline__ double dot3(D3 a, D3 b) {
  double tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z;
  return (tx + ty) + tz;
}
__host__ __device__ __forceinline__ double maxd(double a, double b) { return (a < b) ? b : a; }

constexpr double POS_EPS = 1e-12;
constexpr double STIME_EPS = 1e-12;
constexpr double FLT_MIN_D = 1.17549435082228750797e-38;

// src/util.c:449-462 (MCell3 distinguishable, == libbng distinguishable_p)
__host__ __device__ __forceinline__ bool distinguishable(double a, double b, double eps) {
  double c = fabs(a - b);
  a = fabs(a);
  if (a < 1) a = 1;
  b = fabs(b);
  if (b < a) eps *= a;
  else eps *= b;
  return c > eps;
}

__host__ __device__ __forceinline__ bool in_partition(const PartitionParams& P, D3 p) {
  return p.x >= P.ox && p.y >= P.oy && p.z >= P.oz && p.x < P.qx && p.y < P.qy && p.z < P.qz;
}

// Partition::get_subpart_3d_indices: IVec3((pos - origin) * rcp), truncating
__host__ __device__ __forceinline__ void subpart_3d(const PartitionParams& P, D3 p, int& ix, int& iy, int& iz) {
  ix = (int)((p.x - P.ox) * P.sp_rcp);
  iy = (int)((p.y - P.oy) * P.sp_rcp);
  iz = (int)((p.z - P.oz) * P.sp_rcp);
}
__host__ __device__ __forceinline__ uint32_t subpart_of(const PartitionParams& P, int ix, int iy, int iz) {
  return (uint32_t)ix + (uint32_t)iy * P.n + (uint32_t)iz * P.n2;
}
__host__ __device__ __forceinline__ uint32_t subpart_index(const PartitionParams& P, D3 p) {
  int ix, iy, iz;
  subpart_3d(P, p, ix, iy, iz);
  return subpart_of(P, ix, iy, iz);
}

// collision_utils.inl:48-114
__device__ __forceinline__ D3 displacement_up_to_partition_boundary(const PartitionParams& P, D3 pos, D3 d) {
  D3 nz = d;
  if (nz.x == 0) nz.x = FLT_MIN_D;
  if (nz.y == 0) nz.y = FLT_MIN_D;
  if (nz.z == 0) nz.z = FLT_MIN_D;
  D3 edges = {P.ox + (nz.x > 0 ? 1.0 : 0.0) * P.edge, P.oy + (nz.y > 0 ? 1.0 : 0.0) * P.edge,
              P.oz + (nz.z > 0 ? 1.0 : 0.0) * P.edge};
  D3 diff = {edges.x - pos.x, edges.y - pos.y, edges.z - pos.z};
  if (fabs(diff.x) < POS_EPS || fabs(diff.y) < POS_EPS || fabs(diff.z) < POS_EPS) return {0, 0, 0};
  double tx = diff.x / nz.x, ty = diff.y / nz.y, tz = diff.z / nz.z;
  double hit = 1;
  if (tx >= 0 && tx < ty && tx <= tz) hit = tx;
  else if (ty >= 0 && ty <= tz) hit = ty;
  else if (tz >= 0) hit = tz;
  double s = hit - STIME_EPS;
  return {d.x * s, d.y * s, d.z * s};
}

// The subpartition walk of collect_crossed_subparts
// (collision_utils_subparts.inl:127-300), produced one subpartition at a time.
struct SubpartWalk64 {
  D3 cp;       // point where the walk entered the current subpartition
  D3 disp;     // displacement of the walk (not changed by REDO)
  D3 rcp;      // 1 / guarded displacement
  int ci[3];
  int add[3];
  int urb[3];
  uint32_t cur, dest;
  int steps;

  __device__ __forceinline__ void init(const PartitionParams& P, D3 pos, uint32_t start, D3 d) {
    cp = pos;
    disp = d;
    D3 nz = d;
    if (nz.x == 0) nz.x = FLT_MIN_D;
    if (nz.y == 0) nz.y = FLT_MIN_D;
    if (nz.z == 0) nz.z = FLT_MIN_D;
    urb[0] = nz.x > 0; urb[1] = nz.y > 0; urb[2] = nz.z > 0;
    add[0] = urb[0] ? 1 : -1; add[1] = urb[1] ? 1 : -1; add[2] = urb[2] ? 1 : -1;
    rcp = {1.0 / nz.x, 1.0 / nz.y, 1.0 / nz.z};
    // start is the subpartition of pos in every caller
    subpart_3d(P, pos, ci[0], ci[1], ci[2]);
    D3 dst = {pos.x + d.x, pos.y + d.y, pos.z + d.z};
    int dx, dy, dz;
    subpart_3d(P, dst, dx, dy, dz);
    dest = subpart_of(P, dx, dy, dz);
    cur = start;
    steps = 0;
  }

  // Advances to the next crossed subpartition; false when the walk is over.
  __device__ __forceinline__ bool next(const PartitionParams& P, uint32_t& err) {
    if (cur == dest) return false;
    if (++steps > 3 * (int)P.n + 3) {
      err |= ERR_DDA;
      return false;
    }
    double ex = P.ox + (double)ci[0] * P.sp_len + (double)urb[0] * P.sp_len;
    double ey = P.oy + (double)ci[1] * P.sp_len + (double)urb[1] * P.sp_len;
    double ez = P.oz + (double)ci[2] * P.sp_len + (double)urb[2] * P.sp_len;
    double tx = (ex - cp.x) * rcp.x;
    double ty = (ey - cp.y) * rcp.y;
    double tz = (ez - cp.z) * rcp.z;
    if (tx < ty && tx <= tz) {
      cp = {cp.x + disp.x * tx, cp.y + disp.y * tx, cp.z + disp.z * tx};
      ci[0] += add[0];
      if (ci[0] < 0 || ci[0] >= (int)P.n) { cur = dest; return false; }
    } else if (ty <= tz) {
      cp = {cp.x + disp.x * ty, cp.y + disp.y * ty, cp.z + disp.z * ty};
      ci[1] += add[1];
      if (ci[1] < 0 || ci[1] >= (int)P.n) { cur = dest; return false; }
    } else {
      cp = {cp.x + disp.x * tz, cp.y + disp.y * tz, cp.z + disp.z * tz};
      ci[2] += add[2];
      if (ci[2] < 0 || ci[2] >= (int)P.n) { cur = dest; return false; }
    }
    cur = subpart_of(P, ci[0], ci[1], ci[2]);
    return true;
  }
};

enum : int { WALL_MISS = 0, WALL_FRONT = 1, WALL_BACK = 2, WALL_REDO = 3 };

__device__ __forceinline__ double abs_max_2vec(D3 a, D3 b) {
  double mx = maxd(fabs(a.x), fabs(b.x)), my = maxd(fabs(a.y), fabs(b.y)), mz = maxd(fabs(a.z), fabs(b.z));
  return maxd(maxd(mx, my), mz);
}

// collision_utils.inl:568-603 (FP64)
__device__ __forceinline__ void jump_away_line(D3 p, double k, D3 A, D3 B, D3 n, PhiloxStream& rng, D3& v) {
  D3 e = {B.x - A.x, B.y - A.y, B.z - A.z};
  double elen2 = dot3(e, e);
  double le_1 = 1.0 / sqrt(elen2);
  e = {e.x * le_1, e.y * le_1, e.z * le_1};
  D3 f;
  f.x = n.y * e.z - n.z * e.y;
  f.y = n.z * e.x - n.x * e.z;
  f.z = n.x * e.y - n.y * e.x;
  double mf = maxd(maxd(fabs(f.x), fabs(f.y)), fabs(f.z));
  double tiny = POS_EPS * (abs_max_2vec(p, v) + 1.0) / (k * mf);
  if ((rng.u32() & 1u) == 0) tiny = -tiny;
  v.x -= tiny * f.x;
  v.y -= tiny * f.y;
  v.z -= tiny * f.z;
}

// collision_utils.inl:629-812 (FP64, update_move = true)
__device__ __forceinline__ int collide_wall64(const WallRec64& w, D3 pos, D3& move, double& t, D3& hit,
                                              PhiloxStream& rng) {
  D3 n = {w.nx, w.ny, w.nz};
  double dp = dot3(n, pos);
  double dv = dot3(n, move);
  double dd = dp - w.d;
  double d_eps;
  if (dd > 0) {
    d_eps = POS_EPS;
    if (dd < d_eps) d_eps = 0.5 * dd;
    if (dd + dv > d_eps) return WALL_MISS;
  } else {
    d_eps = -POS_EPS;
    if (dd > d_eps) d_eps = 0.5 * dd;
    if (dd < 0 && dd + dv < d_eps) return WALL_MISS;
  }
  double a;
  if (dd == 0) {
    if (dv != 0) return WALL_MISS;
    a = (abs_max_2vec(pos, move) + 1.0) * POS_EPS;
    if ((rng.u32() & 1u) == 0) a = -a;
    move = {move.x - a * n.x, move.y - a * n.y, move.z - a * n.z};
    return WALL_REDO;
  }
  a = 1.0 / dv;
  a *= -dd;
  t = a;
  hit = {pos.x + a * move.x, pos.y + a * move.y, pos.z + a * move.z};
  D3 local = {hit.x - w.p0x, hit.y - w.p0y, hit.z - w.p0z};
  double b = dot3(local, d3(w.ux, w.uy, w.uz));
  double c = dot3(local, d3(w.vx, w.vy, w.vz));
  double f;
  if (w.uv2v < 0) {
    c = -c;
    f = -w.uv2v;
  } else {
    f = w.uv2v;
  }
  if (c > 0) {
    double g = b * f;
    double h = c * w.uv2u;
    if (g > h) {
      if (c * w.uv1u + g < h + w.uv1u * w.uv2v) return dv > 0 ? WALL_BACK : WALL_FRONT;
      if (!distinguishable(c * w.uv1u + g, h + w.uv1u * w.uv2v, POS_EPS)) {
        jump_away_line(pos, a, d3(w.p1x, w.p1y, w.p1z), d3(w.p2x, w.p2y, w.p2z), n, rng, move);
        return WALL_REDO;
      }
      return WALL_MISS;
    }
    if (!distinguishable(g, h, POS_EPS)) {
      jump_away_line(pos, a, d3(w.p2x, w.p2y, w.p2z), d3(w.p0x, w.p0y, w.p0z), n, rng, move);
      return WALL_REDO;
    }
    return WALL_MISS;
  }
  if (!distinguishable(c, 0.0, POS_EPS)) {
    jump_away_line(pos, a, d3(w.p0x, w.p0y, w.p0z), d3(w.p1x, w.p1y, w.p1z), n, rng, move);
    return WALL_REDO;
  }
  return WALL_MISS;
}

// collide_wall evaluated without moving anything: true exactly when the
// reference returns WALL_MISS (a hit or a re-draw case returns false and
// consumes no random draw).
__device__ __forceinline__ bool wall_certain_miss64(const WallRec64& w, D3 pos, D3 move) {
  D3 n = {w.nx, w.ny, w.nz};
  double dp = dot3(n, pos);
  double dv = dot3(n, move);
  double dd = dp - w.d;
  double d_eps;
  if (dd > 0) {
    d_eps = POS_EPS;
    if (dd < d_eps) d_eps = 0.5 * dd;
    if (dd + dv > d_eps) return true;
  } else {
    d_eps = -POS_EPS;
    if (dd > d_eps) d_eps = 0.5 * dd;
    if (dd < 0 && dd + dv < d_eps) return true;
  }
  if (dd == 0) return dv != 0;
  double a = 1.0 / dv;
  a *= -dd;
  D3 hit = {pos.x + a * move.x, pos.y + a * move.y, pos.z + a * move.z};
  D3 local = {hit.x - w.p0x, hit.y - w.p0y, hit.z - w.p0z};
  double b = dot3(local, d3(w.ux, w.uy, w.uz));
  double c = dot3(local, d3(w.vx, w.vy, w.vz));
  double f;
  if (w.uv2v < 0) {
    c = -c;
    f = -w.uv2v;
  } else {
    f = w.uv2v;
  }
  if (c > 0) {
    double g = b * f;
    double h = c * w.uv2u;
    if (g > h) {
      if (c * w.uv1u + g < h + w.uv1u * w.uv2v) return false;
      return distinguishable(c * w.uv1u + g, h + w.uv1u * w.uv2v, POS_EPS);
    }
    return distinguishable(g, h, POS_EPS);
  }
  return distinguishable(c, 0.0, POS_EPS);
}

struct WallHit {
  uint32_t wall;
  int type;
  double t;
  D3 pos;
};

// get_closest_wall_collision (collision_utils.inl:819-914, FP64): nearest hit
// inside `subpart`; a REDO restarts the scan of this subpartition with the
// re-drawn displacement.
__device__ __forceinline__ bool closest_wall_in_subpart64(const PartitionParams& P, const WallRec64* __restrict__ walls,
                                                          const uint32_t* __restrict__ off,
                                                          const uint32_t* __restrict__ idx,
                                                          const uint32_t* __restrict__ box, uint32_t subpart,
                                                          uint32_t last_hit, D3 pos, D3& move, PhiloxStream& rng,
                                                          WallHit& best, uint32_t& err) {
  const uint32_t k0 = off[subpart], k1 = off[subpart + 1];
  if (k0 == k1) return false;
  int redo = 0;
  bool found;
  for (;;) {
    found = false;
    double closest = 1e20;  // TIME_FOREVER
    bool restarted = false;
    const uint32_t sb = segment_box(P, subpart, pos.x, pos.y, pos.z, move.x, move.y, move.z);
    for (uint32_t k = next_overlap(box, k0, k1, sb); k < k1; k = next_overlap(box, k + 1, k1, sb)) {
      uint32_t wi = idx[k];
      if (wi == last_hit) continue;
      double t;
      D3 hit;
      int ct = collide_wall64(walls[wi], pos, move, t, hit, rng);
      if (ct == WALL_REDO) {
        restarted = true;
        break;
      }
      if (ct != WALL_MISS) {
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

// One step of a volume molecule that cannot react: Brownian displacement and
// the ray-trace/reflect loop of diffuse_vol_molecule
// (diffuse_react_event.cpp:367-618 with only reflective walls), for a full
// time step (t_steps = 1 internal unit, the molecule's own time_step).
// Walls act by the species' surface classes: reflect, cross (transparent) or
// absorb; returns false when the molecule was absorbed.
__device__ __forceinline__ bool diffuse_reflect64(const PartitionParams& P, const WallRec64* __restrict__ walls,
                                                  const uint32_t* __restrict__ off,
                                                  const uint32_t* __restrict__ idx,
                                                  const uint32_t* __restrict__ box, const SurfaceClasses& SC,
                                                  uint32_t species, D3& pos, D3 disp, PhiloxStream& rng,
                                                  uint32_t& err, uint32_t* trace_walls, uint32_t& trace_n,
                                                  uint32_t& n_reflect, uint32_t& n_cross) {
  uint32_t last_hit = kInvalid;
  int seg;
  for (seg = 0; seg < kMaxSegments; ++seg) {
    D3 end = {pos.x + disp.x, pos.y + disp.y, pos.z + disp.z};
    D3 part_disp = in_partition(P, end) ? disp : displacement_up_to_partition_boundary(P, pos, disp);
    int sx, sy, sz;
    subpart_3d(P, pos, sx, sy, sz);
    if (sx < 0 || sy < 0 || sz < 0 || sx >= (int)P.n || sy >= (int)P.n || sz >= (int)P.n) {
      err |= ERR_ESCAPED;  // never index the wall lists outside the grid
      return true;
    }
    uint32_t start = subpart_of(P, sx, sy, sz);
    WallHit hit;
    bool found = closest_wall_in_subpart64(P, walls, off, idx, box, start, last_hit, pos, disp, rng, hit, err);
    if (!found) {
      // the walk only runs when the move leaves the start subpartition
      D3 dst = {pos.x + part_disp.x, pos.y + part_disp.y, pos.z + part_disp.z};
      int dx, dy, dz;
      subpart_3d(P, dst, dx, dy, dz);
      if (subpart_of(P, dx, dy, dz) != start) {
        SubpartWalk64 walk;
        walk.init(P, pos, start, part_disp);
        while (walk.next(P, err)) {
          if (closest_wall_in_subpart64(P, walls, off, idx, box, walk.cur, last_hit, pos, disp, rng, hit, err)) {
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
      // test_intersect's two draws (rxn_utils.inl:593-626), then destroyed
      (void)rng.dbl();
      (void)rng.dbl();
      pos = hit.pos;
      return false;
    }
    const WallRec64& w = walls[hit.wall];
    const double t_hit = hit.t;
    pos = hit.pos;
    last_hit = hit.wall;
    const double s = 1.0 - t_hit;
    if (act == SC_TRANSPARENT) {
      // cross_transparent_wall (diffuse_react_event.cpp:3071-3094)
      disp = {disp.x * s, disp.y * s, disp.z * s};
      ++n_cross;
      continue;
    }
    // reflect_from_wall (collision_utils.inl:1711-1747)
    double rf = -2.0 * dot3(disp, d3(w.nx, w.ny, w.nz));
    disp = {(disp.x + rf * w.nx) * s, (disp.y + rf * w.ny) * s, (disp.z + rf * w.nz) * s};
    ++n_reflect;
  }
  if (seg == kMaxSegments) err |= ERR_SEGMENTS;
  if (!in_partition(P, pos)) err |= ERR_ESCAPED;
  return true;
}

}  // namespace mcg
