// Surface-molecule diffusion on the wall grids (diffuse_surf_molecule,
// src4/diffuse_react_event.cpp:1071-1246, and ray_trace_surf, :1578-1725).
// Surface molecules live on tiles (the tile word holds species and
// orientation, tile_id the molecule id) with their position in the wall's
// uv frame; a step draws a 2-D displacement (pick_surf_displacement,
// diffusion_utils.inl:60-93), traces it across the shared edges of the
// surface net (reflecting off edges no other wall shares), and moves the
// molecule to the tile under its end point if that tile is free, else
// draws again (at most 1 + SURFACE_DIFFUSION_RETRIES times).
#pragma once
#include "path64.h"
#include "react.h"
#include "rng.h"
#include "types.h"

namespace mcg {

// One edge of a wall in the surface net (Edge, src4/wall.h:32-89): the two
// walls it joins (bwd = kInvalid: not shared) and the uv transform from the
// forward wall's frame to the backward wall's.  48 bytes.
struct WallEdgeRec {
  double cos_t, sin_t, tu, tv;
  uint32_t fwd, bwd;
  uint32_t pad0, pad1;
};
static_assert(sizeof(WallEdgeRec) == 48, "edge record");

constexpr int kSurfRetries = 10;  // SURFACE_DIFFUSION_RETRIES (src/mcell_structs_shared.h:61)
constexpr int kEdgeWithin = 3, kEdgeCannotTell = 4;
constexpr int kMaxEdgeCrossings = 4096;  // walls one 2-D step may cross before it is refused

__device__ __host__ __forceinline__ double det2(double au, double av, double bu, double bv) { return au * bv - av * bu; }

// defines.h:733-763
__device__ __host__ __forceinline__ bool distinguishable_uv(double au, double av, double bu, double bv, double eps) {
  double c = fabs(au), d;
  d = fabs(av); if (d > c) c = d;
  d = fabs(bu); if (d > c) c = d;
  d = fabs(bv); if (d > c) c = d;
  double cc = fabs(au - bu);
  d = fabs(av - bv); if (d > cc) cc = d;
  if (c < eps) c = eps;
  return c * eps < cc;
}

// find_edge_point (geometry_utils.inl:215-288): the edge (0, 1, 2) the move
// from (lu, lv) by (du, dv) leaves the wall through, with the point on it;
// kEdgeWithin when the move ends inside, kEdgeCannotTell when ambiguous.
__device__ __host__ __forceinline__ int find_edge_point(const WallRec64& w, double lu, double lv, double du,
                                                        double dv, double& eu, double& ev) {
  const double lxd = det2(lu, lv, du, dv);
  const double lxc1 = -lv * w.uv1u;
  const double dxc1 = -dv * w.uv1u;
  double f, s, t;
  if (dxc1 < -POS_EPS || dxc1 > POS_EPS) {
    f = 1 / dxc1;
    s = -lxd * f;
    if (0 < s && s < 1 && f > 0) {
      t = -lxc1 * f;
      if (POS_EPS < t && t < 1) {
        eu = lu + t * du;
        ev = lv + t * dv;
        return 0;
      } else if (t > 1 + POS_EPS) {
        return kEdgeWithin;
      }
    }
  }
  const double lxc2 = det2(lu, lv, w.uv2u, w.uv2v);
  const double dxc2 = det2(du, dv, w.uv2u, w.uv2v);
  if (dxc2 < -POS_EPS || dxc2 > POS_EPS) {
    f = 1 / dxc2;
    s = 1 + lxd * f;
    if (0 < s && s < 1 && f < 0) {
      t = -lxc2 * f;
      if (POS_EPS < t && t < 1) {
        eu = lu + t * du;
        ev = lv + t * dv;
        return 2;
      } else if (t > 1 + POS_EPS) {
        return kEdgeWithin;
      }
    }
  }
  f = dxc2 - dxc1;
  if (f < -POS_EPS || f > POS_EPS) {
    f = 1 / f;
    s = -(lxd + dxc1) * f;
    if (0 < s && s < 1 && f > 0) {
      t = (w.uv1u * w.uv2v + lxc1 - lxc2) * f;
      if (POS_EPS < t && t < 1) {
        eu = lu + t * du;
        ev = lv + t * dv;
        return 1;
      } else if (t > 1 + POS_EPS) {
        return kEdgeWithin;
      }
    }
  }
  return kEdgeCannotTell;
}

// traverse_surface (geometry_utils.inl:300-340): (lu, lv) in the frame of
// the wall across edge `which` of wall `wi`; kInvalid when not shared.
__device__ __host__ __forceinline__ uint32_t traverse_surface(const WallEdgeRec& e, uint32_t wi, double lu, double lv,
                                                              double& nu, double& nv) {
  if (e.fwd == wi) {
    const double tu = e.cos_t * lu + e.sin_t * lv;
    const double tv = -e.sin_t * lu + e.cos_t * lv;
    nu = tu + e.tu;
    nv = tv + e.tv;
    return e.bwd;
  }
  const double tu = lu - e.tu, tv = lv - e.tv;
  nu = e.cos_t * tu - e.sin_t * tv;
  nv = e.sin_t * tu + e.cos_t * tv;
  return e.fwd;
}

// uv2grid_tile_index (grid_utils.inl:120-191); kInvalid outside the grid.
__device__ __host__ __forceinline__ uint32_t uv2grid(const WallRec64& w, const WallGrid& g, double u, double v) {
  if (g.num_tiles == 1) return 0;
  const uint32_t tile_idx_mid = g.num_tiles - 2 * g.n_axis + 1;
  if (!distinguishable_uv(u, v, 0.0, 0.0, POS_EPS)) return tile_idx_mid;
  if (!distinguishable_uv(u, v, w.uv1u, 0.0, POS_EPS)) return 0;
  if (!distinguishable_uv(u, v, w.uv2u, w.uv2v, POS_EPS)) return g.num_tiles - 1;
  const double striploc = v * g.strip_width_rcp;
  int strip = (int)striploc;
  const double striprem = striploc - strip;
  strip = (int)g.n_axis - strip - 1;
  const double u0 = v * g.vert2_slope;
  const double u1_u0 = w.uv1u - v * g.fullslope;
  const double stripeloc = ((u - u0) / u1_u0) * (strip + (1 - striprem));
  const int stripe = (int)stripeloc;
  const double striperem = stripeloc - stripe;
  const int flip = (striperem < 1 - striprem) ? 0 : 1;
  const int idx = strip * strip + 2 * stripe + flip;
  return (idx < 0 || (uint32_t)idx >= g.num_tiles) ? kInvalid : (uint32_t)idx;
}

// ray_trace_surf (diffuse_react_event.cpp:1578-1725) without region
// borders: the wall the move from (u, v) on wall wi ends on and the end
// point in its frame; kInvalid when an edge crossing is ambiguous (the
// caller draws again).
__device__ __host__ __forceinline__ uint32_t ray_trace_surf(const WallRec64* __restrict__ walls,
                                                            const WallEdgeRec* __restrict__ edges, uint32_t wi,
                                                            double u, double v, double du, double dv, double& nu,
                                                            double& nv, uint32_t& err) {
  double pu = u, pv = v, qu = du, qv = dv;
  for (int guard = 0; guard < kMaxEdgeCrossings; ++guard) {
    const WallRec64& w = walls[wi];
    double bu = 0, bv = 0;
    const int e = find_edge_point(w, pu, pv, qu, qv, bu, bv);
    if (e == kEdgeCannotTell) return kInvalid;
    if (e == kEdgeWithin) {
      nu = pu + qu;
      nv = pv + qv;
      return wi;
    }
    const double ou = pu, ov = pv;
    const WallEdgeRec& ed = edges[3 * wi + e];
    double tu = 0, tv = 0;
    const uint32_t target = (ed.bwd == kInvalid) ? kInvalid : traverse_surface(ed, wi, ou, ov, tu, tv);
    if (target != kInvalid) {
      double mu, mv;
      traverse_surface(ed, wi, ou + qu, ov + qv, mu, mv);
      pu = tu;
      pv = tv;
      qu = mu - pu;
      qv = mv - pv;
      wi = target;
      continue;
    }
    // no neighbouring wall: reflect off the edge
    double nd_u = qu - (bu - ou), nd_v = qv - (bv - ov);
    if (e == 0) {
      nd_v *= -1.0;
    } else {
      double ru = e == 1 ? -w.uv2v : w.uv2v;
      double rv = e == 1 ? w.uv2u - w.uv1u : -w.uv2u;
      double f = 1.0 / sqrt(ru * ru + rv * rv);
      ru *= f;
      rv *= f;
      f = 2.0 * (nd_u * ru + nd_v * rv);
      nd_u -= f * ru;
      nd_v -= f * rv;
    }
    pu = bu;
    pv = bv;
    qu = nd_u;
    qv = nd_v;
  }
  err |= ERR_SEGMENTS;
  return kInvalid;
}

// grid2uv (grid_utils.inl:233-250): the centre of tile `index`.
__device__ __host__ __forceinline__ void grid2uv(const WallRec64& w, const WallGrid& g, uint32_t index, double& u,
                                                 double& v) {
  const int root = (int)sqrt((double)index);
  const int rootrem = (int)index - root * root;
  const int k = (int)g.n_axis - root - 1;
  const int j = rootrem / 2;
  const int i = rootrem - 2 * j;
  const double over3n = 1 / (double)(3 * g.n_axis);
  u = ((double)(3 * j + i + 1)) * over3n * w.uv1u + ((double)(3 * k + i + 1)) * over3n * w.uv2u;
  v = ((double)(3 * k + i + 1)) * over3n * w.uv2v;
}

// pick_surf_displacement (diffusion_utils.inl:60-93): Marsaglia's polar
// method on 16-bit halves of one draw, rejected until 1e-12 <= f <= 1.
__device__ __forceinline__ void surf_displacement(PhiloxStream& rng, double scale, double& du, double& dv) {
  const double one_over_2_to_16th = 1.52587890625e-5;
  double au, av, f;
  do {
    const uint32_t n = rng.u32();
    au = 2 * one_over_2_to_16th * (n & 0xFFFF) - 1;
    av = 2 * one_over_2_to_16th * (n >> 16) - 1;
    f = au * au + av * av;
  } while ((f < POS_EPS) || (f > 1));
  const double s = sqrt(-log(f) / f) * scale;
  du = au * s;
  dv = av * s;
}

}  // namespace mcg
