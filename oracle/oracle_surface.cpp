// TEST INFRASTRUCTURE ONLY (the oracle; see oracle_core.h).
//
// Surface-molecule diffusion, restated from the reference in FP64 and its
// evaluation order:
//   surface_net, compatible_edges          src4/geometry.cpp:50-97, 256-353
//   Edge::reinit_edge_constants             src4/wall.cpp:134-235
//   find_edge_point, traverse_surface       src4/geometry_utils.inl:215-342
//   uv2grid_tile_index, grid2uv             src4/grid_utils.inl:120-191, 233-250
//   pick_surf_displacement                  src4/diffusion_utils.inl:60-93
//   move_sm_on_same_triangle / _to_new_     src4/diffusion_utils.inl:453-548
//   ray_trace_surf                          src4/diffuse_react_event.cpp:1578-1725
//   diffuse_surf_molecule                   src4/diffuse_react_event.cpp:1071-1246
// Region borders (reflect_absorb_inside_out / _outside_in) and surface-surface
// reactions are not on this path: surface species here have no surface-class
// or surface-surface reactions.

#include <map>
#include <stdexcept>
#include <tuple>

#include "oracle_core.h"

namespace orc {

bool distinguishable_vec3(V3 a, V3 b, double eps);

namespace {

struct V2 {
  double u, v;
};
inline double determinant2(V2 a, V2 b) { return a.u * b.v - a.v * b.u; }
inline double dot2(V2 a, V2 b) { return a.u * b.u + a.v * b.v; }
inline double len2_squared(V2 a) { return a.u * a.u + a.v * a.v; }

// defines.h:733-763
bool distinguishable_vec2(V2 a, V2 b, double eps) {
  double c = std::fabs(a.u), d;
  d = std::fabs(a.v); if (d > c) c = d;
  d = std::fabs(b.u); if (d > c) c = d;
  d = std::fabs(b.v); if (d > c) c = d;
  double cc = std::fabs(a.u - b.u);
  d = std::fabs(a.v - b.v); if (d > cc) cc = d;
  if (c < eps) c = eps;
  return c * eps < cc;
}

constexpr int kEdgeWithin = 3, kEdgeCannotTell = 4;
constexpr int kSurfaceDiffusionRetries = 10;  // src/mcell_structs_shared.h:61

}  // namespace

// Edge::reinit_edge_constants (wall.cpp:134-235): the transform from the
// forward wall's uv frame to the backward wall's, flattened along the edge
// that starts at the forward wall's vertex i.
static void edge_constants(const World& w, WallEdge& e, uint32_t i) {
  const Wall& wf = w.walls[e.fwd];
  const Wall& wb = w.walls[e.bwd];
  const uint32_t j = (i + 1) % 3;
  const V3 wf0 = w.verts[wf.v[0]], wfi = w.verts[wf.v[i]], wfj = w.verts[wf.v[j]];
  const V3 wb0 = w.verts[wb.v[0]];
  const V3 di0 = wfi - wf0;
  V2 of{dot(di0, wf.unit_u), dot(di0, wf.unit_v)};
  const V3 dj0 = wfj - wf0;
  V2 tff{dot(dj0, wf.unit_u) - of.u, dot(dj0, wf.unit_v) - of.v};
  if (!(std::fabs(len2_squared(tff) - 0.0) >= POS_EPS)) throw std::runtime_error("malformed edge");
  const double df = 1 / std::sqrt(len2_squared(tff));
  V2 ef{tff.u * df, tff.v * df};
  V2 ff{-ef.v, ef.u};
  const V3 dib = wfi - wb0;
  V2 ob{dot(dib, wb.unit_u), dot(dib, wb.unit_v)};
  const V3 djb = wfj - wb0;
  V2 tfb{dot(djb, wb.unit_u) - ob.u, dot(djb, wb.unit_v) - ob.v};
  if (!(std::fabs(len2_squared(tfb) - 0.0) >= POS_EPS)) throw std::runtime_error("malformed edge");
  const double db = 1 / std::sqrt(len2_squared(tfb));
  V2 eb{tfb.u * db, tfb.v * db};
  V2 fb{-eb.v, eb.u};
  const double m00 = ef.u * eb.u + ff.u * fb.u;
  const double m01 = ef.v * eb.u + ff.v * fb.u;
  const double m10 = ef.u * eb.v + ff.u * fb.v;
  const double m11 = ef.v * eb.v + ff.v * fb.v;
  V2 q = ob;
  q.u -= m00 * of.u + m01 * of.v;
  q.v -= m10 * of.u + m11 * of.v;
  e.cos_t = m00;
  e.sin_t = m01;
  e.tu = q.u;
  e.tv = q.v;
}

// surface_net (geometry.cpp:256-353) per object: the edges of its walls (in
// wall order) that two walls share with opposite vertex order are linked,
// the first wall being the transform's forward side.  Edges are matched by
// their vertices' coordinates; an edge of more than two walls (the
// reference's refine_edge_pairs) is refused.
void World::build_surface_net() {
  edges.assign(walls.size(), {});
  using Key = std::tuple<double, double, double, double, double, double>;
  std::map<std::pair<uint32_t, Key>, std::vector<std::pair<uint32_t, uint32_t>>> groups;
  std::vector<std::pair<uint32_t, Key>> order;
  for (uint32_t wi = 0; wi < walls.size(); ++wi) {
    const Wall& w = walls[wi];
    for (uint32_t j = 0; j < 3; ++j) {
      const V3 a = verts[w.v[j]], b = verts[w.v[(j + 1) % 3]];
      Key k = std::make_tuple(a.x, a.y, a.z, b.x, b.y, b.z);
      Key kr = std::make_tuple(b.x, b.y, b.z, a.x, a.y, a.z);
      if (kr < k) k = kr;
      auto key = std::make_pair(w.object_id, k);
      auto& g = groups[key];
      if (g.empty()) order.push_back(key);
      g.push_back({wi, j});
    }
  }
  for (const auto& key : order) {
    const auto& g = groups[key];
    if (g.size() > 2) throw std::runtime_error("an edge shared by more than two walls of one object");
    if (g.size() < 2) continue;
    const uint32_t wa = g[0].first, ea = g[0].second, wb = g[1].first, eb = g[1].second;
    // compatible_edges (geometry.cpp:50-97): opposite directions, different
    // third vertices
    const Wall& A = walls[wa];
    const Wall& B = walls[wb];
    const V3 a0 = verts[A.v[ea]], a1 = verts[A.v[(ea + 1) % 3]], a2 = verts[A.v[(ea + 2) % 3]];
    const V3 b0 = verts[B.v[eb]], b1 = verts[B.v[(eb + 1) % 3]], b2 = verts[B.v[(eb + 2) % 3]];
    auto same = [](V3 p, V3 q) { return p.x == q.x && p.y == q.y && p.z == q.z; };
    if (!(same(a0, b1) && same(a1, b0) && !same(a2, b2))) continue;
    WallEdge e;
    e.fwd = wa;
    e.bwd = wb;
    edge_constants(*this, e, ea);
    edges[wa][ea] = e;
    edges[wb][eb] = e;
  }
}

// find_edge_point (geometry_utils.inl:215-288): the edge the move from loc
// by disp leaves the wall through (0, 1, 2; edge point set), 3 inside the
// wall, 4 cannot tell.
int World::find_edge_point(uint32_t wi, double lu, double lv, double du, double dv, double& eu,
                           double& ev) const {
  const Wall& here = walls[wi];
  const V2 loc{lu, lv}, disp{du, dv}, uv2{here.uv2_u, here.uv2_v};
  const double lxd = determinant2(loc, disp);
  const double lxc1 = -loc.v * here.uv_vert1_u;
  const double dxc1 = -disp.v * here.uv_vert1_u;
  double f, s, t;
  if (dxc1 < -POS_EPS || dxc1 > POS_EPS) {
    f = 1 / dxc1;
    s = -lxd * f;
    if (0 < s && s < 1 && f > 0) {
      t = -lxc1 * f;
      if (POS_EPS < t && t < 1) {
        eu = loc.u + t * disp.u;
        ev = loc.v + t * disp.v;
        return 0;
      } else if (t > 1 + POS_EPS) {
        return kEdgeWithin;
      }
    }
  }
  const double lxc2 = determinant2(loc, uv2);
  const double dxc2 = determinant2(disp, uv2);
  if (dxc2 < -POS_EPS || dxc2 > POS_EPS) {
    f = 1 / dxc2;
    s = 1 + lxd * f;
    if (0 < s && s < 1 && f < 0) {
      t = -lxc2 * f;
      if (POS_EPS < t && t < 1) {
        eu = loc.u + t * disp.u;
        ev = loc.v + t * disp.v;
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
      t = (here.uv_vert1_u * here.uv2_v + lxc1 - lxc2) * f;
      if (POS_EPS < t && t < 1) {
        eu = loc.u + t * disp.u;
        ev = loc.v + t * disp.v;
        return 1;
      } else if (t > 1 + POS_EPS) {
        return kEdgeWithin;
      }
    }
  }
  return kEdgeCannotTell;
}

// traverse_surface (geometry_utils.inl:300-340): loc in the frame of the
// wall across edge `which`; INVALID_U32 when the edge is not shared (the
// reference's default Edge; newloc is then meaningless).
uint32_t World::traverse_surface(uint32_t wi, double lu, double lv, int which, double& nu, double& nv) const {
  const WallEdge& e = edges[wi][which];
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

// uv2grid_tile_index (grid_utils.inl:120-191)
uint32_t World::uv2grid(double u, double v, uint32_t wi) const {
  const Wall& w = walls[wi];
  const Grid& g = grids[wi];
  if (g.num_tiles == 1) return 0;
  const uint32_t tile_idx_0 = 0;
  const uint32_t tile_idx_mid = g.num_tiles - 2 * g.n_axis + 1;
  const uint32_t tile_idx_last = g.num_tiles - 1;
  const V2 p{u, v};
  if (!distinguishable_vec2(p, V2{0, 0}, POS_EPS)) return tile_idx_mid;
  if (!distinguishable_vec2(p, V2{w.uv_vert1_u, 0}, POS_EPS)) return tile_idx_0;
  if (!distinguishable_vec2(p, V2{w.uv2_u, w.uv2_v}, POS_EPS)) return tile_idx_last;
  const double i = u, j = v;
  const double striploc = j * g.strip_width_rcp;
  int strip = (int)striploc;
  const double striprem = striploc - strip;
  strip = (int)g.n_axis - strip - 1;
  const double u0 = j * g.vert2_slope;
  const double u1_u0 = w.uv_vert1_u - j * g.fullslope;
  const double stripeloc = ((i - u0) / u1_u0) * (strip + (1 - striprem));
  const int stripe = (int)stripeloc;
  const double striperem = stripeloc - stripe;
  const int flip = (striperem < 1 - striprem) ? 0 : 1;
  const int idx = strip * strip + 2 * stripe + flip;
  if (idx < 0 || (uint32_t)idx >= g.num_tiles) throw std::runtime_error("uv2grid: tile index out of range");
  return (uint32_t)idx;
}

// grid2uv (grid_utils.inl:233-250): the tile's centre
void World::grid2uv(uint32_t wi, uint32_t index, double& u, double& v) const {
  const Wall& w = walls[wi];
  const Grid& g = grids[wi];
  const int root = (int)std::sqrt((double)index);
  const int rootrem = (int)index - root * root;
  const int k = (int)g.n_axis - root - 1;
  const int j = rootrem / 2;
  const int i = rootrem - 2 * j;
  const double over3n = 1 / (double)(3 * g.n_axis);
  u = ((double)(3 * j + i + 1)) * over3n * w.uv_vert1_u + ((double)(3 * k + i + 1)) * over3n * w.uv2_u;
  v = ((double)(3 * k + i + 1)) * over3n * w.uv2_v;
}

// ray_trace_surf (diffuse_react_event.cpp:1578-1725) without region
// borders: the wall and uv position the move ends at, INVALID_U32 when an
// edge crossing is ambiguous (the caller draws again).
uint32_t World::ray_trace_surf(uint32_t id, double du, double dv, double& nu, double& nv) {
  const Mol& sm = mols[id];
  uint32_t this_wall = sm.wall;
  V2 this_pos{sm.u, sm.v};
  V2 this_disp{du, dv};
  for (;;) {
    V2 boundary{0, 0};
    const int edge = find_edge_point(this_wall, this_pos.u, this_pos.v, this_disp.u, this_disp.v, boundary.u,
                                     boundary.v);
    if (edge == kEdgeCannotTell) return INVALID_U32;
    if (edge == kEdgeWithin) {
      nu = this_pos.u + this_disp.u;
      nv = this_pos.v + this_disp.v;
      return this_wall;
    }
    const V2 old_pos = this_pos;
    const uint32_t target = traverse_surface(this_wall, old_pos.u, old_pos.v, edge, this_pos.u, this_pos.v);
    if (target != INVALID_U32) {
      V2 d{old_pos.u + this_disp.u, old_pos.v + this_disp.v};
      V2 tmp;
      traverse_surface(this_wall, d.u, d.v, edge, tmp.u, tmp.v);
      this_disp = V2{tmp.u - this_pos.u, tmp.v - this_pos.v};
      this_wall = target;
      continue;
    }
    // no neighbouring wall: reflect off the edge
    const Wall& w = walls[this_wall];
    V2 nd{this_disp.u - (boundary.u - old_pos.u), this_disp.v - (boundary.v - old_pos.v)};
    if (edge == 0) {
      nd.v *= -1.0;
    } else {
      V2 r = edge == 1 ? V2{-w.uv2_v, w.uv2_u - w.uv_vert1_u} : V2{w.uv2_v, -w.uv2_u};
      double f = 1.0 / std::sqrt(len2_squared(r));
      r.u *= f;
      r.v *= f;
      f = 2.0 * dot2(nd, r);
      nd.u -= f * r.u;
      nd.v -= f * r.v;
    }
    this_pos = boundary;
    this_disp = nd;
  }
}

// diffuse_surf_molecule (diffuse_react_event.cpp:1071-1246) for a species
// that can diffuse: up to 1 + SURFACE_DIFFUSION_RETRIES displacements
// (pick_surf_displacement) until one ends on a free tile (or the same one).
void World::diffuse_surf_molecule(uint32_t id, double& max_time) {
  if (edges.size() != walls.size()) build_surface_net();
  const Species& sp = species[mols[id].species];
  double steps = 0.0, t_steps = 0.0;
  t_steps = sp.time_step > max_time ? max_time : sp.time_step;
  if (sp.space_step > 0) {
    if (sp.time_step > max_time) {
      steps = max_time / sp.time_step;
      if (steps < EPS) {
        t_steps = EPS * sp.time_step;
        steps = EPS;
      }
    } else {
      steps = 1.0;
    }
    const double space_factor = steps == 1.0 ? sp.space_step : sp.space_step * std::sqrt(steps);
    DrawSource& r = rng();
    for (int attempt = kSurfaceDiffusionRetries + 1; attempt > 0; --attempt) {
      // pick_surf_displacement (diffusion_utils.inl:60-93)
      const double one_over_2_to_16th = 1.52587890625e-5;
      double au, av, f;
      do {
        const uint32_t n = r.u32();
        au = 2 * one_over_2_to_16th * (n & 0xFFFF) - 1;
        av = 2 * one_over_2_to_16th * (n >> 16) - 1;
        f = au * au + av * av;
      } while ((f < POS_EPS) || (f > 1));
      const double normal_factor = std::sqrt(-std::log(f) / f);
      const double s = normal_factor * space_factor;
      const double du = au * s, dv = av * s;
      double nu = 0, nv = 0;
      const uint32_t nw = ray_trace_surf(id, du, dv, nu, nv);
      if (nw == INVALID_U32) {
        ++surf_retries;
        continue;
      }
      Mol& m = mols[id];
      Grid& g = grids[m.wall];
      if (nw == m.wall) {
        // move_sm_on_same_triangle (diffusion_utils.inl:453-489)
        const uint32_t nt = uv2grid(nu, nv, nw);
        if (nt != m.tile) {
          if (g.tile_mol[nt] != INVALID_U32) {
            ++surf_retries;
            continue;
          }
          g.tile_mol[m.tile] = INVALID_U32;
          g.tile_mol[nt] = id;
          m.tile = nt;
        }
        m.u = nu;
        m.v = nv;
      } else {
        // move_sm_to_new_triangle (diffusion_utils.inl:505-548)
        init_grid(nw);
        Grid& ng = grids[nw];
        Grid& og = grids[m.wall];
        const uint32_t nt = uv2grid(nu, nv, nw);
        if (ng.tile_mol[nt] != INVALID_U32) {
          ++surf_retries;
          continue;
        }
        ++surf_moves_between_walls;
        og.tile_mol[m.tile] = INVALID_U32;
        ng.tile_mol[nt] = id;
        m.tile = nt;
        m.wall = nw;
        m.u = nu;
        m.v = nv;
      }
      break;
    }
  }
  max_time = t_steps;
}

}  // namespace orc
