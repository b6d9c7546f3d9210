// TEST INFRASTRUCTURE ONLY (CPU oracle; see oracle_core.h).
// A C interface over the oracle so that tests (ctypes) and the CPU-baseline
// leg of bench.py can drive it.
#include <cstring>
#include <iomanip>
#include <sstream>
#include <chrono>
#include <stdexcept>
#include <string>

#include "oracle_core.h"

using namespace orc;

static thread_local std::string g_err;

extern "C" {

const char* orc_error(void) { return g_err.c_str(); }

void* orc_create(double time_step, double grid_density, double interaction_radius, double partition_dim,
                 double subpart_dim, int has_bimol, const double* bbox_um) {
  UserConfig u;
  u.time_step = time_step;
  u.surface_grid_density = grid_density;
  u.interaction_radius = interaction_radius;
  u.partition_dimension = partition_dim;
  u.subpartition_dimension = subpart_dim;
  u.has_bimol_vol_rxns = has_bimol;
  Config c = derive_config(u, bbox_um);
  return new World(c);
}

void orc_destroy(void* w) { delete static_cast<World*>(w); }

// out_d: time_unit, length_unit, rcp_length_unit, grid_density, rxn_radius_3d,
// partition_edge_length, origin xyz, opposite xyz, subpart_edge_length, rcp
// out_u: n, n2, n3, use_expanded_list
void orc_get_config(void* wp, double* out_d, uint32_t* out_u) {
  const Config& c = static_cast<World*>(wp)->cfg;
  double d[] = {c.time_unit, c.length_unit, c.rcp_length_unit, c.grid_density, c.rxn_radius_3d,
                c.partition_edge_length, c.origin.x, c.origin.y, c.origin.z, c.opposite.x, c.opposite.y,
                c.opposite.z, c.subpart_edge_length, c.subpart_edge_length_rcp};
  for (int i = 0; i < 14; ++i) out_d[i] = d[i];
  out_u[0] = c.n; out_u[1] = c.n2; out_u[2] = c.n3; out_u[3] = (uint32_t)c.use_expanded_list;
}

uint32_t orc_add_vertices_um(void* wp, const double* xyz, uint32_t n) {
  World* w = static_cast<World*>(wp);
  uint32_t first = (uint32_t)w->verts.size();
  for (uint32_t i = 0; i < n; ++i) w->add_vertex_um({xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]});
  return first;
}

void orc_add_walls(void* wp, const uint32_t* tri, uint32_t n, uint32_t vert_offset, uint32_t object_id,
                   uint32_t surf_class) {
  World* w = static_cast<World*>(wp);
  for (uint32_t i = 0; i < n; ++i)
    w->add_wall(vert_offset + tri[3 * i], vert_offset + tri[3 * i + 1], vert_offset + tri[3 * i + 2], object_id,
                surf_class);
}

void orc_finalize_walls(void* wp) { static_cast<World*>(wp)->finalize_walls(); }

// Surface classes 1..n: action[(c - 1) * num_species + s] = front | back << 2
// (0 reflect, 1 transparent, 2 absorptive).  Species with any non-reflect
// action get SPF_CAN_VOLWALL.  Call after the species are added.
void orc_set_surface_classes(void* wp, uint32_t n, const uint8_t* action) {
  World* w = static_cast<World*>(wp);
  size_t ns = w->species.size();
  w->num_surface_classes = n;
  w->surf_action.assign(action, action + n * ns);
  for (size_t s = 0; s < ns; ++s)
    for (uint32_t c = 0; c < n; ++c)
      if (action[c * ns + s] != 0) w->species[s].flags |= SPF_CAN_VOLWALL;
}

// Sets the surface class of walls [first, first + n).
void orc_set_wall_classes(void* wp, uint32_t first, uint32_t n, uint32_t cls) {
  World* w = static_cast<World*>(wp);
  for (uint32_t i = first; i < first + n && i < w->walls.size(); ++i) w->walls[i].surf_class = cls;
}

// transparent crossings, absorbed molecules
void orc_get_wall_stats(void* wp, uint64_t* out) {
  World* w = static_cast<World*>(wp);
  out[0] = w->transparent_crossings;
  out[1] = w->absorbed;
}

uint32_t orc_num_walls(void* wp) { return (uint32_t)static_cast<World*>(wp)->walls.size(); }

// 16 doubles per wall: normal xyz, unit_u xyz, unit_v xyz, d, uv_vert1_u,
// uv2_u, uv2_v, area, 2 spare
void orc_get_walls(void* wp, double* out) {
  World* w = static_cast<World*>(wp);
  for (size_t i = 0; i < w->walls.size(); ++i) {
    const Wall& x = w->walls[i];
    double* o = out + 16 * i;
    o[0] = x.normal.x; o[1] = x.normal.y; o[2] = x.normal.z;
    o[3] = x.unit_u.x; o[4] = x.unit_u.y; o[5] = x.unit_u.z;
    o[6] = x.unit_v.x; o[7] = x.unit_v.y; o[8] = x.unit_v.z;
    o[9] = x.d; o[10] = x.uv_vert1_u; o[11] = x.uv2_u; o[12] = x.uv2_v; o[13] = x.area;
    o[14] = 0; o[15] = 0;
  }
}

// vertex indices of every wall (3 per wall)
void orc_get_wall_vertices(void* wp, uint32_t* out) {
  World* w = static_cast<World*>(wp);
  for (size_t i = 0; i < w->walls.size(); ++i)
    for (int k = 0; k < 3; ++k) out[3 * i + k] = w->walls[i].v[k];
}

void orc_get_vertices(void* wp, double* out) {
  World* w = static_cast<World*>(wp);
  for (size_t i = 0; i < w->verts.size(); ++i) {
    out[3 * i] = w->verts[i].x; out[3 * i + 1] = w->verts[i].y; out[3 * i + 2] = w->verts[i].z;
  }
}

// CSR of walls per subpartition: off has n3 + 1 entries; returns #entries
uint64_t orc_get_subpart_walls(void* wp, uint32_t* off, uint32_t* idx) {
  World* w = static_cast<World*>(wp);
  if (off) for (size_t i = 0; i < w->sp_wall_off.size(); ++i) off[i] = w->sp_wall_off[i];
  if (idx) for (size_t i = 0; i < w->sp_wall_idx.size(); ++i) idx[i] = w->sp_wall_idx[i];
  return w->sp_wall_idx.size();
}

uint32_t orc_add_species(void* wp, double D, int is_vol, uint32_t flags, double custom_time_step) {
  return static_cast<World*>(wp)->add_species(D, is_vol != 0, flags, custom_time_step);
}

// out: D, space_step, time_step, flags
void orc_get_species(void* wp, uint32_t s, double* out) {
  const Species& sp = static_cast<World*>(wp)->species[s];
  out[0] = sp.D; out[1] = sp.space_step; out[2] = sp.time_step; out[3] = (double)sp.flags;
}

int orc_add_bimol_rxn(void* wp, uint32_t a, uint32_t b, double rate, const uint32_t* products, uint32_t nprod) {
  std::vector<uint32_t> p(products, products + nprod);
  return static_cast<World*>(wp)->add_bimol_rxn(a, b, rate, p);
}

// out: pb_factor, max_fixed_p
void orc_add_unimol_rxn(void* wp, uint32_t species, double rate, const uint32_t* products, uint32_t nprod) {
  std::vector<uint32_t> p(products, products + nprod);
  static_cast<World*>(wp)->add_unimol_rxn(species, rate, p);
}

uint64_t orc_unimol_reactions(void* wp) { return static_cast<World*>(wp)->unimol_reactions; }

// products: n entries of species, orientation (+1, -1, 0) and surface flag
int orc_add_volsurf_rxn(void* wp, uint32_t vol, uint32_t surf, int geom_vol, int geom_surf, double rate, uint32_t n,
                        const uint32_t* species, const int32_t* orients, const uint32_t* is_surf) {
  std::vector<VolSurfRxn::Product> prods;
  for (uint32_t q = 0; q < n; ++q) prods.push_back({species[q], orients[q], is_surf[q] != 0});
  return static_cast<World*>(wp)->add_volsurf_rxn(vol, surf, geom_vol, geom_surf, rate, prods);
}

// out: pb_factor, max_fixed_p
void orc_get_volsurf_rxn(void* wp, int r, double* out) {
  const VolSurfRxn& x = static_cast<World*>(wp)->vs_rxns[r];
  out[0] = x.pb_factor;
  out[1] = x.max_fixed_p;
}

// uv: 2 doubles per molecule (its position on the wall), or null for the
// tiles' centres
int orc_add_surface_molecules(void* wp, uint32_t species, uint32_t n, const uint32_t* walls, const uint32_t* tiles,
                              const int32_t* orients, const double* uv) {
  World* w = static_cast<World*>(wp);
  try {
    for (uint32_t i = 0; i < n; ++i)
      w->add_surface_molecule(species, walls[i], tiles[i], orients[i], uv ? uv + 2 * i : nullptr);
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return 0;
}

// Grid of a wall (initialized on demand): out = n_axis, num_tiles,
// strip_width_rcp, vert2_slope, fullslope, binding_factor, vert0_u, vert0_v
void orc_grid_info(void* wp, uint32_t wall, double* out) {
  World* w = static_cast<World*>(wp);
  w->init_grid(wall);
  const Grid& g = w->grids[wall];
  double v[8] = {(double)g.n_axis, (double)g.num_tiles, g.strip_width_rcp, g.vert2_slope, g.fullslope,
                 g.binding_factor, g.vert0_u, g.vert0_v};
  std::memcpy(out, v, sizeof(v));
}

int64_t orc_xyz2grid(void* wp, uint32_t wall, const double* pos) {
  World* w = static_cast<World*>(wp);
  w->init_grid(wall);
  try {
    return (int64_t)w->xyz2grid({pos[0], pos[1], pos[2]}, wall);
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// surface molecules per tile of a wall: out[k] = molecule id or 0xFFFFFFFF
void orc_get_tiles(void* wp, uint32_t wall, uint32_t* out) {
  World* w = static_cast<World*>(wp);
  if (wall >= w->grids.size()) return;
  const Grid& g = w->grids[wall];
  for (uint32_t k = 0; k < g.num_tiles; ++k) out[k] = g.tile_mol[k];
}

uint64_t orc_volsurf_reactions(void* wp) { return static_cast<World*>(wp)->volsurf_reactions; }

// surface molecules' wall, tile and uv position (indexed by id; volume
// molecules get 0xFFFFFFFF walls)
void orc_get_surface_state(void* wp, uint32_t* walls, uint32_t* tiles, double* uv) {
  World* w = static_cast<World*>(wp);
  for (size_t i = 0; i < w->mols.size(); ++i) {
    const Mol& m = w->mols[i];
    const bool s = (m.flags & MF_SURFACE) != 0;
    walls[i] = s ? m.wall : 0xFFFFFFFFu;
    tiles[i] = s ? m.tile : 0xFFFFFFFFu;
    uv[2 * i] = s ? m.u : 0.0;
    uv[2 * i + 1] = s ? m.v : 0.0;
  }
}

// out: moves between walls, retried displacements
void orc_surface_stats(void* wp, uint64_t* out) {
  World* w = static_cast<World*>(wp);
  out[0] = w->surf_moves_between_walls;
  out[1] = w->surf_retries;
}

// Per wall and edge: the neighbouring wall (0xFFFFFFFF: none) and the edge
// transform (cos, sin, translate u, v; forward flag) of the surface net.
int orc_surface_net(void* wp, uint32_t* nb, double* tr) {
  World* w = static_cast<World*>(wp);
  try {
    w->build_surface_net();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
  for (size_t i = 0; i < w->walls.size(); ++i)
    for (int k = 0; k < 3; ++k) {
      const WallEdge& e = w->edges[i][k];
      nb[3 * i + k] = e.fwd == i ? e.bwd : e.fwd;
      double* t = tr + 5 * (3 * i + k);
      t[0] = e.cos_t; t[1] = e.sin_t; t[2] = e.tu; t[3] = e.tv; t[4] = e.fwd == i ? 1.0 : 0.0;
    }
  return 0;
}

// find_edge_point and traverse_surface for unit tests: out = edge, then
// the edge point (u, v), then the wall across and the point in its frame
void orc_edge_point(void* wp, uint32_t wall, const double* loc, const double* disp, double* out) {
  World* w = static_cast<World*>(wp);
  if (w->edges.size() != w->walls.size()) w->build_surface_net();
  double eu = 0, ev = 0;
  const int e = w->find_edge_point(wall, loc[0], loc[1], disp[0], disp[1], eu, ev);
  out[0] = e; out[1] = eu; out[2] = ev;
  out[3] = -1; out[4] = 0; out[5] = 0;
  if (e < 3) {
    double nu = 0, nv = 0;
    const uint32_t t = w->traverse_surface(wall, eu, ev, e, nu, nv);
    out[3] = t == 0xFFFFFFFFu ? -1.0 : (double)t;
    out[4] = nu; out[5] = nv;
  }
}

uint32_t orc_uv2grid(void* wp, uint32_t wall, double u, double v) {
  World* w = static_cast<World*>(wp);
  w->init_grid(wall);
  return w->uv2grid(u, v, wall);
}

void orc_grid2uv(void* wp, uint32_t wall, uint32_t tile, double* out) {
  World* w = static_cast<World*>(wp);
  w->init_grid(wall);
  w->grid2uv(wall, tile, out[0], out[1]);
}

void orc_get_rxn(void* wp, int r, double* out) {
  const BimolRxn& x = static_cast<World*>(wp)->rxns[r];
  out[0] = x.pb_factor; out[1] = x.max_fixed_p;
}

int orc_add_molecules(void* wp, uint32_t species, uint32_t n, const double* pos_lu) {
  World* w = static_cast<World*>(wp);
  try {
    for (uint32_t i = 0; i < n; ++i)
      w->add_volume_molecule(species, {pos_lu[3 * i], pos_lu[3 * i + 1], pos_lu[3 * i + 2]}, (double)w->iteration);
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return 0;
}

void orc_set_rng(void* wp, int mode, uint64_t seed, uint32_t replica) {
  World* w = static_cast<World*>(wp);
  if (mode == DRAW_ISAAC) w->set_rng_isaac((uint32_t)seed);
  else w->set_rng_philox(seed, replica);
}

void orc_set_snapshot(void* wp, int on) { static_cast<World*>(wp)->snapshot = on != 0; }
void orc_set_trace(void* wp, int on) { static_cast<World*>(wp)->trace = on != 0; }
void orc_set_barrier_interval(void* wp, uint32_t b) { static_cast<World*>(wp)->barrier_interval = b; }
void orc_set_barriers(void* wp, const uint64_t* b, uint32_t n) {
  static_cast<World*>(wp)->barrier_list.assign(b, b + n);
}

// reaction records: u32 x4 (iteration, mover, partner, pathway) + rxn index,
// f64 x4 (time, pos xyz)
uint64_t orc_get_rxn_log(void* wp, uint32_t* u, int32_t* rx, double* d) {
  World* w = static_cast<World*>(wp);
  if (u)
    for (size_t i = 0; i < w->rxn_log.size(); ++i) {
      const RxnRecord& r = w->rxn_log[i];
      u[4 * i] = r.iteration; u[4 * i + 1] = r.mover_id; u[4 * i + 2] = r.partner_id; u[4 * i + 3] = (uint32_t)r.pathway;
      rx[i] = r.rxn;
      d[4 * i] = r.time; d[4 * i + 1] = r.pos.x; d[4 * i + 2] = r.pos.y; d[4 * i + 3] = r.pos.z;
    }
  return w->rxn_log.size();
}
void orc_clear_rxn_log(void* wp) { static_cast<World*>(wp)->rxn_log.clear(); }

// exact_disk for a collision at loc of a mover moving by mv with target id
double orc_exact_disk(void* wp, const double* loc, const double* mv, uint32_t target) {
  World* w = static_cast<World*>(wp);
  return w->exact_disk({loc[0], loc[1], loc[2]}, {mv[0], mv[1], mv[2]}, w->cfg.rxn_radius_3d, w->mols[target],
                       w->mols[target]);
}

// Runs n iterations; returns 0, or -1 with orc_error() set.  *seconds gets
// the wall-clock time of the iteration loop alone.
int orc_run(void* wp, uint32_t n, double* seconds) {
  World* w = static_cast<World*>(wp);
  auto t0 = std::chrono::steady_clock::now();
  try {
    for (uint32_t i = 0; i < n; ++i) w->run_iteration();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return 0;
}

uint64_t orc_iteration(void* wp) { return static_cast<World*>(wp)->iteration; }
uint32_t orc_num_molecules(void* wp) { return (uint32_t)static_cast<World*>(wp)->mols.size(); }

// counted volumes: per wall inside / outside index (INVALID_U32 for both:
// not a counted object)
void orc_set_counted_volumes(void* wp, uint32_t num_walls, const uint32_t* inside, const uint32_t* outside) {
  World* w = static_cast<World*>(wp);
  w->cv_in.assign(inside, inside + num_walls);
  w->cv_out.assign(outside, outside + num_walls);
}
void orc_set_molecule_volumes(void* wp, uint32_t n, const uint32_t* ids, const uint32_t* cvols) {
  World* w = static_cast<World*>(wp);
  for (uint32_t i = 0; i < n; ++i) w->mols[ids[i]].cvol = cvols[i];
}
void orc_get_molecule_volumes(void* wp, uint32_t* out) {
  World* w = static_cast<World*>(wp);
  for (size_t i = 0; i < w->mols.size(); ++i) out[i] = w->mols[i].cvol;
}

void orc_get_molecules(void* wp, uint32_t* species, uint32_t* flags, double* pos, uint32_t* subpart,
                       double* diffusion_time) {
  World* w = static_cast<World*>(wp);
  for (size_t i = 0; i < w->mols.size(); ++i) {
    const Mol& m = w->mols[i];
    if (species) species[i] = m.species;
    if (flags) flags[i] = m.flags;
    if (pos) { pos[3 * i] = m.pos.x; pos[3 * i + 1] = m.pos.y; pos[3 * i + 2] = m.pos.z; }
    if (subpart) subpart[i] = m.subpart;
    if (diffusion_time) diffusion_time[i] = m.diffusion_time;
  }
}

void orc_get_birthdays(void* wp, double* out) {
  World* w = static_cast<World*>(wp);
  for (size_t i = 0; i < w->mols.size(); ++i) out[i] = w->mols[i].birthday;
}

// hit records: u32 x5 (mol_id, iteration, wall, type, subpart_after), time
uint64_t orc_get_hits(void* wp, uint32_t* u, double* t) {
  World* w = static_cast<World*>(wp);
  if (u)
    for (size_t i = 0; i < w->hits.size(); ++i) {
      const HitRecord& h = w->hits[i];
      u[5 * i] = h.mol_id; u[5 * i + 1] = h.iteration; u[5 * i + 2] = h.wall; u[5 * i + 3] = h.type;
      u[5 * i + 4] = h.subpart_after;
      if (t) t[i] = h.time;
    }
  return w->hits.size();
}

void orc_clear_hits(void* wp) { static_cast<World*>(wp)->hits.clear(); }

// ray_voxel_tests, ray_polygon_tests, ray_polygon_colls, mol_wall_reflections,
// vol_mol_vol_mol_collisions, diffuse_3d_calls, redo_count, rng draws, partition clips
void orc_get_stats(void* wp, uint64_t* out) {
  World* w = static_cast<World*>(wp);
  out[0] = w->ray_voxel_tests; out[1] = w->ray_polygon_tests; out[2] = w->ray_polygon_colls;
  out[3] = w->mol_wall_reflections; out[4] = w->vol_mol_vol_mol_collisions; out[5] = w->diffuse_3d_calls;
  out[6] = w->redo_count; out[7] = w->isaac.draws; out[8] = w->partition_clips;
}

// ------------------------------------------------------------ unit functions
void orc_isaac64_u32(uint32_t seed, uint64_t n, uint32_t* out) {
  Isaac64 s;
  s.init(seed);
  for (uint64_t i = 0; i < n; ++i) out[i] = s.u32();
}

void orc_gauss_isaac(uint32_t seed, uint64_t n, double* out) {
  IsaacSource s;
  s.s.init(seed);
  for (uint64_t i = 0; i < n; ++i) out[i] = rng_gauss(s);
}

void orc_philox_block(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { Philox::block(ctr, key, out); }

// Gaussians from a stream of externally given uint32 draws; returns the
// number of draws consumed (or -1 when the tape ran out).
int64_t orc_gauss_from_tape(const uint32_t* tape, uint64_t ntape, uint64_t n, double* out) {
  struct Tape : DrawSource {
    const uint32_t* t; uint64_t n, i = 0; bool over = false;
    uint32_t u32() override { if (i >= n) { over = true; return 0; } return t[i++]; }
  } tp;
  tp.t = tape;
  tp.n = ntape;
  for (uint64_t i = 0; i < n; ++i) {
    out[i] = rng_gauss(tp);
    if (tp.over) return -1;
  }
  return (int64_t)tp.i;
}

void orc_philox_stream(uint64_t seed, uint32_t id, uint32_t iteration, uint32_t replica, uint32_t pass, uint64_t n,
                       uint32_t* out) {
  PhiloxSource p;
  p.reset(seed, id, iteration, replica, pass);
  for (uint64_t i = 0; i < n; ++i) out[i] = p.u32();
}

// collide_wall without move update (REDO reported as 3, no draw consumed):
// returns the CollType; t and hit written when a hit is computed.
int orc_collide_wall(void* wp, const double* pos, uint32_t wall, const double* move, double* t, double* hit) {
  World* w = static_cast<World*>(wp);
  V3 mv{move[0], move[1], move[2]};
  V3 h{0, 0, 0};
  double tt = 0;
  CollType ct = w->collide_wall({pos[0], pos[1], pos[2]}, wall, false, mv, tt, h);
  *t = tt;
  hit[0] = h.x; hit[1] = h.y; hit[2] = h.z;
  return (int)ct;
}

// collide_wall (no move update) for n rays against one wall; types[i] gets
// the CollType (3 = REDO case, no draw consumed)
void orc_collide_wall_batch(void* wp, uint32_t wall, uint64_t n, const double* pos, const double* move,
                            int32_t* types) {
  World* w = static_cast<World*>(wp);
  for (uint64_t i = 0; i < n; ++i) {
    V3 mv{move[3 * i], move[3 * i + 1], move[3 * i + 2]};
    V3 h{0, 0, 0};
    double tt = 0;
    types[i] = (int32_t)w->collide_wall({pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}, wall, true, mv, tt, h);
  }
}

// subpartitions crossed by a move (walls list, ordered); returns count
uint32_t orc_crossed_subparts(void* wp, const double* pos, const double* disp, uint32_t* out, uint32_t cap) {
  World* w = static_cast<World*>(wp);
  Mol m{};
  m.pos = {pos[0], pos[1], pos[2]};
  m.subpart = w->subpart_index(m.pos);
  std::vector<uint32_t> ws, ms;
  w->collect_crossed_subparts(m, {disp[0], disp[1], disp[2]}, false, true, ws, ms);
  uint32_t k = 0;
  for (; k < ws.size() && k < cap; ++k) out[k] = ws[k];
  return (uint32_t)ws.size();
}

// collide_wall as the ray trace calls it (update_move: a REDO case returns
// 3 and draws its jitter sign from the world's stream) for n rays against
// one wall; t and hit are written for FRONT (1) and BACK (2).
void orc_collide_wall_rays(void* wp, uint32_t wall, uint64_t n, const double* pos, const double* move,
                           int32_t* types, double* t, double* hit) {
  World* w = static_cast<World*>(wp);
  for (uint64_t i = 0; i < n; ++i) {
    V3 mv{move[3 * i], move[3 * i + 1], move[3 * i + 2]};
    V3 h{0, 0, 0};
    double tt = 0;
    types[i] = (int32_t)w->collide_wall({pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}, wall, true, mv, tt, h);
    t[i] = tt;
    hit[3 * i] = h.x; hit[3 * i + 1] = h.y; hit[3 * i + 2] = h.z;
  }
}

// test_bimolecular for each of n given draws u[i]; out[i] = pathway or -1
void orc_test_bimolecular_u(double max_fixed_p, const double* cum_probs, uint32_t num_pathways, double scaling,
                            uint64_t n, const double* u, int32_t* out) {
  std::vector<double> cum(cum_probs, cum_probs + num_pathways);
  for (uint64_t i = 0; i < n; ++i) out[i] = test_bimolecular_u(max_fixed_p, cum, scaling, u[i]);
}

uint32_t orc_subpart_index(void* wp, const double* pos) {
  return static_cast<World*>(wp)->subpart_index({pos[0], pos[1], pos[2]});
}

// How C++ iostreams print a count (the checker of the count writer): form 0,
// a default stream (CountItem::write_as_dat, count_buffer.cpp:30-32); form 1,
// std::scientific at precision 8 (write_gdat_value's base, :35-61).
int orc_format_count(double d, int form, char* out, int cap) {
  std::ostringstream os;
  if (form == 1) os << std::scientific << std::setprecision(8);
  os << d;
  const std::string s = os.str();
  if ((int)s.size() >= cap) return -1;
  std::memcpy(out, s.c_str(), s.size() + 1);
  return (int)s.size();
}

}  // extern "C"
