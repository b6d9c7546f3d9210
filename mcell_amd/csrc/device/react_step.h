// One diffusion step of a volume molecule that can react with volume
// molecules: the loop of diffuse_vol_molecule (diffuse_react_event.cpp:
// 367-618) with ray_trace_vol's wall and molecule collisions, processed in
// time order, against the cell-sorted snapshot of the iteration's start.
// The first successful reaction test ends the molecule's step with a
// proposal; whether it happens is decided afterwards (one reactant can be
// claimed by several movers).
#pragma once
#include "displace.h"
#include "exact_disk.h"
#include "filter32.h"
#include "path32.h"
#include "light.h"
#include "react.h"

namespace mcg {

struct ReactParams {
  PartitionParams P;
  const WallRec64* __restrict__ walls;
  const uint32_t* __restrict__ off;
  const uint32_t* __restrict__ idx;
  const FilterWall* __restrict__ ent32;  // FP32 records in list order (FP32 form)
  const WallRec64* __restrict__ ent64;   // FP64 records in list order
  const uint32_t* __restrict__ ent_box;  // sub-cell box of every list entry
  CellGrid G;
  const RxnRec* __restrict__ rxns;
  uint32_t rxn_stride;                   // reaction classes per replica table (0: one shared table)
  const int32_t* __restrict__ pair_rxn;  // [ns * ns], -1 = none
  uint32_t num_species;
  uint64_t seed;
  uint32_t iteration;
  FilterParams F;
  float band;
  SurfaceClasses SC;
  const UnimolRec* __restrict__ unimol;  // per species (num_pathways 0: none)
  // surface molecules on wall grids
  const WallGrid* __restrict__ grids;    // per wall, or null without surface molecules
  const uint32_t* __restrict__ tiles;    // [replica * total_tiles + tile]: tile word or kTileEmpty
  const uint32_t* __restrict__ tile_id;
  const uint32_t* __restrict__ tile_consumed;
  uint32_t total_tiles;
  // per output slot: reactions already carried out in this step by a mover
  // that stayed (its class keeps it); the mover's repeated step passes
  // through that many successful tests
  const uint32_t* __restrict__ kept;
  const uint32_t* __restrict__ tile_wall;  // per tile (of one replica): its wall
  // counted volumes: per wall the index inside / outside its object
  // (kInvalid: not a counted object), or null
  const uint32_t* __restrict__ cv_in;
  const uint32_t* __restrict__ cv_out;
  // reaction counts (SimulationStats::inc_rxn_occurred,
  // Partition::inc_rxn_in_volume_occured_count): per (replica, counted
  // volume, class pathway slot); slots: bimolecular class * 4 + pathway,
  // then rc_unimol + species * 4 + pathway, then rc_vs + class * 4 + pathway
  unsigned long long* rxn_counts;
  uint32_t rc_slots, rc_unimol, rc_vs, rc_volumes;
  // per output slot: 1 + the tile slot a volume product was created on (it
  // does not react with that tile on its first hit there), or 0
  uint32_t* __restrict__ born;
  const VolSurfRec* __restrict__ vs_rxns;
  const int32_t* __restrict__ pair_vs;   // [vol species * ns + surf species], -1 = none
  unsigned long long* ctr;               // CTR_* statistics of the call
  // molecules whose step ended before the iteration's end without a due
  // unimolecular reaction (time steps below one iteration): their slots,
  // for the next phase's moves
  uint32_t* __restrict__ clist;
  uint32_t* __restrict__ ccount;
  // diffusion-time offsets of surface molecules per tile slot (null unless
  // a surface species diffuses): a surface product moves from its
  // reaction time on
  double* __restrict__ ttoff;
  double* __restrict__ tbirth;  // their birthdays (absolute), alongside ttoff
};

// xyz2grid_tile_index (grid_utils.inl:48-118) for a point on the wall.
__device__ __forceinline__ uint32_t xyz2grid(const WallRec64& w, const WallGrid& g, D3 v) {
  if (g.num_tiles == 1) return 0;
  const uint32_t tile_idx_mid = g.num_tiles - 2 * g.n_axis + 1;
  if (!exd::distinguishable_vec3(v, d3(w.p0x, w.p0y, w.p0z), POS_EPS)) return tile_idx_mid;
  if (!exd::distinguishable_vec3(v, d3(w.p1x, w.p1y, w.p1z), POS_EPS)) return g.num_tiles - 1;
  if (!exd::distinguishable_vec3(v, d3(w.p2x, w.p2y, w.p2z), POS_EPS)) return 0;
  const double i = dot3(v, d3(w.ux, w.uy, w.uz)) - g.vert0_u;
  const double j = dot3(v, d3(w.vx, w.vy, w.vz)) - g.vert0_v;
  const double striploc = j * g.strip_width_rcp;
  int strip = (int)striploc;
  const double striprem = striploc - strip;
  strip = (int)g.n_axis - strip - 1;
  const double u0 = j * g.vert2_slope;
  const double u1_u0 = w.uv1u - j * g.fullslope;
  const double stripeloc = ((i - u0) / u1_u0) * (strip + (1 - striprem));
  const int stripe = (int)stripeloc;
  const double striperem = stripeloc - stripe;
  const int flip = (striperem < 1 - striprem) ? 0 : 1;
  int idx = strip * strip + 2 * stripe + flip;
  if (idx < 0) idx = 0;
  if ((uint32_t)idx >= g.num_tiles) idx = (int)g.num_tiles - 1;  // the reference asserts this cannot happen
  return (uint32_t)idx;
}

// Molecule collisions processed in one ray segment before the step reports
// ERR_COLLISIONS (a runaway guard: a segment meets a few partners at most).
constexpr int kMaxCollisions = 4096;

// The first wall the segment pos + disp meets (ray_trace_vol's wall search,
// diffuse_react_event.cpp:713-734): the start subpartition's walls, then
// those of the subpartitions the segment crosses.  `skip` is the wall the
// segment starts on.
__device__ __forceinline__ bool segment_first_wall(const ReactParams& R, D3 pos, D3& disp, uint32_t skip,
                                                   PhiloxStream& rng, WallHit& hit, uint32_t& err) {
  const D3 end = {pos.x + disp.x, pos.y + disp.y, pos.z + disp.z};
  const D3 part_disp = in_partition(R.P, end) ? disp : displacement_up_to_partition_boundary(R.P, pos, disp);
  int sx, sy, sz;
  subpart_3d(R.P, pos, sx, sy, sz);
  if (sx < 0 || sy < 0 || sz < 0 || sx >= (int)R.P.n || sy >= (int)R.P.n || sz >= (int)R.P.n) {
    err |= ERR_ESCAPED;
    return false;
  }
  const uint32_t start = subpart_of(R.P, sx, sy, sz);
  if (closest_wall_in_subpart64(R.P, R.walls, R.off, R.idx, R.ent_box, start, skip, pos, disp, rng, hit, err))
    return true;
  const D3 dst = {pos.x + part_disp.x, pos.y + part_disp.y, pos.z + part_disp.z};
  int dx, dy, dz;
  subpart_3d(R.P, dst, dx, dy, dz);
  if (subpart_of(R.P, dx, dy, dz) == start) return false;
  SubpartWalk64 walk;
  walk.init(R.P, pos, start, part_disp);
  while (walk.next(R.P, err))
    if (closest_wall_in_subpart64(R.P, R.walls, R.off, R.idx, R.ent_box, walk.cur, skip, pos, disp, rng, hit, err))
      return true;
  return false;
}

// The stream pass of the wall search that places a volume product (its
// edge re-draws): Philox(seed, product id, iteration, replica, kAuxPass).
constexpr uint32_t kAuxPass = 0xFFu;

// Where a volume product of a volume-surface reaction starts
// (update_vol_mol_after_rxn_with_surf_mol, diffuse_react_event.cpp:2291-
// 2320, with tiny_diffuse_3D, diffusion_utils.inl:551-595): the hit point p
// on wall wi moved by 2 * 16 POS_EPS along the wall normal to the side
// `orient` names, or half the way to the first other wall on that move.
// rng only serves the wall search's edge re-draws.
__device__ __forceinline__ D3 volume_product_position(const ReactParams& R, D3 p, uint32_t wi, int orient,
                                                      PhiloxStream& rng, uint32_t& err) {
  const WallRec64& w = R.walls[wi];
  const double bump = orient > 0 ? 16.0 * POS_EPS : -16.0 * POS_EPS;
  const D3 d0 = {2.0 * bump * w.nx, 2.0 * bump * w.ny, 2.0 * bump * w.nz};
  D3 disp = d0;
  WallHit hit;
  if (segment_first_wall(R, p, disp, wi, rng, hit, err) && (hit.t <= 1.0 || fabs(hit.t - 1.0) < 1e-12)) {
    const double h = hit.t * 0.5;
    return {p.x + d0.x * h, p.y + d0.y * h, p.z + d0.z * h};
  }
  return {p.x + d0.x, p.y + d0.y, p.z + d0.z};
}

struct Proposal {
  double time;    // absolute time of the collision
  double x, y, z; // collision point
  uint32_t mover_slot, partner_slot;
  uint32_t mover_id, partner_id;
  uint32_t rxn, pathway;
  uint32_t replica;
  uint32_t pad;   // walls reflected off (16 bits) | crossed (15 bits) << 16 | kPadMoverIsB
  uint32_t cvol;  // the mover's counted volume at the reaction
  uint32_t pad2;
};
constexpr uint32_t kPadMoverIsB = 1u << 31;  // the mover is the class's reactant b

// Closest wall in one subpartition (FP64; the entry-box pre-test first).
__device__ __forceinline__ bool closest_wall(const ReactParams& R, uint32_t s, uint32_t last_hit, D3 pos, D3& disp,
                                             PhiloxStream& rng, WallHit& hit, uint32_t& err) {
  return closest_wall_in_subpart64(R.P, R.walls, R.off, R.idx, R.ent_box, s, last_hit, pos, disp, rng, hit, err);
}

// The snapshot range of species t's molecules in the k-th subpartition of the
// set (empty outside the grid: no molecules there).
__device__ __forceinline__ void set_range(const CellGrid& G, const SubpartSet& set, int k, uint32_t replica,
                                          uint32_t t, uint32_t& j0, uint32_t& j1) {
  const uint32_t s = set.at(k);
  const int gx = (int)(s & 1023u), gy = (int)((s >> 10) & 1023u), gz = (int)(s >> 20);
  uint32_t c;
  if (!cell_of(G, gx, gy, gz, c)) {
    j0 = j1 = 0;
    return;
  }
  const uint32_t b = bin_of(G, replica, t, c);
  j0 = G.start[b];
  j1 = G.start[b + 1];
}

// The next molecule collision after key (t_last, id_last) in time order
// (ties: larger partner id first, diffuse_react_event.cpp:354-358), strictly
// before t_limit.  Scans the partner lists of the collected subpartitions.
template <typename PosT>
__device__ __forceinline__ bool next_collision(const ReactParams& R, const Snapshot<PosT>& S, const SubpartSet& set,
                                               uint32_t replica, uint32_t my_species, uint32_t my_id, D3 pos,
                                               D3 disp, double t_limit, double t_last, uint32_t id_last,
                                               bool first, double& t_best, uint32_t& slot_best, D3& pos_best,
                                               int32_t& rxn_best, uint32_t& candidates) {
  const double sigma2 = R.P.rxn_radius * R.P.rxn_radius;
  bool found = false;
  uint32_t id_best = 0;
  candidates = 0;
  // partners farther than the radius from the segment's box cannot collide
  const double R1 = R.P.rxn_radius;
  const double ex = pos.x + disp.x, ey = pos.y + disp.y, ez = pos.z + disp.z;
  const PartnerBox<PosT> box((pos.x < ex ? pos.x : ex) - R1, (pos.x < ex ? ex : pos.x) + R1,
                             (pos.y < ey ? pos.y : ey) - R1, (pos.y < ey ? ey : pos.y) + R1,
                             (pos.z < ez ? pos.z : ez) - R1, (pos.z < ez ? ez : pos.z) + R1);
  // species by species; the next subpartition's range is loaded while the
  // current one's partners are tested, and partner positions two at a time
  // (the result does not depend on the order: the key is a total order)
  constexpr int kGroup = 2;
  for (uint32_t t = 0; t < R.G.ns && set.n > 0; ++t) {
    const int32_t rx = R.pair_rxn[my_species * R.num_species + t];
    if (rx < 0) continue;
    uint32_t n0, n1;
    set_range(R.G, set, 0, replica, t, n0, n1);
    for (int k = 0; k < set.n; ++k) {
      const uint32_t j0 = n0, j1 = n1;
      if (k + 1 < set.n) set_range(R.G, set, k + 1, replica, t, n0, n1);
      for (uint32_t jb = j0; jb < j1; jb += kGroup) {
        PosT gx[kGroup], gy[kGroup], gz[kGroup];
#pragma unroll
        for (int u = 0; u < kGroup; ++u) {
          const uint32_t j = jb + u < j1 ? jb + u : j1 - 1;
          gx[u] = S.x[j];
          gy[u] = S.y[j];
          gz[u] = S.z[j];
        }
#pragma unroll
        for (int u = 0; u < kGroup; ++u) {
          const uint32_t j = jb + u;
          if (j >= j1) break;
          if (box.outside(gx[u], gy[u], gz[u])) continue;
          double tt;
          D3 cp;
          if (!collide_mol64(pos, disp, D3{(double)gx[u], (double)gy[u], (double)gz[u]}, sigma2, tt, cp)) continue;
          if (!(tt < t_limit)) continue;
          const uint32_t pid = S.id[j];
          // consumed partners are gone, except those consumed by this
          // mover's own earlier reactions of the step (it stayed and
          // repeats the step up to there: consumed holds consumer id + 1)
          const uint32_t cj = S.consumed[j];
          if (pid == my_id || (cj != 0 && cj != my_id + 1)) continue;
          // strictly after the last processed key
          if (!first && !(tt > t_last || (tt == t_last && pid < id_last))) continue;
          ++candidates;
          if (!found || tt < t_best || (tt == t_best && pid > id_best)) {
            found = true;
            t_best = tt;
            id_best = pid;
            slot_best = j;
            pos_best = cp;
            rxn_best = rx;
          }
        }
      }
    }
  }
  return found;
}

__device__ __forceinline__ bool cmp_eq_d(double a, double b, double eps) { return fabs(a - b) < eps; }
__device__ __forceinline__ bool cmp_lt_d(double a, double b, double eps) { return a < b && !cmp_eq_d(a, b, eps); }

// Result of one mover's step.
struct MoverEnd {
  D3 pos;          // end position; with a proposal, the start of the segment it was found in
  uint32_t n_reflect, n_cross;  // walls reflected off and crossed (statistics)
  double t_steps;  // time the step covered (units of the time step), before reflections
  double utime;    // unimolecular reaction time after the step (kTimeInvalid: none)
  bool near_plane; // FP32 form: the end point lies near a wall plane
  bool absorbed;   // destroyed by an absorptive surface class
  bool exd;        // deferred: a collision needs exact_disk (the caller repeats the step with it)
  uint32_t due;    // 0, or 1 + the stream pass that draws the unimolecular pathway
  uint32_t cvol;   // counted volume (in: at the start; out: at the end or the reaction)
  bool cont;       // the molecule diffuses again in this iteration (its next stream pass)
  bool mature;     // the step set MOLECULE_FLAG_MATURE
};

// exact_disk is exactly 1 unless an entry of the collision point's
// subpartition has a box overlapping the point's +-R cube (the walls it
// would process; exact_disk_dev returns 1 when it processes none).
__device__ __forceinline__ bool disk_meets_walls(const ReactParams& R, D3 cp) {
  const double r = R.P.rxn_radius;
  const uint32_t csp = subpart_index(R.P, cp);
  const uint32_t db = region_box(R.P, csp, cp.x - r, cp.y - r, cp.z - r, cp.x + r, cp.y + r, cp.z + r);
  const uint32_t e1 = R.off[csp + 1];
  return next_overlap(R.ent_box, R.off[csp], e1, db) < e1;
}

// The molecule's unimolecular time at the start of its step
// (diffuse_single_molecule, diffuse_react_event.cpp:213-241): a newly created
// molecule draws its lifetime first (time_of_unimol, rxn_utils.inl:720-735).
__device__ __forceinline__ double start_unimol_time(const UnimolRec* __restrict__ unimol, uint32_t species,
                                                    uint32_t tag, double utime, double start_time,
                                                    PhiloxStream& rng) {
  const UnimolRec& u = unimol[species];
  if (u.num_pathways == 0) return kTimeInvalid;
  if (!(tag_flags(tag) & MF_SCHEDULE_UNIMOL)) return utime;
  const double k_tot = u.cum[u.num_pathways - 1];
  const double p = rng.dbl();
  if (k_tot <= 0 || !distinguishable(p, 0.0, 1e-12)) return start_time + kTimeForever;
  return start_time + (-log(p) / k_tot);
}

// The step of one mover (diffuse_vol_molecule, diffuse_react_event.cpp:367-
// 618) against the snapshot: displacement clipped to max_time = barrier -
// toff, ray segments with wall reflections, and for species with volume
// reactions the molecule collisions of each segment in time order, each
// reaction test drawing from the mover's stream.  The first successful test
// ends the step with a proposal (return true).
// kPoolN is the vertex pool of exact_disk in this form.  0: a collision
// whose disk meets walls stops the step with out.exd set (nothing written)
// and a later pass repeats the step with a pool; 0 < kPoolN < exd::kPool:
// exact_disk with kPoolN records, and a disk that overflows them stops the
// step the same way; exd::kPool: the last pass, where running out is an
// error.  The pool is private memory per lane, so the small forms keep the
// scratch of the kernels that run them small.
//
// The molecule's stream is pass n of its n-th step in the iteration (the
// tag's pass bits; MF_SKIP_DRAW: the pathway draw of a unimolecular
// reaction that kept it was the pass's first draw).  birth: the molecule's
// birthday (read only while a species with a time step above one iteration
// has not matured).  pass_through: the
// step's successful bimolecular tests already carried out (the mover
// stayed); they are passed as the reference's ray trace goes on after them
// (collide_and_react_with_vol_mol returns false for a kept mover,
// diffuse_react_event.cpp:463-470, 823-827).
template <typename PosT, int kPoolN>
__device__ __forceinline__ bool react_mover(const ReactParams& R, const Snapshot<PosT>& S, const SpeciesRec& sp,
                                            const ZigguratTables& T, D3 pos, uint32_t idv, uint32_t tag, double toff,
                                            double utime, double birth, double barrier, uint32_t pass_through,
                                            uint32_t born_tile, Proposal& prop, MoverEnd& out, uint32_t& err,
                                            uint32_t* set_lds) {
  // The full step tests its candidate walls in FP64 in both forms: after the
  // entry-box pre-test a mover here meets few walls, and the FP32 filter in
  // front of them cost more than it saved.  In the FP32 form the end point
  // is then always rounded to FP32 on its FP64 side of nearby wall planes.
  constexpr bool kF32 = sizeof(PosT) == 4;
  const uint32_t rep = tag_replica(tag), spc = tag_species(tag);
  const uint32_t pass = (tag_flags(tag) & MF_PASS_MASK) >> MF_PASS_SHIFT;
  PhiloxStream rng;
  rng.init(R.seed, idv, R.iteration, rep, pass);
  if (tag_flags(tag) & MF_SKIP_DRAW) (void)rng.u32();
  uint32_t succ = 0;
  double elapsed = (double)R.iteration + toff;
  const double t_start = elapsed;
  out.near_plane = kF32;
  out.absorbed = false;
  out.exd = false;
  out.cont = false;
  out.mature = false;
  out.due = 0;
  out.t_steps = 0;
  out.n_reflect = 0;
  out.n_cross = 0;
  out.pos = pos;
  out.utime = start_unimol_time(R.unimol, spc, tag, utime, elapsed, rng);
  if (out.utime != kTimeInvalid && !(tag_flags(tag) & MF_SCHEDULE_UNIMOL) && out.utime <= elapsed) {
    out.due = pass + 1;  // due at the start: react_unimol_single_molecule before any move
    return false;
  }
  // get_max_time (diffuse_react_event.cpp:164-198)
  double max_time = barrier - toff;
  if (out.utime != kTimeInvalid && out.utime < elapsed + max_time) max_time = out.utime - elapsed;
  if (!(tag_flags(tag) & MF_MATURE) && sp.time_step > 1.0) {
    // a new molecule with a long time step lengthens its steps gradually
    const double age_dependent_max_time = 1.0 + 0.2 * (elapsed - birth);
    if (max_time > age_dependent_max_time) max_time = age_dependent_max_time;
    if (age_dependent_max_time > sp.time_step) out.mature = true;
  }
  double t_steps;
  D3 disp;
  const double r_rate_factor = clipped_displacement(rng, T, sp.space_step, sp.time_step, max_time, t_steps, disp);
  out.t_steps = t_steps;
  uint32_t last_hit = kInvalid;
  const bool can_react = (sp.flags & SPF_CAN_VOLVOL) && !(sp.flags & SPF_TARGET_ONLY);
  SubpartSet set;
  set.v = set_lds;
  set.stride = blockDim.x;
  set.clear();
  int seg;
  for (seg = 0; seg < kMaxSegments; ++seg) {
    const D3 end = {pos.x + disp.x, pos.y + disp.y, pos.z + disp.z};
    const D3 part_disp = in_partition(R.P, end) ? disp : displacement_up_to_partition_boundary(R.P, pos, disp);
    int sx, sy, sz;
    subpart_3d(R.P, pos, sx, sy, sz);
    if (sx < 0 || sy < 0 || sz < 0 || sx >= (int)R.P.n || sy >= (int)R.P.n || sz >= (int)R.P.n) {
      err |= ERR_ESCAPED;
      out.pos = pos;
      return false;
    }
    const uint32_t start = subpart_of(R.P, sx, sy, sz);
    const D3 dst = {pos.x + part_disp.x, pos.y + part_disp.y, pos.z + part_disp.z};
    int dx, dy, dz;
    subpart_3d(R.P, dst, dx, dy, dz);
    const uint32_t dest = subpart_of(R.P, dx, dy, dz);
    WallHit hit;
    uint32_t hit_sp = start;
    bool found = closest_wall(R, start, last_hit, pos, disp, rng, hit, err);
    if (!found && dest != start) {
      SubpartWalk64 walk;
      walk.init(R.P, pos, start, part_disp);
      while (walk.next(R.P, err)) {
        if (closest_wall(R, walk.cur, last_hit, pos, disp, rng, hit, err)) {
          found = true;
          hit_sp = walk.cur;
          break;
        }
      }
    }
    if (can_react) {
      // ray_trace_vol (diffuse_react_event.cpp:735-772): the subpartitions of
      // the move up to the wall hit (the whole move when the hit is in the
      // destination subpartition), then every partner registered in them
      const D3 mol_disp = (found && hit_sp != dest) ? D3{hit.pos.x - pos.x, hit.pos.y - pos.y, hit.pos.z - pos.z}
                                                    : part_disp;
      collect_molecule_subparts(R.P, pos, start, mol_disp, set, err);
      const double t_limit = found ? hit.t : 1e300;
      double t_last = 0;
      uint32_t id_last = 0;
      bool first = true;
      for (int guard = 0;; ++guard) {
        if (guard == kMaxCollisions) {
          err |= ERR_COLLISIONS;
          break;
        }
        double tc;
        uint32_t slot;
        D3 cp;
        int32_t rx;
        uint32_t candidates;
        if (!next_collision(R, S, set, rep, spc, idv, pos, disp, t_limit, t_last, id_last, first, tc, slot, cp, rx,
                            candidates))
          break;
        first = false;
        t_last = tc;
        id_last = S.id[slot];
        if (tc < STIME_EPS) {  // is_immediate_collision (collision_utils.inl:814-816)
          if (candidates == 1) break;
          continue;
        }
        double factor = 1.0;
        if (disk_meets_walls(R, cp)) {
          if constexpr (kPoolN == 0) {
            out.exd = true;
            return false;
          } else {
            bool full = false;
            const D3 target = {(double)S.x[slot], (double)S.y[slot], (double)S.z[slot]};
            factor = exact_disk_dev<kPoolN>(R.P, R.walls, R.off, R.idx, R.ent_box, cp, disp, R.P.rxn_radius, target,
                                            &full);
            if (full) {
              if constexpr (kPoolN >= exd::kPool) {
                err |= ERR_EXD;
              } else {  // the pool ran out: a pass with a larger one repeats the step
                out.exd = true;
                return false;
              }
            }
          }
        }
        if (factor < 0) {  // the disk is occluded
          if (candidates == 1) break;
          continue;
        }
        const RxnRec& rr = R.rxns[rep * R.rxn_stride + rx];
        const int pw = test_bimolecular(rr, factor * r_rate_factor, rng);
        if (pw >= 0 && succ < pass_through) {
          ++succ;  // carried out already: the mover stayed, its step goes on
        } else if (pw >= 0 && S.consumed[slot] == 0) {  // (else: consumed by this mover's earlier reaction)
          prop.time = elapsed + t_steps * tc;
          prop.x = cp.x;
          prop.y = cp.y;
          prop.z = cp.z;
          prop.partner_slot = slot;
          prop.mover_id = idv;
          prop.partner_id = id_last;
          prop.rxn = (uint32_t)rx;
          prop.pathway = (uint32_t)pw;
          prop.replica = rep;
          // outcome_products_random's reactant order (diffuse_react_event.cpp:
          // 2528-2547): the mover is reactant b unless both match in order
          prop.pad = (spc == rr.reactant_a && tag_species(S.tag[slot]) == rr.reactant_b) ? 0u : kPadMoverIsB;
          prop.cvol = out.cvol;
          // ray_trace_vol commits the move before the collisions are
          // processed when the segment hit no wall (diffuse_react_event.cpp:773-777)
          out.pos = found ? pos : D3{pos.x + disp.x, pos.y + disp.y, pos.z + disp.z};
          return true;
        }
        if (candidates == 1) break;  // that was the last collision of the segment
      }
    }
    if (!found) {
      // disp as the wall search left it (an edge re-draw may have moved it)
      pos = {pos.x + disp.x, pos.y + disp.y, pos.z + disp.z};
      out.near_plane = kF32;
      break;
    }
    // collide_and_react_with_surf_mol (diffuse_react_event.cpp:505-527, 845-975):
    // the surface molecule on the tile the hit point falls in.  A tile this
    // mover consumed in a reaction it stayed through is met again, up to
    // that reaction, by the repeat of its step.
    bool flip_through = false;
    if ((sp.flags & SPF_CAN_VOLSURF) && R.grids && R.grids[hit.wall].num_tiles) {
      const WallGrid& g = R.grids[hit.wall];
      const uint32_t slot = rep * R.total_tiles + g.tile_base + xyz2grid(R.walls[hit.wall], g, hit.pos);
      const uint32_t tw = R.tiles[slot];
      const uint32_t tc = R.tile_consumed[slot];
      const bool live = tc == 0 || (tc == idv + 1 && succ < pass_through);
      if (tw != kTileEmpty && live && born_tile == slot + 1) {
        born_tile = 0;  // created on this tile: no rebinding on the first hit (:878-885)
      } else if (tw != kTileEmpty && live) {
        const int32_t ri = R.pair_vs[spc * R.num_species + tile_species(tw)];
        if (ri >= 0) {
          const VolSurfRec& vs = R.vs_rxns[ri];
          // trigger_bimolecular's orientation rule (rxn_utils.inl:58-84)
          const int ga = vs.geom_vol, gb = vs.geom_surf;
          const int orient_vol = hit.type == WALL_FRONT ? 1 : -1;
          if (ga == 0 || gb == 0 || (ga + gb) * (ga - gb) != 0 || orient_vol * tile_orient(tw) * ga * gb > 0) {
            RxnRec rr;
            rr.max_fixed_p = vs.max_fixed_p;
            rr.max_p_f = vs.max_p_f;
            rr.num_pathways = vs.num_pathways;
            for (int k = 0; k < kMaxPathways; ++k) rr.cum_probs[k] = vs.cum_probs[k];
            const int pw = test_bimolecular(rr, r_rate_factor / g.binding_factor, rng);
            if (pw >= 0 && succ < pass_through) {
              // carried out already and the mover stayed (outcome_products_
              // random, :2690-2720): the orientation draws again, then the
              // mover passes through the wall (RX_FLIP, :940-971) when its
              // product is oriented other than the rule's volume reactant,
              // else it goes on as after no reaction
              ++succ;
              uint32_t drawn = 0;
              for (uint32_t q = 0; q < vs.num_products[pw]; ++q)
                if (vs.product_orient[pw][q] == 0 && (rng.u32() & 1u)) drawn |= 1u << q;
              if (vs.keep_vol[pw]) {  // (always: a reaction that consumed the mover ended its step)
                const uint32_t kq = vs.keep_vol[pw] - 1;
                const int flip = (vs.geom_surf != 0 && tile_orient(tw) != vs.geom_surf) ? -1 : 1;
                const int o = vs.product_orient[pw][kq] != 0 ? vs.product_orient[pw][kq] * flip
                                                             : (((drawn >> kq) & 1u) ? 1 : -1);
                flip_through = o != vs.geom_vol;
              }
            } else if (pw >= 0) {
              prop.time = elapsed + t_steps * hit.t;
              prop.x = hit.pos.x;
              prop.y = hit.pos.y;
              prop.z = hit.pos.z;
              prop.partner_slot = kTileBit | slot;
              prop.mover_id = idv;
              prop.partner_id = R.tile_id[slot];
              prop.rxn = kVolSurfBit | (uint32_t)ri;
              // products without an orientation in the rule draw one each,
              // in the rule's order (outcome_products_random, :2620-2625)
              uint32_t drawn = 0;
              for (uint32_t q = 0; q < vs.num_products[pw]; ++q)
                if (vs.product_orient[pw][q] == 0 && (rng.u32() & 1u)) drawn |= 1u << q;
              prop.pathway = (uint32_t)pw | (drawn << 8);
              prop.replica = rep;
              prop.pad = 0;
              prop.cvol = out.cvol;
              out.pos = pos;
              return true;
            }
          }
        }
      }
    }
    // collide_and_react_with_walls (diffuse_react_event.cpp:529-561, 991-1067);
    // a flip crosses the wall as a transparent one would
    const uint32_t act = flip_through ? SC_TRANSPARENT : R.SC.of(hit.wall, spc, hit.type == WALL_FRONT);
    if (act == SC_ABSORB) {
      (void)rng.dbl();  // test_intersect's two draws (rxn_utils.inl:593-626)
      (void)rng.dbl();
      out.pos = hit.pos;
      out.absorbed = true;
      return false;
    }
    elapsed += t_steps * hit.t;
    const WallRec64& w = R.walls[hit.wall];
    const double t_hit = hit.t;
    pos = hit.pos;
    t_steps *= (1.0 - t_hit);
    last_hit = hit.wall;
    const double s = 1.0 - t_hit;
    if (act == SC_TRANSPARENT) {
      // cross_transparent_wall (diffuse_react_event.cpp:3071-3094) with
      // update_counted_volume_id_when_crossing_wall (collision_utils.inl:
      // 1637-1694): a front hit enters the wall's object, a back hit leaves it
      if (R.cv_in && R.cv_in[hit.wall] != kInvalid)
        out.cvol = hit.type == WALL_FRONT ? R.cv_in[hit.wall] : R.cv_out[hit.wall];
      if (t_steps < 1e-12) t_steps = 1e-12;
      disp = {disp.x * s, disp.y * s, disp.z * s};
      if (!flip_through) ++out.n_cross;
      continue;
    }
    // reflect_from_wall (collision_utils.inl:1711-1747)
    const double rf = -2.0 * dot3(disp, d3(w.nx, w.ny, w.nz));
    disp = {(disp.x + rf * w.nx) * s, (disp.y + rf * w.ny) * s, (disp.z + rf * w.nz) * s};
    ++out.n_reflect;
  }
  if (seg == kMaxSegments) err |= ERR_SEGMENTS;
  if (!in_partition(R.P, pos)) err |= ERR_ESCAPED;
  out.pos = pos;
  // diffuse_single_molecule's rescheduling (diffuse_react_event.cpp:294-
  // 319): diffusion_time += max_time; the molecule is called again in this
  // iteration when its unimolecular reaction is due before the iteration's
  // end or its diffusion time is still before it; that call reacts first
  // when the reaction time is not after the new diffusion time
  // (:213-221), else it diffuses on
  const double t_new = t_start + out.t_steps;
  const double t_end = (double)R.iteration + 1.0;
  if (!(sp.space_step > 0)) {
    // a molecule that cannot diffuse is scheduled at its reaction time
    // (:320-333)
    if (out.utime != kTimeInvalid && out.utime < t_end) out.due = pass + 2;
  } else if ((out.utime != kTimeInvalid && out.utime < t_end) || cmp_lt_d(t_new, t_end, 1e-12)) {
    if (out.utime != kTimeInvalid && out.utime <= t_new) out.due = pass + 2;
    else out.cont = true;
  }
  return false;
}

}  // namespace mcg
