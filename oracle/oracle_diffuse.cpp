// TEST INFRASTRUCTURE ONLY (CPU oracle; see oracle_core.h).
// The per-iteration diffuse-react step: src4/diffuse_react_event.cpp:53-1067
// and the collision/diffusion/reaction utilities it includes.
#include <algorithm>
#include <stdexcept>

#include "oracle_core.h"

namespace orc {

// ------------------------------------------------------------ reactant sets
std::vector<uint32_t> World::partner_species_of(uint32_t s) const {
  std::vector<uint32_t> out;
  size_t n = species.size();
  for (size_t t = 0; t < n; ++t)
    if (pair_rxn[s * n + t] >= 0) out.push_back((uint32_t)t);
  return out;
}

void World::reactants_insert(uint32_t mover_species, uint32_t subpart, uint32_t id) {
  auto& per_sp = reactants[mover_species];
  if (per_sp.empty()) per_sp.resize(cfg.n3);
  auto& v = per_sp[subpart];
  auto it = std::lower_bound(v.begin(), v.end(), id);
  if (it == v.end() || *it != id) v.insert(it, id);
}

void World::reactants_erase(uint32_t mover_species, uint32_t subpart, uint32_t id) {
  auto& per_sp = reactants[mover_species];
  if (per_sp.empty()) return;
  auto& v = per_sp[subpart];
  auto it = std::lower_bound(v.begin(), v.end(), id);
  if (it != v.end() && *it == id) v.erase(it);
}

// Partition::change_vol_reactants_map_from_orig_to_current
// (src4/partition.h:351-403): molecule m is a candidate for every mover
// species that has a bimolecular reaction with m's species.
void World::register_reactant(const Mol& m, bool add) {
  if (!(species[m.species].flags & SPF_CAN_VOLVOL)) return;
  for (uint32_t s : partner_species_of(m.species)) {
    if (species[s].flags & SPF_TARGET_ONLY) continue;
    if (add) reactants_insert(s, m.reactant_subpart, m.id);
    else reactants_erase(s, m.reactant_subpart, m.id);
  }
}

// Partition::add_volume_molecule (src4/partition.h:555-597), release at time t.
uint32_t World::add_volume_molecule(uint32_t sp, V3 pos, double time) {
  if (!in_partition(pos)) throw std::runtime_error("cannot create molecule outside partition");
  Mol m{};
  m.id = (uint32_t)mols.size();
  m.species = sp;
  m.flags = 0;
  m.pos = pos;
  m.subpart = subpart_index(pos);
  m.reactant_subpart = m.subpart;
  m.diffusion_time = time;
  m.birthday = time;
  m.unimol_rxn_time = TIME_INVALID;
  m.flags |= MF_SCHEDULE_UNIMOL;  // set for every new molecule (release_event.cpp, outcome_products_random)
  mols.push_back(m);
  schedulable.push_back(m.id);
  register_reactant(mols.back(), true);
  return m.id;
}

void World::start_stream(const Mol& m, uint32_t pass) {
  if (mode == DRAW_PHILOX) philox.reset(philox_seed, m.id, (uint32_t)iteration, replica, pass);
}

// diffusion_utils / rxn_utils.inl:720-735 (time_of_unimol) through
// pick_unimol_rxn_class_and_set_rxn_time (diffuse_react_event.cpp:1731-1755):
// one draw, lifetime -ln(p) / k_tot, none when p is indistinguishable from 0.
void World::pick_unimol_time(Mol& m, double current_time) {
  if (m.species >= unimol.size() || unimol[m.species].cum.empty()) {
    m.unimol_rxn_time = TIME_INVALID;
    return;
  }
  double k_tot = unimol[m.species].cum.back();
  double p = rng().dbl();
  double t;
  if (k_tot <= 0 || !distinguishable(p, 0, EPS)) t = TIME_FOREVER;
  else t = -std::log(p) / k_tot;
  m.unimol_rxn_time = current_time + t;
}

// react_unimol_single_molecule + outcome_unimolecular
// (diffuse_react_event.cpp:1760-1826, 2939-3003): the pathway (a draw only with
// several), products at the molecule's position, created at the scheduled
// time, then the molecule is destroyed unless the pathway keeps it.  Returns
// whether it survived.
bool World::react_unimol(uint32_t id) {
  Mol& m = mols[id];
  double t = m.unimol_rxn_time;
  const UnimolRxn& u = unimol[m.species];
  int pw = 0;
  if (u.cum.size() > 1) {
    double match = rng().dbl() * u.cum.back();
    pw = 0;
    while (pw + 1 < (int)u.cum.size() && match > u.cum[pw]) ++pw;
  }
  V3 pos = m.pos;
  const uint32_t sp = m.species;  // `m` dangles once products are added
  std::vector<uint32_t> prods = u.products[pw];
  // outcome_unimolecular + outcome_products_random (diffuse_react_event.cpp:
  // 2686-2727, 2939-3003): a product of the reactant's species is the
  // reactant itself, which stays and diffuses on with a new lifetime
  const bool keep = std::find(prods.begin(), prods.end(), sp) != prods.end();
  bool used = false;
  for (uint32_t ps : prods) {
    if (keep && !used && ps == sp) {
      used = true;
      continue;
    }
    uint32_t nid = add_volume_molecule(ps, pos, t);
    mols[nid].cvol = mols[id].cvol;  // the volume the reactant is in
    if (cmp_lt(t, event_time + 1.0, EPS)) new_diffuse_actions.push_back({nid});
  }
  ++unimol_reactions;
  rxn_log.push_back({(uint32_t)iteration, id, INVALID_U32, (int32_t)(rxns.size() + sp), pw, t, pos});
  Mol& mm = mols[id];
  if (keep) {
    mm.flags |= MF_SCHEDULE_UNIMOL;
    return true;
  }
  mm.flags |= MF_DEFUNCT;
  register_reactant(mm, false);
  return false;
}

// ------------------------------------------------------------ one iteration
// DiffuseReactEvent::step + diffuse_molecules (diffuse_react_event.cpp:53-161)
// with event_time = iteration and time_up_to_next_barrier = 1.
void World::run_iteration() {
  event_time = (double)iteration;
  // Calendar::get_time_up_to_next_barrier (scheduler.cpp:212-242): the next
  // count event strictly after this iteration, at most 100 iterations away
  if (!barrier_list.empty()) {
    // the first listed barrier after this iteration; none: TIME_FOREVER
    time_up_to_next_barrier = 100.0;
    for (uint64_t b : barrier_list)
      if (b > iteration) {
        time_up_to_next_barrier = std::min((double)(b - iteration), 100.0);
        break;
      }
  } else {
    uint64_t b = barrier_interval ? barrier_interval : 1;
    uint64_t next_barrier = (iteration / b + 1) * b;
    time_up_to_next_barrier = std::min((double)(next_barrier - iteration), 100.0);
  }
  calls_this_iteration.assign(mols.size(), 0);
  if (snapshot) {
    snapshot_pos.resize(mols.size());
    for (size_t i = 0; i < mols.size(); ++i) snapshot_pos[i] = mols[i].pos;
  }
  // Partition::get_molecules_ready_for_diffusion (partition.h:232-246)
  std::vector<uint32_t> ready;
  double time_it_end = (double)iteration + 1;
  for (uint32_t id : schedulable) {
    const Mol& m = mols[id];
    if (!(m.flags & MF_DEFUNCT) && cmp_lt(m.diffusion_time, time_it_end, EPS)) ready.push_back(id);
  }
  std::vector<uint32_t> delayed;
  for (uint32_t id : ready) {
    double delay = mols[id].diffusion_time - event_time;
    if (cmp_eq(delay, 0.0)) diffuse_single_molecule(id);
    else delayed.push_back(id);
  }
  for (uint32_t id : delayed) diffuse_single_molecule(id);
  for (size_t i = 0; i < new_diffuse_actions.size(); ++i) {
    born_wall = new_diffuse_actions[i].born_wall;
    born_tile = new_diffuse_actions[i].born_tile;
    diffuse_single_molecule(new_diffuse_actions[i].id);
    born_wall = born_tile = INVALID_U32;
  }
  new_diffuse_actions.clear();
  if (snapshot) {
    // reactant sets follow the final positions once the iteration is over
    for (auto& per : reactants)
      for (auto& v : per) v.clear();
    for (Mol& m : mols) {
      m.reactant_subpart = m.subpart;
      if (!(m.flags & MF_DEFUNCT)) register_reactant(m, true);
    }
  }
  // drop defunct ids from the schedulable list (DefragmentationEvent)
  schedulable.erase(std::remove_if(schedulable.begin(), schedulable.end(),
                                   [&](uint32_t id) { return (mols[id].flags & MF_DEFUNCT) != 0; }),
                    schedulable.end());
  ++iteration;
}

// diffuse_react_event.cpp:164-198
double World::get_max_time(Mol& m) {
  const Species& sp = species[m.species];
  double time_from_event_start = m.diffusion_time - event_time;
  double max_time = time_up_to_next_barrier - time_from_event_start;
  if (m.unimol_rxn_time != TIME_INVALID && m.unimol_rxn_time < m.diffusion_time + max_time)
    max_time = m.unimol_rxn_time - m.diffusion_time;
  if (!(m.flags & MF_MATURE) && sp.time_step > 1.0) {
    double age_dependent_max_time = 1.0 + 0.2 * (m.diffusion_time - m.birthday);
    if (max_time > age_dependent_max_time) max_time = age_dependent_max_time;
    if (age_dependent_max_time > sp.time_step) m.flags |= MF_MATURE;
  }
  return max_time;
}

// diffuse_react_event.cpp:201-337 (volume molecules).  In the per-molecule
// draw mode the n-th call for a molecule within an iteration reads stream
// pass n.
void World::diffuse_single_molecule(uint32_t id) {
  Mol& m0 = mols[id];
  if (m0.flags & MF_DEFUNCT) return;
  if (calls_this_iteration.size() < mols.size()) calls_this_iteration.resize(mols.size(), 0);
  start_stream(m0, calls_this_iteration[id]++);
  double diffusion_start_time = m0.diffusion_time;
  if (m0.unimol_rxn_time != TIME_INVALID && m0.unimol_rxn_time <= diffusion_start_time) {
    if (!react_unimol(id)) return;
  }
  if (mols[id].flags & MF_SCHEDULE_UNIMOL) {
    mols[id].flags &= ~MF_SCHEDULE_UNIMOL;
    pick_unimol_time(mols[id], diffusion_start_time);
  }
  double max_time = get_max_time(mols[id]);
  // (m0 dangles once a reaction that kept the molecule added products)
  if (mols[id].flags & MF_SURFACE) diffuse_surf_molecule(id, max_time);
  else if (species[mols[id].species].is_vol) diffuse_vol_molecule(id, max_time, diffusion_start_time);
  Mol& m = mols[id];
  if (m.flags & MF_DEFUNCT) return;
  const Species& sp = species[m.species];
  double event_end_time = (double)iteration + 1;
  if (sp.space_step > 0) {  // can_diffuse
    m.diffusion_time += max_time;
    if ((m.unimol_rxn_time != TIME_INVALID && m.unimol_rxn_time < event_end_time) ||
        cmp_lt(m.diffusion_time, event_time + 1.0, EPS)) {  // before_this_iterations_end
      new_diffuse_actions.push_back({m.id});
    } else {
      double rounded = std::round(m.diffusion_time);
      if (cmp_eq(m.diffusion_time, rounded, SQRT_EPS)) m.diffusion_time = rounded;
    }
  } else if (m.unimol_rxn_time != TIME_INVALID) {
    // cannot diffuse: scheduled for its unimolecular reaction
    m.diffusion_time = m.unimol_rxn_time;
    if (m.unimol_rxn_time < event_end_time) new_diffuse_actions.push_back({m.id});
  } else {
    m.diffusion_time = TIME_FOREVER;
  }
}

// sort_collisions_by_time (diffuse_react_event.cpp:341-364); ties between two
// vol-vol collisions go by descending partner id.  Other ties keep the order
// in which the collisions were found (a wall hit is found first).
static void sort_collisions_by_time(std::vector<Collision>& c) {
  std::stable_sort(c.begin(), c.end(), [](const Collision& l, const Collision& r) {
    if (l.time < r.time) return true;
    if (l.time > r.time) return false;
    if (l.type == CollType::VOLVOL && r.type == CollType::VOLVOL) return l.partner > r.partner;
    return false;
  });
}

// diffusion_utils.inl:113-119, :366-432 (unclamped branch)
void World::diffuse_vol_molecule(uint32_t id, double& max_time, double diffusion_start_time) {
  const Species& sp = species[mols[id].species];
  if (!(sp.space_step > 0)) return;
  ++diffuse_3d_calls;
  DrawSource& r = rng();
  double steps = 1.0, t_steps = 1.0, rate_factor = 1.0, r_rate_factor = 1.0;
  t_steps = steps * sp.time_step;
  if (t_steps > max_time) {
    t_steps = max_time;
    steps = max_time / sp.time_step;
  }
  if (steps < EPS) {
    steps = EPS;
    t_steps = EPS * sp.time_step;
  }
  V3 disp;
  double scale;
  if (steps == 1.0) {
    scale = sp.space_step;
    r_rate_factor = rate_factor = 1.0;
  } else {
    rate_factor = std::sqrt(steps);
    r_rate_factor = 1.0 / rate_factor;
    scale = rate_factor * sp.space_step;
  }
  disp.x = scale * rng_gauss(r) * 0.70710678118654752440;
  disp.y = scale * rng_gauss(r) * 0.70710678118654752440;
  disp.z = scale * rng_gauss(r) * 0.70710678118654752440;
  max_time = t_steps;

  // ray-trace loop (diffuse_react_event.cpp:423-572)
  std::vector<Collision> collisions;
  bool was_defunct = false;
  uint32_t last_hit = INVALID_U32;
  double elapsed = diffusion_start_time;
  bool can_vol_react = (sp.flags & SPF_CAN_VOLVOL) != 0;
  int state;
  do {
    state = ray_trace_vol(id, can_vol_react, last_hit, disp, collisions);
    if (collisions.size() > 1) sort_collisions_by_time(collisions);
    for (Collision& c : collisions) {
      if (c.type == CollType::VOLVOL) {
        ++vol_mol_vol_mol_collisions;
        if (c.time < STIME_EPS) continue;  // is_immediate_collision
        if (collide_and_react_with_vol_mol(c, disp, t_steps, r_rate_factor, elapsed)) {
          was_defunct = true;
          break;
        }
      } else {
        // collide_and_react_with_surf_mol first (diffuse_react_event.cpp:505-527)
        if ((species[mols[id].species].flags & SPF_CAN_VOLSURF) && c.wall < grids.size() &&
            grids[c.wall].num_tiles) {
          const int res = collide_and_react_with_surf_mol(c, t_steps, r_rate_factor, elapsed);
          if (res == 1) {
            was_defunct = true;
            break;
          }
          if (res == 0) {
            // the mover stayed and passes through the wall (:947-971)
            cross_transparent_wall(c, mols[id], disp, t_steps, elapsed, last_hit);
            break;
          }
        }
        Mol& vm = mols[id];
        // collide_and_react_with_walls (diffuse_react_event.cpp:529-549, 991-1067)
        if (species[vm.species].flags & SPF_CAN_VOLWALL) {
          int act = wall_action(walls[c.wall], vm.species, c.type);
          if (act == SC_TRANSPARENT) {
            cross_transparent_wall(c, vm, disp, t_steps, elapsed, last_hit);
            ++transparent_crossings;
            if (trace) hits.push_back({vm.id, (uint32_t)iteration, c.wall, (uint32_t)c.type, vm.subpart, c.time});
            break;
          }
          if (act == SC_ABSORB) {
            // test_intersect (rxn_utils.inl:593-626) with the absorptive class's
            // unbounded rate: two draws, then outcome_intersect destroys the molecule
            DrawSource& r = rng();
            (void)r.dbl();
            (void)r.dbl();
            vm.flags |= MF_DEFUNCT;
            register_reactant(vm, false);
            ++absorbed;
            if (trace) hits.push_back({vm.id, (uint32_t)iteration, c.wall, (uint32_t)c.type, vm.subpart, c.time});
            was_defunct = true;
            break;
          }
        }
        elapsed += t_steps * c.time;
        reflect_from_wall(c, vm, disp, t_steps, last_hit);
        ++mol_wall_reflections;
        if (trace) hits.push_back({vm.id, (uint32_t)iteration, c.wall, (uint32_t)c.type, vm.subpart, c.time});
        break;
      }
    }
  } while (state != 0 && !was_defunct);

  if (!was_defunct) {
    Mol& vm = mols[id];
    if (!in_partition(vm.pos)) throw std::runtime_error("molecule escaped the partition");
    if (!snapshot && vm.subpart != vm.reactant_subpart) {
      // Partition::update_molecule_reactants_map (partition.h:406-418)
      register_reactant(vm, false);
      vm.reactant_subpart = vm.subpart;
      register_reactant(vm, true);
    }
  }
}

// ray_trace_vol (diffuse_react_event.cpp:627-780).  Returns 0 = FINISHED,
// 1 = RAY_TRACE_HIT_WALL.
int World::ray_trace_vol(uint32_t id, bool can_vol_react, uint32_t last_hit, V3& remaining,
                         std::vector<Collision>& collisions) {
  ++ray_voxel_tests;
  int res = 0;
  collisions.clear();
  Mol vm = mols[id];  // copy: the position is committed below
  V3 part_disp = remaining;
  if (!in_partition(vm.pos + remaining)) {
    part_disp = displacement_up_to_partition_boundary(vm.pos, remaining);
    ++partition_clips;
  }

  std::vector<uint32_t> walls_sp, mols_sp;
  uint32_t last_subpart = collect_crossed_subparts(vm, part_disp, can_vol_react, true, walls_sp, mols_sp);
  V3 disp_up_to_wall = remaining;
  bool wall_hit_in_last_subpart = false;
  if (!walls_sp.empty()) {
    Collision closest{};
    for (uint32_t s : walls_sp) {
      if (get_closest_wall_collision(vm, s, last_hit, remaining, disp_up_to_wall, closest)) {
        collisions.push_back(closest);
        res = 1;
        wall_hit_in_last_subpart = (last_subpart == s);
        break;
      }
    }
  }
  if (can_vol_react) {
    if (res == 1 && !wall_hit_in_last_subpart) {
      walls_sp.clear();
      mols_sp.clear();
      collect_crossed_subparts(vm, disp_up_to_wall, true, false, walls_sp, mols_sp);
    }
    for (uint32_t s : mols_sp) {
      const auto& per_sp = reactants[vm.species];
      if (per_sp.empty()) continue;
      for (uint32_t pid : per_sp[s]) {
        double t;
        V3 p;
        if (collide_mol(vm, remaining, pid, t, p)) {
          int rx = pair_rxn[vm.species * species.size() + mols[pid].species];
          if (rx < 0) continue;
          Collision c{};
          c.type = CollType::VOLVOL;
          c.mover = vm.id;
          c.time = t;
          c.pos = p;
          c.partner = pid;
          c.rxn = rx;
          c.wall = INVALID_U32;
          collisions.push_back(c);
        }
      }
    }
  }
  if (res == 0) {
    Mol& m = mols[id];
    m.pos = m.pos + remaining;
    m.subpart = subpart_index(m.pos);
  }
  return res;
}

// update_vol_mol_after_rxn_with_surf_mol + tiny_diffuse_3D
// (diffuse_react_event.cpp:2291-2320, diffusion_utils.inl:551-595): the
// product at its creation point moves by 2 * 16 POS_EPS along the wall's
// normal to the side `orient` names, or half the way to the first other
// wall on that move.  In the per-molecule draw mode the wall search's edge
// re-draws read the product's pass 255 (the reference uses its auxiliary
// generator); the mover's own stream is left where it was.
V3 World::volume_product_position(uint32_t id, uint32_t wall, int orient) {
  const double bump = orient > 0 ? 16.0 * POS_EPS : -16.0 * POS_EPS;
  const V3 disp = v3(2.0 * bump) * walls[wall].normal;
  const V3 start = mols[id].pos;
  V3 temp = disp;
  std::vector<Collision> collisions;
  const PhiloxSource saved = philox;
  if (mode == DRAW_PHILOX) philox.reset(philox_seed, id, (uint32_t)iteration, replica, 0xFFu);
  ray_trace_vol(id, false, wall, temp, collisions);
  if (mode == DRAW_PHILOX) philox = saved;
  V3 nd = disp;
  for (const Collision& c : collisions)
    if (c.type != CollType::VOLVOL && (c.time <= 1.0 || cmp_eq(c.time, 1.0, EPS))) nd = disp * v3(c.time * 0.5);
  return start + nd;
}

// collision_utils.inl:48-114
V3 World::displacement_up_to_partition_boundary(V3 pos, V3 displacement) const {
  V3 dnz = displacement;
  guard_zero_div(dnz);
  V3 dir{dnz.x > 0 ? 1.0 : 0.0, dnz.y > 0 ? 1.0 : 0.0, dnz.z > 0 ? 1.0 : 0.0};
  V3 edges = cfg.origin + dir * v3(cfg.partition_edge_length);
  V3 diff = edges - pos;
  double hit_time = 1;
  if (std::fabs(diff.x) < POS_EPS || std::fabs(diff.y) < POS_EPS || std::fabs(diff.z) < POS_EPS) return v3(0);
  V3 t = diff / dnz;
  if (t.x >= 0 && t.x < t.y && t.x <= t.z) hit_time = t.x;
  else if (t.y >= 0 && t.y <= t.z) hit_time = t.y;
  else if (t.z >= 0) hit_time = t.z;
  return displacement * v3(hit_time - STIME_EPS);
}

// collision_utils_subparts.inl:38-122
static void collect_neighbors(const World& w, V3 pos, const int idx[3], double rxn_radius, double sp_len,
                              std::vector<uint32_t>& out) {
  auto ins = [&](uint32_t s) {
    if (std::find(out.begin(), out.end(), s) == out.end()) out.push_back(s);
  };
  double part_len = w.cfg.partition_edge_length;
  V3 rel = pos - w.cfg.origin;
  V3 plus = rel + v3(rxn_radius);
  V3 minus = rel - v3(rxn_radius);
  V3 boundary = V3{(double)idx[0], (double)idx[1], (double)idx[2]} * v3(sp_len);
  int xd = 0, yd = 0, zd = 0;
  if (minus.x < boundary.x && minus.x > 0.0) { ins(w.subpart_from_3d(idx[0] - 1, idx[1], idx[2])); xd = -1; }
  else if (plus.x > boundary.x + sp_len && plus.x < part_len) { ins(w.subpart_from_3d(idx[0] + 1, idx[1], idx[2])); xd = 1; }
  if (minus.y < boundary.y && minus.y > 0.0) { ins(w.subpart_from_3d(idx[0], idx[1] - 1, idx[2])); yd = -1; }
  else if (plus.y > boundary.y + sp_len && plus.y < part_len) { ins(w.subpart_from_3d(idx[0], idx[1] + 1, idx[2])); yd = 1; }
  if (minus.z < boundary.z && minus.z > 0.0) { ins(w.subpart_from_3d(idx[0], idx[1], idx[2] - 1)); zd = -1; }
  else if (plus.z > boundary.z + sp_len && plus.z < part_len) { ins(w.subpart_from_3d(idx[0], idx[1], idx[2] + 1)); zd = 1; }
  if (xd && yd) ins(w.subpart_from_3d(idx[0] + xd, idx[1] + yd, idx[2]));
  if (xd && zd) ins(w.subpart_from_3d(idx[0] + xd, idx[1], idx[2] + zd));
  if (yd && zd) ins(w.subpart_from_3d(idx[0], idx[1] + yd, idx[2] + zd));
  if (xd && yd && zd) ins(w.subpart_from_3d(idx[0] + xd, idx[1] + yd, idx[2] + zd));
}

// collision_utils_subparts.inl:127-300 (3-D DDA over the subpartition grid)
uint32_t World::collect_crossed_subparts(const Mol& vm, V3 displacement, bool for_mols, bool for_walls,
                                         std::vector<uint32_t>& walls_sp, std::vector<uint32_t>& mols_sp) const {
  auto ins = [&](uint32_t s) {
    if (std::find(mols_sp.begin(), mols_sp.end(), s) == mols_sp.end()) mols_sp.push_back(s);
  };
  uint32_t cur = vm.subpart;
  if (for_walls) walls_sp.push_back(cur);
  if (for_mols) ins(cur);
  V3 dest = vm.pos + displacement;
  V3 dnz = displacement;
  guard_zero_div(dnz);
  int urb[3] = {dnz.x > 0 ? 1 : 0, dnz.y > 0 ? 1 : 0, dnz.z > 0 ? 1 : 0};
  int src[3], dst[3];
  subpart_3d_from_index(cur, src);
  subpart_3d(dest, dst);
  double r_nb = cfg.rxn_radius_3d * POS_SQRT2;
  double sp_len = cfg.subpart_edge_length;
  if (for_mols && cfg.use_expanded_list) {
    std::vector<uint32_t> tmp = mols_sp;
    collect_neighbors(*this, vm.pos, src, r_nb, sp_len, mols_sp);
  }
  uint32_t dest_index = subpart_from_3d(dst[0], dst[1], dst[2]);
  if (cur != dest_index) {
    int add[3] = {urb[0] == 0 ? -1 : 1, urb[1] == 0 ? -1 : 1, urb[2] == 0 ? -1 : 1};
    V3 cp = vm.pos;
    int ci[3] = {src[0], src[1], src[2]};
    uint32_t cidx;
    V3 rcp = v3(1.0) / dnz;
    int guard = 0;
    do {
      V3 edges = cfg.origin + V3{(double)ci[0], (double)ci[1], (double)ci[2]} * v3(sp_len) +
                 V3{(double)urb[0], (double)urb[1], (double)urb[2]} * v3(sp_len);
      V3 diff = edges - cp;
      V3 t = diff * rcp;
      if (t.x < t.y && t.x <= t.z) {
        cp = cp + displacement * v3(t.x);
        ci[0] += add[0];
        if (!in_range(ci[0])) break;
      } else if (t.y <= t.z) {
        cp = cp + displacement * v3(t.y);
        ci[1] += add[1];
        if (!in_range(ci[1])) break;
      } else {
        cp = cp + displacement * v3(t.z);
        ci[2] += add[2];
        if (!in_range(ci[2])) break;
      }
      cidx = subpart_from_3d(ci[0], ci[1], ci[2]);
      if (for_walls) walls_sp.push_back(cidx);
      if (for_mols) ins(cidx);
      if (for_mols && cfg.use_expanded_list) collect_neighbors(*this, cp, ci, r_nb, sp_len, mols_sp);
      if (++guard > 3 * (int)cfg.n + 3) throw std::runtime_error("DDA did not terminate");
    } while (cidx != dest_index);
  }
  if (for_mols && cfg.use_expanded_list) collect_neighbors(*this, dest, dst, r_nb, sp_len, mols_sp);
  return dest_index;
}

// collision_utils.inl:568-603 (FP64 branch); jitter sign from one uint32 draw
void World::jump_away_line(V3 p, double k, V3 A, V3 B, V3 n, V3& v) {
  V3 e = B - A;
  double elen2 = dot(e, e);
  double le_1 = 1.0 / std::sqrt(elen2);
  e = e * v3(le_1);
  V3 f;
  f.x = n.y * e.z - n.z * e.y;
  f.y = n.z * e.x - n.x * e.z;
  f.z = n.x * e.y - n.y * e.x;
  V3 af{std::fabs(f.x), std::fabs(f.y), std::fabs(f.z)};
  double tiny = POS_EPS * (abs_max_2vec(p, v) + 1.0) / (k * max3(af));
  if ((rng().u32() & 1) == 0) tiny = -tiny;
  v.x -= tiny * f.x;
  v.y -= tiny * f.y;
  v.z -= tiny * f.z;
}

// collision_utils.inl:629-812
CollType World::collide_wall(V3 pos, uint32_t wi, bool update_move, V3& move, double& t, V3& hit) {
  ++ray_polygon_tests;
  const Wall& w = walls[wi];
  double dp = dot(w.normal, pos);
  double dv = dot(w.normal, move);
  double dd = dp - w.d;
  double d_eps;
  if (dd > 0) {
    d_eps = POS_EPS;
    if (dd < d_eps) d_eps = 0.5 * dd;
    if (dd + dv > d_eps) return CollType::WALL_MISS;
  } else {
    d_eps = -POS_EPS;
    if (dd > d_eps) d_eps = 0.5 * dd;
    if (dd < 0 && dd + dv < d_eps) return CollType::WALL_MISS;
  }
  double a;
  if (dd == 0) {
    if (dv != 0) return CollType::WALL_MISS;
    if (update_move) {
      a = (abs_max_2vec(pos, move) + 1.0) * POS_EPS;
      if ((rng().u32() & 1) == 0) a = -a;
      move = move - v3(a) * w.normal;
      ++redo_count;
      return CollType::WALL_REDO;
    }
    return CollType::WALL_MISS;
  }
  a = 1.0 / dv;
  a *= -dd;
  t = a;
  hit = pos + a * move;
  V3 v0 = verts[w.v[0]];
  V3 local = hit - v0;
  double b = dot(local, w.unit_u);
  double c = dot(local, w.unit_v);
  double f;
  if (w.uv2_v < 0) {
    c = -c;
    f = -w.uv2_v;
  } else {
    f = w.uv2_v;
  }
  if (c > 0) {
    double g = b * f;
    double h = c * w.uv2_u;
    if (g > h) {
      if (c * w.uv_vert1_u + g < h + w.uv_vert1_u * w.uv2_v) {
        return dv > 0 ? CollType::WALL_BACK : CollType::WALL_FRONT;
      } else if (!distinguishable(c * w.uv_vert1_u + g, h + w.uv_vert1_u * w.uv2_v, POS_EPS)) {
        if (update_move) {
          jump_away_line(pos, a, verts[w.v[1]], verts[w.v[2]], w.normal, move);
          ++redo_count;
          return CollType::WALL_REDO;
        }
        return CollType::WALL_MISS;
      }
      return CollType::WALL_MISS;
    } else if (!distinguishable(g, h, POS_EPS)) {
      if (update_move) {
        jump_away_line(pos, a, verts[w.v[2]], v0, w.normal, move);
        ++redo_count;
        return CollType::WALL_REDO;
      }
      return CollType::WALL_MISS;
    }
    return CollType::WALL_MISS;
  } else if (!distinguishable(c, 0.0, POS_EPS)) {
    if (update_move) {
      jump_away_line(pos, a, v0, verts[w.v[1]], w.normal, move);
      ++redo_count;
      return CollType::WALL_REDO;
    }
    return CollType::WALL_MISS;
  }
  return CollType::WALL_MISS;
}

// collision_utils.inl:819-914 (FP64 branch)
bool World::get_closest_wall_collision(const Mol& vm, uint32_t subpart, uint32_t last_hit, V3& displacement,
                                       V3& disp_up_to_wall, Collision& closest) {
  int redo_guard = 0;
restart:
  bool found = false;
  double closest_t = TIME_FOREVER;
  for (uint32_t k = sp_wall_off[subpart]; k < sp_wall_off[subpart + 1]; ++k) {
    uint32_t wi = sp_wall_idx[k];
    if (wi == last_hit) continue;
    double t;
    V3 hit;
    CollType ct = collide_wall(vm.pos, wi, true, displacement, t, hit);
    if (ct == CollType::WALL_REDO) {
      if (++redo_guard > 10000) throw std::runtime_error("REDO did not terminate");
      goto restart;
    }
    if (ct != CollType::WALL_MISS) {
      ++ray_polygon_colls;
      if (subpart_index(hit) != subpart) continue;
      if (t < closest_t) {
        found = true;
        closest_t = t;
        closest = Collision{ct, vm.id, t, hit, wi, INVALID_U32, -1};
      }
    }
  }
  if (found) disp_up_to_wall = closest.pos - vm.pos;
  return found;
}

// collision_utils.inl:464-515
bool World::collide_mol(const Mol& vm, V3 displacement, uint32_t pid, double& t, V3& cpos) const {
  const Mol& other = mols[pid];
  if (other.flags & MF_DEFUNCT) return false;
  V3 opos = snapshot ? snapshot_pos[pid] : other.pos;
  V3 dir = opos - vm.pos;
  double d = dot(dir, displacement);
  if (d < 0) return false;
  double movelen2 = dot(displacement, displacement);
  if (d > movelen2) return false;
  double dirlen2 = dot(dir, dir);
  double sigma2 = cfg.rxn_radius_3d * cfg.rxn_radius_3d;
  if (movelen2 * dirlen2 - d * d > movelen2 * sigma2) return false;
  if (vm.id == other.id) return false;
  t = d / movelen2;
  cpos = vm.pos + t * displacement;
  return true;
}

// The surface class of a wall decides what a hit does to a species:
// trigger_intersect / find_mol_reactions_with_surf_classes
// (rxn_utils.inl:120-287) look the class up by the wall's regions and the hit
// orientation (front = WALL_FRONT).
int World::wall_action(const Wall& w, uint32_t sp, CollType type) const {
  if (w.surf_class == 0 || w.surf_class > num_surface_classes) return SC_REFLECT;
  uint8_t a = surf_action[(size_t)(w.surf_class - 1) * species.size() + sp];
  return type == CollType::WALL_FRONT ? (a & 3) : ((a >> 2) & 3);
}

// cross_transparent_wall (diffuse_react_event.cpp:3007-3099), the branch for
// walls that bound no compartment: move to the hit point, keep the direction
// and scale the remaining displacement and time.
void World::cross_transparent_wall(const Collision& c, Mol& vm, V3& disp, double& t_steps, double& elapsed,
                                   uint32_t& last_hit) {
  // update_counted_volume_id_when_crossing_wall (collision_utils.inl:1637-
  // 1694) for objects that do not intersect: a front hit goes inside
  if (!cv_in.empty() && cv_in[c.wall] != INVALID_U32)
    vm.cvol = c.type == CollType::WALL_FRONT ? cv_in[c.wall] : cv_out[c.wall];
  vm.pos = c.pos;
  vm.subpart = subpart_index(vm.pos);
  double t_smash = c.time;
  disp = disp * v3(1.0 - t_smash);
  elapsed += t_steps * t_smash;
  t_steps *= (1.0 - t_smash);
  if (t_steps < EPS) t_steps = EPS;
  last_hit = c.wall;
}

// collision_utils.inl:1711-1747 (FP64 branch)
void World::reflect_from_wall(const Collision& c, Mol& vm, V3& disp, double& t_steps, uint32_t& last_hit) {
  const Wall& w = walls[c.wall];
  double t_reflect = c.time;
  vm.pos = c.pos;
  vm.subpart = subpart_index(vm.pos);
  t_steps *= (1.0 - t_reflect);
  last_hit = c.wall;
  double reflect_factor = -2.0 * dot(disp, w.normal);
  disp = (disp + v3(reflect_factor) * w.normal) * v3(1.0 - t_reflect);
}

}  // namespace orc
