// Volume-volume reactions: the reaction-class ABI and the host loop of a
// reactive iteration (kernels in kernels_react.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <map>
#include <cmath>
#include <cstring>

#include "engine.h"

using namespace mcg;

namespace mcg {

namespace {

// Box of subpartitions that reactive molecules may occupy: the geometry's
// and the added molecules' bounding box, one subpartition of margin.
void setup_grid(mcg_ctx* c) {
  const PartitionParams& P = c->P;
  double lo[3] = {c->mol_lo[0], c->mol_lo[1], c->mol_lo[2]};
  double hi[3] = {c->mol_hi[0], c->mol_hi[1], c->mol_hi[2]};
  for (size_t v = 0; v + 2 < c->vert_host.size(); v += 3)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], c->vert_host[v + a]);
      hi[a] = std::max(hi[a], c->vert_host[v + a]);
    }
  const double o[3] = {P.ox, P.oy, P.oz};
  CellGrid& G = c->rx.grid;
  for (int a = 0; a < 3; ++a) {
    int l = 0, h = (int)P.n - 1;
    if (lo[a] <= hi[a]) {
      l = std::max(0, (int)((lo[a] - o[a]) * P.sp_rcp) - 1);
      h = std::min((int)P.n - 1, (int)((hi[a] - o[a]) * P.sp_rcp) + 1);
    }
    G.lo[a] = l;
    G.dim[a] = h - l + 1;
  }
  G.per_replica = (uint32_t)G.dim[0] * (uint32_t)G.dim[1] * (uint32_t)G.dim[2];
  G.ns = c->num_species;
  // sort bins: (replica, cell, species)
  const uint64_t bins = (uint64_t)G.per_replica * c->cfg.num_replicas * G.ns;
  c->rx.num_cells = bins < 0xFFFFFFF0ull ? (uint32_t)bins : 0u;
}

template <typename PosT>
ReactPipeline<PosT> make_pipeline(mcg_ctx* c, PosT* x, PosT* y, PosT* z, PosT* sx, PosT* sy, PosT* sz) {
  ReactState& r = c->rx;
  ReactPipeline<PosT> p{};
  p.R.P = c->P;
  p.R.walls = c->walls64.p;
  p.R.off = c->off.p;
  p.R.idx = c->idx.p;
  p.R.ent32 = reinterpret_cast<const FilterWall*>(c->ent32.p);
  p.R.ent64 = c->ent64.p;
  p.R.ent_box = c->ent_box.p;
  p.queue = SlowQueue{c->queue.p, c->qcount.p, 0};
  p.R.G = r.grid;
  p.R.G.start = r.cell_start.p;
  p.R.rxns = r.rxns_d.p;
  p.R.rxn_stride = c->per_replica ? (uint32_t)r.rxns.size() : 0u;
  p.R.pair_rxn = r.pair_d.p;
  p.R.num_species = c->num_species;
  p.R.seed = c->cfg.seed;
  p.R.F = make_filter_params(c->max_abs_coord);
  p.R.band = 2.0f * p.R.F.e_dd + 8.0f * p.R.F.eps * c->max_abs_coord;
  p.R.SC = c->surface_classes();
  p.R.ctr = c->counters.p;
  p.species = c->species_table();
  // without unimolecular classes no molecule has a reaction time: the
  // arrays are neither copied nor read
  const bool unimol = std::any_of(r.unimol.begin(), r.unimol.end(), [](const UnimolRec& u) { return u.num_pathways > 0; });
  const bool cv = r.num_volumes > 0;
  p.M = RMolView<PosT>{x, y, z, c->id.p, c->tag.p, r.toff.p, unimol ? r.utime.p : nullptr, cv ? r.cvol.p : nullptr,
                       r.birth.p};
  p.Sm = RMolView<PosT>{sx,      sy,      sz, r.sid.p, r.stag.p, r.stoff.p, unimol ? r.sutime.p : nullptr,
                        cv ? r.scvol.p : nullptr, r.birth.p ? r.sbirth.p : nullptr};
  p.R.cv_in = cv ? r.cv_in.p : nullptr;
  p.R.cv_out = cv ? r.cv_out.p : nullptr;
  p.R.rxn_counts = r.rxn_counts.p;
  p.R.rc_slots = r.rc_slots;
  p.R.rc_unimol = r.rc_unimol;
  p.R.rc_vs = r.rc_vs;
  p.R.rc_volumes = r.rc_volumes;
  p.R.unimol = r.unimol_d.p;
  const bool surf = r.total_tiles && !r.vs.empty();
  p.R.grids = surf ? r.grids_d.p : nullptr;
  p.R.tiles = r.tiles.p;
  p.R.tile_id = r.tile_id.p;
  p.R.tile_consumed = r.tile_consumed.p;
  p.R.total_tiles = r.total_tiles;
  p.R.vs_rxns = r.vs_d.p;
  p.R.pair_vs = r.pair_vs_d.p;
  p.tiles = r.tiles.p;
  p.tile_id = r.tile_id.p;
  p.tile_owner = r.tile_owner.p;
  p.tile_consumed = r.tile_consumed.p;
  p.key = r.key.p;
  p.rank = r.rank.p;
  p.cell_count = r.cell_count.p;
  p.cell_start = r.cell_start.p;
  p.consumed = r.consumed.p;
  p.owner = r.owner.p;
  p.prop = r.prop.p;
  p.kept = r.kept.p;
  p.aprop = r.aprop.p;
  p.R.kept = r.kept.p;
  p.R.born = r.born.p;
  p.R.tile_wall = r.tile_wall.p;
  p.plist = r.plist.p;
  p.alist = r.alist.p;
  p.alist_sorted = r.alist_sorted.p;
  p.rlist0 = r.rlist0.p;
  p.rlist1 = r.rlist1.p;
  p.mlist = r.mlist.p;
  p.ulist = r.ulist.p;
  p.clist = r.clist.p;
  p.R.clist = r.clist.p;
  p.R.ttoff = r.surf_on ? r.ttoff.p : nullptr;
  p.R.tbirth = r.surf_on ? r.tbirth.p : nullptr;
  p.R.ccount = r.hdr.p + kHdrContinue;
  p.keys_a = r.keys_a.p;
  p.keys_b = r.keys_b.p;
  p.cnt = r.cnt.p;
  p.pre = r.pre.p;
  p.head = r.head.p;
  p.next_id = r.next_id.p;
  p.hdr = r.hdr.p;
  p.scan_tmp = r.scan_tmp.p;
  p.scan_tmp_bytes = r.scan_tmp.n;
  return p;
}

int read_hdr(mcg_ctx* c) {
  HIP_TRY(hipMemcpyAsync(c->rx.hdr_host, c->rx.hdr.p, kHdrWords * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MCG_OK;
}

// The drain's form for the next iteration: the flag's, the decided one, or
// (while undecided) the trials' alternation.
uint32_t drain_form_for(const mcg_ctx* c) {
  if (c->cfg.flags & MCG_FLAG_DRAIN_INLINE) return 1;
  if (c->cfg.flags & MCG_FLAG_DRAIN_DEFER) return 2;
  return c->rx.drain_form ? c->rx.drain_form : 1 + (c->rx.form_trials & 1);
}

// Books the move phase's time of an undecided form's trial.  The first trial
// of each form is not counted (its kernels' first launch loads their code);
// after three counted trials of each the faster form is kept.
void drain_form_timed(ReactState& r, uint32_t form, float ms) {
  constexpr uint32_t kTrials = 8;
  if (r.drain_form) return;
  if (r.form_trials >= 2) {
    r.form_ms[form - 1] += ms;
    ++r.form_n[form - 1];
  }
  if (++r.form_trials == kTrials)
    r.drain_form = r.form_ms[0] * r.form_n[1] <= r.form_ms[1] * r.form_n[0] ? 1u : 2u;
}

// Claims, acceptance and repeated moves until every proposal of the phase
// is settled.  max_props bounds the phase's proposals (its movers); the
// settling kernels read the actual count on the device, so the header is
// read once per round, after the claims are settled.
template <typename PosT>
int settle_rounds(mcg_ctx* c, ReactPipeline<PosT>& p, bool snapshot_src, double barrier, uint32_t* err,
                  uint32_t& rounds, uint32_t max_props) {
  uint32_t* h = c->rx.hdr_host;
  for (uint32_t r = 0;; ++r) {
    if (r > 100000) return fail(MCG_E_LIMIT, "reaction settling did not converge");
    uint32_t* rl = (r & 1) ? p.rlist1 : p.rlist0;
    uint32_t* rc = p.hdr + ((r & 1) ? kHdrRerun1 : kHdrRerun0);
    HIP_TRY(p.settle(c->stream, max_props, rl, rc));
    int st = read_hdr(c);
    if (st) return st;
    if (h[kHdrProposals] == 0) return MCG_OK;
    ++rounds;
    const uint32_t nr = h[(r & 1) ? kHdrRerun1 : kHdrRerun0];
    if (nr == 0) return MCG_OK;
    HIP_TRY(p.move(c->stream, snapshot_src, rl, rc, nr, barrier, false, err));
    max_props = nr;
  }
}

template <typename PosT>
int run_iterations(mcg_ctx* c, ReactPipeline<PosT>& p, uint32_t first, uint32_t num, double time_up,
                   mcg_step_stats* stats, double& move_ms) {
  ReactState& r = c->rx;
  uint32_t* h = r.hdr_host;
  const bool propose_only = (c->cfg.flags & MCG_FLAG_PROPOSE_ONLY) != 0;
  const uint64_t cap = c->cfg.molecule_capacity;
  uint64_t reactions = 0, steps = 0;
  uint32_t rounds = 0;
  r.log_count = 0;
  HIP_TRY(launch_fill_u32(c->stream, p.hdr, 0u, kHdrWords));
  HIP_TRY(launch_fill_u64(c->stream, c->counters.p, 0ull, CTR_WORDS));
  // the kernels report errors into the header, which every iteration reads
  uint32_t* const err = p.hdr + kHdrErr;
  move_ms = 0;
  for (uint32_t k = 0; k < num; ++k) {
    const uint32_t it = first + k;
    // the scheduler's time up to the next barrier from this iteration,
    // capped at DIFFUSION_TIME_UPPER_LIMIT (scheduler.cpp:212-242, 303-307;
    // diffuse_react_event.h:37)
    const double barrier = std::min(time_up - (double)k, 100.0);
    p.R.iteration = it;
    const uint32_t n_in = (uint32_t)c->n;
    HIP_TRY(p.sort(c->stream, n_in, r.num_cells, err));
    HIP_TRY(launch_copy_u32(c->stream, p.cell_start + r.num_cells, p.hdr + kHdrNumOut));
    HIP_TRY(launch_copy_u32(c->stream, p.cell_start + r.num_cells, p.hdr + kHdrSnapshot));
    HIP_TRY(launch_fill_u32(c->stream, p.hdr + kHdrAccepted, 0u, 1u));
    const uint32_t form = drain_form_for(c);
    p.drain_inline = form == 1;
    HIP_TRY(hipEventRecord(c->evm0, c->stream));
    HIP_TRY(p.move(c->stream, true, nullptr, p.cell_start + r.num_cells, n_in, barrier, propose_only, err));
    HIP_TRY(hipEventRecord(c->evm1, c->stream));
    if (propose_only) {
      HIP_TRY(p.log_proposals(c->stream, n_in, it, r.log.p, p.hdr + kHdrLog, (uint32_t)r.log_cap));
    } else {
      bool snapshot_src = true;
      uint32_t phase_movers = n_in;
      for (;;) {
        int st = settle_rounds(c, p, snapshot_src, barrier, err, rounds, phase_movers);
        if (st) return st;
        if (h[kHdrUnimolDue]) {
          HIP_TRY(p.fire_unimol(c->stream, h[kHdrUnimolDue]));
          HIP_TRY(launch_fill_u32(c->stream, p.hdr + kHdrUnimolDue, 0u, 1u));
          st = read_hdr(c);
          if (st) return st;
        }
        // the next phase's movers: the products that move in this
        // iteration, then the movers that diffuse again (the phase's moves
        // are over, so the header's count of those is final)
        const uint32_t na = h[kHdrAccepted], nc = h[kHdrContinue];
        if (na == 0 && nc == 0) break;
        HIP_TRY(launch_fill_u32(c->stream, p.hdr + kHdrProductMovers, 0u, 1u));
        if (na) {
          reactions += na;
          const uint64_t log_base = std::min<uint64_t>(r.log_count, 0xFFFFFFFFull);
          HIP_TRY(p.products(c->stream, na, (uint32_t)cap, r.log.p, (uint32_t)log_base, (uint32_t)r.log_cap,
                             snapshot_src, err));
          r.log_count += na;
          HIP_TRY(launch_fill_u32(c->stream, p.hdr + kHdrAccepted, 0u, 1u));
        }
        HIP_TRY(p.continuations(c->stream, nc));
        uint32_t nm = nc;
        if (na) {
          st = read_hdr(c);
          if (st) return st;
          nm = h[kHdrProductMovers];
        }
        if (nm == 0) break;
        snapshot_src = false;
        phase_movers = nm;
        steps += nm;
        HIP_TRY(p.move(c->stream, false, p.mlist, p.hdr + kHdrProductMovers, nm, barrier, false, err));
      }
      // the surface molecules' steps follow the volume phases
      if (r.surf_on)
        if (int st = surface_phase(c, it, barrier, err)) return st;
    }
    // the settling loop ends on a header read; the proposal log does not
    if (propose_only) {
      int st = read_hdr(c);
      if (st) return st;
    }
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, c->evm0, c->evm1));
    move_ms += ms;
    if (n_in) drain_form_timed(r, form, ms);
    r.last_form = form;
    const uint32_t e = h[kHdrErr];
    if (e) {
      if (e & ERR_ESCAPED) return fail(MCG_E_ESCAPED, "a molecule escaped the partition");
      if (e & ERR_GRID) return fail(MCG_E_LIMIT, "a molecule left the subpartition box of the reaction grid");
      if (e & ERR_QUEUE) return fail(MCG_E_CAPACITY, "molecule capacity exceeded by reaction products");
      return fail(MCG_E_LIMIT, "per-step limit exceeded (error bits " + std::to_string(e) + ")");
    }
    // k_bump_ids keeps the used slots within the capacity
    c->n = std::min<uint64_t>(h[kHdrNumOut], cap);
    steps += h[kHdrSnapshot];
  }
  if (propose_only) r.log_count = h[kHdrLog];
  if (stats) {
    unsigned long long ctr[CTR_WORDS] = {0, 0};
    HIP_TRY(hipMemcpy(ctr, c->counters.p, sizeof(ctr), hipMemcpyDeviceToHost));
    stats->molecule_steps = steps;
    stats->reactions = reactions;
    stats->wall_hits = ctr[CTR_REFLECTIONS];
    stats->transparent_crossings = ctr[CTR_CROSSINGS];
    stats->settle_rounds = rounds;
    stats->drain_form = r.last_form;
    // movers of the last iteration that took the full step
    std::vector<uint32_t> q(kNumQueueShards, 0);
    if (num && c->n)
      HIP_TRY(hipMemcpy(q.data(), c->qcount.p, kNumQueueShards * 4, hipMemcpyDeviceToHost));
    uint64_t hard = 0;
    for (uint32_t v : q) hard += v;
    stats->slow_path_last = (uint32_t)hard;
  }
  return MCG_OK;
}

}  // namespace

// Buffers of the reactive path, allocated once (by the first reaction-class call).
int ensure_react_buffers(mcg_ctx* c) {
  ReactState& r = c->rx;
  if (r.on) return MCG_OK;
  const size_t cap = c->cfg.molecule_capacity;
  const uint32_t ns = c->num_species;
  bool bad = r.pair_d.alloc((size_t)ns * ns) || r.toff.alloc(cap) || r.utime.alloc(cap) || r.sutime.alloc(cap) ||
             r.sid.alloc(cap) || r.stag.alloc(cap) || r.stoff.alloc(cap) || r.key.alloc(cap) ||
             r.rank.alloc(cap) || r.consumed.alloc(cap) || r.owner.alloc(cap) || r.prop.alloc(cap) ||
             r.plist.alloc(cap) || r.alist.alloc(cap) || r.alist_sorted.alloc(cap) || r.rlist0.alloc(cap) ||
             r.rlist1.alloc(cap) || r.mlist.alloc(cap) || r.ulist.alloc(cap) || r.clist.alloc(cap) || r.keys_a.alloc(cap) ||
             r.keys_b.alloc(cap) || r.cnt.alloc(cap + 1) || r.pre.alloc(cap + 1) ||
             r.head.alloc(c->cfg.num_replicas) || r.next_id.alloc(c->cfg.num_replicas) || r.hdr.alloc(kHdrWords) ||
             r.unimol_d.alloc(ns) || r.rxns_d.alloc(1);
  if (!bad)
    bad = c->fp32() ? (r.sxf.alloc(cap) || r.syf.alloc(cap) || r.szf.alloc(cap))
                    : (r.sx.alloc(cap) || r.sy.alloc(cap) || r.sz.alloc(cap));
  if (bad) return fail(MCG_E_NOMEM, "reaction buffers");
  if (!r.hdr_host) HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&r.hdr_host), kHdrWords * 4, hipHostMallocDefault));
  std::vector<int32_t> none((size_t)ns * ns, -1);
  if (r.pair.empty()) r.pair = none;
  HIP_TRY(hipMemcpy(r.pair_d.p, r.pair.data(), r.pair.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(r.unimol_d.p, 0, ns * sizeof(UnimolRec)));
  HIP_TRY(hipMemset(r.toff.p, 0, cap * sizeof(double)));
  // molecules added so far have no unimolecular time yet
  std::vector<double> inv(std::max<size_t>(c->n, 1), kTimeInvalid);
  if (c->n) HIP_TRY(hipMemcpy(r.utime.p, inv.data(), c->n * sizeof(double), hipMemcpyHostToDevice));
  if (r.next_id_host.size() != c->cfg.num_replicas) r.next_id_host.assign(c->cfg.num_replicas, 0);
  r.next_id_dirty = true;
  r.grid_dirty = true;
  r.on = true;
  return MCG_OK;
}

int ensure_step_buffers(mcg_ctx* c) {
  if (!c->custom_steps()) return MCG_OK;
  if (int st = ensure_react_buffers(c)) return st;
  ReactState& r = c->rx;
  if (c->long_steps() && !r.birth.p) {
    const size_t cap = c->cfg.molecule_capacity;
    if (r.birth.alloc(cap) || r.sbirth.alloc(cap)) return fail(MCG_E_NOMEM, "birthday buffers");
    // molecules added so far were released at the clock (molecule.h:139-141)
    if (c->n) {
      std::vector<double> b(c->n, (double)c->clock);
      HIP_TRY(hipMemcpy(r.birth.p, b.data(), c->n * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  return MCG_OK;
}

int react_step(mcg_ctx* c, uint32_t first_iteration, uint32_t num_iterations, double time_up_to_next_barrier,
               mcg_step_stats* stats) {
  if (c->cfg.flags & MCG_FLAG_TRACE_HITS) return fail(MCG_E_UNSUPPORTED, "hit traces need a context without reactions");
  ReactState& r = c->rx;
  auto t0 = std::chrono::steady_clock::now();
  if (r.grid_dirty) {
    setup_grid(c);
    if (r.num_cells == 0) return fail(MCG_E_LIMIT, "reaction grid too large (cells x replicas x species)");
    if (r.cell_count.alloc((size_t)r.num_cells + 1) || r.cell_start.alloc((size_t)r.num_cells + 1))
      return fail(MCG_E_NOMEM, "cell grid");
    size_t tmp = ReactPipeline<double>::scan_bytes(r.num_cells, (uint32_t)c->cfg.molecule_capacity);
    if (tmp > r.scan_tmp.n && r.scan_tmp.alloc(tmp)) return fail(MCG_E_NOMEM, "scan scratch");
    r.grid_dirty = false;
  }
  if (r.next_id_dirty) {
    HIP_TRY(hipMemcpy(r.next_id.p, r.next_id_host.data(), r.next_id_host.size() * 4, hipMemcpyHostToDevice));
    r.next_id_dirty = false;
  }
  if (int st = ensure_surface_diffusion(c)) return st;
  if (int st = sync_tiles_to_device(c)) return st;
  if (!r.rxn_counts.p) {
    // the classes are fixed from the first step on
    r.rc_unimol = (uint32_t)r.rxns.size() * 4;
    r.rc_vs = r.rc_unimol + c->num_species * 4;
    r.rc_slots = r.rc_vs + (uint32_t)r.vs.size() * 4;
    r.rc_volumes = std::max<uint32_t>(r.num_volumes, 1);
    const size_t n = (size_t)c->cfg.num_replicas * r.rc_volumes * r.rc_slots;
    if (r.rxn_counts.alloc(std::max<size_t>(n, 1))) return fail(MCG_E_NOMEM, "reaction counts");
    HIP_TRY(hipMemset(r.rxn_counts.p, 0, std::max<size_t>(n, 1) * 8));
  }
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  double move_ms = 0;
  int st;
  if (c->fp32()) {
    auto p = make_pipeline<float>(c, c->xf.p, c->yf.p, c->zf.p, r.sxf.p, r.syf.p, r.szf.p);
    st = run_iterations(c, p, first_iteration, num_iterations, time_up_to_next_barrier, stats, move_ms);
  } else {
    auto p = make_pipeline<double>(c, c->x.p, c->y.p, c->z.p, r.sx.p, r.sy.p, r.sz.p);
    st = run_iterations(c, p, first_iteration, num_iterations, time_up_to_next_barrier, stats, move_ms);
  }
  if (st) return st;
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  HIP_TRY(hipEventSynchronize(c->ev1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_kernel_ms = num_iterations ? move_ms / num_iterations : 0;
  c->last_launches = num_iterations;
  if (stats) {
    stats->kernel_ms = ms;
    stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats->kernel_launches = num_iterations;
  }
  return MCG_OK;
}

int count_species(mcg_ctx* c, uint64_t* counts, bool with_surface) {
  const size_t nc = (size_t)c->cfg.num_replicas * c->num_species;
  DevBuf<uint64_t> d;
  if (d.alloc(std::max<size_t>(nc, 1))) return fail(MCG_E_NOMEM, "count buffer");
  return count_species_into(c, counts, d.p, with_surface, true);
}

int count_species_into(mcg_ctx* c, uint64_t* counts, uint64_t* dcounts, bool with_surface, bool allow_alloc) {
  const size_t nc = (size_t)c->cfg.num_replicas * c->num_species;
  ReactState& r = c->rx;
  if (with_surface)
    if (int st = sync_tiles_to_device(c, allow_alloc)) return st;
  const uint32_t tiles = with_surface ? r.total_tiles : 0u;
  // cleared by a kernel (no memset node if a caller captures the stream)
  HIP_TRY(launch_fill_u64(c->stream, reinterpret_cast<unsigned long long*>(dcounts), 0ull, (uint32_t)nc));
  if (c->n || tiles) {
    if (c->fp32()) {
      ReactPipeline<float> p{};
      p.M = RMolView<float>{c->xf.p, c->yf.p, c->zf.p, c->id.p, c->tag.p, nullptr, nullptr, nullptr};
      p.R.total_tiles = tiles;
      p.tiles = r.tiles.p;
      HIP_TRY(p.count_species(c->stream, (uint32_t)c->n, c->num_species, c->cfg.num_replicas, dcounts));
    } else {
      ReactPipeline<double> p{};
      p.M = RMolView<double>{c->x.p, c->y.p, c->z.p, c->id.p, c->tag.p, nullptr, nullptr, nullptr};
      p.R.total_tiles = tiles;
      p.tiles = r.tiles.p;
      HIP_TRY(p.count_species(c->stream, (uint32_t)c->n, c->num_species, c->cfg.num_replicas, dcounts));
    }
  }
  HIP_TRY(hipMemcpyAsync(counts, dcounts, nc * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MCG_OK;
}

}  // namespace mcg

extern "C" {

namespace {

// Splits a pathway's products into the reactants it keeps and the new
// molecules (outcome_products_random, diffuse_react_event.cpp:2551-2565,
// 2686-2727): reactant a is kept when its species is a product and takes
// the first such product; reactant b likewise takes the next product of its
// species (for a == b, b stays only when the species appears twice).
// Returns kKeepA | kKeepB; b = kInvalid for a unimolecular class.
uint32_t split_products(uint32_t a, uint32_t b, const uint32_t* prods, uint32_t n, uint32_t* out, uint32_t& n_out) {
  bool keep_a = false, keep_b = false;
  uint32_t n_a = 0;
  for (uint32_t q = 0; q < n; ++q) {
    keep_a |= prods[q] == a;
    keep_b |= prods[q] == b;
    n_a += prods[q] == a;
  }
  if (a == b) keep_b = n_a > 1;
  bool used_a = false, used_b = false;
  n_out = 0;
  for (uint32_t q = 0; q < n; ++q) {
    const uint32_t ps = prods[q];
    if (keep_a && !used_a && ps == a) {
      used_a = true;
      continue;
    }
    if (keep_b && !used_b && ps == b) {
      used_b = true;
      continue;
    }
    out[n_out++] = ps;
  }
  return (keep_a ? kKeepA : 0u) | (keep_b ? kKeepB : 0u);
}

// The per-slot buffers of classes that keep a reactant, allocated once.
int ensure_keep_buffers(mcg_ctx* c) {
  ReactState& r = c->rx;
  if (r.kept.p) return MCG_OK;
  const size_t cap = c->cfg.molecule_capacity;
  if (r.kept.alloc(cap) || r.aprop.alloc(cap)) return fail(MCG_E_NOMEM, "buffers of classes that keep a reactant");
  HIP_TRY(hipMemset(r.kept.p, 0, cap * sizeof(uint32_t)));
  return MCG_OK;
}

}  // namespace

int mcg_set_bimol_rxns(mcg_ctx* c, const mcg_bimol_rxn* rxns, uint32_t n) {
  if (!c || (n && !rxns)) return fail(MCG_E_INVALID, "null argument");
  MCG_ALIVE(c);
  if (c->num_species == 0) return fail(MCG_E_STATE, "species must be set first");
  if (c->stepped) return fail(MCG_E_STATE, "reaction classes must be set before the first step");
  if (!c->rx.rxns.empty()) return fail(MCG_E_STATE, "reaction classes are already set");
  if (c->per_replica) return fail(MCG_E_STATE, "reaction classes must be set before per-replica parameters");
  if (n == 0) return MCG_OK;
  const uint32_t ns = c->num_species;
  std::vector<RxnRec> h(n);
  std::vector<int32_t> pair((size_t)ns * ns, -1);
  std::vector<SpeciesRec> sp = c->species_host;
  bool any_keep = false;
  for (uint32_t i = 0; i < n; ++i) {
    const mcg_bimol_rxn& x = rxns[i];
    if (x.reactant_a >= ns || x.reactant_b >= ns) return fail(MCG_E_INVALID, "reactant species out of range");
    if (x.num_pathways < 1 || x.num_pathways > (uint32_t)kMaxPathways) return fail(MCG_E_INVALID, "1..4 pathways");
    if (pair[x.reactant_a * ns + x.reactant_b] >= 0) return fail(MCG_E_INVALID, "two classes for one reactant pair");
    RxnRec& r = h[i];
    std::memset(&r, 0, sizeof(r));
    double last = 0;
    for (uint32_t k = 0; k < x.num_pathways; ++k) {
      if (!(x.cum_probs[k] >= last)) return fail(MCG_E_INVALID, "cumulative probabilities must not decrease");
      last = x.cum_probs[k];
      r.cum_probs[k] = x.cum_probs[k];
      if (x.num_products[k] > (uint32_t)kMaxProducts) return fail(MCG_E_INVALID, "at most 3 products per pathway");
      for (uint32_t q = 0; q < x.num_products[k]; ++q)
        if (x.products[k][q] >= ns) return fail(MCG_E_INVALID, "product species out of range");
      r.keep[k] = split_products(x.reactant_a, x.reactant_b, x.products[k], x.num_products[k], r.products[k],
                                 r.num_products[k]);
      any_keep |= r.keep[k] != 0;
    }
    r.num_pathways = x.num_pathways;
    r.max_fixed_p = r.cum_probs[x.num_pathways - 1];
    r.max_p_f = (float)r.max_fixed_p;
    r.reactant_a = x.reactant_a;
    r.reactant_b = x.reactant_b;
    pair[x.reactant_a * ns + x.reactant_b] = (int32_t)i;
    pair[x.reactant_b * ns + x.reactant_a] = (int32_t)i;
    sp[x.reactant_a].flags |= SPF_CAN_VOLVOL;
    sp[x.reactant_b].flags |= SPF_CAN_VOLVOL;
  }
  ReactState& r = c->rx;
  int st = ensure_react_buffers(c);
  if (st) return st;
  if (any_keep && (st = ensure_keep_buffers(c))) return st;
  if (r.rxns_d.alloc(n)) return fail(MCG_E_NOMEM, "reaction buffers");
  HIP_TRY(hipMemcpy(r.rxns_d.p, h.data(), n * sizeof(RxnRec), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(r.pair_d.p, pair.data(), pair.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->species.p, sp.data(), ns * sizeof(SpeciesRec), hipMemcpyHostToDevice));
  c->species_host = sp;
  r.rxns = h;
  r.pair = pair;
  return MCG_OK;
}

int mcg_set_unimol_rxns(mcg_ctx* c, const mcg_unimol_rxn* rxns, uint32_t n) {
  if (!c || (n && !rxns)) return fail(MCG_E_INVALID, "null argument");
  MCG_ALIVE(c);
  if (c->num_species == 0) return fail(MCG_E_STATE, "species must be set first");
  if (c->stepped) return fail(MCG_E_STATE, "reaction classes must be set before the first step");
  if (c->per_replica) return fail(MCG_E_STATE, "reaction classes must be set before per-replica parameters");
  const uint32_t ns = c->num_species;
  std::vector<UnimolRec> u(ns);
  std::memset(u.data(), 0, ns * sizeof(UnimolRec));
  bool any_keep = false;
  for (uint32_t i = 0; i < n; ++i) {
    const mcg_unimol_rxn& x = rxns[i];
    if (x.reactant >= ns) return fail(MCG_E_INVALID, "reactant species out of range");
    if (c->species_host[x.reactant].flags & SPF_SURFACE)
      return fail(MCG_E_UNSUPPORTED, "unimolecular classes of surface species are not on this path");
    if (u[x.reactant].num_pathways) return fail(MCG_E_INVALID, "one unimolecular class per species");
    if (x.num_pathways < 1 || x.num_pathways > (uint32_t)kMaxPathways) return fail(MCG_E_INVALID, "1..4 pathways");
    double last = 0;
    for (uint32_t k = 0; k < x.num_pathways; ++k) {
      if (!(x.cum_rates[k] >= last)) return fail(MCG_E_INVALID, "cumulative rates must not decrease");
      last = x.cum_rates[k];
      if (x.num_products[k] > (uint32_t)kMaxProducts) return fail(MCG_E_INVALID, "at most 3 products per pathway");
      for (uint32_t q = 0; q < x.num_products[k]; ++q)
        if (x.products[k][q] >= ns) return fail(MCG_E_INVALID, "product species out of range");
      u[x.reactant].keep[k] = split_products(x.reactant, kInvalid, x.products[k], x.num_products[k],
                                             u[x.reactant].products[k], u[x.reactant].num_products[k]) & kKeepA;
      any_keep |= u[x.reactant].keep[k] != 0;
      u[x.reactant].cum[k] = x.cum_rates[k];
    }
    u[x.reactant].num_pathways = x.num_pathways;
  }
  int st = ensure_react_buffers(c);
  if (st) return st;
  if (any_keep && (st = ensure_keep_buffers(c))) return st;
  c->rx.unimol = u;
  HIP_TRY(hipMemcpy(c->rx.unimol_d.p, u.data(), ns * sizeof(UnimolRec), hipMemcpyHostToDevice));
  return MCG_OK;
}

int mcg_set_replica_params(mcg_ctx* c, uint32_t replica, const mcg_species* species, const mcg_bimol_rxn* rxns) {
  if (!c || !species) return fail(MCG_E_INVALID, "null argument");
  MCG_ALIVE(c);
  if (c->num_species == 0) return fail(MCG_E_STATE, "species must be set first");
  if (replica >= c->cfg.num_replicas) return fail(MCG_E_INVALID, "replica out of range");
  const uint32_t ns = c->num_species, R = c->cfg.num_replicas;
  ReactState& r = c->rx;
  const size_t nrx = r.rxns.size();
  if (nrx && !rxns) return fail(MCG_E_INVALID, "the replica's reaction classes are required");
  for (uint32_t s = 0; s < ns; ++s) {
    if (!(species[s].space_step >= 0)) return fail(MCG_E_INVALID, "negative space step");
    if (!valid_time_step(species[s].time_step)) return fail(MCG_E_INVALID, kTimeStepRange);
  }
  std::vector<RxnRec> rep_rx(nrx);
  for (size_t i = 0; i < nrx; ++i) {
    const mcg_bimol_rxn& x = rxns[i];
    RxnRec q = r.rxns[i];
    if (x.reactant_a != q.reactant_a || x.reactant_b != q.reactant_b || x.num_pathways != q.num_pathways)
      return fail(MCG_E_INVALID, "per-replica classes must keep the base table's reactants and pathways");
    double last = 0;
    for (uint32_t k = 0; k < q.num_pathways; ++k) {
      if (x.num_products[k] > (uint32_t)kMaxProducts) return fail(MCG_E_INVALID, "per-replica products differ");
      uint32_t xp[kMaxProducts], nxp = 0;
      const uint32_t xkeep = split_products(x.reactant_a, x.reactant_b, x.products[k], x.num_products[k], xp, nxp);
      if (nxp != q.num_products[k] || xkeep != q.keep[k]) return fail(MCG_E_INVALID, "per-replica products differ");
      for (uint32_t m = 0; m < nxp; ++m)
        if (xp[m] != q.products[k][m]) return fail(MCG_E_INVALID, "per-replica products differ");
      if (!(x.cum_probs[k] >= last)) return fail(MCG_E_INVALID, "cumulative probabilities must not decrease");
      last = x.cum_probs[k];
      q.cum_probs[k] = x.cum_probs[k];
    }
    q.max_fixed_p = q.cum_probs[q.num_pathways - 1];
    q.max_p_f = (float)q.max_fixed_p;
    rep_rx[i] = q;
  }
  if (!c->per_replica) {
    c->species_replicas.clear();
    for (uint32_t k = 0; k < R; ++k)
      c->species_replicas.insert(c->species_replicas.end(), c->species_host.begin(), c->species_host.end());
    r.rxns_replicas.clear();
    for (uint32_t k = 0; k < R; ++k) r.rxns_replicas.insert(r.rxns_replicas.end(), r.rxns.begin(), r.rxns.end());
    if (c->species.alloc((size_t)R * ns)) return fail(MCG_E_NOMEM, "species tables");
    if (nrx && r.rxns_d.alloc((size_t)R * nrx)) return fail(MCG_E_NOMEM, "reaction tables");
    c->per_replica = true;
  }
  for (uint32_t s = 0; s < ns; ++s) {
    SpeciesRec& d = c->species_replicas[(size_t)replica * ns + s];
    d.space_step = species[s].space_step;
    d.time_step = species[s].time_step;
  }
  for (size_t i = 0; i < nrx; ++i) r.rxns_replicas[(size_t)replica * nrx + i] = rep_rx[i];
  HIP_TRY(hipMemcpy(c->species.p, c->species_replicas.data(), c->species_replicas.size() * sizeof(SpeciesRec),
                    hipMemcpyHostToDevice));
  if (nrx)
    HIP_TRY(hipMemcpy(r.rxns_d.p, r.rxns_replicas.data(), r.rxns_replicas.size() * sizeof(RxnRec),
                      hipMemcpyHostToDevice));
  return ensure_step_buffers(c);
}

int mcg_set_rxn_log(mcg_ctx* c, uint64_t capacity) {
  if (!c) return fail(MCG_E_INVALID, "null context");
  MCG_ALIVE(c);
  if (capacity > 0xFFFFFFFFull) return fail(MCG_E_INVALID, "log capacity too large");
  if (c->rx.log.alloc(capacity)) return fail(MCG_E_NOMEM, "reaction log");
  c->rx.log_cap = capacity;
  c->rx.log_count = 0;
  return MCG_OK;
}

int mcg_get_rxn_log(mcg_ctx* c, mcg_rxn_event* out, uint64_t capacity, uint64_t* count) {
  if (!c) return fail(MCG_E_INVALID, "null context");
  MCG_ALIVE(c);
  static_assert(sizeof(mcg_rxn_event) == sizeof(RxnEvent), "event layout");
  const uint64_t kept = std::min(c->rx.log_count, c->rx.log_cap);
  if (count) *count = c->rx.log_count;
  if (out && kept) {
    if (capacity < kept) return fail(MCG_E_CAPACITY, "output capacity too small");
    HIP_TRY(hipMemcpy(out, c->rx.log.p, kept * sizeof(RxnEvent), hipMemcpyDeviceToHost));
  }
  return MCG_OK;
}

int mcg_count_species(mcg_ctx* c, uint64_t* counts) {
  if (!c || !counts) return fail(MCG_E_INVALID, "null argument");
  MCG_ALIVE(c);
  if (c->num_species == 0) return fail(MCG_E_STATE, "species must be set first");
  return count_species(c, counts);
}

int mcg_set_counted_volumes(mcg_ctx* c, uint32_t num_volumes, const uint32_t* wall_inside,
                            const uint32_t* wall_outside) {
  if (!c || (num_volumes && (!wall_inside || !wall_outside))) return fail(MCG_E_INVALID, "null argument");
  MCG_ALIVE(c);
  if (!c->have_geometry) return fail(MCG_E_STATE, "the geometry must be set first");
  if (c->stepped) return fail(MCG_E_STATE, "counted volumes must be set before the first step");
  ReactState& r = c->rx;
  if (r.num_volumes) return fail(MCG_E_STATE, "counted volumes are already set");
  if (num_volumes == 0) return MCG_OK;
  const size_t nw = c->walls_host.size();
  for (size_t w = 0; w < nw; ++w) {
    const bool none = wall_inside[w] == kInvalid && wall_outside[w] == kInvalid;
    if (!none && (wall_inside[w] >= num_volumes || wall_outside[w] >= num_volumes))
      return fail(MCG_E_INVALID, "counted volume index out of range (intersecting objects are not on this path)");
  }
  int st = ensure_react_buffers(c);
  if (st) return st;
  const size_t cap = c->cfg.molecule_capacity;
  if (r.cvol.alloc(cap) || r.scvol.alloc(cap) || r.cv_in.alloc(std::max<size_t>(nw, 1)) ||
      r.cv_out.alloc(std::max<size_t>(nw, 1)))
    return fail(MCG_E_NOMEM, "counted volume buffers");
  // molecules added before: outside every counted object (index 0)
  HIP_TRY(hipMemset(r.cvol.p, 0, cap * 4));
  if (nw) {
    HIP_TRY(hipMemcpy(r.cv_in.p, wall_inside, nw * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(r.cv_out.p, wall_outside, nw * 4, hipMemcpyHostToDevice));
  }
  r.num_volumes = num_volumes;
  return MCG_OK;
}

int mcg_count_species_by_volume(mcg_ctx* c, uint64_t* counts) {
  if (!c || !counts) return fail(MCG_E_INVALID, "null argument");
  MCG_ALIVE(c);
  ReactState& r = c->rx;
  if (r.num_volumes == 0) return fail(MCG_E_STATE, "no counted volumes are set");
  const size_t nc = (size_t)c->cfg.num_replicas * r.num_volumes * c->num_species;
  if (nc > 0xFFFFFFFFull) return fail(MCG_E_LIMIT, "too many counters");
  DevBuf<uint64_t> d;
  if (d.alloc(std::max<size_t>(nc, 1))) return fail(MCG_E_NOMEM, "count buffer");
  HIP_TRY(launch_fill_u64(c->stream, reinterpret_cast<unsigned long long*>(d.p), 0ull, (uint32_t)nc));
  HIP_TRY(launch_count_by_volume(c->stream, r.cvol.p, c->tag.p, (uint32_t)c->n, c->num_species, c->cfg.num_replicas,
                                 r.num_volumes, d.p));
  HIP_TRY(hipMemcpyAsync(counts, d.p, nc * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MCG_OK;
}

int mcg_get_rxn_counts(mcg_ctx* c, uint64_t* counts, uint32_t* dims) {
  if (!c) return fail(MCG_E_INVALID, "null context");
  MCG_ALIVE(c);
  ReactState& r = c->rx;
  const uint32_t nv = std::max<uint32_t>(r.num_volumes, 1);
  const uint32_t slots = ((uint32_t)r.rxns.size() + c->num_species + (uint32_t)r.vs.size()) * 4;
  if (dims) {
    dims[0] = c->cfg.num_replicas;
    dims[1] = nv;
    dims[2] = slots;
  }
  if (!counts) return MCG_OK;
  const size_t n = (size_t)c->cfg.num_replicas * nv * slots;
  if (!r.rxn_counts.p) {  // no step has run
    std::fill(counts, counts + n, 0ull);
    return MCG_OK;
  }
  HIP_TRY(hipMemcpyAsync(counts, r.rxn_counts.p, n * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MCG_OK;
}

int mcg_count_species_noalloc(mcg_ctx* c, uint64_t* counts) {
  if (!c || !counts) return fail(MCG_E_INVALID, "null argument");
  MCG_ALIVE(c);
  if (c->num_species == 0) return fail(MCG_E_STATE, "species must be set first");
  const uint64_t nc = (uint64_t)c->cfg.num_replicas * c->num_species;
  if (nc > kCountScratchMax)
    return fail(MCG_E_LIMIT, "replicas x species = " + std::to_string(nc) + " exceeds the " +
                                 std::to_string(kCountScratchMax) +
                                 " counters of the allocation-free count (use mcg_count_species)");
  return count_species_into(c, counts, c->count_scratch, true, false);
}

}  // extern "C"

namespace mcg {
namespace {

// Grid::initialize (src4/wall.cpp:38-74) from the wall's constants.
WallGrid make_wall_grid(const WallRec64& w, uint32_t tile_base) {
  WallGrid g{};
  uint32_t n_axis = (uint32_t)std::ceil(std::sqrt(w.area));
  if (n_axis < 1) n_axis = 1;
  g.n_axis = n_axis;
  g.num_tiles = n_axis * n_axis;
  g.tile_base = tile_base;
  g.strip_width_rcp = 1 / (w.uv2v / ((double)n_axis));
  g.vert2_slope = w.uv2u / w.uv2v;
  g.fullslope = w.uv1u / w.uv2v;
  g.binding_factor = ((double)g.num_tiles) / w.area;
  g.vert0_u = w.p0x * w.ux + w.p0y * w.uy + w.p0z * w.uz;
  g.vert0_v = w.p0x * w.vx + w.p0y * w.vy + w.p0z * w.vz;
  return g;
}

// The host tile arrays laid out for `total` tiles per replica (new grids
// appended after the existing tiles), molecules kept in their slots.
void relayout_tiles(const ReactState& r, uint32_t R, uint64_t total, std::vector<uint32_t>& nt,
                    std::vector<uint32_t>& nid, std::vector<double>& nu, std::vector<double>& nv,
                    std::vector<double>& nto, std::vector<double>& nb) {
  nt.assign(total * R, kTileEmpty);
  nid.assign(total * R, kInvalid);
  nu.assign(total * R, 0.0);
  nv.assign(total * R, 0.0);
  nto.assign(total * R, 0.0);
  nb.assign(total * R, 0.0);
  const bool pos = r.tu_host.size() == (size_t)r.total_tiles * R;
  for (uint32_t rep = 0; rep < R; ++rep)
    for (uint32_t t = 0; t < r.total_tiles; ++t) {
      const size_t a = rep * total + t, b = (size_t)rep * r.total_tiles + t;
      nt[a] = r.tiles_host[b];
      nid[a] = r.tile_id_host[b];
      if (pos) {
        nu[a] = r.tu_host[b];
        nv[a] = r.tv_host[b];
        nto[a] = r.ttoff_host[b];
        nb[a] = r.tbirth_host[b];
      }
    }
}

}  // namespace

// Brings the host copy of the tiles up to date with the device's.
int download_tiles(mcg_ctx* c) {
  ReactState& r = c->rx;
  if (!r.tiles_on_device || r.tiles_dirty) return MCG_OK;
  const size_t nt = (size_t)r.total_tiles * c->cfg.num_replicas;
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(r.tiles_host.data(), r.tiles.p, nt * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(r.tile_id_host.data(), r.tile_id.p, nt * 4, hipMemcpyDeviceToHost));
  if (r.surf_on && r.tu.n == nt) {
    HIP_TRY(hipMemcpy(r.tu_host.data(), r.tu.p, nt * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(r.tv_host.data(), r.tv.p, nt * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(r.ttoff_host.data(), r.ttoff.p, nt * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(r.tbirth_host.data(), r.tbirth.p, nt * 8, hipMemcpyDeviceToHost));
  }
  return MCG_OK;
}

int sync_tiles_to_device(mcg_ctx* c, bool allow_alloc) {
  ReactState& r = c->rx;
  if (!r.tiles_dirty) return MCG_OK;
  const size_t nt = (size_t)r.total_tiles * c->cfg.num_replicas;
  if (nt > 0xFFFFFFFFull) return fail(MCG_E_LIMIT, "too many tiles (tiles x replicas >= 2^32)");
  if (!allow_alloc && (r.tiles.n != nt || r.grids_d.n != r.grids_host.size()))
    return fail(MCG_E_STATE, "surface molecules on new walls need tile buffers: run mcg_step or "
                             "mcg_count_species once before the allocation-free count");
  if (r.tiles.n != nt) {
    if (r.tiles.alloc(nt) || r.tile_id.alloc(nt) || r.tile_owner.alloc(nt) || r.tile_consumed.alloc(nt))
      return fail(MCG_E_NOMEM, "tile buffers");
  }
  if (r.grids_d.n != r.grids_host.size() && r.grids_d.alloc(r.grids_host.size()))
    return fail(MCG_E_NOMEM, "grid buffer");
  HIP_TRY(hipMemcpy(r.grids_d.p, r.grids_host.data(), r.grids_host.size() * sizeof(WallGrid), hipMemcpyHostToDevice));
  if (r.tile_wall.n != r.total_tiles) {
    if (!allow_alloc) return fail(MCG_E_STATE, "surface molecules on new walls need tile buffers");
    std::vector<uint32_t> tw(r.total_tiles, kInvalid);
    for (size_t w = 0; w < r.grids_host.size(); ++w)
      for (uint32_t t = 0; t < r.grids_host[w].num_tiles; ++t) tw[r.grids_host[w].tile_base + t] = (uint32_t)w;
    if (r.tile_wall.alloc(r.total_tiles)) return fail(MCG_E_NOMEM, "tile buffers");
    if (r.total_tiles) HIP_TRY(hipMemcpy(r.tile_wall.p, tw.data(), tw.size() * 4, hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMemcpy(r.tiles.p, r.tiles_host.data(), nt * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(r.tile_id.p, r.tile_id_host.data(), nt * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(r.tile_owner.p, 0xFF, nt * 4));
  HIP_TRY(hipMemset(r.tile_consumed.p, 0, nt * 4));
  if (r.surf_on) {
    if (r.tu.n != nt) {
      if (!allow_alloc) return fail(MCG_E_STATE, "surface molecules on new walls need tile buffers");
      if (r.tu.alloc(nt) || r.tv.alloc(nt) || r.ttoff.alloc(nt) || r.tbirth.alloc(nt) || r.spu.alloc(nt) ||
          r.spv.alloc(nt) || r.sclaim.alloc(nt) || r.sprop.alloc(nt) || r.splist.alloc(nt) ||
          r.slose0.alloc(nt) || r.slose1.alloc(nt) || r.sdone.alloc(nt) || r.scont0.alloc(nt) ||
          r.scont1.alloc(nt))
        return fail(MCG_E_NOMEM, "surface phase buffers");
      HIP_TRY(launch_fill_claims(c->stream, r.sclaim.p, (uint32_t)nt));
    }
    if (r.sdone.n == nt) {
      // molecules keep their slots' stamps only within an iteration
      HIP_TRY(launch_fill_done(c->stream, r.sdone.p, (uint32_t)nt));
    }
    HIP_TRY(hipMemcpy(r.tu.p, r.tu_host.data(), nt * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(r.tv.p, r.tv_host.data(), nt * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(r.ttoff.p, r.ttoff_host.data(), nt * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(r.tbirth.p, r.tbirth_host.data(), nt * 8, hipMemcpyHostToDevice));
  }
  r.tiles_dirty = false;
  r.tiles_on_device = true;
  return MCG_OK;
}

namespace {

// Edge::reinit_edge_constants (src4/wall.cpp:134-235): the uv transform
// from the forward wall's frame to the backward wall's across the edge
// that starts at the forward wall's vertex i.
bool edge_constants(const WallRec64& wf, const WallRec64& wb, uint32_t i, WallEdgeRec& e) {
  const double fv[3][3] = {{wf.p0x, wf.p0y, wf.p0z}, {wf.p1x, wf.p1y, wf.p1z}, {wf.p2x, wf.p2y, wf.p2z}};
  const uint32_t j = (i + 1) % 3;
  auto dotu = [](const double* a, const double* o, const WallRec64& w) {
    return (a[0] - o[0]) * w.ux + (a[1] - o[1]) * w.uy + (a[2] - o[2]) * w.uz;
  };
  auto dotv = [](const double* a, const double* o, const WallRec64& w) {
    return (a[0] - o[0]) * w.vx + (a[1] - o[1]) * w.vy + (a[2] - o[2]) * w.vz;
  };
  const double b0[3] = {wb.p0x, wb.p0y, wb.p0z};
  const double ofu = dotu(fv[i], fv[0], wf), ofv = dotv(fv[i], fv[0], wf);
  const double ffu = dotu(fv[j], fv[0], wf) - ofu, ffv = dotv(fv[j], fv[0], wf) - ofv;
  if (!(std::fabs(ffu * ffu + ffv * ffv) >= POS_EPS)) return false;
  const double df = 1 / std::sqrt(ffu * ffu + ffv * ffv);
  const double efu = ffu * df, efv = ffv * df, hfu = -efv, hfv = efu;
  const double obu = dotu(fv[i], b0, wb), obv = dotv(fv[i], b0, wb);
  const double fbu = dotu(fv[j], b0, wb) - obu, fbv = dotv(fv[j], b0, wb) - obv;
  if (!(std::fabs(fbu * fbu + fbv * fbv) >= POS_EPS)) return false;
  const double db = 1 / std::sqrt(fbu * fbu + fbv * fbv);
  const double ebu = fbu * db, ebv = fbv * db, hbu = -ebv, hbv = ebu;
  const double m00 = efu * ebu + hfu * hbu, m01 = efv * ebu + hfv * hbu;
  const double m10 = efu * ebv + hfu * hbv, m11 = efv * ebv + hfv * hbv;
  e.cos_t = m00;
  e.sin_t = m01;
  e.tu = obu - (m00 * ofu + m01 * ofv);
  e.tv = obv - (m10 * ofu + m11 * ofv);
  return true;
}

// surface_net (src4/geometry.cpp:256-353) over all walls: an edge of two
// walls traversed in opposite vertex order (compatible_edges, :50-97) links
// them, the wall of smaller index being the forward side.  Walls are
// matched by their vertices' coordinates, over the whole mesh (the
// reference links walls of one object; objects that share no edge give the
// same net).
int build_surface_net(const std::vector<WallRec64>& walls, std::vector<WallEdgeRec>& out) {
  WallEdgeRec none{};
  none.fwd = none.bwd = kInvalid;
  out.assign(3 * walls.size(), none);
  using Key = std::array<double, 6>;
  std::map<Key, std::vector<std::pair<uint32_t, uint32_t>>> groups;
  auto vtx = [&](uint32_t w, uint32_t k) -> std::array<double, 3> {
    const WallRec64& r = walls[w];
    if (k == 0) return {r.p0x, r.p0y, r.p0z};
    if (k == 1) return {r.p1x, r.p1y, r.p1z};
    return {r.p2x, r.p2y, r.p2z};
  };
  for (uint32_t w = 0; w < walls.size(); ++w)
    for (uint32_t j = 0; j < 3; ++j) {
      const auto a = vtx(w, j), b = vtx(w, (j + 1) % 3);
      Key k{a[0], a[1], a[2], b[0], b[1], b[2]}, kr{b[0], b[1], b[2], a[0], a[1], a[2]};
      groups[std::min(k, kr)].push_back({w, j});
    }
  for (const auto& kv : groups) {
    const auto& g = kv.second;
    if (g.size() > 2) return fail(MCG_E_UNSUPPORTED, "an edge shared by more than two walls (surface diffusion)");
    if (g.size() < 2) continue;
    const uint32_t wa = std::min(g[0].first, g[1].first) == g[0].first ? 0 : 1;
    const uint32_t fa = g[wa].first, ea = g[wa].second, fb = g[1 - wa].first, eb = g[1 - wa].second;
    if (!(vtx(fa, ea) == vtx(fb, (eb + 1) % 3) && vtx(fa, (ea + 1) % 3) == vtx(fb, eb) &&
          !(vtx(fa, (ea + 2) % 3) == vtx(fb, (eb + 2) % 3))))
      continue;
    WallEdgeRec e{};
    e.fwd = fa;
    e.bwd = fb;
    if (!edge_constants(walls[fa], walls[fb], ea, e)) return fail(MCG_E_INVALID, "malformed wall pair at an edge");
    out[3 * fa + ea] = e;
    out[3 * fb + eb] = e;
  }
  return MCG_OK;
}

}  // namespace

int ensure_surface_diffusion(mcg_ctx* c) {
  ReactState& r = c->rx;
  if (r.surf_on) return MCG_OK;
  bool on = false;
  for (const SpeciesRec& s : c->per_replica ? c->species_replicas : c->species_host)
    if ((s.flags & SPF_SURFACE) && s.space_step > 0) on = true;
  if (!on || r.total_tiles == 0) return MCG_OK;
  if (int st = download_tiles(c)) return st;
  // every wall gets its grid: a molecule may reach any wall of the mesh
  const uint32_t R = c->cfg.num_replicas;
  const uint32_t nw = (uint32_t)c->walls_host.size();
  if (r.grids_host.size() != nw) r.grids_host.assign(nw, WallGrid{});
  std::vector<WallGrid> g = r.grids_host;
  uint64_t total = r.total_tiles;
  for (uint32_t w = 0; w < nw; ++w) {
    if (g[w].num_tiles) continue;
    g[w] = make_wall_grid(c->walls_host[w], (uint32_t)total);
    total += g[w].num_tiles;
    if (total * R > 0xFFFFFFFFull) return fail(MCG_E_LIMIT, "too many tiles (tiles x replicas >= 2^32)");
  }
  std::vector<uint32_t> nt, nid;
  std::vector<double> nu, nv, nto, nb;
  relayout_tiles(r, R, total, nt, nid, nu, nv, nto, nb);
  std::vector<WallEdgeRec> edges;
  if (int st = build_surface_net(c->walls_host, edges)) return st;
  if (r.edges_d.alloc(edges.size())) return fail(MCG_E_NOMEM, "surface net");
  HIP_TRY(hipMemcpy(r.edges_d.p, edges.data(), edges.size() * sizeof(WallEdgeRec), hipMemcpyHostToDevice));
  r.grids_host = g;
  r.tiles_host.swap(nt);
  r.tile_id_host.swap(nid);
  r.tu_host.swap(nu);
  r.tv_host.swap(nv);
  r.ttoff_host.swap(nto);
  r.tbirth_host.swap(nb);
  r.total_tiles = (uint32_t)total;
  r.tiles_dirty = true;
  r.surf_on = true;
  return MCG_OK;
}

int surface_phase(mcg_ctx* c, uint32_t iteration, double barrier, uint32_t* err) {
  ReactState& r = c->rx;
  uint32_t* h = r.hdr_host;
  const uint32_t R = c->cfg.num_replicas;
  SurfPhase P{};
  P.walls = c->walls64.p;
  P.grids = r.grids_d.p;
  P.edges = r.edges_d.p;
  P.tile_wall = r.tile_wall.p;
  P.species = c->species_table();
  P.tiles = r.tiles.p;
  P.tile_id = r.tile_id.p;
  P.tile_consumed = r.tile_consumed.p;
  P.tu = r.tu.p;
  P.tv = r.tv.p;
  P.ttoff = r.ttoff.p;
  P.birth = r.tbirth.p;
  P.claim = r.sclaim.p;
  P.prop = r.sprop.p;
  P.pu = r.spu.p;
  P.pv = r.spv.p;
  P.plist = r.splist.p;
  P.done = r.sdone.p;
  P.hdr = r.hdr.p;
  P.err = err;
  P.total_tiles = r.total_tiles;
  P.num_replicas = R;
  P.seed = c->cfg.seed;
  P.iteration = iteration;
  // never 0xFFFFFFF0, the high bits of a slot no iteration has stamped
  P.stamp_base = ((iteration + 1u) % 0x0FFFFFFFu) << 4;
  P.barrier = barrier;
  const uint32_t n_all = r.total_tiles * R;
  uint32_t* lists[2] = {r.slose0.p, r.slose1.p};
  const uint32_t words[2] = {kHdrSurfLose0, kHdrSurfLose1};
  uint32_t* conts[2] = {r.scont0.p, r.scont1.p};
  const uint32_t cwords[2] = {kHdrSurfCont0, kHdrSurfCont1};
  // pass k: the molecules' k-th steps of the iteration; the first over every
  // slot, the others over the molecules the previous pass listed
  const uint32_t* in_list = nullptr;
  const uint32_t* in_count = nullptr;
  uint32_t n_in = n_all;
  for (uint32_t pass = 0; pass < kSurfPasses; ++pass) {
    P.pass = pass;
    P.clist = conts[pass & 1];
    P.ccount = r.hdr.p + cwords[pass & 1];
    HIP_TRY(launch_zero_word(c->stream, P.ccount));
    HIP_TRY(surf_round(c->stream, P, in_list, in_count, n_in, lists[0], r.hdr.p + words[0]));
    for (uint32_t round = 0;; ++round) {
      if (int st = read_hdr(c)) return st;
      const uint32_t nl = h[words[round & 1]];
      if (nl == 0) break;
      if (round > 100000) return fail(MCG_E_LIMIT, "surface moves did not settle");
      HIP_TRY(surf_round(c->stream, P, lists[round & 1], r.hdr.p + words[round & 1], nl, lists[(round + 1) & 1],
                         r.hdr.p + words[(round + 1) & 1]));
    }
    n_in = h[cwords[pass & 1]];
    if (n_in == 0) return MCG_OK;
    in_list = conts[pass & 1];
    in_count = r.hdr.p + cwords[pass & 1];
  }
  return MCG_OK;  // a molecule listed past the last pass raised ERR_PASSES
}

}  // namespace mcg

extern "C" {

int mcg_set_volsurf_rxns(mcg_ctx* c, const mcg_volsurf_rxn* rxns, uint32_t n) {
  if (!c || (n && !rxns)) return fail(MCG_E_INVALID, "null argument");
  MCG_ALIVE(c);
  if (c->num_species == 0) return fail(MCG_E_STATE, "species must be set first");
  if (c->stepped) return fail(MCG_E_STATE, "reaction classes must be set before the first step");
  if (c->per_replica) return fail(MCG_E_STATE, "reaction classes must be set before per-replica parameters");
  ReactState& r = c->rx;
  if (!r.vs.empty()) return fail(MCG_E_STATE, "volume-surface reaction classes are already set");
  if (n == 0) return MCG_OK;
  const uint32_t ns = c->num_species;
  std::vector<VolSurfRec> h(n);
  std::vector<int32_t> pair((size_t)ns * ns, -1);
  std::vector<SpeciesRec> sp = c->species_host;
  auto orient_ok = [](int32_t o) { return o >= -1 && o <= 1; };
  for (uint32_t i = 0; i < n; ++i) {
    const mcg_volsurf_rxn& x = rxns[i];
    VolSurfRec& v = h[i];
    std::memset(&v, 0, sizeof(v));
    if (x.vol_species >= ns || x.surf_species >= ns) return fail(MCG_E_INVALID, "reactant species out of range");
    if (pair[x.vol_species * ns + x.surf_species] >= 0)
      return fail(MCG_E_INVALID, "one class per (volume, surface) species pair");
    if (!orient_ok(x.geom_vol) || !orient_ok(x.geom_surf)) return fail(MCG_E_INVALID, "orientations are 0, +1 or -1");
    if (x.num_pathways < 1 || x.num_pathways > (uint32_t)kMaxPathways) return fail(MCG_E_INVALID, "1..4 pathways");
    double last = 0;
    for (uint32_t k = 0; k < x.num_pathways; ++k) {
      if (!(x.cum_probs[k] >= last)) return fail(MCG_E_INVALID, "cumulative probabilities must not decrease");
      last = x.cum_probs[k];
      v.cum_probs[k] = x.cum_probs[k];
      const uint32_t np = x.num_products[k];
      if (np > (uint32_t)kMaxProducts) return fail(MCG_E_INVALID, "at most 3 products per pathway");
      uint32_t new_surf = 0;
      for (uint32_t q = 0; q < np; ++q) {
        const uint32_t ps = x.products[k][q];
        const bool surf = (x.surf_products[k] >> q) & 1u;
        if (ps >= ns || ps > 0xFFu) return fail(MCG_E_INVALID, "product species out of range");
        if (!orient_ok(x.product_orient[k][q])) return fail(MCG_E_INVALID, "orientations are 0, +1 or -1");
        v.products[k][q] = ps;
        v.product_orient[k][q] = x.product_orient[k][q];
        if (surf) v.surf_mask[k] |= 1u << q;
        // the first surface product of the surface reactant's species is that reactant
        if (surf && ps == x.surf_species && !v.keep_surf[k]) {
          v.keep_surf[k] = q + 1;
          continue;
        }
        // and the first volume product of the volume reactant's species is
        // the volume reactant, which goes on with its step
        if (!surf && ps == x.vol_species && !v.keep_vol[k]) {
          v.keep_vol[k] = q + 1;
          continue;
        }
        if (surf) ++new_surf;
      }
      if (new_surf > (v.keep_surf[k] ? 0u : 1u))
        return fail(MCG_E_UNSUPPORTED, "surface products beyond the reactant's tile need a neighbour-tile search");
      v.num_products[k] = np;
      v.num_new[k] = np - (v.keep_surf[k] ? 1u : 0u) - (v.keep_vol[k] ? 1u : 0u);
    }
    v.num_pathways = x.num_pathways;
    v.max_fixed_p = v.cum_probs[x.num_pathways - 1];
    v.max_p_f = (float)v.max_fixed_p;
    v.geom_vol = x.geom_vol;
    v.geom_surf = x.geom_surf;
    v.surf_species = x.surf_species;
    pair[x.vol_species * ns + x.surf_species] = (int32_t)i;
    sp[x.vol_species].flags |= SPF_CAN_VOLSURF;
  }
  int st = ensure_react_buffers(c);
  if (st) return st;
  if (r.vs_d.alloc(n) || r.pair_vs_d.alloc(pair.size())) return fail(MCG_E_NOMEM, "reaction buffers");
  bool vol_products = false, keep_vol = false;
  for (const VolSurfRec& v : h)
    for (uint32_t k = 0; k < v.num_pathways; ++k) {
      keep_vol |= v.keep_vol[k] != 0;
      for (uint32_t q = 0; q < v.num_products[k]; ++q)
        vol_products |= !((v.surf_mask[k] >> q) & 1u) && v.keep_vol[k] != q + 1;
    }
  if (keep_vol && (st = ensure_keep_buffers(c))) return st;
  if (vol_products && !r.born.p) {
    const size_t cap = c->cfg.molecule_capacity;
    if (r.born.alloc(cap)) return fail(MCG_E_NOMEM, "reaction buffers");
    HIP_TRY(hipMemset(r.born.p, 0, cap * 4));
  }
  HIP_TRY(hipMemcpy(r.vs_d.p, h.data(), n * sizeof(VolSurfRec), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(r.pair_vs_d.p, pair.data(), pair.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->species.p, sp.data(), ns * sizeof(SpeciesRec), hipMemcpyHostToDevice));
  c->species_host = sp;
  r.vs = h;
  r.pair_vs = pair;
  return MCG_OK;
}

int mcg_add_surface_molecules(mcg_ctx* c, uint32_t replica, const uint32_t* ids, const uint32_t* species,
                              const uint32_t* walls, const uint32_t* tiles, const int32_t* orients, const double* uv,
                              const double* dtimes, const uint32_t* flags, const double* birthdays, uint64_t count) {
  if (!c || (count && (!ids || !species || !walls || !tiles || !orients))) return fail(MCG_E_INVALID, "null argument");
  MCG_ALIVE(c);
  if (!c->have_geometry) return fail(MCG_E_STATE, "surface molecules need the geometry first");
  if (c->num_species == 0) return fail(MCG_E_STATE, "species must be set first");
  if (replica >= c->cfg.num_replicas) return fail(MCG_E_INVALID, "replica out of range");
  if (count == 0) return MCG_OK;
  ReactState& r = c->rx;
  const uint32_t R = c->cfg.num_replicas;
  const uint32_t nw = (uint32_t)c->walls_host.size();
  for (uint64_t i = 0; i < count; ++i) {
    if (walls[i] >= nw) return fail(MCG_E_INVALID, "wall index out of range");
    if (species[i] >= c->num_species) return fail(MCG_E_INVALID, "species id out of range");
    if (orients[i] != 1 && orients[i] != -1) return fail(MCG_E_INVALID, "surface molecules are oriented +1 or -1");
    if (dtimes && !(dtimes[i] >= (double)c->clock)) return fail(MCG_E_INVALID, "diffusion time before the engine's clock");
    if (flags && (flags[i] & ~MCG_MOLF_MATURE)) return fail(MCG_E_INVALID, "surface molecules take the mature flag only");
    if (birthdays && birthdays[i] != kTimeInvalid && !(birthdays[i] <= (dtimes ? dtimes[i] : (double)c->clock)))
      return fail(MCG_E_INVALID, "birthday after the diffusion time");
  }
  if (int st = download_tiles(c)) return st;
  // grids for walls without one, appended after the existing tiles
  if (r.grids_host.size() != nw) r.grids_host.assign(nw, WallGrid{});
  uint64_t total = r.total_tiles;
  std::vector<WallGrid> g = r.grids_host;
  for (uint64_t i = 0; i < count; ++i) {
    WallGrid& w = g[walls[i]];
    if (w.num_tiles) continue;
    w = make_wall_grid(c->walls_host[walls[i]], (uint32_t)total);
    total += w.num_tiles;
    if (total * R > 0xFFFFFFFFull) return fail(MCG_E_LIMIT, "too many tiles (tiles x replicas >= 2^32)");
  }
  for (uint64_t i = 0; i < count; ++i)
    if (tiles[i] >= g[walls[i]].num_tiles) return fail(MCG_E_INVALID, "tile index out of range");
  std::vector<uint32_t> nt, nid;
  std::vector<double> nu, nv, nto, nb;
  relayout_tiles(r, R, total, nt, nid, nu, nv, nto, nb);
  for (uint64_t i = 0; i < count; ++i) {
    const size_t slot = replica * total + g[walls[i]].tile_base + tiles[i];
    if (nt[slot] != kTileEmpty) return fail(MCG_E_INVALID, "tile already occupied");
    nt[slot] = make_tile(species[i], orients[i]) | ((flags && (flags[i] & MCG_MOLF_MATURE)) ? kTileMature : 0u);
    nid[slot] = ids[i];
    // the position on the wall: given, else the tile's centre (grid2uv)
    if (uv) {
      nu[slot] = uv[2 * i];
      nv[slot] = uv[2 * i + 1];
    } else {
      grid2uv(c->walls_host[walls[i]], g[walls[i]], tiles[i], nu[slot], nv[slot]);
    }
    nto[slot] = dtimes ? dtimes[i] - (double)c->clock : 0.0;
    nb[slot] = (birthdays && birthdays[i] != kTimeInvalid) ? birthdays[i] : (double)c->clock;
  }
  // the id counters (the device holds them once reactive steps have run)
  if (r.on && !r.next_id_dirty && r.next_id_host.size() == R)
    HIP_TRY(hipMemcpy(r.next_id_host.data(), r.next_id.p, R * 4, hipMemcpyDeviceToHost));
  if (r.next_id_host.size() != R) r.next_id_host.assign(R, 0);
  for (uint64_t i = 0; i < count; ++i) r.next_id_host[replica] = std::max(r.next_id_host[replica], ids[i] + 1);
  r.next_id_dirty = true;
  r.grids_host = g;
  r.tiles_host.swap(nt);
  r.tile_id_host.swap(nid);
  r.tu_host.swap(nu);
  r.tv_host.swap(nv);
  r.ttoff_host.swap(nto);
  r.tbirth_host.swap(nb);
  r.total_tiles = (uint32_t)total;
  r.tiles_dirty = true;
  return MCG_OK;
}

int mcg_get_surface_molecules(mcg_ctx* c, uint32_t* ids, uint32_t* species, uint32_t* replica, uint32_t* walls,
                              uint32_t* tiles, int32_t* orients, double* uv, double* dtimes, uint32_t* flags,
                              double* birthdays, uint64_t capacity, uint64_t* count) {
  if (!c) return fail(MCG_E_INVALID, "null context");
  MCG_ALIVE(c);
  ReactState& r = c->rx;
  if (int st = download_tiles(c)) return st;
  uint64_t k = 0;
  for (uint32_t rep = 0; rep < c->cfg.num_replicas; ++rep)
    for (uint32_t w = 0; w < (uint32_t)r.grids_host.size(); ++w) {
      const WallGrid& g = r.grids_host[w];
      for (uint32_t t = 0; t < g.num_tiles; ++t) {
        const size_t slot = (size_t)rep * r.total_tiles + g.tile_base + t;
        const uint32_t tw = r.tiles_host[slot];
        if (tw == kTileEmpty) continue;
        if (k < capacity) {
          if (ids) ids[k] = r.tile_id_host[slot];
          if (species) species[k] = tile_species(tw);
          if (replica) replica[k] = rep;
          if (walls) walls[k] = w;
          if (tiles) tiles[k] = t;
          if (orients) orients[k] = tile_orient(tw);
          if (uv) {
            uv[2 * k] = r.tu_host[slot];
            uv[2 * k + 1] = r.tv_host[slot];
          }
          if (dtimes) dtimes[k] = (double)c->clock + r.ttoff_host[slot];
          if (flags) flags[k] = (tw & kTileMature) ? MCG_MOLF_MATURE : 0u;
          if (birthdays) birthdays[k] = r.tbirth_host[slot];
        }
        ++k;
      }
    }
  if (count) *count = k;
  return MCG_OK;
}

}  // extern "C"
