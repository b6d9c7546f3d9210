// The engine context and the C ABI declared in include/mcg.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "engine.h"
#include "geometry.h"

using namespace mcg;

namespace mcg {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int check_alive(const mcg_ctx* c) {
  if (c && c->failed) return fail(MCG_E_STATE, "the context failed in an earlier step (" + c->fail_msg + ")");
  return MCG_OK;
}

}  // namespace mcg

mcg_ctx::~mcg_ctx() {
  if (rx.hdr_host) (void)hipHostFree(rx.hdr_host);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (evm0) (void)hipEventDestroy(evm0);
  if (evm1) (void)hipEventDestroy(evm1);
  if (stream) (void)hipStreamDestroy(stream);
}

namespace {

// The FP32 position of an FP64 point, kept strictly on the point's side of
// every wall plane of its subpartition (a rounded point on a plane would let
// the first move leave through the wall: collide_wall misses a move that
// starts on the plane, collision_utils.inl:647-665).
bool round_same_side_host(const mcg_ctx* c, double px, double py, double pz, float& fx, float& fy, float& fz) {
  const PartitionParams& P = c->P;
  fx = (float)px;
  fy = (float)py;
  fz = (float)pz;
  const int sx = (int)((px - P.ox) * P.sp_rcp), sy = (int)((py - P.oy) * P.sp_rcp), sz = (int)((pz - P.oz) * P.sp_rcp);
  if (sx < 0 || sy < 0 || sz < 0 || sx >= (int)P.n || sy >= (int)P.n || sz >= (int)P.n) return true;
  const uint32_t s = (uint32_t)sx + (uint32_t)sy * P.n + (uint32_t)sz * P.n2;
  for (int pass = 0; pass < 8; ++pass) {
    bool moved = false;
    for (uint32_t k = c->off_host[s]; k < c->off_host[s + 1]; ++k) {
      const WallRec64& w = c->walls_host[c->idx_host[k]];
      const double sd = ((w.nx * px + w.ny * py) + w.nz * pz) - w.d;
      const double sf = ((w.nx * (double)fx + w.ny * (double)fy) + w.nz * (double)fz) - w.d;
      if (sd == 0 || (sd > 0 && sf > 0) || (sd < 0 && sf < 0)) continue;
      const float dir = sd > 0 ? 1.0f : -1.0f;
      if (w.nx * dir > 0) fx = std::nextafter(fx, 3.4e38f);
      else if (w.nx * dir < 0) fx = std::nextafter(fx, -3.4e38f);
      if (w.ny * dir > 0) fy = std::nextafter(fy, 3.4e38f);
      else if (w.ny * dir < 0) fy = std::nextafter(fy, -3.4e38f);
      if (w.nz * dir > 0) fz = std::nextafter(fz, 3.4e38f);
      else if (w.nz * dir < 0) fz = std::nextafter(fz, -3.4e38f);
      moved = true;
    }
    if (!moved) return true;
  }
  return false;
}

}  // namespace

extern "C" {

const char* mcg_last_error(void) { return g_err.c_str(); }
int mcg_abi_version(void) { return MCG_ABI_VERSION; }

int mcg_create(const mcg_config* cfg, mcg_ctx** out) {
  if (!cfg || !out) return fail(MCG_E_INVALID, "null argument");
  *out = nullptr;
  if (cfg->num_subparts_per_edge == 0 || cfg->num_subparts_per_edge > 1024)
    return fail(MCG_E_INVALID, "num_subparts_per_edge out of range");
  if (!(cfg->subpart_edge_length > 0) || !(cfg->partition_edge_length > 0))
    return fail(MCG_E_INVALID, "non-positive partition or subpartition length");
  if (cfg->precision != MCG_PREC_FP64 && cfg->precision != MCG_PREC_FP32_FALLBACK)
    return fail(MCG_E_INVALID, "unknown precision");
  if (cfg->num_replicas == 0 || cfg->num_replicas > 65536) return fail(MCG_E_INVALID, "num_replicas out of range");
  if (cfg->molecule_capacity == 0 || cfg->molecule_capacity >= 0xFFFFFFFFull)
    return fail(MCG_E_INVALID, "molecule_capacity out of range");
  // MCell4Converter's check (mcell4_converter.cpp:121-127): the expanded
  // neighbour collection of a subpartition reaches one neighbour only
  if (!(cfg->rxn_radius_3d * 1.41421356238 < cfg->subpart_edge_length / 2))
    return fail(MCG_E_INVALID, "reaction radius multiplied by sqrt(2) must be less than half of the subpartition edge");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (ndev == 0) return fail(MCG_E_HIP, "no HIP device");
  auto* c = new mcg_ctx();
  c->cfg = *cfg;
  PartitionParams& P = c->P;
  P.ox = cfg->origin[0]; P.oy = cfg->origin[1]; P.oz = cfg->origin[2];
  P.edge = cfg->partition_edge_length;
  P.qx = P.ox + P.edge; P.qy = P.oy + P.edge; P.qz = P.oz + P.edge;  // partition.cpp:57
  P.sp_len = cfg->subpart_edge_length;
  P.sp_rcp = cfg->subpart_edge_length_rcp;
  P.rxn_radius = cfg->rxn_radius_3d;
  P.n = cfg->num_subparts_per_edge;
  P.n2 = P.n * P.n;
  P.use_expanded_list = cfg->use_expanded_list ? 1u : 0u;
  auto cleanup_fail = [&](int code) {
    delete c;
    return code;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess ||
      (e = hipEventCreate(&c->evm0)) != hipSuccess || (e = hipEventCreate(&c->evm1)) != hipSuccess)
    return cleanup_fail(fail(MCG_E_HIP, std::string("stream/event creation: ") + hipGetErrorString(e)));
  size_t cap = cfg->molecule_capacity;
  bool pos_fail = c->fp32() ? (c->xf.alloc(cap) || c->yf.alloc(cap) || c->zf.alloc(cap))
                            : (c->x.alloc(cap) || c->y.alloc(cap) || c->z.alloc(cap));
  if (pos_fail || c->id.alloc(cap) || c->tag.alloc(cap) || c->err.alloc(1) ||
      c->queue.alloc((size_t)kNumQueueShards * shard_capacity((uint32_t)cap, kNumQueueShards / 2)) ||
      c->qcount.alloc(kNumQueueShards) || c->counters.alloc(CTR_WORDS) ||
      c->mqueue.alloc((size_t)kNumQueueShards * shard_capacity((uint32_t)cap)) || c->mqcount.alloc(kNumQueueShards))
    return cleanup_fail(fail(MCG_E_NOMEM, "molecule buffers"));
  if ((e = count_scratch_address(&c->count_scratch)) != hipSuccess)
    return cleanup_fail(fail(MCG_E_HIP, std::string("count array address: ") + hipGetErrorString(e)));
  c->clock = cfg->start_iteration;
  double m = 0;
  for (double v : {P.ox, P.oy, P.oz, P.qx, P.qy, P.qz}) m = std::max(m, std::fabs(v));
  c->max_abs_coord = (float)m;
  if (cfg->flags & MCG_FLAG_TRACE_HITS) {
    if (c->trace_wall.alloc(cap * kTraceHits) || c->trace_count.alloc(cap))
      return cleanup_fail(fail(MCG_E_NOMEM, "trace buffers"));
  }
  if ((e = hipMemsetAsync(c->err.p, 0, sizeof(uint32_t), c->stream)) != hipSuccess)
    return cleanup_fail(fail(MCG_E_HIP, hipGetErrorString(e)));
  *out = c;
  return MCG_OK;
}

int mcg_destroy(mcg_ctx* ctx) {
  if (!ctx) return MCG_OK;
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
  return MCG_OK;
}

int mcg_set_geometry(mcg_ctx* c, const double* vertices, uint32_t nv, const uint32_t* walls, uint32_t nw,
                     const uint32_t* wall_class) {
  if (!c || (nv && !vertices) || (nw && !walls)) return fail(MCG_E_INVALID, "null argument");
  MCG_ALIVE(c);
  for (uint32_t i = 0; i < nw; ++i) {
    if (walls[3 * i] >= nv || walls[3 * i + 1] >= nv || walls[3 * i + 2] >= nv)
      return fail(MCG_E_INVALID, "wall vertex index out of range");
  }
  const PartitionParams& P = c->P;
  for (uint32_t i = 0; i < nv; ++i) {
    const double* v = vertices + 3 * i;
    if (!(v[0] >= P.ox && v[1] >= P.oy && v[2] >= P.oz && v[0] < P.qx && v[1] < P.qy && v[2] < P.qz))
      return fail(MCG_E_INVALID, "vertex outside the partition");
  }
  c->vert_host.assign(vertices, vertices + 3 * (size_t)nv);
  c->wall_class_host.assign(nw, 0u);
  if (wall_class) c->wall_class_host.assign(wall_class, wall_class + nw);
  for (uint32_t v : c->wall_class_host)
    if (c->sc_on && v > c->num_surface_classes) return fail(MCG_E_INVALID, "wall class without a surface class");
  if (c->wall_class_d.alloc(std::max<size_t>(nw, 1))) return fail(MCG_E_NOMEM, "wall classes");
  if (nw)
    HIP_TRY(hipMemcpy(c->wall_class_d.p, c->wall_class_host.data(), nw * 4, hipMemcpyHostToDevice));
  c->rx.grid_dirty = true;
  c->rx.reset_drain_form();
  c->walls_host.resize(nw);
  for (uint32_t i = 0; i < nw; ++i)
    c->walls_host[i] = wall_constants(vertices + 3 * walls[3 * i], vertices + 3 * walls[3 * i + 1],
                                      vertices + 3 * walls[3 * i + 2]);
  bin_walls(P, c->walls_host, c->off_host, c->idx_host);
  std::vector<WallRec32> w32(nw);
  for (uint32_t i = 0; i < nw; ++i) w32[i] = wall_to_fp32(c->walls_host[i]);
  if (c->walls64.alloc(std::max<size_t>(nw, 1)) || c->walls32.alloc(std::max<size_t>(nw, 1)) ||
      c->off.alloc(c->off_host.size()) || c->idx.alloc(std::max<size_t>(c->idx_host.size(), 1)))
    return fail(MCG_E_NOMEM, "wall buffers");
  if (nw) {
    HIP_TRY(hipMemcpy(c->walls64.p, c->walls_host.data(), nw * sizeof(WallRec64), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->walls32.p, w32.data(), nw * sizeof(WallRec32), hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMemcpy(c->off.p, c->off_host.data(), c->off_host.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (!c->idx_host.empty())
    HIP_TRY(hipMemcpy(c->idx.p, c->idx_host.data(), c->idx_host.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  // wall records repeated in list order: a subpartition's walls are one
  // contiguous read, without the index indirection
  size_t ne = c->idx_host.size();
  std::vector<WallRec64> e64(std::max<size_t>(ne, 1));
  std::vector<WallRec32> e32(std::max<size_t>(ne, 1));
  for (size_t k = 0; k < ne; ++k) {
    e64[k] = c->walls_host[c->idx_host[k]];
    e32[k] = w32[c->idx_host[k]];
  }
  // sub-cell box of every entry: the wall's bounding box within the entry's subpartition
  std::vector<uint32_t> eb;
  entry_boxes(P, c->walls_host, c->off_host, c->idx_host, eb);
  if (c->ent64.alloc(e64.size()) || c->ent32.alloc(e32.size()) || c->ent_box.alloc(eb.size()))
    return fail(MCG_E_NOMEM, "wall entry buffers");
  HIP_TRY(hipMemcpy(c->ent64.p, e64.data(), e64.size() * sizeof(WallRec64), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->ent32.p, e32.data(), e32.size() * sizeof(WallRec32), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->ent_box.p, eb.data(), eb.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  c->have_geometry = true;
  return MCG_OK;
}

int mcg_set_species(mcg_ctx* c, const mcg_species* sp, uint32_t ns) {
  if (!c || !sp || ns == 0 || ns > 256) return fail(MCG_E_INVALID, "bad species table");
  MCG_ALIVE(c);
  std::vector<SpeciesRec> h(ns);
  for (uint32_t i = 0; i < ns; ++i) {
    if (!(sp[i].space_step >= 0)) return fail(MCG_E_INVALID, "negative space step");
    if (!valid_time_step(sp[i].time_step)) return fail(MCG_E_INVALID, kTimeStepRange);
    h[i].space_step = sp[i].space_step;
    h[i].time_step = sp[i].time_step;
    h[i].flags = sp[i].flags;
  }
  if (c->rx.on || c->per_replica)
    return fail(MCG_E_STATE, "species must be set before reaction classes and per-replica parameters");
  if (c->species.alloc(ns)) return fail(MCG_E_NOMEM, "species buffer");
  HIP_TRY(hipMemcpy(c->species.p, h.data(), ns * sizeof(SpeciesRec), hipMemcpyHostToDevice));
  c->species_host = h;
  c->num_species = ns;
  return MCG_OK;
}

int mcg_add_volume_molecules(mcg_ctx* c, uint32_t replica, const uint32_t* ids, const uint32_t* species,
                             const double* pos, const double* dtimes, const double* utimes, const uint32_t* flags,
                             const uint32_t* cvols, const double* birthdays, uint64_t count) {
  if (!c || (count && (!ids || !species || !pos))) return fail(MCG_E_INVALID, "null argument");
  MCG_ALIVE(c);
  if (replica >= c->cfg.num_replicas) return fail(MCG_E_INVALID, "replica out of range");
  if (c->n + count > c->cfg.molecule_capacity) return fail(MCG_E_CAPACITY, "molecule capacity exceeded");
  const PartitionParams& P = c->P;
  const double clock = (double)c->clock;
  std::vector<double> hx(count), hy(count), hz(count), htoff(count, 0.0), hut(count, kTimeInvalid),
      hb(count, clock);
  std::vector<uint32_t> ht(count);
  bool any_toff = false, any_time = false;
  for (uint64_t i = 0; i < count; ++i) {
    const double* p = pos + 3 * i;
    if (!(p[0] >= P.ox && p[1] >= P.oy && p[2] >= P.oz && p[0] < P.qx && p[1] < P.qy && p[2] < P.qz))
      return fail(MCG_E_INVALID, "molecule outside the partition (partition.h:559-563)");
    if (species[i] >= c->num_species) return fail(MCG_E_INVALID, "species id out of range");
    hx[i] = p[0]; hy[i] = p[1]; hz[i] = p[2];
    uint32_t f = MF_SCHEDULE_UNIMOL;
    if (flags) {
      if (flags[i] & ~(MCG_MOLF_SCHEDULE_UNIMOL | MCG_MOLF_MATURE)) return fail(MCG_E_INVALID, "unknown molecule flag");
      f = (flags[i] & MCG_MOLF_SCHEDULE_UNIMOL) ? MF_SCHEDULE_UNIMOL : 0u;
      any_time |= f == 0;
      if (flags[i] & MCG_MOLF_MATURE) f |= MF_MATURE;
    }
    if (birthdays && birthdays[i] != kTimeInvalid) {
      // Molecule::birthday: not after the molecule's diffusion time
      if (!(birthdays[i] <= (dtimes ? dtimes[i] : clock))) return fail(MCG_E_INVALID, "birthday after the diffusion time");
      hb[i] = birthdays[i];
    }
    if (dtimes) {
      // Molecule::diffusion_time as an offset past the clock (exact: both lie
      // within one small range of whole iterations)
      if (!(dtimes[i] >= clock)) return fail(MCG_E_INVALID, "diffusion time before the engine's clock");
      htoff[i] = dtimes[i] >= kTimeForever ? HUGE_VAL : dtimes[i] - clock;
      if (htoff[i] != 0.0) {
        any_toff = true;
        f |= MF_TOFF;
      }
    }
    if (utimes) {
      hut[i] = utimes[i];
      any_time |= utimes[i] != kTimeInvalid;
    }
    if (cvols && c->rx.num_volumes && cvols[i] >= c->rx.num_volumes)
      return fail(MCG_E_INVALID, "counted volume index out of range");
    ht[i] = make_tag(species[i], f, replica);
  }
  // offsets and drawn lifetimes live in the reactive path's arrays, and
  // so do birthdays (species time steps other than one iteration)
  if (any_toff || any_time)
    if (int st = ensure_react_buffers(c)) return st;
  if (int st = ensure_step_buffers(c)) return st;
  if (c->rx.on && count) {
    // the device holds the id counters once steps have run
    if (!c->rx.next_id_dirty)
      HIP_TRY(hipMemcpy(c->rx.next_id_host.data(), c->rx.next_id.p, c->rx.next_id_host.size() * 4,
                        hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpyAsync(c->rx.toff.p + c->n, htoff.data(), count * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipMemcpyAsync(c->rx.utime.p + c->n, hut.data(), count * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
    if (c->rx.num_volumes) {
      std::vector<uint32_t> hcv(count, 0u);
      if (cvols) std::copy(cvols, cvols + count, hcv.begin());
      HIP_TRY(hipMemcpyAsync(c->rx.cvol.p + c->n, hcv.data(), count * 4, hipMemcpyHostToDevice, c->stream));
    }
    if (c->rx.birth.p)
      HIP_TRY(hipMemcpyAsync(c->rx.birth.p + c->n, hb.data(), count * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  if (c->rx.next_id_host.size() != c->cfg.num_replicas) c->rx.next_id_host.assign(c->cfg.num_replicas, 0);
  for (uint64_t i = 0; i < count; ++i) {
    c->rx.next_id_host[replica] = std::max(c->rx.next_id_host[replica], ids[i] + 1);
    for (int a = 0; a < 3; ++a) {
      c->mol_lo[a] = std::min(c->mol_lo[a], pos[3 * i + a]);
      c->mol_hi[a] = std::max(c->mol_hi[a], pos[3 * i + a]);
    }
  }
  if (count) {
    c->rx.next_id_dirty = true;
    c->rx.grid_dirty = true;
    c->rx.reset_drain_form();
  }
  std::vector<float> fx, fy, fz;
  if (c->fp32()) {
    if (!c->have_geometry) return fail(MCG_E_STATE, "the FP32 form needs the geometry before the molecules");
    fx.resize(count); fy.resize(count); fz.resize(count);
    for (uint64_t i = 0; i < count; ++i) {
      if (!round_same_side_host(c, hx[i], hy[i], hz[i], fx[i], fy[i], fz[i]))
        return fail(MCG_E_INVALID, "molecule cannot be rounded to FP32 on its side of the nearby wall planes");
      if (!(fx[i] < (float)P.qx && fy[i] < (float)P.qy && fz[i] < (float)P.qz))
        return fail(MCG_E_INVALID, "molecule rounds onto the partition boundary in FP32");
    }
  }
  if (count) {
    size_t o = c->n;
    if (c->fp32()) {
      HIP_TRY(hipMemcpyAsync(c->xf.p + o, fx.data(), count * 4, hipMemcpyHostToDevice, c->stream));
      HIP_TRY(hipMemcpyAsync(c->yf.p + o, fy.data(), count * 4, hipMemcpyHostToDevice, c->stream));
      HIP_TRY(hipMemcpyAsync(c->zf.p + o, fz.data(), count * 4, hipMemcpyHostToDevice, c->stream));
    } else {
      HIP_TRY(hipMemcpyAsync(c->x.p + o, hx.data(), count * 8, hipMemcpyHostToDevice, c->stream));
      HIP_TRY(hipMemcpyAsync(c->y.p + o, hy.data(), count * 8, hipMemcpyHostToDevice, c->stream));
      HIP_TRY(hipMemcpyAsync(c->z.p + o, hz.data(), count * 8, hipMemcpyHostToDevice, c->stream));
    }
    HIP_TRY(hipMemcpyAsync(c->id.p + o, ids, count * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->tag.p + o, ht.data(), count * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  c->n += count;
  return MCG_OK;
}

int mcg_get_next_molecule_ids(mcg_ctx* c, uint32_t* next_ids) {
  if (!c || !next_ids) return fail(MCG_E_INVALID, "null argument");
  MCG_ALIVE(c);
  ReactState& r = c->rx;
  if (r.next_id_host.size() != c->cfg.num_replicas) r.next_id_host.assign(c->cfg.num_replicas, 0);
  // the device holds the counters once reactive steps have run
  if (r.on && !r.next_id_dirty) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(r.next_id_host.data(), r.next_id.p, r.next_id_host.size() * 4, hipMemcpyDeviceToHost));
  }
  std::memcpy(next_ids, r.next_id_host.data(), r.next_id_host.size() * 4);
  return MCG_OK;
}

int mcg_set_next_molecule_ids(mcg_ctx* c, const uint32_t* next_ids) {
  if (!c || !next_ids) return fail(MCG_E_INVALID, "null argument");
  MCG_ALIVE(c);
  std::vector<uint32_t> cur(c->cfg.num_replicas);
  if (int st = mcg_get_next_molecule_ids(c, cur.data())) return st;
  for (uint32_t k = 0; k < c->cfg.num_replicas; ++k) c->rx.next_id_host[k] = std::max(cur[k], next_ids[k]);
  c->rx.next_id_dirty = true;
  return MCG_OK;
}

namespace {

// Marks the context failed after a step whose kernels ran into an error.
int step_failed(mcg_ctx* c, int st) {
  if (st != MCG_OK && st != MCG_E_INVALID && st != MCG_E_STATE) {
    c->failed = true;
    c->fail_msg = g_err;
  }
  return st;
}

int free_step(mcg_ctx* c, uint32_t first_iteration, uint32_t num_iterations, mcg_step_stats* stats) {
  auto t0 = std::chrono::steady_clock::now();
  MolView64 m{c->x.p, c->y.p, c->z.p, c->id.p, c->tag.p};
  MolView32 mf{c->xf.p, c->yf.p, c->zf.p, c->id.p, c->tag.p};
  WallView64 W{c->walls64.p, c->off.p, c->idx.p, c->ent64.p, c->ent32.p, c->ent_box.p};
  HitTrace tr{c->trace_wall.p, c->trace_count.p, (c->cfg.flags & MCG_FLAG_TRACE_HITS) ? 1u : 0u};
  uint32_t n = (uint32_t)c->n;
  SlowQueue Q{c->queue.p, c->qcount.p, 0};
  SlowQueue MQ{c->mqueue.p, c->mqcount.p, 0};
  HIP_TRY(launch_fill_u64(c->stream, c->counters.p, 0ull, CTR_WORDS));
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  // every molecule has its diffusion time at the iteration start and the
  // species time step 1, so the barrier (>= 1 ahead) never clips a step
  for (uint32_t k = 0; k < num_iterations; ++k) {
    if (c->fp32())
      HIP_TRY(launch_diffuse_reflect32(c->stream, mf, n, c->species_table(), c->surface_classes(), c->P, W,
                                       c->walls32.p, c->max_abs_coord, c->cfg.seed, first_iteration + k, MQ, Q,
                                       c->err.p, tr, c->counters.p));
    else
      HIP_TRY(launch_diffuse_reflect64(c->stream, m, n, c->species_table(), c->surface_classes(), c->P, W,
                                       c->cfg.seed, first_iteration + k, MQ, Q, c->err.p, tr, c->counters.p));
  }
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  uint32_t err = 0, slow = 0;
  unsigned long long ctr[CTR_WORDS] = {0, 0};
  std::vector<uint32_t> shard_counts(kNumQueueShards, 0);
  HIP_TRY(hipMemcpyAsync(&err, c->err.p, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(ctr, c->counters.p, sizeof(ctr), hipMemcpyDeviceToHost, c->stream));
  if (num_iterations && n)
    HIP_TRY(hipMemcpyAsync(shard_counts.data(), c->qcount.p, kNumQueueShards * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (uint32_t v : shard_counts) slow += v;
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_kernel_ms = num_iterations ? ms / num_iterations : 0;
  c->last_launches = num_iterations;
  if (stats) {
    stats->molecule_steps = (uint64_t)n * num_iterations;
    stats->wall_hits = ctr[CTR_REFLECTIONS];
    stats->transparent_crossings = ctr[CTR_CROSSINGS];
    stats->reactions = 0;
    stats->kernel_ms = ms;
    stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats->kernel_launches = 3 * num_iterations;
    stats->slow_path_last = slow;
  }
  if (err) {
    HIP_TRY(launch_fill_u32(c->stream, c->err.p, 0u, 1u));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (err & ERR_ESCAPED) return fail(MCG_E_ESCAPED, "a molecule escaped the partition");
    return fail(MCG_E_LIMIT, "per-step limit exceeded (error bits " + std::to_string(err) + ")");
  }
  return MCG_OK;
}

}  // namespace

int mcg_step(mcg_ctx* c, uint32_t first_iteration, uint32_t num_iterations, double time_up_to_next_barrier,
             mcg_step_stats* stats) {
  if (!c) return fail(MCG_E_INVALID, "null context");
  MCG_ALIVE(c);
  if (!c->have_geometry || c->num_species == 0) return fail(MCG_E_STATE, "geometry and species must be set first");
  // DiffuseReactEvent::set_barrier_time_for_next_execution's contract
  // (diffuse_react_event.h:142-150): a whole number > 0
  if (!(time_up_to_next_barrier >= 1.0) || time_up_to_next_barrier != std::floor(time_up_to_next_barrier))
    return fail(MCG_E_INVALID, "time_up_to_next_barrier must be a whole number >= 1");
  if ((double)num_iterations > time_up_to_next_barrier)
    return fail(MCG_E_INVALID, "a call must not run past its barrier (num_iterations > time_up_to_next_barrier)");
  if (first_iteration != c->clock)
    return fail(MCG_E_INVALID, "first_iteration " + std::to_string(first_iteration) + " is not the engine's clock " +
                                   std::to_string(c->clock));
  if (stats) std::memset(stats, 0, sizeof(*stats));
  // species time steps other than one iteration and surface molecules
  // that diffuse take the reactive path
  if (int e = ensure_step_buffers(c)) return e;
  if (c->surface_diffusion() && c->rx.total_tiles)
    if (int e = ensure_react_buffers(c)) return e;
  c->stepped = true;
  int st = c->rx.on ? react_step(c, first_iteration, num_iterations, time_up_to_next_barrier, stats)
                    : free_step(c, first_iteration, num_iterations, stats);
  if (st == MCG_OK) c->clock = (uint64_t)first_iteration + num_iterations;
  return step_failed(c, st);
}

int mcg_set_surface_classes(mcg_ctx* c, uint32_t num_classes, const uint8_t* action) {
  if (!c || (num_classes && !action)) return fail(MCG_E_INVALID, "null argument");
  MCG_ALIVE(c);
  if (c->num_species == 0 || !c->have_geometry) return fail(MCG_E_STATE, "geometry and species must be set first");
  if (c->per_replica) return fail(MCG_E_STATE, "surface classes must be set before per-replica parameters");
  const uint32_t ns = c->num_species;
  for (uint32_t v : c->wall_class_host)
    if (v > num_classes) return fail(MCG_E_INVALID, "a wall names a surface class beyond num_classes");
  bool on = false;
  std::vector<SpeciesRec> sp = c->species_host;
  for (uint32_t k = 0; k < num_classes; ++k)
    for (uint32_t s = 0; s < ns; ++s) {
      const uint8_t a = action[(size_t)k * ns + s];
      if ((a & 3u) > MCG_WALL_ABSORB || ((a >> 2) & 3u) > MCG_WALL_ABSORB || (a >> 4))
        return fail(MCG_E_INVALID, "bad surface-class action");
      if (a) {
        on = true;
        sp[s].flags |= SPF_CAN_VOLWALL;
      }
    }
  if (c->sc_action_d.alloc(std::max<size_t>((size_t)num_classes * ns, 1))) return fail(MCG_E_NOMEM, "class table");
  if (num_classes) HIP_TRY(hipMemcpy(c->sc_action_d.p, action, (size_t)num_classes * ns, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->species.p, sp.data(), ns * sizeof(SpeciesRec), hipMemcpyHostToDevice));
  c->species_host = sp;
  c->num_surface_classes = num_classes;
  c->sc_on = on;
  return MCG_OK;
}

int mcg_sync(mcg_ctx* c) {
  if (!c) return fail(MCG_E_INVALID, "null context");
  MCG_ALIVE(c);
  HIP_TRY(hipStreamSynchronize(c->stream));
  return MCG_OK;
}

int mcg_num_molecules(mcg_ctx* c, uint64_t* count) {
  if (!c || !count) return fail(MCG_E_INVALID, "null argument");
  MCG_ALIVE(c);
  if ((!c->rx.on && !c->sc_on) || c->num_species == 0) {
    *count = c->n;
    return MCG_OK;
  }
  std::vector<uint64_t> counts((size_t)c->cfg.num_replicas * c->num_species);
  int st = count_species(c, counts.data(), false);
  if (st != MCG_OK) return st;
  uint64_t t = 0;
  for (uint64_t v : counts) t += v;
  *count = t;
  return MCG_OK;
}

int mcg_get_volume_molecules(mcg_ctx* c, uint32_t* ids, uint32_t* species, uint32_t* replica, double* pos,
                             uint32_t* subparts, double* dtimes, double* utimes, uint32_t* flags, uint32_t* cvols,
                             double* birthdays, uint64_t capacity, uint64_t* count) {
  if (!c) return fail(MCG_E_INVALID, "null context");
  MCG_ALIVE(c);
  const uint64_t n = c->n;
  std::vector<double> hx(n), hy(n), hz(n), htoff, hut, hb;
  std::vector<uint32_t> ht(n), hid(n), hcv;
  if (n && c->fp32()) {
    std::vector<float> fx(n), fy(n), fz(n);
    HIP_TRY(hipMemcpyAsync(fx.data(), c->xf.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(fy.data(), c->yf.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(fz.data(), c->zf.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (uint64_t i = 0; i < n; ++i) { hx[i] = fx[i]; hy[i] = fy[i]; hz[i] = fz[i]; }
  }
  // unimolecular times are kept only while a unimolecular class exists
  const bool have_ut = c->rx.on && std::any_of(c->rx.unimol.begin(), c->rx.unimol.end(),
                                               [](const UnimolRec& u) { return u.num_pathways > 0; });
  if (n) {
    if (!c->fp32()) {
      HIP_TRY(hipMemcpyAsync(hx.data(), c->x.p, n * 8, hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipMemcpyAsync(hy.data(), c->y.p, n * 8, hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipMemcpyAsync(hz.data(), c->z.p, n * 8, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipMemcpyAsync(ht.data(), c->tag.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(hid.data(), c->id.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    if (dtimes && c->rx.on) {
      htoff.resize(n);
      HIP_TRY(hipMemcpyAsync(htoff.data(), c->rx.toff.p, n * 8, hipMemcpyDeviceToHost, c->stream));
    }
    if (utimes && have_ut) {
      hut.resize(n);
      HIP_TRY(hipMemcpyAsync(hut.data(), c->rx.utime.p, n * 8, hipMemcpyDeviceToHost, c->stream));
    }
    if (cvols && c->rx.num_volumes) {
      hcv.resize(n);
      HIP_TRY(hipMemcpyAsync(hcv.data(), c->rx.cvol.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    }
    if (birthdays && c->rx.birth.p) {
      hb.resize(n);
      HIP_TRY(hipMemcpyAsync(hb.data(), c->rx.birth.p, n * 8, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  uint64_t live = 0;
  for (uint64_t i = 0; i < n; ++i) live += (tag_flags(ht[i]) & MF_DEFUNCT) ? 0 : 1;
  if (count) *count = live;
  if (capacity < live) return fail(MCG_E_CAPACITY, "output capacity too small");
  const PartitionParams& P = c->P;
  const double clock = (double)c->clock;
  uint64_t k = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t f = tag_flags(ht[i]);
    if (f & MF_DEFUNCT) continue;
    if (ids) ids[k] = hid[i];
    if (species) species[k] = tag_species(ht[i]);
    if (replica) replica[k] = tag_replica(ht[i]);
    if (pos) { pos[3 * k] = hx[i]; pos[3 * k + 1] = hy[i]; pos[3 * k + 2] = hz[i]; }
    if (subparts) {
      int ix = (int)((hx[i] - P.ox) * P.sp_rcp), iy = (int)((hy[i] - P.oy) * P.sp_rcp),
          iz = (int)((hz[i] - P.oz) * P.sp_rcp);
      subparts[k] = (uint32_t)ix + (uint32_t)iy * P.n + (uint32_t)iz * P.n2;
    }
    if (dtimes) {
      // an infinite offset: scheduled never (TIME_FOREVER)
      const double off = ((f & MF_TOFF) && !htoff.empty()) ? htoff[i] : 0.0;
      dtimes[k] = std::isinf(off) ? kTimeForever : clock + off;
    }
    if (utimes) utimes[k] = hut.empty() ? kTimeInvalid : hut[i];
    if (flags) flags[k] = ((f & MF_SCHEDULE_UNIMOL) ? MCG_MOLF_SCHEDULE_UNIMOL : 0u) | ((f & MF_MATURE) ? MCG_MOLF_MATURE : 0u);
    if (cvols) cvols[k] = hcv.empty() ? 0u : hcv[i];
    if (birthdays) birthdays[k] = hb.empty() ? kTimeInvalid : hb[i];
    ++k;
  }
  return MCG_OK;
}

int mcg_get_walls(mcg_ctx* c, double* wc, uint32_t* offsets, uint32_t* indices, uint64_t* num_entries) {
  if (!c) return fail(MCG_E_INVALID, "null context");
  MCG_ALIVE(c);
  if (wc)
    for (size_t i = 0; i < c->walls_host.size(); ++i) {
      const WallRec64& w = c->walls_host[i];
      double* o = wc + 16 * i;
      double v[16] = {w.nx, w.ny, w.nz, w.ux, w.uy, w.uz, w.vx, w.vy, w.vz, w.d, w.uv1u, w.uv2u, w.uv2v, w.area, 0, 0};
      std::memcpy(o, v, sizeof(v));
    }
  if (offsets) std::memcpy(offsets, c->off_host.data(), c->off_host.size() * 4);
  if (indices) std::memcpy(indices, c->idx_host.data(), c->idx_host.size() * 4);
  if (num_entries) *num_entries = c->idx_host.size();
  return MCG_OK;
}

int mcg_get_hit_trace(mcg_ctx* c, uint32_t* walls, uint32_t* counts, uint64_t capacity) {
  if (!c) return fail(MCG_E_INVALID, "null context");
  MCG_ALIVE(c);
  if (!(c->cfg.flags & MCG_FLAG_TRACE_HITS)) return fail(MCG_E_STATE, "created without MCG_FLAG_TRACE_HITS");
  if (c->rx.on) return fail(MCG_E_UNSUPPORTED, "hit traces are recorded without reaction classes only");
  if (capacity < c->n) return fail(MCG_E_CAPACITY, "output capacity too small");
  if (c->n) {
    if (walls) HIP_TRY(hipMemcpy(walls, c->trace_wall.p, c->n * kTraceHits * 4, hipMemcpyDeviceToHost));
    if (counts) HIP_TRY(hipMemcpy(counts, c->trace_count.p, c->n * 4, hipMemcpyDeviceToHost));
  }
  return MCG_OK;
}

int mcg_last_kernel_time(mcg_ctx* c, double* avg_ms, uint32_t* launches) {
  if (!c) return fail(MCG_E_INVALID, "null context");
  MCG_ALIVE(c);
  if (avg_ms) *avg_ms = c->last_kernel_ms;
  if (launches) *launches = c->last_launches;
  return MCG_OK;
}

}  // extern "C"
