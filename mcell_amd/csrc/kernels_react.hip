// HIP kernels of one diffuse-react iteration with reaction classes
// (volume-volume, volume-surface and unimolecular).
//
// The reference moves molecules one after another, each seeing the others
// where they are at that moment (diffuse_react_event.cpp:53-161).  Here all
// molecules of an iteration move at once against a snapshot of the
// iteration's start, and reactions are settled afterwards:
//
//   k_cell_keys / scan / k_scatter   counting sort of the live molecules by
//                                    bin = (replica, species, subpartition)
//                                    into the snapshot
//   k_move_light                     every mover: the displacement; a move that
//                                    certainly meets no wall commits here, after
//                                    up to two molecule collisions resolved in
//                                    place; the others are queued (wall movers
//                                    and partner movers in separate halves)
//   k_move_drain, k_move             the full step of react_step.h for queued
//                                    movers (and later for reruns and product
//                                    movers); a mover whose reaction test
//                                    succeeds stops with a proposal and claims
//                                    itself and its partner (atomicMin of its id)
//   k_accept                         a proposal is accepted when it holds both
//                                    claims; its reactants are consumed
//   k_requeue                        claims are released; movers whose proposal
//                                    lost and that were not consumed move again
//                                    (same stream, consumed partners skipped)
//   k_unimol_fire                    due unimolecular events of unconsumed molecules
//   products                         accepted reactions in (replica, mover id)
//                                    order get product ids per replica; products
//                                    move for the rest of the step as new movers;
//                                    a surface product takes its receptor's tile
//
// The smallest mover id among open proposals always wins its claims, so every
// round settles at least one proposal.  Molecules leave the snapshot order in
// the output arrays (slot j of the snapshot writes slot j of the output);
// consumed molecules stay as defunct holes until the next sort drops them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include <hipcub/hipcub.hpp>

#include "device/displace.h"
#include "device/light.h"
#include "device/react_step.h"
#include "device/schedule.h"
#include "kernels_react.h"

namespace mcg {

static __constant__ ZigguratTables c_zig_r = MCG_ZIGGURAT_TABLES_INIT;

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------- cell sort --
template <typename PosT>
__global__ __launch_bounds__(256) void k_cell_keys(RMolView<PosT> m, uint32_t n, PartitionParams P, CellGrid G,
                                                   uint32_t* __restrict__ key, uint32_t* __restrict__ rank,
                                                   uint32_t* __restrict__ cell_count, uint32_t* __restrict__ err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t c = kInvalid;
  if (i < n) {
    const uint32_t tag = m.tag[i];
    if (!(tag_flags(tag) & MF_DEFUNCT)) {
      int sx, sy, sz;
      subpart_3d(P, D3{(double)m.x[i], (double)m.y[i], (double)m.z[i]}, sx, sy, sz);
      if (!cell_of(G, sx, sy, sz, c)) {
        atomicOr(err, ERR_GRID);
        c = kInvalid;
      } else {
        c = bin_of(G, tag_replica(tag), tag_species(tag), c);
      }
    }
    key[i] = c;
  }
  // the molecules arrive mostly in bin order, so neighbouring lanes often
  // share a bin: one atomic per run of equal bins in the wave, each lane's
  // rank = the run's base + its place in the run
  const int lane = __lane_id();
  const uint32_t prev = __shfl_up(c, 1);
  const bool head = lane == 0 || prev != c;
  const uint64_t heads = __ballot(head);
  const uint64_t upto = (2ull << lane) - 1;  // lanes 0..lane
  const int my_head = 63 - __clzll((long long)(heads & upto));
  const uint64_t after = heads & ~upto;
  const int next_head = after ? __ffsll((long long)after) - 1 : 64;
  uint32_t base = 0;
  if (head && c != kInvalid) base = atomicAdd(cell_count + c, (uint32_t)(next_head - lane));
  base = __shfl(base, my_head);
  if (i < n && c != kInvalid) rank[i] = base + (uint32_t)(lane - my_head);
}

template <typename PosT>
__global__ __launch_bounds__(256) void k_scatter(RMolView<PosT> m, uint32_t n, const uint32_t* __restrict__ key,
                                                 const uint32_t* __restrict__ rank,
                                                 const uint32_t* __restrict__ cell_start, RMolView<PosT> s) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = key[i];
  if (c == kInvalid) return;
  const uint32_t j = cell_start[c] + rank[i];
  s.x[j] = m.x[i];
  s.y[j] = m.y[i];
  s.z[j] = m.z[i];
  s.id[j] = m.id[i];
  const uint32_t tag = m.tag[i];
  s.tag[j] = tag;
  if (tag_flags(tag) & MF_TOFF) s.toff[j] = m.toff[i];
  if (s.utime) s.utime[j] = m.utime[i];
  if (s.cvol) s.cvol[j] = m.cvol[i];
  if (s.birth) s.birth[j] = m.birth[i];
}

// ----------------------------------------------------------------- moves --
// FP32 form: the end point in FP32, kept on the FP64 point's side of nearby
// wall planes.  Returns the error bit of a rounding that did not settle.
__device__ __forceinline__ uint32_t store_pos(const ReactParams& R, double* x, double* y, double* z, uint32_t j, D3 p,
                                              bool) {
  x[j] = p.x;
  y[j] = p.y;
  z[j] = p.z;
  return 0;
}
__device__ __forceinline__ uint32_t store_pos(const ReactParams& R, float* x, float* y, float* z, uint32_t j, D3 p,
                                              bool near_plane) {
  float fx, fy, fz;
  uint32_t e = 0;
  if (near_plane) {
    if (!round_to_fp32_same_side(R.P, R.walls, R.off, R.idx, R.ent_box, p.x, p.y, p.z, &fx, &fy, &fz))
      e = ERR_ROUND32;
  } else {
    fx = (float)p.x;
    fy = (float)p.y;
    fz = (float)p.z;
  }
  x[j] = fx;
  y[j] = fy;
  z[j] = fz;
  return e;
}

// The snapshot range of species t's molecules in the row of cells
// (x0 .. x0 + nx, y, z) of replica rep (bins of one row are consecutive).
__device__ __forceinline__ void row_range(const CellGrid& G, uint32_t rep, uint32_t t, int x0, int y, int z, int nx,
                                          uint32_t& k0, uint32_t& k1) {
  uint32_t c0;
  if (!cell_of(G, x0, y, z, c0)) {
    k0 = k1 = 0;
    return;
  }
  const uint32_t b0 = bin_of(G, rep, t, c0);
  k0 = G.start[b0];
  k1 = G.start[b0 + (uint32_t)nx + 1];
}

// Diffusion-time offsets are zero for every molecule except products before
// their next barrier; the offset array is only read and written for tags
// that carry MF_TOFF.
__device__ __forceinline__ double load_toff(const double* __restrict__ toff, uint32_t tag, uint32_t j) {
  return (tag_flags(tag) & MF_TOFF) ? toff[j] : 0.0;
}
__device__ __forceinline__ uint32_t store_toff(double* __restrict__ toff, uint32_t j, double v, uint32_t tag) {
  if (v != 0.0) {
    toff[j] = v;
    return tag | (MF_TOFF << 8);
  }
  return tag & ~(MF_TOFF << 8);
}

__device__ __forceinline__ void stage_zig(ZigguratTables& T) {
  const uint32_t* zs = reinterpret_cast<const uint32_t*>(&c_zig_r);
  uint32_t* zd = reinterpret_cast<uint32_t*>(&T);
  for (int k = threadIdx.x; k < (int)(sizeof(ZigguratTables) / 4); k += blockDim.x) zd[k] = zs[k];
  __syncthreads();
}

// Wave-aggregated append of slot j for the lanes with `want` set.
__device__ __forceinline__ void append_slot(bool want, uint32_t j, uint32_t* __restrict__ list,
                                            uint32_t* __restrict__ count) {
  const uint64_t mask = __ballot(want);
  if (!mask) return;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
  base = __shfl(base, leader);
  if (want) list[base + __popcll(mask & ((1ull << lane) - 1))] = j;
}

// The full step of the mover in slot j (react_step.h): writes its end state,
// or records its proposal and claims.  Returns whether it proposed.  The
// reflections and crossings of a committed step are added to n_reflect /
// n_cross; a proposal carries those before its reaction in its pad word
// (counted when it is accepted: a losing mover repeats its step).
// A step that needs a larger exact_disk pool than kPoolN (react_mover)
// writes nothing and sets `deferred` (the caller queues the mover for a pass
// with a larger pool).
template <typename PosT, int kPoolN>
__device__ __forceinline__ bool move_one(const ReactParams& R, const Snapshot<PosT>& S, const RMolView<PosT>& src,
                                         const RMolView<PosT>& out, uint32_t j, SpeciesTable species,
                                         const ZigguratTables& T, double barrier, uint32_t propose_only,
                                         Proposal* __restrict__ prop, uint32_t* __restrict__ owner,
                                         uint32_t* __restrict__ tile_owner, uint32_t* __restrict__ err_out,
                                         bool& due, bool& deferred, uint32_t* set_lds, uint32_t& n_reflect,
                                         uint32_t& n_cross) {
  const uint32_t tag = src.tag[j];
  const uint32_t idv = src.id[j];
  const double toff = load_toff(src.toff, tag, j);
  const double utime = src.utime ? src.utime[j] : kTimeInvalid;
  const double birth = src.birth ? src.birth[j] : 0.0;
  const SpeciesRec sp = species.of(tag);
  D3 pos = {(double)src.x[j], (double)src.y[j], (double)src.z[j]};
  uint32_t err = 0;
  MoverEnd e;
  e.pos = pos;
  e.t_steps = 0;
  e.utime = utime;
  e.near_plane = false;
  e.absorbed = false;
  e.exd = false;
  e.due = 0;
  e.n_reflect = 0;
  e.n_cross = 0;
  e.cvol = src.cvol ? src.cvol[j] : kInvalid;
  e.cont = false;
  e.mature = false;
  Proposal p;
  bool proposed = false;
  const bool moves = sp.space_step > 0 || R.unimol[tag_species(tag)].num_pathways > 0;
  const uint32_t pass_through = R.kept ? R.kept[j] : 0u;
  const uint32_t born = R.born ? R.born[j] : 0u;
  if (moves)
    proposed = react_mover<PosT, kPoolN>(R, S, sp, T, pos, idv, tag, toff, utime, birth, barrier, pass_through, born,
                                         p, e, err, set_lds);
  deferred = e.exd;
  if (deferred) {
    due = false;
    if (err) atomicOr(err_out, err);
    return false;
  }
  due = e.due != 0;
  if (due) {
    // the unimolecular event, settled after the phase's bimolecular claims
    p.time = e.utime;
    p.x = e.pos.x;
    p.y = e.pos.y;
    p.z = e.pos.z;
    p.mover_slot = j;
    p.partner_slot = kInvalid;
    p.mover_id = idv;
    p.partner_id = kInvalid;
    p.rxn = kUnimolBit | tag_species(tag);
    p.pathway = e.due - 1;  // the stream pass of the pathway draw
    p.replica = tag_replica(tag);
    p.pad = 0;
    p.cvol = e.cvol;
    prop[j] = p;
  }
  if (proposed) {
    p.mover_slot = j;
    p.pad = (p.pad & kPadMoverIsB) | (e.n_reflect > 0xFFFFu ? 0xFFFFu : e.n_reflect) |
            ((e.n_cross > 0x7FFFu ? 0x7FFFu : e.n_cross) << 16);
    prop[j] = p;
    if (!propose_only) {
      atomicMin(owner + j, idv);
      if (p.partner_slot & kTileBit) atomicMin(tile_owner + (p.partner_slot & ~kTileBit), idv);
      else atomicMin(owner + p.partner_slot, idv);
    }
  }
  if (!proposed || propose_only) {
    err |= store_pos(R, out.x, out.y, out.z, j, e.pos, e.near_plane);
    out.id[j] = idv;
    uint32_t t2 = (tag & ~(MF_STEP_FLAGS << 8)) | (e.mature ? (MF_MATURE << 8) : 0u);
    if (e.cont && !propose_only) {
      // the next step of this iteration reads the stream's next pass and
      // starts at the new diffusion time
      const uint32_t pass = ((tag_flags(tag) & MF_PASS_MASK) >> MF_PASS_SHIFT) + 1;
      if (pass > (MF_PASS_MASK >> MF_PASS_SHIFT)) {
        err |= ERR_PASSES;
      } else {
        t2 |= (pass << MF_PASS_SHIFT) << 8;
        R.clist[atomicAdd(R.ccount, 1u)] = j;
      }
      t2 = store_toff(out.toff, j, (((double)R.iteration + toff) + e.t_steps) - (double)R.iteration, t2);
    } else {
      t2 = store_toff(out.toff, j,
                      sp.space_step > 0 ? next_toff(R.iteration, toff, e.t_steps) : immobile_toff(R.iteration, e.utime),
                      t2);
    }
    out.tag[j] = e.absorbed ? (t2 | (MF_DEFUNCT << 8)) : t2;
    if (out.utime) out.utime[j] = e.utime;
    if (out.cvol) out.cvol[j] = e.cvol;
    if (out.birth) out.birth[j] = birth;
    if (born) R.born[j] = 0;  // the step that could use it is over
    n_reflect += e.n_reflect;
    n_cross += e.n_cross;
  }
  if (err) atomicOr(err_out, err);
  return proposed;
}

// One mover per thread.  Movers are snapshot slots [0, *count) when list is
// null, else slots list[0, *count).  src holds the movers' start state (the
// snapshot, or the output arrays for products); out receives end states.
// Movers whose exact_disk needs more than kPoolN vertex records are appended
// to xlist (kPoolN < exd::kPool).
// The full-step kernels hold ~60 KB of LDS per workgroup, so at most two
// workgroups share a CU (2 waves per SIMD); their registers are capped at
// what two waves per SIMD allow (256), the rest spilling to scratch, which
// measured ~10 % faster than the uncapped one-wave form.
template <typename PosT, int kPoolN>
__global__ __launch_bounds__(256, 2) void k_move(ReactParams R, Snapshot<PosT> S, RMolView<PosT> src,
                                              RMolView<PosT> out, const uint32_t* __restrict__ list,
                                              const uint32_t* __restrict__ count,
                                              SpeciesTable species, double barrier,
                                              uint32_t propose_only, Proposal* __restrict__ prop,
                                              uint32_t* __restrict__ plist, uint32_t* __restrict__ pcount,
                                              uint32_t* __restrict__ owner, uint32_t* __restrict__ tile_owner,
                                              uint32_t* __restrict__ err_out, uint32_t* __restrict__ ulist,
                                              uint32_t* __restrict__ ucount, uint32_t* __restrict__ xlist,
                                              uint32_t* __restrict__ xcount) {
  __shared__ ZigguratTables T;
  __shared__ uint32_t sets[SubpartSet::kCap * 256];
  const uint32_t n = *count;
  if ((uint64_t)blockIdx.x * blockDim.x >= n) return;  // (uniform per block)
  stage_zig(T);
  uint32_t n_reflect = 0, n_cross = 0;
  // grid-stride: the launch is sized by a host bound on the count, so a
  // small grid walks a list whose actual length is known only on the device
  for (uint32_t b = blockIdx.x * blockDim.x; b < n; b += gridDim.x * blockDim.x) {
    const uint32_t q = b + threadIdx.x;
    bool proposed = false, due = false, deferred = false;
    uint32_t j = 0;
    if (q < n) {
      j = list ? list[q] : q;
      proposed = move_one<PosT, kPoolN>(R, S, src, out, j, species, T, barrier, propose_only, prop, owner,
                                        tile_owner, err_out, due, deferred, sets + threadIdx.x, n_reflect, n_cross);
    }
    append_slot(proposed, j, plist, pcount);
    append_slot(due, j, ulist, ucount);
    if constexpr (kPoolN < exd::kPool) append_slot(deferred, j, xlist, xcount);
  }
  wave_count(R.ctr + CTR_REFLECTIONS, n_reflect);
  wave_count(R.ctr + CTR_CROSSINGS, n_cross);
}

// The full step for every queued mover (the shards read as one list,
// grid-stride; consecutive threads take consecutive slots of a shard, which
// came from neighbouring cells).
template <typename PosT, int kPoolN>
__global__ __launch_bounds__(256, 2) void k_move_drain(ReactParams R, Snapshot<PosT> S, RMolView<PosT> src,
                                                    RMolView<PosT> out, SlowQueue Q, SpeciesTable species, double barrier,
                                                    uint32_t propose_only, Proposal* __restrict__ prop,
                                                    uint32_t* __restrict__ plist, uint32_t* __restrict__ pcount,
                                                    uint32_t* __restrict__ owner, uint32_t* __restrict__ tile_owner,
                                                    uint32_t* __restrict__ err_out, uint32_t* __restrict__ ulist,
                                                    uint32_t* __restrict__ ucount, uint32_t* __restrict__ xlist,
                                                    uint32_t* __restrict__ xcount) {
  __shared__ ZigguratTables T;
  __shared__ uint32_t sets[SubpartSet::kCap * 256];
  __shared__ uint32_t pre[kQueueShards + 1];
  __shared__ uint32_t part[256];
  const uint32_t total = queue_prefix(Q.count, Q.nshards, pre, part);
  if ((uint64_t)blockIdx.x * blockDim.x >= total) return;
  stage_zig(T);
  uint32_t n_reflect = 0, n_cross = 0;
  for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
    const uint32_t q = base + threadIdx.x;
    bool proposed = false, due = false, deferred = false;
    uint32_t j = 0;
    if (q < total) {
      j = queue_at(Q.slots, Q.shard_cap, pre, q);
      proposed = move_one<PosT, kPoolN>(R, S, src, out, j, species, T, barrier, propose_only, prop, owner,
                                        tile_owner, err_out, due, deferred, sets + threadIdx.x, n_reflect, n_cross);
    }
    append_slot(proposed, j, plist, pcount);
    append_slot(due, j, ulist, ucount);
    append_slot(deferred, j, xlist, xcount);
  }
  wave_count(R.ctr + CTR_REFLECTIONS, n_reflect);
  wave_count(R.ctr + CTR_CROSSINGS, n_cross);
}

// The common step, register-light: movers whose straight move certainly
// meets no wall commit pos + disp here (the reference's ray trace would find
// no wall and commit the same point, diffuse_react_event.cpp:773-777) unless
// they pass a reaction partner within the reaction radius; the others go to
// the queue for the full step, which re-draws the same displacement from the
// same stream.  The partner scan covers every cell of the segment's box
// widened by the radius, so it finds every partner collide_mol accepts (a
// superset of the reference's candidates).  One or two collisions with
// partners in the start or end subpartition, which the reference always
// collects, are processed here in the reference's order when no wall is near
// the collision point: exact_disk is then exactly 1, and a failed
// test_bimolecular draw leaves the move as it was.  Registers are capped at
// 128 (4 waves per SIMD); one register more costs a wave.
template <typename PosT>
__global__ __launch_bounds__(256, 4) void k_move_light(ReactParams R, Snapshot<PosT> S, RMolView<PosT> out,
                                                    const uint32_t* __restrict__ count, SpeciesTable species,
                                                    double barrier, SlowQueue Q, uint32_t* __restrict__ err_out) {
  constexpr bool kF32 = sizeof(PosT) == 4;
  __shared__ ZigguratTables T;
  stage_zig(T);
  const uint32_t n = *count;
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  bool hard = false, partner_hard = false;
  uint32_t err = 0;
  if (j < n) {
    const uint32_t tag = S.tag[j];
    const uint32_t idv = S.id[j];
    const double toff = load_toff(S.toff, tag, j);
    const double utime = S.utime ? S.utime[j] : kTimeInvalid;
    const SpeciesRec sp = species.of(tag);
    const D3 pos = {(double)S.x[j], (double)S.y[j], (double)S.z[j]};
    // molecules with a unimolecular class take the full step when their
    // lifetime is still to be drawn or ends within the step's time budget
    const bool unimol_now = R.unimol[tag_species(tag)].num_pathways > 0 &&
                            ((tag_flags(tag) & MF_SCHEDULE_UNIMOL) || utime < (double)R.iteration + barrier);
    // time steps below one iteration (more than one step in it) and
    // long time steps still lengthening (get_max_time's maturity rule)
    // take the full step
    const bool steps_apart =
        sp.time_step < 1.0 || (sp.time_step > 1.0 && !(tag_flags(tag) & MF_MATURE));
    if (!ready_this_iteration(R.iteration, toff)) {
      // scheduled for a later iteration (Partition::get_molecules_ready_for_
      // diffusion, partition.h:232-246): stays where it is
      store_pos(R, out.x, out.y, out.z, j, pos, false);
      out.id[j] = idv;
      out.tag[j] = store_toff(out.toff, j, toff_past_iteration(R.iteration, toff), tag);
      if (out.utime) out.utime[j] = utime;
      if (out.cvol) out.cvol[j] = S.cvol[j];
      if (out.birth) out.birth[j] = S.birth[j];
    } else if (unimol_now || steps_apart) {
      hard = true;
    } else if (sp.space_step > 0) {
      PhiloxStream rng;
      rng.init(R.seed, idv, R.iteration, tag_replica(tag), 0);
      double t_steps;
      D3 disp;
      const double r_rate_factor =
          clipped_displacement(rng, T, sp.space_step, sp.time_step, barrier - toff, t_steps, disp);
      const D3 end = {pos.x + disp.x, pos.y + disp.y, pos.z + disp.z};
      int lo[3], hi[3];
      bool clear = move_index_box(R.P, pos, end, lo, hi);
      bool near_plane = false;
      if (clear) {
        if (kF32) clear = all_walls_miss32(R.P, R.ent32, R.off, R.ent_box, R.F, R.band, pos, disp, lo, hi, near_plane);
        else clear = all_walls_miss64(R.P, R.ent64, R.off, R.ent_box, pos, disp, lo, hi);
      }
      const bool walls_clear = clear;
      if (clear && (sp.flags & SPF_CAN_VOLVOL) && !(sp.flags & SPF_TARGET_ONLY)) {
        // partners within the reaction radius of the segment: every one lies
        // in the cells of the segment's box widened by the radius
        const double sigma2 = R.P.rxn_radius * R.P.rxn_radius;
        const double o[3] = {R.P.ox, R.P.oy, R.P.oz};
        const double a[3] = {pos.x, pos.y, pos.z}, b[3] = {end.x, end.y, end.z};
        int clo[3], chi[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double mn = (a[k] < b[k] ? a[k] : b[k]) - R.P.rxn_radius;
          const double mx = (a[k] < b[k] ? b[k] : a[k]) + R.P.rxn_radius;
          int l = (int)floor((mn - o[k]) * R.P.sp_rcp), h = (int)floor((mx - o[k]) * R.P.sp_rcp);
          clo[k] = l < R.G.lo[k] ? R.G.lo[k] : l;
          chi[k] = h > R.G.lo[k] + R.G.dim[k] - 1 ? R.G.lo[k] + R.G.dim[k] - 1 : h;
        }
        const uint32_t my_sp = tag_species(tag), rep = tag_replica(tag);
        // the segment's box widened by the radius: partners outside it cannot collide
        const double R1 = R.P.rxn_radius;
        const PartnerBox<PosT> box((a[0] < b[0] ? a[0] : b[0]) - R1, (a[0] < b[0] ? b[0] : a[0]) + R1,
                                   (a[1] < b[1] ? a[1] : b[1]) - R1, (a[1] < b[1] ? b[1] : a[1]) + R1,
                                   (a[2] < b[2] ? a[2] : b[2]) - R1, (a[2] < b[2] ? b[2] : a[2]) + R1);
        // up to two collisions are resolved here when both partners sit in the
        // start or end subpartition (cells every collection of the reference holds)
        int ss[3], se[3];
        subpart_3d(R.P, pos, ss[0], ss[1], ss[2]);
        subpart_3d(R.P, end, se[0], se[1], se[2]);
        int ncand = 0;
        double ct[2];
        uint32_t cid[2];
        int32_t crx[2];
        D3 ccp[2];
        bool cok[2];
        // one species' partners in a row of cells along x are one range of the
        // snapshot; the next row's range is loaded while the current row's
        // partners are tested, and partner positions are loaded kGroup at a
        // time, so that the dependent loads overlap (the candidates' order
        // does not matter: at most two are kept and they are sorted below)
        constexpr int kGroup = kF32 ? 4 : 2;
        const int nx = chi[0] - clo[0];
        const bool rows = nx >= 0 && chi[1] >= clo[1] && chi[2] >= clo[2];
        for (uint32_t t = 0; t < R.G.ns && rows && clear; ++t) {
          const int32_t rx = R.pair_rxn[my_sp * R.num_species + t];
          if (rx < 0) continue;
          int y = clo[1], z = clo[2];
          uint32_t n0, n1;
          row_range(R.G, rep, t, clo[0], y, z, nx, n0, n1);
          while (clear) {
            const uint32_t k0 = n0, k1 = n1;
            const bool last = y == chi[1] && z == chi[2];
            if (!last) {
              if (++y > chi[1]) {
                y = clo[1];
                ++z;
              }
              row_range(R.G, rep, t, clo[0], y, z, nx, n0, n1);
            }
            for (uint32_t kb = k0; kb < k1 && clear; kb += kGroup) {
              PosT gx[kGroup], gy[kGroup], gz[kGroup];
#pragma unroll
              for (int u = 0; u < kGroup; ++u) {
                const uint32_t k = kb + u < k1 ? kb + u : k1 - 1;
                gx[u] = S.x[k];
                gy[u] = S.y[k];
                gz[u] = S.z[k];
              }
#pragma unroll
              for (int u = 0; u < kGroup; ++u) {
                const uint32_t k = kb + u;
                if (k >= k1 || !clear) break;
                if (box.outside(gx[u], gy[u], gz[u])) continue;
                const double qx = (double)gx[u], qy = (double)gy[u], qz = (double)gz[u];
                double tc;
                D3 cp;
                if (!collide_mol64(pos, disp, D3{qx, qy, qz}, sigma2, tc, cp)) continue;
                if (k == j) continue;
                if (ncand == 2) {
                  clear = false;
                  break;
                }
                int q3[3];
                subpart_3d(R.P, D3{qx, qy, qz}, q3[0], q3[1], q3[2]);
                ct[ncand] = tc;
                cid[ncand] = S.id[k];
                crx[ncand] = rx;
                ccp[ncand] = cp;
                cok[ncand] = (q3[0] == ss[0] && q3[1] == ss[1] && q3[2] == ss[2]) ||
                             (q3[0] == se[0] && q3[1] == se[1] && q3[2] == se[2]);
                ++ncand;
              }
            }
            if (last) break;
          }
        }
        if (clear && ncand > 0) {
          // the reference's order: time, then the larger partner id first
          if (ncand == 2 && (ct[1] < ct[0] || (ct[1] == ct[0] && cid[1] > cid[0]))) {
            const double tt = ct[0]; ct[0] = ct[1]; ct[1] = tt;
            const uint32_t ti = cid[0]; cid[0] = cid[1]; cid[1] = ti;
            const int32_t tr = crx[0]; crx[0] = crx[1]; crx[1] = tr;
            const D3 tp = ccp[0]; ccp[0] = ccp[1]; ccp[1] = tp;
            const bool tb = cok[0]; cok[0] = cok[1]; cok[1] = tb;
          }
          for (int q = 0; q < ncand && clear; ++q) {
            if (!cok[q]) {
              clear = false;
              break;
            }
            if (ct[q] < STIME_EPS) continue;  // is_immediate_collision (collision_utils.inl:814-816)
            // exact_disk is exactly 1 when no wall of the collision point's
            // subpartition comes within R of it (no wall is processed)
            int qx, qy, qz;
            subpart_3d(R.P, ccp[q], qx, qy, qz);
            const uint32_t csp = subpart_of(R.P, qx, qy, qz);
            const uint32_t db = region_box(R.P, csp, ccp[q].x - R1, ccp[q].y - R1, ccp[q].z - R1, ccp[q].x + R1,
                                           ccp[q].y + R1, ccp[q].z + R1);
            const uint32_t e1 = R.off[csp + 1];
            if (next_overlap<kLightOverlapGroup>(R.ent_box, R.off[csp], e1, db) < e1) {
              clear = false;
              break;
            }
            // test_bimolecular with factor 1: a reaction goes to the full step,
            // which re-draws the same numbers and proposes it
            if (test_bimolecular(R.rxns[rep * R.rxn_stride + crx[q]], r_rate_factor, rng) >= 0) clear = false;
          }
        }
      }
      if (clear) {
        // an end point next to a wall plane keeps its side in FP32
        err |= store_pos(R, out.x, out.y, out.z, j, end, near_plane);
        out.id[j] = idv;
        out.tag[j] = store_toff(out.toff, j, next_toff(R.iteration, toff, t_steps), tag & ~(MF_STEP_FLAGS << 8));
        if (out.utime) out.utime[j] = utime;
        if (out.cvol) out.cvol[j] = S.cvol[j];
        if (out.birth) out.birth[j] = S.birth[j];
      } else {
        hard = true;
        partner_hard = walls_clear;
      }
    } else {
      store_pos(R, out.x, out.y, out.z, j, pos, false);
      out.id[j] = idv;
      out.tag[j] = store_toff(out.toff, j, immobile_toff(R.iteration, utime), tag & ~(MF_STEP_FLAGS << 8));
      if (out.utime) out.utime[j] = utime;
      if (out.cvol) out.cvol[j] = S.cvol[j];
      if (out.birth) out.birth[j] = S.birth[j];
    }
  }
  // movers queued for walls (and unimolecular events) and movers queued for
  // partners go to separate halves of the shards, so that the full step's
  // lanes mostly follow the same path
  if (err) atomicOr(err_out, err);
  constexpr uint32_t kHalf = kQueueShards / 2;
  const uint32_t s0 = blockIdx.x % kHalf, s1 = kHalf + blockIdx.x % kHalf;
  wave_append(hard && !partner_hard, j, Q.slots + (size_t)s0 * Q.shard_cap, Q.count + s0);
  wave_append(hard && partner_hard, j, Q.slots + (size_t)s1 * Q.shard_cap, Q.count + s1);
}

// Appends slot j (or, with aprop, a copy of prop[j]) to the accepted list.
// Movers that stay after a reaction propose again in the same phase, so
// with such classes the accepted proposals are copied out of prop.
__device__ __forceinline__ void append_accepted(bool acc, uint32_t j, const Proposal* __restrict__ prop,
                                                Proposal* __restrict__ aprop, uint32_t* __restrict__ alist,
                                                uint32_t* __restrict__ acount) {
  const uint64_t mask = __ballot(acc);
  if (!mask) return;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(acount, (uint32_t)__popcll(mask));
  base = __shfl(base, leader);
  if (acc) {
    const uint32_t a = base + __popcll(mask & ((1ull << lane) - 1));
    if (aprop) {
      aprop[a] = prop[j];
      alist[a] = a;
    } else {
      alist[a] = j;
    }
  }
}

// A proposal whose mover and partner both carry its claim is accepted.
template <typename PosT>
__global__ __launch_bounds__(256) void k_accept(const uint32_t* __restrict__ plist, const uint32_t* __restrict__ pcount,
                                                Proposal* __restrict__ prop, const uint32_t* __restrict__ owner,
                                                const uint32_t* __restrict__ tile_owner,
                                                uint32_t* __restrict__ consumed, uint32_t* __restrict__ tile_consumed,
                                                uint32_t* __restrict__ out_tag,
                                                const RxnRec* __restrict__ rxns, uint32_t rxn_stride,
                                                const VolSurfRec* __restrict__ vs_rxns,
                                                uint32_t* __restrict__ kept, Proposal* __restrict__ aprop,
                                                uint32_t* __restrict__ alist,
                                                uint32_t* __restrict__ acount, unsigned long long* __restrict__ ctr) {
  const uint32_t n = *pcount;
  uint32_t n_reflect = 0, n_cross = 0;
  for (uint32_t q0 = blockIdx.x * blockDim.x; q0 < n; q0 += gridDim.x * blockDim.x) {
  const uint32_t q = q0 + threadIdx.x;
  bool acc = false;
  uint32_t j = 0;
  if (q < n) {
    j = plist[q];
    const Proposal& p = prop[j];
    const bool tile = (p.partner_slot & kTileBit) != 0;
    const uint32_t ps = p.partner_slot & ~kTileBit;
    if (owner[j] == p.mover_id && (tile ? tile_owner[ps] : owner[ps]) == p.mover_id) {
      acc = true;
      // outcome_bimolecular (diffuse_react_event.cpp:1833-1895): the
      // reactants are consumed unless the pathway keeps them (bit 0: the
      // mover, bit 1: the partner, from the class's a / b bits); a surface
      // reactant's tile goes to its product.  consumed holds the consuming
      // mover's id + 1.
      uint32_t keep = 0;
      if (!tile) {
        const uint32_t k = rxns[p.replica * rxn_stride + p.rxn].keep[p.pathway];
        keep = (p.pad & kPadMoverIsB) ? ((k >> 1) & 1u) | ((k & 1u) << 1) : k;
      } else if (vs_rxns[p.rxn & ~kVolSurfBit].keep_vol[p.pathway & kPathwayMask]) {
        keep = 1u;
      }
      if (keep & 1u) {
        // the mover stays: its step is repeated through this reaction and
        // goes on (the repeat counts its walls)
        ++kept[j];
      } else {
        n_reflect += p.pad & 0xFFFFu;  // the walls the mover met before the reaction
        n_cross += (p.pad >> 16) & 0x7FFFu;
        consumed[j] = p.mover_id + 1;
        out_tag[j] = out_tag[j] | (MF_DEFUNCT << 8);
      }
      if (tile) {
        // the surface reactant stays when a product of its species keeps it;
        // a consumed tile names its consumer (a kept mover's repeated step
        // passes through it)
        if (!vs_rxns[p.rxn & ~kVolSurfBit].keep_surf[p.pathway & kPathwayMask]) tile_consumed[ps] = p.mover_id + 1;
      } else if (!(keep & 2u)) {
        consumed[ps] = p.mover_id + 1;
        out_tag[ps] = out_tag[ps] | (MF_DEFUNCT << 8);
      }
    }
  }
  append_accepted(acc, j, prop, aprop, alist, acount);
  }
  wave_count(ctr + CTR_REFLECTIONS, n_reflect);
  wave_count(ctr + CTR_CROSSINGS, n_cross);
}

// Unimolecular events of the phase (react_unimol_single_molecule +
// outcome_unimolecular, diffuse_react_event.cpp:1760-1826, 2939-3003): every
// due molecule that no accepted bimolecular reaction consumed reacts at its
// scheduled time; the pathway is drawn from the molecule's stream when the
// class has several (which_unimolecular, rxn_utils.inl:772-783).
// A pathway that keeps the reactant leaves it alive: the products phase
// lets it diffuse on from its reaction time in the stream pass of the
// pathway draw (outcome_unimolecular, diffuse_react_event.cpp:2990-3003, and
// the rest of diffuse_single_molecule); pad records that pass (bits 0-7)
// and whether the draw was made (bit 8).
__global__ __launch_bounds__(256) void k_unimol_fire(ReactParams R, const uint32_t* __restrict__ ulist,
                                                     const uint32_t* __restrict__ ucount,
                                                     Proposal* __restrict__ prop, uint32_t* __restrict__ consumed,
                                                     uint32_t* __restrict__ out_tag, Proposal* __restrict__ aprop,
                                                     uint32_t* __restrict__ alist, uint32_t* __restrict__ acount) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  bool fire = false;
  uint32_t j = 0;
  if (q < *ucount) {
    j = ulist[q];
    if (!consumed[j]) {
      Proposal& p = prop[j];
      const UnimolRec& u = R.unimol[p.rxn & ~kUnimolBit];
      uint32_t pw = 0;
      if (u.num_pathways > 1) {
        PhiloxStream rng;
        rng.init(R.seed, p.mover_id, R.iteration, p.replica, p.pathway);
        const double match = rng.dbl() * u.cum[u.num_pathways - 1];
        pw = (uint32_t)binary_search_double(u.cum, match, (int)u.num_pathways - 1, 1);
      }
      p.pad = p.pathway | (u.num_pathways > 1 ? 0x100u : 0u);
      p.pathway = pw;
      if (!u.keep[pw]) {
        consumed[j] = p.mover_id + 1;
        out_tag[j] = out_tag[j] | (MF_DEFUNCT << 8);
      }
      fire = true;
    }
  }
  append_accepted(fire, j, prop, aprop, alist, acount);
}

// Releases the claims of every proposal of the round and queues the movers
// of lost proposals that are still alive.
__global__ __launch_bounds__(256) void k_requeue(const uint32_t* __restrict__ plist,
                                                 const uint32_t* __restrict__ pcount,
                                                 const Proposal* __restrict__ prop, uint32_t* __restrict__ owner,
                                                 uint32_t* __restrict__ tile_owner,
                                                 const uint32_t* __restrict__ consumed, uint32_t* __restrict__ rlist,
                                                 uint32_t* __restrict__ rcount) {
  const uint32_t n = *pcount;
  for (uint32_t q0 = blockIdx.x * blockDim.x; q0 < n; q0 += gridDim.x * blockDim.x) {
  const uint32_t q = q0 + threadIdx.x;
  bool again = false;
  uint32_t j = 0;
  if (q < n) {
    j = plist[q];
    owner[j] = kInvalid;
    const uint32_t ps = prop[j].partner_slot;
    if (ps & kTileBit) tile_owner[ps & ~kTileBit] = kInvalid;
    else owner[ps] = kInvalid;
    again = consumed[j] == 0;
  }
  const uint64_t mask = __ballot(again);
  if (mask) {
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(rcount, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (again) rlist[base + __popcll(mask & ((1ull << lane) - 1))] = j;
  }
  }
}

// ---------------------------------------------------------------- products --
// The order in which the reference would carry the phase's reactions out,
// which numbers their products: per replica, by the diffusion call they
// happen in (pass: a unimolecular event due after the molecule's first move
// happens in its second call, which the reference queues behind every
// first call), then the movers in the order diffuse_molecules takes them
// (diffuse_react_event.cpp:68-124) -- in the iteration's first phase those
// starting at the iteration's start, then those starting later (late = 1),
// each by id; in later phases new products before reactants that stayed --
// and a mover's reactions in the order they were accepted (the sort is
// stable).  key = replica << 36 | pass << 33 | late << 32 | mover id.
__device__ __forceinline__ uint32_t key_replica(uint64_t key) { return (uint32_t)(key >> 36); }

__global__ __launch_bounds__(256) void k_acc_keys(const uint32_t* __restrict__ alist, uint32_t n,
                                                  const Proposal* __restrict__ prop, const uint32_t* __restrict__ stag,
                                                  const double* __restrict__ stoff, const uint32_t* __restrict__ hdr,
                                                  uint32_t first_phase, uint64_t* __restrict__ keys) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const Proposal& p = prop[alist[k]];
  const uint32_t j = p.mover_slot;
  const bool snap = j < hdr[kHdrSnapshot];
  const bool late = first_phase ? (snap && load_toff(stoff, stag[j], j) != 0.0) : snap;
  // a fired unimolecular event keeps its draw's pass in pad (k_unimol_fire)
  const uint32_t pass = (first_phase && (p.rxn & kUnimolBit)) ? min(p.pad & 0xFFu, 7u) : 0u;
  keys[k] = ((uint64_t)p.replica << 36) | ((uint64_t)pass << 33) | ((uint64_t)late << 32) | p.mover_id;
}

__device__ __forceinline__ uint32_t new_products(const RxnRec& r, uint32_t pathway) {
  return r.num_products[pathway];
}

__global__ __launch_bounds__(256) void k_prod_counts(const uint32_t* __restrict__ sorted_slots, uint32_t n,
                                                     const Proposal* __restrict__ prop,
                                                     const RxnRec* __restrict__ rxns, uint32_t rxn_stride,
                                                     const UnimolRec* __restrict__ unimol,
                                                     const VolSurfRec* __restrict__ vs_rxns,
                                                     uint32_t* __restrict__ cnt) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > n) return;
  if (k == n) {
    cnt[k] = 0;
    return;
  }
  const Proposal& p = prop[sorted_slots[k]];
  cnt[k] = (p.rxn & kUnimolBit)    ? unimol[p.rxn & ~kUnimolBit].num_products[p.pathway]
           : (p.rxn & kVolSurfBit) ? vs_rxns[p.rxn & ~kVolSurfBit].num_new[p.pathway & kPathwayMask]
                                   : new_products(rxns[p.replica * rxn_stride + p.rxn], p.pathway);
}

__global__ __launch_bounds__(256) void k_seg_heads(const uint64_t* __restrict__ keys, uint32_t n,
                                                   uint32_t* __restrict__ head) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t rep = key_replica(keys[k]);
  if (k == 0 || key_replica(keys[k - 1]) != rep) head[rep] = k;
}

// Products of accepted reaction k (in (replica, mover id) order) get ids
// next_id[replica] + (products of earlier reactions of that replica) + i and
// output slots base + pre[k] + i.  Products created before the iteration's
// end move in this iteration (diffuse_react_event.cpp:2883-2886).
// The counter of one carried-out reaction (the volume the mover was in; 0
// without counted volumes), or ~0.
__device__ __forceinline__ uint64_t reaction_counter(const ReactParams& R, const Proposal& p, uint32_t rep) {
  uint32_t slot;
  if (p.rxn & kVolSurfBit) slot = R.rc_vs + (p.rxn & ~kVolSurfBit) * 4 + (p.pathway & kPathwayMask);
  else if (p.rxn & kUnimolBit) slot = R.rc_unimol + (p.rxn & ~kUnimolBit) * 4 + p.pathway;
  else slot = p.rxn * 4 + p.pathway;
  if (!R.rxn_counts || slot >= R.rc_slots) return ~0ull;
  const uint32_t v = (R.rc_volumes > 1 && p.cvol < R.rc_volumes) ? p.cvol : 0u;
  return ((uint64_t)rep * R.rc_volumes + v) * R.rc_slots + slot;
}

// Adds one to counter idx of every lane (~0: none) with one atomic per
// distinct counter in the wave (a phase's reactions share a few counters).
__device__ __forceinline__ void wave_add_counters(unsigned long long* __restrict__ base, uint64_t idx) {
  uint64_t active = __ballot(idx != ~0ull);
  while (active) {
    const int leader = __ffsll((long long)active) - 1;
    const uint64_t li = __shfl((unsigned long long)idx, leader);
    const uint64_t same = __ballot(idx == li) & active;
    if (__lane_id() == leader) atomicAdd(base + li, (unsigned long long)__popcll(same));
    active &= ~same;
  }
}

template <typename PosT>
__global__ __launch_bounds__(256) void k_products(ReactParams R, const uint32_t* __restrict__ sorted_slots,
                                                  const uint64_t* __restrict__ keys, uint32_t n,
                                                  const Proposal* __restrict__ prop, const uint32_t* __restrict__ pre,
                                                  const uint32_t* __restrict__ head,
                                                  const uint32_t* __restrict__ next_id, const uint32_t* __restrict__ hdr,
                                                  uint32_t capacity, RMolView<PosT> out, uint32_t* __restrict__ consumed,
                                                  uint32_t* __restrict__ owner, uint32_t* __restrict__ mlist,
                                                  uint32_t* __restrict__ mcount, RxnEvent* __restrict__ log,
                                                  uint32_t log_base, uint32_t log_cap, uint32_t* __restrict__ err,
                                                  uint32_t* __restrict__ tiles, uint32_t* __restrict__ tile_id,
                                                  uint32_t* __restrict__ tile_consumed, uint32_t* __restrict__ kept) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  // reaction counts, with every lane of the wave present
  wave_add_counters(R.rxn_counts, k < n ? reaction_counter(R, prop[sorted_slots[k]], key_replica(keys[k])) : ~0ull);
  if (k >= n) return;
  const Proposal& p = prop[sorted_slots[k]];
  if (p.rxn & kVolSurfBit) {
    // outcome_products_random (diffuse_react_event.cpp:2446-2933) for V + S:
    // new products in the rule's order get ids and output slots; a new
    // surface product takes S's tile (its output slot stays a hole), a kept
    // S takes its product's orientation, a volume product starts next to
    // the wall and moves for the rest of the step
    const VolSurfRec& vs = R.vs_rxns[p.rxn & ~kVolSurfBit];
    const uint32_t pw = p.pathway & kPathwayMask, drawn = p.pathway >> 8;
    const uint32_t rep = key_replica(keys[k]);
    const uint32_t first_id = next_id[rep] + pre[k] - pre[head[rep]];
    const uint32_t ts = p.partner_slot & ~kTileBit;
    const uint32_t wall = R.tile_wall[ts - rep * R.total_tiles];
    // a given orientation flips when S is not oriented as in the rule (:2636-2658)
    const int flip = (vs.geom_surf != 0 && tile_orient(tiles[ts]) != vs.geom_surf) ? -1 : 1;
    const bool moves_now = cmp_lt_d(p.time, (double)R.iteration + 1.0, 1e-12);
    uint32_t i = 0;
    for (uint32_t q = 0; q < vs.num_products[pw]; ++q) {
      const int o = vs.product_orient[pw][q] != 0 ? vs.product_orient[pw][q] * flip : (((drawn >> q) & 1u) ? 1 : -1);
      const uint32_t ps = vs.products[pw][q];
      if (vs.keep_surf[pw] == q + 1) {
        tiles[ts] = make_tile(ps, o) | (tiles[ts] & kTileMature);
        continue;
      }
      if (vs.keep_vol[pw] == q + 1) continue;  // the mover, which goes on with its step
      const uint32_t id = first_id + i;
      const uint32_t slot = hdr[kHdrNumOut] + pre[k] + i;
      ++i;
      if (slot >= capacity) {
        atomicOr(err, ERR_QUEUE);
        break;
      }
      owner[slot] = kInvalid;
      if (kept) kept[slot] = 0;
      if ((vs.surf_mask[pw] >> q) & 1u) {
        tiles[ts] = make_tile(ps, o);
        tile_id[ts] = id;
        tile_consumed[ts] = 0;  // the product reacts with later movers
        // it takes the receptor's position on the tile (GridPos::
        // make_with_mol_pos, diffuse_react_event.cpp:2040-2050) and diffuses
        // from its reaction time on
        if (R.ttoff) {
          R.ttoff[ts] = p.time - (double)R.iteration;
          R.tbirth[ts] = p.time;
        }
        out.tag[slot] = make_tag(0, MF_DEFUNCT, rep);
        consumed[slot] = kInvalid;
        if (R.born) R.born[slot] = 0;
        continue;
      }
      PhiloxStream aux;  // the wall search's edge re-draws only
      aux.init(R.seed, id, R.iteration, rep, kAuxPass);
      uint32_t e2 = 0;
      const D3 vp = volume_product_position(R, D3{p.x, p.y, p.z}, wall, o, aux, e2);
      if (store_pos(R, out.x, out.y, out.z, slot, vp, true)) e2 |= ERR_ROUND32;
      if (e2) atomicOr(err, e2);
      out.id[slot] = id;
      out.tag[slot] = store_toff(out.toff, slot, p.time - (double)(R.iteration + (moves_now ? 0u : 1u)),
                                 make_tag(ps, MF_SCHEDULE_UNIMOL, rep));
      if (out.utime) out.utime[slot] = kTimeInvalid;
      if (out.birth) out.birth[slot] = p.time;
      // update_counted_volume_id_when_crossing_wall with the product's
      // orientation on a counted object's wall (:2775-2781), else the
      // volume the reactant was in
      if (out.cvol)
        out.cvol[slot] = (R.cv_in[wall] != kInvalid) ? (o > 0 ? R.cv_out[wall] : R.cv_in[wall]) : p.cvol;
      consumed[slot] = 0;
      if (R.born) R.born[slot] = moves_now ? ts + 1 : 0u;
      if (moves_now) mlist[atomicAdd(mcount, 1u)] = slot;
    }
    if (log && log_base + k < log_cap) {
      RxnEvent& e = log[log_base + k];
      e.time = p.time;
      e.x = p.x;
      e.y = p.y;
      e.z = p.z;
      e.iteration = R.iteration;
      e.replica = rep;
      e.mover_id = p.mover_id;
      e.partner_id = p.partner_id;
      e.rxn = p.rxn;
      e.pathway = pw;
      e.first_product_id = first_id;
      e.num_products = vs.num_new[pw];
    }
    return;
  }
  const bool uni = (p.rxn & kUnimolBit) != 0;
  const uint32_t* prods = uni ? R.unimol[p.rxn & ~kUnimolBit].products[p.pathway]
                              : R.rxns[p.replica * R.rxn_stride + p.rxn].products[p.pathway];
  const uint32_t np = uni ? R.unimol[p.rxn & ~kUnimolBit].num_products[p.pathway]
                          : R.rxns[p.replica * R.rxn_stride + p.rxn].num_products[p.pathway];
  const uint32_t rep = key_replica(keys[k]);
  const uint32_t base = hdr[kHdrNumOut];
  const uint32_t first_id = next_id[rep] + pre[k] - pre[head[rep]];
  const bool moves_now = cmp_lt_d(p.time, (double)R.iteration + 1.0, 1e-12);
  for (uint32_t i = 0; i < np; ++i) {
    const uint32_t slot = base + pre[k] + i;
    if (slot >= capacity) {
      atomicOr(err, ERR_QUEUE);
      break;
    }
    if (store_pos(R, out.x, out.y, out.z, slot, D3{p.x, p.y, p.z}, true)) atomicOr(err, ERR_ROUND32);
    out.id[slot] = first_id + i;
    out.tag[slot] = store_toff(out.toff, slot, p.time - (double)(R.iteration + (moves_now ? 0u : 1u)),
                               make_tag(prods[i], MF_SCHEDULE_UNIMOL, rep));
    if (out.utime) out.utime[slot] = kTimeInvalid;
    if (out.cvol) out.cvol[slot] = p.cvol;  // where the reaction happened
    if (out.birth) out.birth[slot] = p.time;
    consumed[slot] = 0;
    owner[slot] = kInvalid;
    if (kept) kept[slot] = 0;
    if (moves_now) mlist[atomicAdd(mcount, 1u)] = slot;
  }
  if (uni && R.unimol[p.rxn & ~kUnimolBit].keep[p.pathway] && moves_now) {
    // the reactant stays and diffuses on from its reaction time, drawing a
    // new lifetime first, in the stream pass of the pathway draw after it
    const uint32_t j = p.mover_slot;
    const uint32_t pass = p.pad & 0xFFu;
    if (pass > (MF_PASS_MASK >> MF_PASS_SHIFT)) {
      atomicOr(err, ERR_PASSES);
    } else {
      const uint32_t f = MF_SCHEDULE_UNIMOL | (pass << MF_PASS_SHIFT) | ((p.pad & 0x100u) ? MF_SKIP_DRAW : 0u);
      const uint32_t t = (out.tag[j] & ~(MF_STEP_FLAGS << 8)) | (f << 8);
      out.tag[j] = store_toff(out.toff, j, p.time - (double)R.iteration, t);
      if (out.utime) out.utime[j] = kTimeInvalid;
      mlist[atomicAdd(mcount, 1u)] = j;
    }
  }
  if (log && log_base + k < log_cap) {
    RxnEvent& e = log[log_base + k];
    e.time = p.time;
    e.x = p.x;
    e.y = p.y;
    e.z = p.z;
    e.iteration = R.iteration;
    e.replica = rep;
    e.mover_id = p.mover_id;
    e.partner_id = p.partner_id;
    e.rxn = p.rxn;
    e.pathway = p.pathway;
    e.first_product_id = first_id;
    e.num_products = np;
  }
}

// The used output slots never exceed the capacity: k_products wrote no
// product past it (and reported ERR_QUEUE).
__global__ void k_bump_ids(const uint64_t* __restrict__ keys, uint32_t n, const uint32_t* __restrict__ pre,
                           const uint32_t* __restrict__ head, uint32_t* __restrict__ next_id,
                           uint32_t* __restrict__ hdr, uint32_t capacity) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t rep = key_replica(keys[k]);
  if (k == n - 1 || key_replica(keys[k + 1]) != rep) next_id[rep] += pre[k + 1] - pre[head[rep]];
  if (k == n - 1) {
    const uint64_t used = (uint64_t)hdr[kHdrNumOut] + pre[n];
    hdr[kHdrNumOut] = used < capacity ? (uint32_t)used : capacity;
  }
}

// Proposals logged without being carried out (the snapshot parity mode).
__global__ __launch_bounds__(256) void k_log_proposals(const uint32_t* __restrict__ plist,
                                                       const uint32_t* __restrict__ pcount,
                                                       const Proposal* __restrict__ prop, uint32_t iteration,
                                                       RxnEvent* __restrict__ log, uint32_t* __restrict__ log_count,
                                                       uint32_t log_cap) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= *pcount) return;
  const Proposal& p = prop[plist[q]];
  const uint32_t k = atomicAdd(log_count, 1u);
  if (k >= log_cap) return;
  RxnEvent& e = log[k];
  e.time = p.time;
  e.x = p.x;
  e.y = p.y;
  e.z = p.z;
  e.iteration = iteration;
  e.replica = p.replica;
  e.mover_id = p.mover_id;
  e.partner_id = p.partner_id;
  e.rxn = p.rxn;
  e.pathway = p.pathway & kPathwayMask;
  e.first_product_id = kInvalid;
  e.num_products = 0;
}

// Per-(replica, species) counts: a workgroup histogram in LDS when the bins
// fit, flushed with one atomic per non-empty bin; global atomics otherwise.
constexpr uint32_t kCountLdsBins = 4096;

__device__ __forceinline__ void count_add(uint32_t* lds, uint64_t* counts, uint32_t bins, uint32_t b) {
  if (bins <= kCountLdsBins) atomicAdd(lds + b, 1u);
  else atomicAdd((unsigned long long*)(counts + b), 1ull);
}

__device__ __forceinline__ void count_flush(const uint32_t* lds, uint64_t* counts, uint32_t bins) {
  if (bins > kCountLdsBins) return;
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < bins; b += blockDim.x)
    if (lds[b]) atomicAdd((unsigned long long*)(counts + b), (unsigned long long)lds[b]);
}

__device__ __forceinline__ void count_clear(uint32_t* lds, uint32_t bins) {
  if (bins > kCountLdsBins) return;
  for (uint32_t b = threadIdx.x; b < bins; b += blockDim.x) lds[b] = 0;
  __syncthreads();
}

template <typename PosT>
__global__ __launch_bounds__(256) void k_count_species(RMolView<PosT> m, uint32_t n, uint32_t num_species,
                                                       uint32_t num_replicas, uint64_t* __restrict__ counts) {
  __shared__ uint32_t h[kCountLdsBins];
  const uint32_t bins = num_species * num_replicas;
  count_clear(h, bins);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t tag = m.tag[i];
    if (tag_flags(tag) & MF_DEFUNCT) continue;
    const uint32_t s = tag_species(tag), r = tag_replica(tag);
    if (s < num_species && r < num_replicas) count_add(h, counts, bins, r * num_species + s);
  }
  count_flush(h, counts, bins);
}

__global__ __launch_bounds__(256) void k_count_tiles(const uint32_t* __restrict__ tiles, uint32_t total_tiles,
                                                     uint32_t num_replicas, uint32_t num_species,
                                                     uint64_t* __restrict__ counts) {
  __shared__ uint32_t h[kCountLdsBins];
  const uint32_t bins = num_species * num_replicas;
  count_clear(h, bins);
  const uint64_t nt = (uint64_t)total_tiles * num_replicas;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nt; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t w = tiles[i];
    if (w == kTileEmpty) continue;
    const uint32_t r = (uint32_t)(i / total_tiles), sp = tile_species(w);
    if (sp < num_species) count_add(h, counts, bins, r * num_species + sp);
  }
  count_flush(h, counts, bins);
}

// --------------------------------------------------------------- launchers --
static inline dim3 blocks_for(uint32_t n) { return dim3((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }
// The full-step list kernel walks its list grid-stride: at most two
// workgroups per CU fit (its LDS), and every workgroup launched beyond the
// list's actual length (a device count under a host bound) still costs a
// workgroup launch.
static inline dim3 move_grid(uint32_t max_movers) { return dim3(std::min<uint32_t>(blocks_for(max_movers).x, 512u)); }

template <typename PosT>
hipError_t ReactPipeline<PosT>::sort(hipStream_t s, uint32_t n_in, uint32_t num_cells, uint32_t* err) {
  hipLaunchKernelGGL(k_fill_u32, blocks_for(num_cells + 1), dim3(256), 0, s, cell_count, 0u, num_cells + 1);
  if (n_in)
    hipLaunchKernelGGL(k_cell_keys<PosT>, blocks_for(n_in), dim3(256), 0, s, M, n_in, R.P, R.G, key, rank, cell_count,
                       err);
  size_t tmp = scan_tmp_bytes;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(scan_tmp, tmp, cell_count, cell_start, num_cells + 1, s);
  if (e != hipSuccess) return e;
  if (n_in) {
    hipLaunchKernelGGL(k_scatter<PosT>, blocks_for(n_in), dim3(256), 0, s, M, n_in, key, rank, cell_start, Sm);
    // claim words of the snapshot's slots (n_in bounds it), written in order
    hipLaunchKernelGGL(k_fill_u32, blocks_for(n_in), dim3(256), 0, s, consumed, 0u, n_in);
    if (kept) hipLaunchKernelGGL(k_fill_u32, blocks_for(n_in), dim3(256), 0, s, kept, 0u, n_in);
    if (R.born) hipLaunchKernelGGL(k_fill_u32, blocks_for(n_in), dim3(256), 0, s, R.born, 0u, n_in);
    hipLaunchKernelGGL(k_fill_u32, blocks_for(n_in), dim3(256), 0, s, owner, kInvalid, n_in);
  }
  return hipGetLastError();
}

template <typename PosT>
size_t ReactPipeline<PosT>::scan_bytes(uint32_t num_cells, uint32_t capacity) {
  size_t a = 0, b = 0, c = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, a, (uint32_t*)nullptr, (uint32_t*)nullptr, num_cells + 1);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, (uint32_t*)nullptr, (uint32_t*)nullptr, capacity + 1);
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, c, (uint64_t*)nullptr, (uint64_t*)nullptr, (uint32_t*)nullptr,
                                           (uint32_t*)nullptr, capacity);
  return std::max(a, std::max(b, c));
}

// Vertex pools of exact_disk (40 bytes a record, private memory per lane)
// in the full-step passes: the drain runs with kDrainPool records in place
// (drain_inline) or defers every step with a disk that meets walls, the list
// moves (reruns, products) start with kListPool, and the steps they defer
// run again with exd::kPool.  The pool in place costs the drain about a
// quarter of its time even where disks rarely meet walls; deferring costs a
// second, sparse pass over the deferred steps.
constexpr int kDrainPool = 48;
constexpr int kListPool = 48;

template <typename PosT>
hipError_t ReactPipeline<PosT>::move(hipStream_t s, bool snapshot_src, const uint32_t* list, const uint32_t* count,
                                     uint32_t max_movers, double barrier, bool propose_only, uint32_t* err) {
  if (max_movers == 0) return hipSuccess;
  Snapshot<PosT> S{Sm.x, Sm.y, Sm.z, Sm.id, Sm.tag, consumed, Sm.toff, Sm.utime, Sm.cvol, Sm.birth};
  const uint32_t po = propose_only ? 1u : 0u;
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, hdr + kHdrProposals, 0u, 1u);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, hdr + kHdrExd, 0u, 1u);
  // the deferred steps go to a list that is free during this move: mlist
  // until the products are listed, else the rerun list settle() does not
  // read next
  uint32_t* const xlist = (list == nullptr) ? mlist : (list == rlist0 ? rlist1 : rlist0);
  const RMolView<PosT> src = snapshot_src ? Sm : M;
  if (list == nullptr && snapshot_src) {
    // every snapshot slot: light pass, then the queued movers in full
    SlowQueue Q = queue;
    Q.shard_cap = shard_capacity(max_movers, kQueueShards / 2);
    hipLaunchKernelGGL(k_fill_u32, dim3(kQueueShards / 256), dim3(256), 0, s, Q.count, 0u, kQueueShards);
    hipLaunchKernelGGL(k_move_light<PosT>, blocks_for(max_movers), dim3(256), 0, s, R, S, M, count, species, barrier,
                       Q, err);
    Q.nshards = kQueueShards / 2 + std::min<uint32_t>(kQueueShards / 2, blocks_for(max_movers).x);
    if (drain_inline)
      hipLaunchKernelGGL((k_move_drain<PosT, kDrainPool>), dim3(Q.nshards), dim3(256), 0, s, R, S, Sm, M, Q, species,
                         barrier, po, prop, plist, hdr + kHdrProposals, owner, tile_owner, err, ulist,
                         hdr + kHdrUnimolDue, xlist, hdr + kHdrExd);
    else
      hipLaunchKernelGGL((k_move_drain<PosT, 0>), dim3(Q.nshards), dim3(256), 0, s, R, S, Sm, M, Q, species, barrier,
                         po, prop, plist, hdr + kHdrProposals, owner, tile_owner, err, ulist, hdr + kHdrUnimolDue,
                         xlist, hdr + kHdrExd);
  } else {
    hipLaunchKernelGGL((k_move<PosT, kListPool>), move_grid(max_movers), dim3(256), 0, s, R, S, src, M, list, count,
                       species, barrier, po, prop, plist, hdr + kHdrProposals, owner, tile_owner, err, ulist,
                       hdr + kHdrUnimolDue, xlist, hdr + kHdrExd);
  }
  // the deferred steps, with the large pool
  hipLaunchKernelGGL((k_move<PosT, exd::kPool>), move_grid(max_movers), dim3(256), 0, s, R, S, src, M, xlist,
                     hdr + kHdrExd, species, barrier, po, prop, plist, hdr + kHdrProposals, owner, tile_owner, err,
                     ulist, hdr + kHdrUnimolDue, nullptr, nullptr);
  return hipGetLastError();
}

template <typename PosT>
hipError_t ReactPipeline<PosT>::settle(hipStream_t s, uint32_t max_props, uint32_t* rlist, uint32_t* rcount) {
  // grid-stride over the device's proposal count; max_props bounds it
  const dim3 grid = dim3(std::min<uint32_t>(blocks_for(max_props).x, 1024u));
  hipLaunchKernelGGL(k_accept<PosT>, grid, dim3(256), 0, s, plist, hdr + kHdrProposals, prop, owner,
                     tile_owner, consumed, tile_consumed, M.tag, R.rxns, R.rxn_stride, R.vs_rxns, kept, aprop,
                     alist, hdr + kHdrAccepted, R.ctr);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, rcount, 0u, 1u);
  hipLaunchKernelGGL(k_requeue, grid, dim3(256), 0, s, plist, hdr + kHdrProposals, prop, owner,
                     tile_owner, consumed, rlist, rcount);
  return hipGetLastError();
}

template <typename PosT>
hipError_t ReactPipeline<PosT>::fire_unimol(hipStream_t s, uint32_t max_due) {
  if (max_due == 0) return hipSuccess;
  hipLaunchKernelGGL(k_unimol_fire, blocks_for(max_due), dim3(256), 0, s, R, ulist, hdr + kHdrUnimolDue, prop,
                     consumed, M.tag, aprop, alist, hdr + kHdrAccepted);
  return hipGetLastError();
}

template <typename PosT>
hipError_t ReactPipeline<PosT>::products(hipStream_t s, uint32_t n_acc, uint32_t capacity, RxnEvent* log,
                                         uint32_t log_base, uint32_t log_cap, bool first_phase, uint32_t* err) {
  if (n_acc == 0) return hipSuccess;
  // the accepted proposals: copies when movers can stay, else in place
  const Proposal* const P = aprop ? aprop : prop;
  hipLaunchKernelGGL(k_acc_keys, blocks_for(n_acc), dim3(256), 0, s, alist, n_acc, P, Sm.tag, Sm.toff, hdr,
                     first_phase ? 1u : 0u, keys_a);
  size_t tmp = scan_tmp_bytes;
  hipError_t e =
      hipcub::DeviceRadixSort::SortPairs(scan_tmp, tmp, keys_a, keys_b, alist, alist_sorted, n_acc, 0, 64, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_prod_counts, blocks_for(n_acc + 1), dim3(256), 0, s, alist_sorted, n_acc, P, R.rxns,
                     R.rxn_stride, R.unimol, R.vs_rxns, cnt);
  tmp = scan_tmp_bytes;
  e = hipcub::DeviceScan::ExclusiveSum(scan_tmp, tmp, cnt, pre, n_acc + 1, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_seg_heads, blocks_for(n_acc), dim3(256), 0, s, keys_b, n_acc, head);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, hdr + kHdrProductMovers, 0u, 1u);
  hipLaunchKernelGGL(k_products<PosT>, blocks_for(n_acc), dim3(256), 0, s, R, alist_sorted, keys_b, n_acc, P, pre,
                     head, next_id, hdr, capacity, M, consumed, owner, mlist, hdr + kHdrProductMovers, log, log_base,
                     log_cap, err, tiles, tile_id, tile_consumed, kept);
  hipLaunchKernelGGL(k_bump_ids, blocks_for(n_acc), dim3(256), 0, s, keys_b, n_acc, pre, head, next_id, hdr,
                     capacity);
  return hipGetLastError();
}

// dst[*dst_count + i] = src[i] for i < *src_count (the counts are read on
// the device; max bounds the grid).
__global__ __launch_bounds__(256) void k_append_list(const uint32_t* __restrict__ src,
                                                     const uint32_t* __restrict__ src_count,
                                                     uint32_t* __restrict__ dst, const uint32_t* __restrict__ dst_count) {
  const uint32_t n = *src_count, base = *dst_count;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[base + i] = src[i];
}
__global__ void k_take_count(uint32_t* __restrict__ dst_count, uint32_t* __restrict__ src_count) {
  if (threadIdx.x == 0) {
    *dst_count += *src_count;
    *src_count = 0;
  }
}

template <typename PosT>
hipError_t ReactPipeline<PosT>::continuations(hipStream_t s, uint32_t max_cont) {
  if (max_cont == 0) return hipSuccess;
  hipLaunchKernelGGL(k_append_list, dim3(std::min<uint32_t>(blocks_for(max_cont).x, 2048u)), dim3(256), 0, s, clist,
                     hdr + kHdrContinue, mlist, hdr + kHdrProductMovers);
  hipLaunchKernelGGL(k_take_count, dim3(1), dim3(64), 0, s, hdr + kHdrProductMovers, hdr + kHdrContinue);
  return hipGetLastError();
}

template <typename PosT>
hipError_t ReactPipeline<PosT>::log_proposals(hipStream_t s, uint32_t max_props, uint32_t iteration, RxnEvent* log,
                                              uint32_t* log_count, uint32_t log_cap) {
  if (max_props == 0) return hipSuccess;
  hipLaunchKernelGGL(k_log_proposals, blocks_for(max_props), dim3(256), 0, s, plist, hdr + kHdrProposals, prop,
                     iteration, log, log_count, log_cap);
  return hipGetLastError();
}

template <typename PosT>
hipError_t ReactPipeline<PosT>::count_species(hipStream_t s, uint32_t n, uint32_t num_species, uint32_t num_replicas,
                                              uint64_t* counts) {
  if (n)
    hipLaunchKernelGGL(k_count_species<PosT>, dim3(std::min<uint32_t>(blocks_for(n).x, 2048u)), dim3(256), 0, s, M,
                       n, num_species, num_replicas, counts);
  const uint64_t nt = (uint64_t)R.total_tiles * num_replicas;
  if (nt)
    hipLaunchKernelGGL(k_count_tiles, dim3((uint32_t)std::min<uint64_t>((nt + 255) / 256, 2048u)), dim3(256), 0, s,
                       tiles, R.total_tiles, num_replicas, num_species, counts);
  return hipGetLastError();
}

// Live molecules per (replica, counted volume, species); molecules whose
// index is outside [0, num_volumes) are not counted.
__global__ __launch_bounds__(256) void k_count_by_volume(const uint32_t* __restrict__ cvol,
                                                         const uint32_t* __restrict__ tag, uint32_t n,
                                                         uint32_t num_species, uint32_t num_replicas,
                                                         uint32_t num_volumes, uint64_t* __restrict__ counts) {
  __shared__ uint32_t h[kCountLdsBins];
  const uint32_t bins = num_species * num_replicas * num_volumes;
  count_clear(h, bins);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t t = tag[i];
    if (tag_flags(t) & MF_DEFUNCT) continue;
    const uint32_t s = tag_species(t), r = tag_replica(t), v = cvol[i];
    if (s < num_species && r < num_replicas && v < num_volumes)
      count_add(h, counts, bins, (r * num_volumes + v) * num_species + s);
  }
  count_flush(h, counts, bins);
}

hipError_t launch_count_by_volume(hipStream_t s, const uint32_t* cvol, const uint32_t* tag, uint32_t n,
                                  uint32_t num_species, uint32_t num_replicas, uint32_t num_volumes,
                                  uint64_t* counts) {
  if (n)
    hipLaunchKernelGGL(k_count_by_volume, dim3(std::min<uint32_t>(blocks_for(n).x, 2048u)), dim3(256), 0, s, cvol,
                       tag, n, num_species, num_replicas, num_volumes, counts);
  return hipGetLastError();
}

template struct ReactPipeline<double>;
template struct ReactPipeline<float>;

// ------------------------------------------------- allocation-free counts --
// The counters of mcg_count_species_noalloc live in this file-scope device
// array (each GPU the library runs on holds its own copy until the process
// ends).  Host code never passes the array's name to a kernel: it is a
// host-side stand-in, not a device address.  count_scratch_address looks the
// address up once per device (with that device current) and keeps it.
static __device__ __align__(256) uint64_t scratch[kCountScratchMax];

hipError_t count_scratch_address(uint64_t** out) {
  static std::mutex mu;
  static uint64_t* per_device[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (per_device[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    per_device[dev] = static_cast<uint64_t*>(p);
  }
  *out = per_device[dev];
  return hipSuccess;
}

hipError_t launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, uint32_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, blocks_for(n), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch_fill_u64(hipStream_t s, unsigned long long* p, unsigned long long v, uint32_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, blocks_for(n), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

__global__ void k_copy_u32(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) {
  if (threadIdx.x == 0) *dst = *src;
}

hipError_t launch_copy_u32(hipStream_t s, const uint32_t* src, uint32_t* dst) {
  hipLaunchKernelGGL(k_copy_u32, dim3(1), dim3(64), 0, s, src, dst);
  return hipGetLastError();
}

}  // namespace mcg
