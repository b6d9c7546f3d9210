// The engine context behind the C ABI (include/mcg.h); internal to libmcg.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mcg.h"
#include "kernels.h"
#include "kernels_react.h"
#include "kernels_surface.h"

namespace mcg {

// Records the error text for mcg_last_error and returns `code`.
int fail(int code, const std::string& msg);

// MCG_E_STATE when an earlier step left the context's state undefined.
int check_alive(const mcg_ctx* c);
#define MCG_ALIVE(c)                                   \
  do {                                                 \
    if (int st_ = ::mcg::check_alive(c)) return st_;   \
  } while (0)

#define HIP_TRY(expr)                                                                                  \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess) return ::mcg::fail(MCG_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t count) {
    release();
    if (count == 0) return hipSuccess;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
    if (e == hipSuccess) n = count;
    else p = nullptr;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// Buffers of the reactive path (allocated by mcg_set_bimol_rxns).
struct ReactState {
  bool on = false;
  std::vector<RxnRec> rxns;           // base table
  std::vector<RxnRec> rxns_replicas;  // num_replicas * rxns.size() when per replica
  std::vector<int32_t> pair;  // species x species -> rxn index or -1
  DevBuf<RxnRec> rxns_d;
  DevBuf<int32_t> pair_d;
  DevBuf<double> toff;                       // molecules' diffusion-time offsets
  DevBuf<double> utime, sutime;              // unimolecular reaction times (molecules, snapshot)
  std::vector<UnimolRec> unimol;             // per species
  DevBuf<UnimolRec> unimol_d;
  DevBuf<double> sx, sy, sz;                 // snapshot positions (FP64 form)
  DevBuf<float> sxf, syf, szf;               // snapshot positions (FP32 form)
  DevBuf<uint32_t> sid, stag;
  DevBuf<double> stoff;
  DevBuf<uint32_t> key, rank, consumed, owner;
  DevBuf<Proposal> prop;
  // classes that keep a reactant: per-slot counts of the reactions a mover
  // stayed through, and copies of the accepted proposals (allocated with
  // the first such class)
  DevBuf<uint32_t> kept;
  DevBuf<Proposal> aprop;
  DevBuf<uint32_t> born;       // volume products of volume-surface classes: per slot, 1 + creation tile
  // counted volumes (Partition's counted volume indices): per molecule and
  // snapshot slot, and per wall the index inside / outside its object
  // (kInvalid: the wall's object is not counted)
  uint32_t num_volumes = 0;
  DevBuf<uint32_t> cvol, scvol, cv_in, cv_out;
  // reaction counts since the first step (layout: ReactParams::rxn_counts)
  DevBuf<unsigned long long> rxn_counts;
  uint32_t rc_slots = 0, rc_unimol = 0, rc_vs = 0, rc_volumes = 0;
  DevBuf<uint32_t> tile_wall;  // per tile of one replica: its wall
  // species time steps other than one iteration: birthdays per molecule
  // and snapshot slot (Molecule::birthday, kept while a species' time step
  // exceeds one iteration: get_max_time's maturity rule), and the list of
  // movers that diffuse again in the same iteration
  DevBuf<double> birth, sbirth;
  DevBuf<uint32_t> clist;
  DevBuf<uint32_t> plist, alist, alist_sorted, rlist0, rlist1, mlist, ulist;
  DevBuf<uint64_t> keys_a, keys_b;
  DevBuf<uint32_t> cnt, pre, head, next_id;
  DevBuf<uint32_t> hdr;
  DevBuf<uint32_t> cell_count, cell_start;
  DevBuf<unsigned char> scan_tmp;
  CellGrid grid{};
  uint32_t num_cells = 0;
  bool grid_dirty = true;
  std::vector<uint32_t> next_id_host;  // per replica: 1 + largest id seen
  bool next_id_dirty = true;
  uint32_t* hdr_host = nullptr;        // pinned mirror of hdr
  // the drain's form (1 exact_disk in place, 2 deferred): fixed by a flag,
  // else timed on the first iterations after the geometry or the molecules
  // are set (0 while undecided; the trials alternate the forms)
  uint32_t drain_form = 0;
  uint32_t form_trials = 0;
  double form_ms[2] = {0, 0};
  uint32_t form_n[2] = {0, 0};
  uint32_t last_form = 0;
  void reset_drain_form() {
    drain_form = 0;
    form_trials = 0;
    form_ms[0] = form_ms[1] = 0;
    form_n[0] = form_n[1] = 0;
  }
  // volume-surface reaction classes
  std::vector<VolSurfRec> vs;
  std::vector<int32_t> pair_vs;  // vol species x surf species -> class or -1
  DevBuf<VolSurfRec> vs_d;
  DevBuf<int32_t> pair_vs_d;
  // surface molecules: per-wall tile grids, tiles replica-major
  // [replica * total_tiles + grid.tile_base + tile]
  std::vector<WallGrid> grids_host;  // per wall once a wall has a grid
  std::vector<uint32_t> tiles_host, tile_id_host;
  uint32_t total_tiles = 0;
  bool tiles_dirty = false;      // the host copy is newer than the device's
  bool tiles_on_device = false;  // steps may have changed the device copy
  DevBuf<WallGrid> grids_d;
  DevBuf<uint32_t> tiles, tile_id, tile_owner, tile_consumed;
  // surface molecules' positions in their walls' uv frames, diffusion-
  // time offsets and birthdays, per tile slot (host copies; on the device
  // while surface species diffuse), and the surface phase's buffers
  std::vector<double> tu_host, tv_host, ttoff_host, tbirth_host;
  bool surf_on = false;  // some surface species diffuses: grids on every wall, the surface phase runs
  DevBuf<double> tu, tv, ttoff, tbirth, spu, spv;
  DevBuf<unsigned long long> sclaim;
  DevBuf<uint32_t> sprop, splist, slose0, slose1, sdone, scont0, scont1;
  DevBuf<WallEdgeRec> edges_d;
  // reaction log of the last mcg_step call
  DevBuf<RxnEvent> log;
  uint64_t log_cap = 0;
  uint64_t log_count = 0;
};

}  // namespace mcg

struct mcg_ctx {
  mcg_config cfg{};
  mcg::PartitionParams P{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evm0 = nullptr, evm1 = nullptr;

  // geometry
  std::vector<mcg::WallRec64> walls_host;
  std::vector<uint32_t> off_host, idx_host;
  std::vector<double> vert_host;
  mcg::DevBuf<mcg::WallRec64> walls64;
  mcg::DevBuf<mcg::WallRec32> walls32;
  mcg::DevBuf<uint32_t> off, idx;
  mcg::DevBuf<mcg::WallRec64> ent64;  // wall records in list (entry) order
  mcg::DevBuf<mcg::WallRec32> ent32;
  mcg::DevBuf<uint32_t> ent_box;  // sub-cell box of every list entry
  bool have_geometry = false;
  // surface classes
  std::vector<uint32_t> wall_class_host;
  mcg::DevBuf<uint32_t> wall_class_d;
  mcg::DevBuf<uint8_t> sc_action_d;
  uint32_t num_surface_classes = 0;
  bool sc_on = false;
  mcg::SurfaceClasses surface_classes() const {
    return {wall_class_d.p, sc_action_d.p, num_species, sc_on ? 1u : 0u};
  }

  // species: one table, or after mcg_set_replica_params one table per replica
  mcg::DevBuf<mcg::SpeciesRec> species;
  std::vector<mcg::SpeciesRec> species_host;      // base table (num_species)
  std::vector<mcg::SpeciesRec> species_replicas;  // num_replicas * num_species when per replica
  uint32_t num_species = 0;
  bool per_replica = false;
  mcg::SpeciesTable species_table() const { return {species.p, per_replica ? num_species : 0u}; }
  // a species (of any replica) whose time step is not one iteration / is
  // longer than one iteration
  bool custom_steps() const { return any_step([](double t) { return !(t == 1.0); }); }
  bool long_steps() const { return any_step([](double t) { return t > 1.0; }); }
  // a surface species (of any replica) that diffuses
  bool surface_diffusion() const {
    for (const mcg::SpeciesRec& s : per_replica ? species_replicas : species_host)
      if ((s.flags & mcg::SPF_SURFACE) && s.space_step > 0) return true;
    return false;
  }
  template <typename F>
  bool any_step(F f) const {
    for (const mcg::SpeciesRec& s : per_replica ? species_replicas : species_host)
      if (f(s.time_step)) return true;
    return false;
  }

  // molecules (structure of arrays): FP64 or FP32 positions by precision.
  // n counts array slots; with reactions on, consumed molecules stay as
  // defunct slots until the next iteration's sort drops them.
  uint64_t n = 0;
  mcg::DevBuf<double> x, y, z;
  mcg::DevBuf<float> xf, yf, zf;
  mcg::DevBuf<uint32_t> id, tag;
  float max_abs_coord = 0;
  bool fp32() const { return cfg.precision == MCG_PREC_FP32_FALLBACK; }
  double mol_lo[3] = {1e300, 1e300, 1e300}, mol_hi[3] = {-1e300, -1e300, -1e300};

  // diagnostics
  mcg::DevBuf<uint32_t> err;
  mcg::DevBuf<uint32_t> queue, qcount;  // hard-path queue of molecule indices
  mcg::DevBuf<uint32_t> mqueue, mqcount;  // middle-pass queue (free diffusion)
  mcg::DevBuf<uint32_t> trace_wall, trace_count;
  double last_kernel_ms = 0;
  uint32_t last_launches = 0;

  mcg::ReactState rx;
  bool stepped = false;  // reaction classes are fixed once a step has run
  uint64_t clock = 0;    // iteration the next mcg_step starts at (Molecule::diffusion_time base)
  bool failed = false;   // a step failed after its kernels ran: the molecule state is undefined
  std::string fail_msg;
  mcg::DevBuf<unsigned long long> counters;  // CTR_* of the current mcg_step call
  uint64_t* count_scratch = nullptr;  // the library's count array on this context's device

  ~mcg_ctx();
};

namespace mcg {
// Species time steps (internal units): a molecule diffuses at most eight
// times in one iteration (its stream passes), so steps are at least 1/8.
inline bool valid_time_step(double t) { return t >= 0.125 && t <= 1e12; }
constexpr const char* kTimeStepRange = "species time step outside [1/8, 1e12] iterations";
// The reactive iteration loop (engine_react.cpp).
int react_step(mcg_ctx* c, uint32_t first_iteration, uint32_t num_iterations, double time_up_to_next_barrier,
               mcg_step_stats* stats);
// Allocates the reactive path's buffers once (reaction classes of any kind).
int ensure_react_buffers(mcg_ctx* c);
// Surface species that diffuse: grids on every wall, the surface net and
// the surface phase's buffers (engine_react.cpp); the surface phase of
// one iteration.
int ensure_surface_diffusion(mcg_ctx* c);
int download_tiles(mcg_ctx* c);
int surface_phase(mcg_ctx* c, uint32_t iteration, double barrier, uint32_t* err);
// Species time steps other than one iteration run on the reactive path;
// long ones keep birthdays (allocated here, the existing molecules' set to
// the clock).
int ensure_step_buffers(mcg_ctx* c);
// Live molecules per (replica, species) into counts (size num_replicas *
// num_species); surface molecules too when with_surface.
int count_species(mcg_ctx* c, uint64_t* counts, bool with_surface = true);
// The same into the device buffer dcounts (num_replicas * num_species
// entries, cleared here); allocates nothing when allow_alloc is false.
int count_species_into(mcg_ctx* c, uint64_t* counts, uint64_t* dcounts, bool with_surface, bool allow_alloc);
// Uploads the host copy of the tiles when it is newer (before steps and
// counts); with allow_alloc false, MCG_E_STATE when that needs new buffers.
int sync_tiles_to_device(mcg_ctx* c, bool allow_alloc = true);
}  // namespace mcg
