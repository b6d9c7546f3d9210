// Host-side view of the reaction pipeline (kernels_react.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "device/react_step.h"
#include "kernels.h"

namespace mcg {

// Molecule arrays of the reactive path: positions, id, tag and the offset of
// the molecule's diffusion time past the start of its iteration (non-zero
// only for products created inside a step).
template <typename PosT>
struct RMolView {
  PosT* x;
  PosT* y;
  PosT* z;
  uint32_t* id;
  uint32_t* tag;
  double* toff;
  double* utime;  // unimolecular reaction time (absolute) or kTimeInvalid; null without unimolecular classes
  uint32_t* cvol;  // counted volume index (Molecule::v.counted_volume_index); null without counted volumes
  double* birth;   // Molecule::birthday (absolute); null unless a species' time step exceeds one iteration
};

// One carried-out (or, in the snapshot mode, proposed) reaction: 64 bytes,
// the layout of mcg_rxn_event in the C ABI.
struct RxnEvent {
  double time;  // absolute time (iterations)
  double x, y, z;
  uint32_t iteration, replica;
  uint32_t mover_id, partner_id;
  uint32_t rxn, pathway;
  uint32_t first_product_id, num_products;
};
static_assert(sizeof(RxnEvent) == 64, "event layout");

// device counters
enum : uint32_t {
  kHdrNumOut = 0,         // used output slots (snapshot size + products)
  kHdrProposals = 1,      // proposals of the current round
  kHdrAccepted = 2,       // accepted reactions of the current phase
  kHdrProductMovers = 3,  // products that move in this iteration
  kHdrRerun0 = 4,         // movers queued for another round (two lists)
  kHdrRerun1 = 5,
  kHdrSnapshot = 6,       // molecules in the snapshot (the iteration's movers)
  kHdrLog = 7,            // events logged (snapshot mode)
  kHdrUnimolDue = 8,      // molecules whose unimolecular reaction falls in this phase
  kHdrErr = 9,            // error bits (ERR_*) of the call's kernels
  kHdrExd = 10,           // movers the full step deferred to its exact_disk pass
  kHdrContinue = 11,      // movers of the phase that diffuse again in this iteration
  kHdrSurfProps = 12,     // surface phase: moves to other tiles proposed in the round
  kHdrSurfLose0 = 13,     // surface phase: movers whose tile went to a smaller id (two lists)
  kHdrSurfLose1 = 14,
  kHdrSurfCont0 = 15,     // surface phase: molecules that step again in this iteration (two lists)
  kHdrSurfCont1 = 16,
  kHdrWords = 18,
};

template <typename PosT>
struct ReactPipeline {
  ReactParams R;
  SpeciesTable species;
  RMolView<PosT> M;   // molecules: input of the sort, output of the moves
  RMolView<PosT> Sm;  // cell-sorted snapshot
  uint32_t *key, *rank, *cell_count, *cell_start;
  uint32_t *consumed, *owner;
  uint32_t* kept;   // per slot: reactions its mover stayed through this iteration (null: no class keeps a reactant)
  Proposal* aprop;  // accepted proposals (null: read in place from prop)
  uint32_t *tiles, *tile_id, *tile_owner, *tile_consumed;  // surface molecules per (replica, tile)
  Proposal* prop;
  uint32_t *plist, *alist, *alist_sorted, *rlist0, *rlist1, *mlist, *ulist;
  uint32_t* clist;  // slots of the phase's movers that diffuse again (kHdrContinue of them)
  uint64_t *keys_a, *keys_b;
  uint32_t *cnt, *pre, *head, *next_id;
  uint32_t* hdr;
  SlowQueue queue;  // light-pass queue of movers that need the full step
  bool drain_inline;  // the drain runs exact_disk in place (else it defers those steps)
  void* scan_tmp;
  size_t scan_tmp_bytes;

  static size_t scan_bytes(uint32_t num_cells, uint32_t capacity);
  hipError_t sort(hipStream_t s, uint32_t n_in, uint32_t num_cells, uint32_t* err);
  hipError_t move(hipStream_t s, bool snapshot_src, const uint32_t* list, const uint32_t* count, uint32_t max_movers,
                  double barrier, bool propose_only, uint32_t* err);
  hipError_t settle(hipStream_t s, uint32_t max_props, uint32_t* rlist, uint32_t* rcount);
  hipError_t fire_unimol(hipStream_t s, uint32_t max_due);
  hipError_t products(hipStream_t s, uint32_t n_acc, uint32_t capacity, RxnEvent* log, uint32_t log_base,
                      uint32_t log_cap, bool first_phase, uint32_t* err);
  // appends the phase's continuing movers to the next phase's mover list
  // (after its products) and clears their count
  hipError_t continuations(hipStream_t s, uint32_t max_cont);
  hipError_t log_proposals(hipStream_t s, uint32_t max_props, uint32_t iteration, RxnEvent* log, uint32_t* log_count,
                           uint32_t log_cap);
  hipError_t count_species(hipStream_t s, uint32_t n, uint32_t num_species, uint32_t num_replicas, uint64_t* counts);
};

// Allocation-free species counts: at most kCountScratchMax (replica,
// species) counters in a device array of the library; its address on the
// current device (looked up once per device).
constexpr uint32_t kCountScratchMax = 65536;  // 512 KiB per device
constexpr int kMaxDevices = 64;
hipError_t count_scratch_address(uint64_t** out);

hipError_t launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, uint32_t n);
hipError_t launch_count_by_volume(hipStream_t s, const uint32_t* cvol, const uint32_t* tag, uint32_t n,
                                  uint32_t num_species, uint32_t num_replicas, uint32_t num_volumes,
                                  uint64_t* counts);
hipError_t launch_fill_u64(hipStream_t s, unsigned long long* p, unsigned long long v, uint32_t n);
hipError_t launch_copy_u32(hipStream_t s, const uint32_t* src, uint32_t* dst);

}  // namespace mcg
