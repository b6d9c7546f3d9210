// Volume-volume reactions on the GPU: the molecule-collision half of
// ray_trace_vol (diffuse_react_event.cpp:713-772), collide_mol
// (collision_utils.inl:464-515), exact_disk (exact_disk_utils.inl:840-1145)
// and test_bimolecular (rxn_utils.inl:301-414), against a snapshot of the
// start-of-iteration molecules held in cell-sorted order.
#pragma once
#include "path64.h"
#include "types.h"

namespace mcg {

constexpr int kMaxPathways = 4;
constexpr int kMaxProducts = 3;

// One bimolecular reaction class (both reactants volume molecules).
struct RxnRec {
  double max_fixed_p;
  double cum_probs[kMaxPathways];
  float max_p_f;  // rxn_utils.inl:367 holds max_p in a float
  uint32_t num_pathways;
  uint32_t reactant_a, reactant_b;
  uint32_t num_products[kMaxPathways];   // new molecules (kept reactants excluded)
  uint32_t products[kMaxPathways][kMaxProducts];
  uint32_t keep[kMaxPathways];           // kKeepA | kKeepB: reactants that stay
};
// A reactant is kept when its species is among the pathway's products
// (outcome_products_random, diffuse_react_event.cpp:2551-2565, 2686-2727):
// the first product of that species is the reactant itself, not a new
// molecule.  Reactant a takes the first match; for a == b, b is kept only
// when the species appears twice.
constexpr uint32_t kKeepA = 1u, kKeepB = 2u;

// The unimolecular class of one species: cumulative pathway rates per
// internal time unit (k_tot = cum[num_pathways - 1]) and products.
struct UnimolRec {
  double cum[kMaxPathways];
  uint32_t num_pathways;  // 0: the species has no unimolecular reaction
  uint32_t num_products[kMaxPathways];  // new molecules (a kept reactant excluded)
  uint32_t products[kMaxPathways][kMaxProducts];
  uint32_t keep[kMaxPathways];          // 1: the reactant stays and diffuses on
};
constexpr uint32_t kUnimolBit = 0x80000000u;  // Proposal::rxn of a unimolecular event: bit | species

// Surface grid of a wall (Grid::initialize, src4/wall.cpp:38-74); num_tiles
// 0 = the wall has no grid.  Tiles of all grids are numbered from tile_base.
struct WallGrid {
  double strip_width_rcp, vert2_slope, fullslope, binding_factor;
  double vert0_u, vert0_v;
  uint32_t n_axis, num_tiles, tile_base, pad;
};
static_assert(sizeof(WallGrid) == 64, "grid record");

// A volume + surface reaction class: the volume reactant hitting the tile of
// the surface reactant.  The volume reactant is consumed; the surface
// reactant stays (keep_surf: its product index + 1) or its tile goes to the
// one new surface product; volume products start next to the wall.
// Products are in the rule's order; orientation 0 is drawn per reaction.
struct VolSurfRec {
  double max_fixed_p;
  double cum_probs[kMaxPathways];
  float max_p_f;
  uint32_t num_pathways;
  int32_t geom_vol, geom_surf;  // orientations in the rule (0 none, +1, -1)
  uint32_t surf_species;
  uint32_t num_products[kMaxPathways];
  uint32_t products[kMaxPathways][kMaxProducts];
  int32_t product_orient[kMaxPathways][kMaxProducts];
  uint32_t surf_mask[kMaxPathways];  // bit q: product q is a surface species
  uint32_t keep_surf[kMaxPathways];  // 0, or 1 + the product index that is the kept surface reactant
  uint32_t num_new[kMaxPathways];    // new molecules (the kept reactants excluded)
  uint32_t keep_vol[kMaxPathways];   // 0, or 1 + the product index that is the kept volume reactant
};
// Proposal::pathway of a volume-surface reaction: pathway | drawn
// orientation bits << 8 (bit q: product q drawn up)
constexpr uint32_t kPathwayMask = 0xFFu;
constexpr uint32_t kVolSurfBit = 0x40000000u;  // Proposal::rxn of a volume-surface reaction: bit | class
constexpr uint32_t kTileBit = 0x80000000u;     // Proposal::partner_slot naming a tile: bit | tile slot
constexpr uint32_t kTileEmpty = 0xFFFFFFFFu;
// tile word: species (8 bits) | orientation + 1 (2 bits) << 8 | mature
// (MOLECULE_FLAG_MATURE: a surface molecule with a long time step whose
// steps no longer lengthen) << 10 | delayed << 11 (within an iteration's
// surface phase: the molecule's diffusion time did not fall on the
// iteration's start, so its steps come after the others')
constexpr uint32_t kTileMature = 1u << 10;
constexpr uint32_t kTileDelayed = 1u << 11;
__device__ __host__ __forceinline__ uint32_t tile_species(uint32_t w) { return w & 0xFFu; }
__device__ __host__ __forceinline__ int tile_orient(uint32_t w) { return (int)((w >> 8) & 3u) - 1; }
__device__ __host__ __forceinline__ uint32_t make_tile(uint32_t species, int orient) {
  return (species & 0xFFu) | ((uint32_t)(orient + 1) << 8);
}

// Cell grid over the region molecules can occupy: subpartition index box
// [lo, lo + dim) per replica, per_replica = dim.x * dim.y * dim.z cells,
// numbered x fastest.  The snapshot is sorted by bin = (replica, species,
// cell), so a mover reads exactly the partners of the species it reacts with,
// and one species' molecules in a row of cells along x are one contiguous
// range: start[] is the CSR over the bins.
struct CellGrid {
  int lo[3];
  int dim[3];
  uint32_t per_replica;
  uint32_t ns;
  const uint32_t* __restrict__ start;
};

// The replica-local cell of subpartition (gx, gy, gz); false outside the grid.
__device__ __forceinline__ bool cell_of(const CellGrid& G, int gx, int gy, int gz, uint32_t& c) {
  const int x = gx - G.lo[0], y = gy - G.lo[1], z = gz - G.lo[2];
  if (x < 0 || y < 0 || z < 0 || x >= G.dim[0] || y >= G.dim[1] || z >= G.dim[2]) return false;
  c = (uint32_t)x + (uint32_t)G.dim[0] * ((uint32_t)y + (uint32_t)G.dim[1] * (uint32_t)z);
  return true;
}

// The bin of species t in replica-local cell c.
__device__ __forceinline__ uint32_t bin_of(const CellGrid& G, uint32_t replica, uint32_t t, uint32_t c) {
  return (replica * G.ns + t) * G.per_replica + c;
}

// Snapshot of the start-of-iteration molecules, cell-sorted (FP64 positions
// or FP32 positions promoted on load).
template <typename PosT>
struct Snapshot {
  const PosT* __restrict__ x;
  const PosT* __restrict__ y;
  const PosT* __restrict__ z;
  const uint32_t* __restrict__ id;
  const uint32_t* __restrict__ tag;
  const uint32_t* __restrict__ consumed;  // nonzero: taken by an accepted reaction this iteration (its mover's id + 1)
  const double* __restrict__ toff;        // diffusion-time offsets of the snapshot's molecules
  const double* __restrict__ utime;       // unimolecular reaction times (absolute) or kTimeInvalid
  const uint32_t* __restrict__ cvol;      // counted volume indices, or null
  const double* __restrict__ birth;       // birthdays (absolute), or null
};

// A small set of subpartitions (the reference's SubpartIndicesSet), stored
// in the calling workgroup's LDS as packed index triples x | y << 10 | z << 20
// (at most 1024 subpartitions per edge): v[k * stride] is entry k of this
// lane's set (stride = the workgroup size, so that a wave's lanes touch
// consecutive words).
__device__ __forceinline__ uint32_t pack_sp(int x, int y, int z) {
  return (uint32_t)x | ((uint32_t)y << 10) | ((uint32_t)z << 20);
}
struct SubpartSet {
  static constexpr int kCap = 48;
  uint32_t* v;
  uint32_t stride;
  int n;
  bool overflow;
  __device__ __forceinline__ uint32_t at(int k) const { return v[(uint32_t)k * stride]; }
  __device__ __forceinline__ void clear() {
    n = 0;
    overflow = false;
  }
  __device__ __forceinline__ void insert(uint32_t s) {
    for (int k = 0; k < n; ++k)
      if (at(k) == s) return;
    if (n < kCap) v[(uint32_t)(n++) * stride] = s;
    else overflow = true;
  }
};

// collision_utils_subparts.inl:38-122
__device__ __forceinline__ void collect_neighbors(const PartitionParams& P, D3 pos, const int idx[3], double r,
                                                  SubpartSet& out) {
  const double part_len = P.edge;
  D3 rel = {pos.x - P.ox, pos.y - P.oy, pos.z - P.oz};
  D3 plus = {rel.x + r, rel.y + r, rel.z + r};
  D3 minus = {rel.x - r, rel.y - r, rel.z - r};
  D3 b = {(double)idx[0] * P.sp_len, (double)idx[1] * P.sp_len, (double)idx[2] * P.sp_len};
  int xd = 0, yd = 0, zd = 0;
  if (minus.x < b.x && minus.x > 0.0) { out.insert(pack_sp(idx[0] - 1, idx[1], idx[2])); xd = -1; }
  else if (plus.x > b.x + P.sp_len && plus.x < part_len) { out.insert(pack_sp(idx[0] + 1, idx[1], idx[2])); xd = 1; }
  if (minus.y < b.y && minus.y > 0.0) { out.insert(pack_sp(idx[0], idx[1] - 1, idx[2])); yd = -1; }
  else if (plus.y > b.y + P.sp_len && plus.y < part_len) { out.insert(pack_sp(idx[0], idx[1] + 1, idx[2])); yd = 1; }
  if (minus.z < b.z && minus.z > 0.0) { out.insert(pack_sp(idx[0], idx[1], idx[2] - 1)); zd = -1; }
  else if (plus.z > b.z + P.sp_len && plus.z < part_len) { out.insert(pack_sp(idx[0], idx[1], idx[2] + 1)); zd = 1; }
  if (xd && yd) out.insert(pack_sp(idx[0] + xd, idx[1] + yd, idx[2]));
  if (xd && zd) out.insert(pack_sp(idx[0] + xd, idx[1], idx[2] + zd));
  if (yd && zd) out.insert(pack_sp(idx[0], idx[1] + yd, idx[2] + zd));
  if (xd && yd && zd) out.insert(pack_sp(idx[0] + xd, idx[1] + yd, idx[2] + zd));
}

// collect_crossed_subparts with collect_for_molecules (and not for walls):
// the start subpartition, every subpartition the walk enters, and with the
// expanded list the neighbours within rxn_radius*sqrt(2) of the start, of
// each crossing point and of the end point.
__device__ __forceinline__ void collect_molecule_subparts(const PartitionParams& P, D3 pos, uint32_t start,
                                                          D3 disp, SubpartSet& out, uint32_t& err) {
  out.clear();
  int src[3];
  subpart_3d(P, pos, src[0], src[1], src[2]);  // the subpartition `start` names
  out.insert(pack_sp(src[0], src[1], src[2]));
  const double r_nb = P.rxn_radius * 1.41421356238;  // POS_SQRT2, defines.h:148
  if (P.use_expanded_list) collect_neighbors(P, pos, src, r_nb, out);
  SubpartWalk64 walk;
  walk.init(P, pos, start, disp);
  while (walk.next(P, err)) {
    out.insert(pack_sp(walk.ci[0], walk.ci[1], walk.ci[2]));
    if (P.use_expanded_list) collect_neighbors(P, walk.cp, walk.ci, r_nb, out);
  }
  if (P.use_expanded_list) {
    D3 dst = {pos.x + disp.x, pos.y + disp.y, pos.z + disp.z};
    int d[3];
    subpart_3d(P, dst, d[0], d[1], d[2]);
    collect_neighbors(P, dst, d, r_nb, out);
  }
  if (out.overflow) err |= ERR_CELLS;
}

// The box [lo, hi] partners of a segment must lie in (its bounding box
// widened by the reaction radius), compared in the snapshot's own precision:
// for FP32 positions the bounds are rounded outwards, so a partner the FP32
// compare rejects lies outside the FP64 box too.
template <typename PosT>
struct PartnerBox;
template <>
struct PartnerBox<double> {
  double x0, x1, y0, y1, z0, z1;
  __device__ __forceinline__ PartnerBox(double ax0, double ax1, double ay0, double ay1, double az0, double az1)
      : x0(ax0), x1(ax1), y0(ay0), y1(ay1), z0(az0), z1(az1) {}
  __device__ __forceinline__ bool outside(double x, double y, double z) const {
    return x < x0 || x > x1 || y < y0 || y > y1 || z < z0 || z > z1;
  }
};
template <>
struct PartnerBox<float> {
  float x0, x1, y0, y1, z0, z1;
  __device__ __forceinline__ PartnerBox(double ax0, double ax1, double ay0, double ay1, double az0, double az1)
      : x0(__double2float_rd(ax0)), x1(__double2float_ru(ax1)), y0(__double2float_rd(ay0)),
        y1(__double2float_ru(ay1)), z0(__double2float_rd(az0)), z1(__double2float_ru(az1)) {}
  __device__ __forceinline__ bool outside(float x, float y, float z) const {
    return x < x0 || x > x1 || y < y0 || y > y1 || z < z0 || z > z1;
  }
};

// collide_mol (collision_utils.inl:464-515) without the defunct/self checks
__device__ __forceinline__ bool collide_mol64(D3 pos, D3 disp, D3 other, double sigma2, double& t, D3& cpos) {
  D3 dir = {other.x - pos.x, other.y - pos.y, other.z - pos.z};
  double d = dot3(dir, disp);
  if (d < 0) return false;
  double movelen2 = dot3(disp, disp);
  if (d > movelen2) return false;
  double dirlen2 = dot3(dir, dir);
  if (movelen2 * dirlen2 - d * d > movelen2 * sigma2) return false;
  t = d / movelen2;
  cpos = {pos.x + t * disp.x, pos.y + t * disp.y, pos.z + t * disp.z};
  return true;
}

// rxn_utils.inl:301-320
__device__ __forceinline__ int binary_search_double(const double* A, double match, int max_idx, double mult) {
  int min_idx = 0;
  while (max_idx - min_idx > 1) {
    int mid = (max_idx + min_idx) / 2;
    if (match > (A[mid] * mult)) min_idx = mid;
    else max_idx = mid;
  }
  if (match > A[min_idx] * mult) return max_idx;
  return min_idx;
}

// test_bimolecular (rxn_utils.inl:336-414) with local_prob_factor == 0;
// -1 = no reaction.  One uniform draw.
__device__ __forceinline__ int test_bimolecular(const RxnRec& rx, double scaling, PhiloxStream& rng) {
  double prob;
  if (rx.max_fixed_p < scaling) {
    prob = rng.dbl() * scaling;
    if (prob >= rx.max_fixed_p) return -1;
  } else {
    const float max_p = rx.max_p_f;
    if (max_p >= scaling) {
      prob = rng.dbl() * max_p;
    } else {
      prob = rng.dbl() * scaling;
      if (prob >= max_p) return -1;
    }
  }
  return binary_search_double(rx.cum_probs, prob, (int)rx.num_pathways - 1, 1);
}

}  // namespace mcg
