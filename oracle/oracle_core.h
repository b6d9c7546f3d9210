// TEST INFRASTRUCTURE ONLY.  This is the CPU oracle: a restatement, in FP64
// and in the reference's evaluation order, of the MCell4 diffuse-react step
// (src4/diffuse_react_event.cpp and the .inl files it includes).  Only tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it, and only
// as the checker / CPU baseline -- never as part of the product path.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off).  The reference's release
// build uses -march=core2 (CMakeLists.txt:56), which has no FMA, so no
// multiply-add may be contracted here either.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace orc {

// ----------------------------------------------------------------- constants
// MCell3 values of the libbng constants used on this path
// (src/mcell_structs.h:238-240: EPS_C, SQRT_EPS_C, GIGANTIC).  MCell3's
// collide_wall uses EPS_C exactly where MCell4 uses POS_EPS
// (src/wall_util.c:849-872 vs src4/collision_utils.inl:673-705).
constexpr double EPS = 1e-12;
constexpr double SQRT_EPS = 1e-6;
constexpr double POS_EPS = 1e-12;
constexpr double STIME_EPS = 1e-12;
constexpr double DBL_GIGANTIC = 1e140;
constexpr double TIME_FOREVER = 1e20;          // src4/defines.h:145
constexpr double TIME_INVALID = -256;          // src4/defines.h:144
constexpr double MY_PI = 3.14159265358979323846;
constexpr double N_AV = 6.0221417930e23;       // src/mcell_structs_shared.h:19
constexpr double POS_SQRT2 = 1.41421356238;    // src4/defines.h:148
constexpr double PARTITION_EDGE_EXTRA_MARGIN_UM = 0.01;  // src4/defines.h:153
constexpr uint32_t MAX_SUBPARTS_PER_PARTITION = 300;     // src4/defines.h:157
constexpr double FLT_MIN_D = 1.17549435082228750797e-38; // FLT_MIN (guard_zero_div)
constexpr uint32_t INVALID_U32 = 0xFFFFFFFFu;

// ------------------------------------------------------------------- vectors
// glm evaluation order (libs/glm/detail/func_geometric.inl:48-54, :68-79):
// dot = (x*x' + y*y') + z*z', cross = (y*z'-y'*z, z*x'-z'*x, x*y'-x'*y).
struct V3 {
  double x, y, z;
};
inline V3 v3(double a) { return {a, a, a}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline double dot(V3 a, V3 b) {
  double tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z;
  return (tx + ty) + tz;
}
inline V3 cross(V3 x, V3 y) {
  return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
inline double len3_squared(V3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }  // defines.h:202
inline double maxd(double a, double b) { return (a < b) ? b : a; }
inline double mind(double a, double b) { return (b < a) ? b : a; }
inline double max3(V3 v) { return maxd(maxd(v.x, v.y), v.z); }  // glm::compMax
inline double min3(V3 v) { return mind(mind(v.x, v.y), v.z); }
inline double abs_max_2vec(V3 a, V3 b) {  // defines.h:175-180
  V3 m{maxd(std::fabs(a.x), std::fabs(b.x)), maxd(std::fabs(a.y), std::fabs(b.y)),
       maxd(std::fabs(a.z), std::fabs(b.z))};
  return max3(m);
}
// src/util.c:449-462 (MCell3 distinguishable == libbng distinguishable_p for FP64)
inline bool distinguishable(double a, double b, double eps) {
  double c = std::fabs(a - b);
  a = std::fabs(a);
  if (a < 1) a = 1;
  b = std::fabs(b);
  if (b < a) eps *= a; else eps *= b;
  return c > eps;
}
inline bool cmp_eq(double a, double b, double eps = EPS) { return std::fabs(a - b) < eps; }
inline bool cmp_lt(double a, double b, double eps) { return a < b && !cmp_eq(a, b, eps); }
inline void guard_zero_div(V3& v) {  // defines.h:342-352
  if (v.x == 0) v.x = FLT_MIN_D;
  if (v.y == 0) v.y = FLT_MIN_D;
  if (v.z == 0) v.z = FLT_MIN_D;
}

// ----------------------------------------------------------------------- RNG
// ISAAC64 (Jenkins; src/isaac64.c:60-155, src/isaac64.h:40-75) read as uint32
// halves from the top of the result block downwards.
struct Isaac64 {
  uint32_t randcnt = 0;
  uint64_t aa = 0, bb = 0, cc = 0;
  uint64_t randrsl[256];
  uint64_t mm[256];
  uint64_t rngblocks = 0;
  void init(uint32_t seed);
  void generate();
  uint32_t u32() {
    if (randcnt == 0) {
      generate();
      randcnt = 512;
    }
    --randcnt;
    uint32_t v;
    std::memcpy(&v, reinterpret_cast<const unsigned char*>(randrsl) + 4 * randcnt, 4);
    return v;
  }
};

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), written from the paper.
struct Philox {
  static void block(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]);
};

constexpr double DBL32 = 2.3283064365386962890625e-10;  // src/isaac64.h:33

// A draw source: the global ISAAC64 stream of the reference, or the
// per-(molecule, iteration) Philox stream the GPU engine uses.
struct DrawSource {
  virtual ~DrawSource() {}
  virtual uint32_t u32() = 0;
  double dbl() { return DBL32 * (double)u32(); }  // rng_dbl = isaac64_dbl32
};
struct IsaacSource : DrawSource {
  Isaac64 s;
  uint64_t draws = 0;
  uint32_t u32() override { ++draws; return s.u32(); }
};
struct PhiloxSource : DrawSource {
  uint32_t key[2];
  uint32_t ctr[4];  // {block, id, iteration, (replica << 8) | pass}
  uint32_t buf[4];
  int idx = 4;
  uint64_t draws = 0;
  void reset(uint64_t seed, uint32_t id, uint32_t iteration, uint32_t replica, uint32_t pass) {
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32);
    ctr[0] = 0; ctr[1] = id; ctr[2] = iteration; ctr[3] = (replica << 8) | (pass & 0xFFu);
    idx = 4;
    draws = 0;
  }
  uint32_t u32() override {
    if (idx == 4) {
      Philox::block(ctr, key, buf);
      ctr[0]++;
      idx = 0;
    }
    ++draws;
    return buf[idx++];
  }
};

// Marsaglia-Tsang Ziggurat exactly as src/rng.c:173-218.
double rng_gauss(DrawSource& r);

// ------------------------------------------------------------ configuration
struct UserConfig {       // the libmcell Config fields on this path
  double time_step = 1e-6;                 // s   (simulation_setup.yaml)
  double surface_grid_density = 10000;     // /um^2
  double interaction_radius = -1;          // um, <0 = default
  double partition_dimension = 10;         // um
  double subpartition_dimension = 0.5;     // um
  uint32_t seed = 1;
  int has_bimol_vol_rxns = 0;
};

struct Config {           // SimulationConfig + Partition corners
  double time_unit, length_unit, rcp_length_unit, grid_density;
  double rxn_radius_3d;
  double partition_edge_length;
  V3 origin, opposite;
  double subpart_edge_length, subpart_edge_length_rcp;
  uint32_t n, n2, n3;
  int use_expanded_list;
};

// libmcell/api/mcell4_converter.cpp:237-393 + src4/simulation_config.cpp:45-70
// + src4/partition.cpp:57.  bbox_um = {llf.xyz, urb.xyz} over all vertices.
Config derive_config(const UserConfig& u, const double bbox_um[6]);

// -------------------------------------------------------------------- walls
struct Wall {
  uint32_t v[3];
  V3 normal, unit_u, unit_v;
  double d;           // distance_to_origin
  double uv_vert1_u;
  double uv2_u, uv2_v;
  double area;
  uint32_t object_id;
  uint32_t surf_class;  // 0 = reflective (default)
};

// ------------------------------------------------------------------ species
// surface-class actions (reflect is the default of every wall)
enum : int { SC_REFLECT = 0, SC_TRANSPARENT = 1, SC_ABSORB = 2 };

enum : uint32_t {
  SPF_CAN_VOLVOL = 1u << 0,
  SPF_CAN_VOLWALL = 1u << 1,
  SPF_CAN_VOLSURF = 1u << 2,
  SPF_TARGET_ONLY = 1u << 3,
};
struct Species {
  double D;            // cm^2/s
  double space_step;   // internal length units
  double time_step;    // internal time units
  uint32_t flags;
  bool is_vol;
};

// Surface grid of a wall (Grid::initialize, src4/wall.cpp:38-74).
struct Grid {
  uint32_t n_axis = 0, num_tiles = 0;  // num_tiles 0: not initialized
  double strip_width_rcp = 0, vert2_slope = 0, fullslope = 0, binding_factor = 0;
  double vert0_u = 0, vert0_v = 0;
  std::vector<uint32_t> tile_mol;      // molecule id per tile, or INVALID
};

// A wall edge as surface molecules cross it (Edge, src4/wall.h:32-89): the
// two walls it joins (bwd INVALID_U32: not shared) and the uv transform from
// the forward wall's frame to the backward wall's (rotation, translation).
struct WallEdge {
  uint32_t fwd = 0xFFFFFFFFu, bwd = 0xFFFFFFFFu;
  double cos_t = 0, sin_t = 0, tu = 0, tv = 0;
};

// Volume + surface reaction class: the volume reactant hitting a tile
// occupied by the surface reactant; the surface product replaces the
// surface reactant on its tile and the volume reactant is destroyed.
struct VolSurfRxn {
  uint32_t vol, surf;            // species
  int geom_vol, geom_surf;       // orientations in the rule (0 none, +1, -1)
  double rate, pb_factor;
  double max_fixed_p;
  std::vector<double> cum_probs;
  // the products in the rule's order: species, orientation (+1, -1, 0 =
  // none: drawn per reaction), surface species or not
  struct Product {
    uint32_t species;
    int orient;
    bool surf;
  };
  std::vector<Product> products;
};

// A unimolecular reaction class of one species: cumulative pathway rates
// per internal time unit (k_tot = cum.back()), products per pathway.
struct UnimolRxn {
  std::vector<double> cum;
  std::vector<std::vector<uint32_t>> products;
};

struct BimolRxn {
  uint32_t a, b;              // reactant species (a = rule reactant 0)
  double rate;                // M^-1 s^-1
  double pb_factor;
  double max_fixed_p;         // cum_probs.back()
  std::vector<double> cum_probs;
  std::vector<std::vector<uint32_t>> products;  // per pathway: product species
};

// ----------------------------------------------------------------- molecules
enum : uint32_t {
  MF_DEFUNCT = 1u << 0,
  MF_MATURE = 1u << 1,
  MF_SCHEDULE_UNIMOL = 1u << 2,  // MOLECULE_FLAG_SCHEDULE_UNIMOL_RXN: lifetime not drawn yet
  MF_SURFACE = 1u << 3,          // a surface molecule (on a wall tile)
};
struct Mol {
  uint32_t id;
  uint32_t species;
  uint32_t flags;
  uint32_t wall = 0xFFFFFFFFu;  // surface molecules: wall, tile, orientation (+1 up, -1 down)
  uint32_t tile = 0;
  int orient = 0;
  double u = 0, v = 0;        // surface molecules: position in the wall's uv frame (Molecule::s.pos)
  V3 pos;
  uint32_t subpart;
  uint32_t reactant_subpart;  // subpart whose reactant sets hold this molecule
  double diffusion_time;
  double birthday;
  double unimol_rxn_time;
  uint32_t cvol = 0;  // counted volume index (Molecule::v.counted_volume_index)
};

enum class CollType : int { WALL_MISS = 0, WALL_FRONT = 1, WALL_BACK = 2, WALL_REDO = 3, VOLVOL = 4 };

struct Collision {
  CollType type;
  uint32_t mover;
  double time;
  V3 pos;
  uint32_t wall;       // wall collisions
  uint32_t partner;    // vol-vol collisions
  int32_t rxn;         // vol-vol: bimolecular rxn index
};

// One wall hit seen while diffusing a molecule (for the parity trace).
struct HitRecord {
  uint32_t mol_id;
  uint32_t iteration;
  uint32_t wall;
  uint32_t type;        // 1 front, 2 back
  uint32_t subpart_after;
  double time;
};

struct RxnRecord {
  uint32_t iteration;
  uint32_t mover_id;
  uint32_t partner_id;
  int32_t rxn;
  int32_t pathway;
  double time;
  V3 pos;
};

enum DrawMode : int { DRAW_ISAAC = 0, DRAW_PHILOX = 1 };

class World {
 public:
  Config cfg;
  std::vector<V3> verts;
  std::vector<Wall> walls;
  std::vector<uint32_t> sp_wall_off;   // CSR subpart -> walls (ascending wall index)
  std::vector<uint32_t> sp_wall_idx;
  std::vector<Species> species;
  std::vector<BimolRxn> rxns;
  std::vector<int32_t> pair_rxn;        // species x species -> rxn index or -1
  std::vector<UnimolRxn> unimol;        // per species (empty cum: none)
  std::vector<Grid> grids;              // per wall
  std::vector<VolSurfRxn> vs_rxns;
  std::vector<int32_t> pair_vs;         // vol species x surf species -> vs rxn index or -1
  uint64_t volsurf_reactions = 0;
  std::vector<uint32_t> calls_this_iteration;  // per molecule id: diffuse calls (stream pass)
  uint64_t unimol_reactions = 0;
  // surface classes: action of class c (1-based on walls) for species s hit
  // from the front (bits 0-1) and from the back (bits 2-3)
  uint32_t num_surface_classes = 0;
  std::vector<uint8_t> surf_action;     // [(c - 1) * species.size() + s]
  uint64_t transparent_crossings = 0, absorbed = 0;
  std::vector<Mol> mols;                // indexed by id
  uint64_t iteration = 0;

  DrawMode mode = DRAW_ISAAC;
  uint64_t philox_seed = 0;
  uint32_t replica = 0;
  IsaacSource isaac;
  PhiloxSource philox;

  bool trace = false;
  std::vector<HitRecord> hits;
  std::vector<RxnRecord> rxn_log;

  // statistics mirroring SimulationStats (src4/simulation_stats.h:46-155)
  uint64_t ray_voxel_tests = 0, ray_polygon_tests = 0, ray_polygon_colls = 0;
  uint64_t mol_wall_reflections = 0, vol_mol_vol_mol_collisions = 0, diffuse_3d_calls = 0;
  uint64_t redo_count = 0;
  uint64_t partition_clips = 0;  // ray segments clipped at the partition face

  // snapshot mode: each mover sees the start-of-iteration positions and no
  // molecule is consumed while the iteration runs; the first reaction a mover
  // would undergo is recorded instead of executed.
  bool snapshot = false;

  // iterations between scheduler barriers (count events); the time a product
  // may diffuse past its creation is clipped at the next barrier.  A non-empty
  // barrier list (ascending iterations) replaces the periodic rule.
  uint32_t barrier_interval = 1;
  std::vector<uint64_t> barrier_list;

  explicit World(const Config& c);
  void set_rng_isaac(uint32_t seed) { mode = DRAW_ISAAC; isaac.s.init(seed); }
  void set_rng_philox(uint64_t seed, uint32_t replica_) { mode = DRAW_PHILOX; philox_seed = seed; replica = replica_; }

  uint32_t add_vertex_um(V3 p_um);
  void add_wall(uint32_t v0, uint32_t v1, uint32_t v2, uint32_t object_id, uint32_t surf_class);
  void finalize_walls();
  // custom_time_step: seconds (0: the global time step)
  uint32_t add_species(double D_cm2s, bool is_vol, uint32_t extra_flags, double custom_time_step = 0.0);
  int add_bimol_rxn(uint32_t a, uint32_t b, double rate, const std::vector<uint32_t>& products);
  int add_volsurf_rxn(uint32_t vol, uint32_t surf, int geom_vol, int geom_surf, double rate,
                      const std::vector<VolSurfRxn::Product>& products);
  // uv: the molecule's position on the wall (null: the tile's centre,
  // grid2uv); surface molecules of species that diffuse are scheduled at
  // `time` (null time: the current iteration)
  uint32_t add_surface_molecule(uint32_t species, uint32_t wall, uint32_t tile, int orient,
                                const double* uv = nullptr, const double* time = nullptr);
  void init_grid(uint32_t wall);
  uint32_t xyz2grid(V3 v, uint32_t wall) const;
  // surface-molecule diffusion (oracle_surface.cpp)
  std::vector<std::array<WallEdge, 3>> edges;  // per wall, built by build_surface_net
  uint64_t surf_moves_between_walls = 0, surf_retries = 0;
  void build_surface_net();
  uint32_t uv2grid(double u, double v, uint32_t wall) const;
  void grid2uv(uint32_t wall, uint32_t tile, double& u, double& v) const;
  // find_edge_point / traverse_surface (public for unit tests)
  int find_edge_point(uint32_t wall, double lu, double lv, double du, double dv, double& eu, double& ev) const;
  uint32_t traverse_surface(uint32_t wall, double lu, double lv, int edge, double& nu, double& nv) const;
  int collide_and_react_with_surf_mol(const Collision& c, double t_steps, double r_rate_factor, double elapsed);
  // rate in s^-1 per pathway (pb = time_unit, MCell3 src/react_util.c:74-78)
  void add_unimol_rxn(uint32_t species, double rate, const std::vector<uint32_t>& products);
  uint32_t add_volume_molecule(uint32_t species, V3 pos, double time);

  void run_iteration();

  // --- partition index math (src4/partition.h:248-309)
  bool in_partition(V3 p) const {
    return p.x >= cfg.origin.x && p.y >= cfg.origin.y && p.z >= cfg.origin.z &&
           p.x < cfg.opposite.x && p.y < cfg.opposite.y && p.z < cfg.opposite.z;
  }
  void subpart_3d(V3 p, int idx[3]) const {
    V3 rel = p - cfg.origin;
    idx[0] = (int)(rel.x * cfg.subpart_edge_length_rcp);
    idx[1] = (int)(rel.y * cfg.subpart_edge_length_rcp);
    idx[2] = (int)(rel.z * cfg.subpart_edge_length_rcp);
  }
  uint32_t subpart_from_3d(int x, int y, int z) const {
    return (uint32_t)x + (uint32_t)y * cfg.n + (uint32_t)z * cfg.n2;
  }
  uint32_t subpart_index(V3 p) const {
    int i[3];
    subpart_3d(p, i);
    return subpart_from_3d(i[0], i[1], i[2]);
  }
  void subpart_3d_from_index(uint32_t s, int idx[3]) const {
    idx[0] = (int)(s % cfg.n);
    idx[1] = (int)((s / cfg.n) % cfg.n);
    idx[2] = (int)((s / cfg.n2) % cfg.n);
  }
  bool in_range(int i) const { return i >= 0 && i < (int)cfg.n; }

  DrawSource& rng() { return mode == DRAW_ISAAC ? (DrawSource&)isaac : (DrawSource&)philox; }

  // --- the ray-trace building blocks (public for unit tests)
  uint32_t collect_crossed_subparts(const Mol& vm, V3 displacement, bool for_mols, bool for_walls,
                                    std::vector<uint32_t>& walls_sp, std::vector<uint32_t>& mols_sp) const;
  CollType collide_wall(V3 pos, uint32_t wall_index, bool update_move, V3& move, double& t, V3& hit);
  bool get_closest_wall_collision(const Mol& vm, uint32_t subpart, uint32_t last_hit, V3& displacement,
                                  V3& disp_up_to_wall, Collision& closest);
  V3 displacement_up_to_partition_boundary(V3 pos, V3 disp) const;

 private:
  // reactant sets per (mover species, subpart): ascending ids of molecules that
  // the mover species can react with (Partition::get_volume_molecule_reactants)
  std::vector<std::vector<std::vector<uint32_t>>> reactants;  // [species][subpart] (lazily sized)
  std::vector<uint32_t> schedulable;
  // where_created_this_iteration: a volume product of a volume-surface
  // reaction does not react with its creation tile on its first hit there
  struct DiffuseAction {
    uint32_t id;
    uint32_t born_wall = INVALID_U32, born_tile = INVALID_U32;
  };
  uint32_t born_wall = INVALID_U32, born_tile = INVALID_U32;  // of the molecule diffusing now

 public:
  // counted volumes: per wall the index inside / outside its object
  // (INVALID_U32: not a counted object); empty = none
  std::vector<uint32_t> cv_in, cv_out;

 private:
  std::vector<DiffuseAction> new_diffuse_actions;
  std::vector<V3> snapshot_pos;
  double event_time = 0;
  double time_up_to_next_barrier = 1;

  void reactants_insert(uint32_t mover_species, uint32_t subpart, uint32_t id);
  void reactants_erase(uint32_t mover_species, uint32_t subpart, uint32_t id);
  void register_reactant(const Mol& m, bool add);
  std::vector<uint32_t> partner_species_of(uint32_t s) const;
  void start_stream(const Mol& m, uint32_t pass);

  void diffuse_single_molecule(uint32_t id);
  void pick_unimol_time(Mol& m, double current_time);
  bool react_unimol(uint32_t id);
  double get_max_time(Mol& m);
  void diffuse_vol_molecule(uint32_t id, double& max_time, double diffusion_start_time);
  void diffuse_surf_molecule(uint32_t id, double& max_time);
  uint32_t ray_trace_surf(uint32_t id, double du, double dv, double& nu, double& nv);
  int ray_trace_vol(uint32_t id, bool can_vol_react, uint32_t last_hit, V3& remaining_disp,
                    std::vector<Collision>& collisions);
  V3 volume_product_position(uint32_t id, uint32_t wall, int orient);
  void jump_away_line(V3 p, double k, V3 A, V3 B, V3 n, V3& v);
  bool collide_mol(const Mol& vm, V3 displacement, uint32_t partner, double& t, V3& pos) const;
  void reflect_from_wall(const Collision& c, Mol& vm, V3& disp, double& t_steps, uint32_t& last_hit);
  int wall_action(const Wall& w, uint32_t species, CollType type) const;
  void cross_transparent_wall(const Collision& c, Mol& vm, V3& disp, double& t_steps, double& elapsed,
                              uint32_t& last_hit);
  bool collide_and_react_with_vol_mol(const Collision& c, V3& disp, double t_steps, double r_rate_factor,
                                      double elapsed);
  int test_bimolecular(const BimolRxn& rx, double scaling);
  int test_bimolecular_p(double max_fixed_p, const std::vector<double>& cum_probs, double scaling);
  void outcome_bimolecular(const Collision& c, int pathway, double time);
 public:
  double exact_disk(V3 loc, V3 mv, double R, const Mol& moving, const Mol& target) const;
};

// test_bimolecular (rxn_utils.inl:336-414) with its draw u = rng_dbl() given.
int test_bimolecular_u(double max_fixed_p, const std::vector<double>& cum_probs, double scaling, double u);

}  // namespace orc
