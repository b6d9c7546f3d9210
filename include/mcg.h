/*
 * mcg -- MI355X diffuse-react engine for MCell4: the C ABI.
 *
 * These entry points are what a GPU replacement of MCell4's DiffuseReactEvent
 * (src4/diffuse_react_event.h:108-339) calls.  World, Partition, the species,
 * reaction and geometry definitions and the libmcell Model API stay on the
 * host; the event hands their flattened data to the engine once, keeps the
 * molecules resident in HBM, and calls mcg_step once per iteration (or for a
 * run of iterations between two scheduler barriers).
 *
 * All functions return MCG_OK (0) or an error status; mcg_last_error() gives
 * the text of the last error on the calling thread.  Lengths are in the
 * reference's internal units (length_unit = 1/sqrt(surface_grid_density) um,
 * time_unit = time_step s), exactly as Partition and SimulationConfig hold them.
 *
 * Threading: a context is used by one host thread at a time.  The engine runs
 * its kernels on its own HIP stream (default priority).
 */
#ifndef MCG_H_
#define MCG_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCG_ABI_VERSION 9

typedef struct mcg_ctx mcg_ctx;

enum {
  MCG_OK = 0,
  MCG_E_INVALID = 1,     /* bad argument */
  MCG_E_HIP = 2,         /* HIP runtime error */
  MCG_E_CAPACITY = 3,    /* molecule capacity exceeded */
  MCG_E_ESCAPED = 4,     /* a molecule left the partition (reference: fatal error) */
  MCG_E_UNSUPPORTED = 5, /* feature not on this path */
  MCG_E_NOMEM = 6,       /* device allocation failed */
  MCG_E_LIMIT = 7,       /* a per-step safety limit was hit (reflections, re-draws, walk) */
  MCG_E_STATE = 8        /* call out of order (e.g. step before geometry), or the context failed */
};

/* After an mcg_step that returned an error other than MCG_E_INVALID or
 * MCG_E_STATE the molecule state is undefined (the reference terminates the
 * process at this point); every later call on the context except
 * mcg_destroy returns MCG_E_STATE. */

enum {
  MCG_PREC_FP64 = 0,          /* plain FP64: positions, ray-triangle test, reflection */
  MCG_PREC_FP32_FALLBACK = 1  /* FP32 positions and ray test, FP64 near edges / planes */
};

/* SimulationConfig + Partition geometry (src4/simulation_config.h:28-135,
 * src4/partition.h:1112-1204), internal units. */
typedef struct mcg_config {
  double time_unit;              /* SimulationConfig::time_unit [s] */
  double length_unit;            /* SimulationConfig::length_unit [um] */
  double origin[3];              /* Partition::origin_corner */
  double partition_edge_length;  /* SimulationConfig::partition_edge_length */
  double subpart_edge_length;    /* SimulationConfig::subpart_edge_length */
  double subpart_edge_length_rcp;
  uint32_t num_subparts_per_edge;
  int32_t use_expanded_list;     /* SimulationConfig::use_expanded_list */
  double rxn_radius_3d;          /* SimulationConfig::rxn_radius_3d */
  int32_t precision;             /* MCG_PREC_* */
  uint32_t num_replicas;         /* independent instances sharing the mesh (>= 1) */
  uint64_t seed;                 /* key of the per-molecule counter-based streams */
  uint64_t molecule_capacity;    /* volume molecules over all replicas (incl. products) */
  uint32_t flags;                /* MCG_FLAG_* */
  uint32_t start_iteration;      /* iteration of the first mcg_step (World's current iteration when
                                    the event is created; 0 for a new simulation) */
} mcg_config;

#define MCG_FLAG_TRACE_HITS 1u    /* record the wall indices each molecule hits (parity tests) */
#define MCG_FLAG_PROPOSE_ONLY 2u  /* parity mode: every molecule moves against the iteration-start
                                     snapshot and successful reaction tests are logged, not carried
                                     out (a reacting molecule stops where its last ray segment began) */
/* The full step's drain (the movers the light pass queues) runs exact_disk
 * either in place, with a small vertex pool, or defers every step with a
 * disk that meets walls to a pass of its own: the results are the same, the
 * speed depends on how often disks meet walls.  By default the engine times
 * both forms on the first iterations after the geometry or the molecules are
 * set and keeps the faster; these flags fix one. */
#define MCG_FLAG_DRAIN_INLINE 4u
#define MCG_FLAG_DRAIN_DEFER 8u

/* BNG::Species fields on this path (space_step/time_step from
 * Species::update_space_and_time_step, MCell3 src/mcell_species.c:237-300).
 * time_step is 1 for the global step; a custom step lies in [1/8, 1e12]
 * (MCG_E_INVALID otherwise: a molecule diffuses at most eight times in one
 * iteration).  Steps below one iteration diffuse the molecule again within
 * the iteration (diffuse_single_molecule, diffuse_react_event.cpp:294-319);
 * steps above it start short and lengthen with the molecule's age until it
 * is mature (get_max_time, :164-198).  Contexts with such species take the
 * reactive pipeline. */
typedef struct mcg_species {
  double space_step;   /* internal length units */
  double time_step;    /* internal time units */
  uint32_t flags;      /* MCG_SPF_* */
  uint32_t reserved;
} mcg_species;

#define MCG_SPF_CAN_VOLVOL 1u
#define MCG_SPF_CAN_VOLWALL 2u
#define MCG_SPF_CAN_VOLSURF 4u
#define MCG_SPF_TARGET_ONLY 8u
#define MCG_SPF_SURFACE 16u  /* a surface species; with space_step > 0 its molecules diffuse on the mesh */

/* A volume+volume reaction class (BNG::RxnClass for two volume reactants
 * as test_bimolecular and outcome_bimolecular use it, rxn_utils.inl:336-414,
 * diffuse_react_event.cpp:1833-1895).  Probabilities are per collision,
 * cumulative over pathways (RxnClass::cum_probs; the last entry is
 * max_fixed_p).  A product of a reactant's species keeps that reactant
 * (outcome_products_random, diffuse_react_event.cpp:2551-2565,2686-2727);
 * every other product is a new volume molecule created at the collision
 * point, and the reactants not kept are destroyed. */
typedef struct mcg_bimol_rxn {
  uint32_t reactant_a, reactant_b; /* species ids */
  uint32_t num_pathways;           /* 1..4 */
  uint32_t reserved;
  double cum_probs[4];
  uint32_t num_products[4];        /* per pathway, 0..3 */
  uint32_t products[4][3];         /* product species per pathway */
} mcg_bimol_rxn;

/* A unimolecular reaction class of one volume species (BNG::RxnClass with
 * one reactant as pick_unimol_rxn_classes, time_of_unimol and
 * which_unimolecular use it, rxn_utils.inl:696-783).  Rates are cumulative
 * over pathways, per internal time unit (rate [1/s] * time_unit; the last
 * entry is k_tot).  A molecule draws its lifetime -ln(u)/k_tot when it is
 * created; at that time it is replaced by the pathway's products, which are
 * new volume molecules at its position (outcome_unimolecular,
 * diffuse_react_event.cpp:2939-3003). */
typedef struct mcg_unimol_rxn {
  uint32_t reactant;
  uint32_t num_pathways;           /* 1..4 */
  double cum_rates[4];
  uint32_t num_products[4];        /* per pathway, 0..3 */
  uint32_t products[4][3];
} mcg_unimol_rxn;

/* One reaction that happened (or, with MCG_FLAG_PROPOSE_ONLY, was proposed). */
typedef struct mcg_rxn_event {
  double time;                /* absolute time in iterations */
  double pos[3];              /* collision point */
  uint32_t iteration, replica;
  uint32_t mover_id, partner_id;
  uint32_t rxn, pathway;      /* index into the mcg_set_bimol_rxns table; for a unimolecular
                                 reaction 0x80000000 | reactant species (partner_id 0xFFFFFFFF);
                                 for a volume-surface one 0x40000000 | mcg_set_volsurf_rxns index */
  uint32_t first_product_id;  /* products get ids first_product_id + k */
  uint32_t num_products;
} mcg_rxn_event;

/* Per-step counters mirroring SimulationStats (src4/simulation_stats.h). */
typedef struct mcg_step_stats {
  uint64_t molecule_steps;       /* volume molecules diffused, summed over iterations */
  uint64_t wall_hits;            /* reflections off walls (SimulationStats::mol_wall_reflections) */
  uint64_t reactions;
  double kernel_ms;              /* device time of the diffuse kernels (HIP events) */
  double wall_ms;                /* host wall time of the call */
  uint32_t kernel_launches;
  uint32_t slow_path_last;       /* molecules that needed the full ray trace in the last iteration */
  uint64_t transparent_crossings;/* walls crossed through a transparent surface class */
  uint32_t settle_rounds;        /* claim-settling rounds of the reactive path, summed over iterations */
  uint32_t drain_form;           /* drain form of the last iteration: 1 in place, 2 deferred (MCG_FLAG_DRAIN_*) */
} mcg_step_stats;

const char* mcg_last_error(void);
int mcg_abi_version(void);

/* Replaces `new DiffuseReactEvent(world)` at src4/world.cpp:280-282: creates
 * an engine on the current HIP device for the partition described by cfg. */
int mcg_create(const mcg_config* cfg, mcg_ctx** out);
int mcg_destroy(mcg_ctx* ctx);

/* Partition geometry: vertices (Partition::get_geometry_vertex, internal
 * units, xyz interleaved) and walls (vertex index triples, the wall index is
 * the position in the array).  The engine recomputes the wall constants
 * (Wall::initialize_wall_constants, src4/wall.cpp:281-342) and the per
 * subpartition wall lists (Partition::finalize_walls, src4/partition.cpp:91-118).
 * wall_class may be NULL (all walls reflective); otherwise per wall 0 (no
 * surface class) or a 1-based class of mcg_set_surface_classes. */
int mcg_set_geometry(mcg_ctx* ctx, const double* vertices, uint32_t num_vertices, const uint32_t* walls,
                     uint32_t num_walls, const uint32_t* wall_class);

/* Surface classes (the reference's surface-class species with reflective,
 * transparent and absorptive properties, mcell4_converter.cpp:552-560, acting
 * through collide_and_react_with_walls, diffuse_react_event.cpp:991-1067).
 * Walls carry a class through mcg_set_geometry's wall_class (0 = none, else
 * 1..num_classes).  action[(c - 1) * num_species + s] says what class c does
 * to species s hit from the wall's front (bits 0-1: the move runs against the
 * normal) and from its back (bits 2-3): MCG_WALL_REFLECT, MCG_WALL_TRANSPARENT
 * (cross_transparent_wall, :3007-3099) or MCG_WALL_ABSORB (destroyed after
 * test_intersect's two draws).  Species with any other action than reflect get
 * MCG_SPF_CAN_VOLWALL.  Call after mcg_set_geometry and mcg_set_species. */
#define MCG_WALL_REFLECT 0u
#define MCG_WALL_TRANSPARENT 1u
#define MCG_WALL_ABSORB 2u
int mcg_set_surface_classes(mcg_ctx* ctx, uint32_t num_classes, const uint8_t* action);

/* Species table, indexed by species id (< 256). */
int mcg_set_species(mcg_ctx* ctx, const mcg_species* species, uint32_t num_species);

/* Volume-volume reaction classes.  Reactant species get MCG_SPF_CAN_VOLVOL.
 * Molecules of reactive species must stay within the box of subpartitions
 * spanned by the geometry and the molecules added (one subpartition of
 * margin); leaving it is reported as MCG_E_LIMIT. */
int mcg_set_bimol_rxns(mcg_ctx* ctx, const mcg_bimol_rxn* rxns, uint32_t num_rxns);

/* Per-replica parameters for batches of instances that differ in diffusion
 * constants or rates (parameter sweeps over one mesh): replica `replica`
 * moves with these species parameters (num_species entries; the flags stay
 * the base table's) and reacts with these classes (the base table's
 * reactants, pathways and products; only cum_probs may differ).  rxns may be
 * NULL when no reaction classes are set.  Call after mcg_set_species and
 * mcg_set_bimol_rxns; replicas never set keep the base tables.  No
 * counterpart in MCell4, where each instance is its own process. */
int mcg_set_replica_params(mcg_ctx* ctx, uint32_t replica, const mcg_species* species, const mcg_bimol_rxn* rxns);

/* Unimolecular reaction classes (at most one per species).  Reaction classes
 * of every kind are set before the first mcg_step (MCG_E_STATE after it). */
int mcg_set_unimol_rxns(mcg_ctx* ctx, const mcg_unimol_rxn* rxns, uint32_t num_rxns);

/* A volume + surface reaction class, V + S -> products in that reactant
 * order: the volume reactant hits a wall tile holding the surface reactant
 * (collide_and_react_with_surf_mol, diffuse_react_event.cpp:845-975).
 * cum_probs are cumulative reaction probabilities per hit (rate x pb_factor,
 * MCell3 src/react_util.c:132-157) before the tile binding factor; geom_vol
 * and geom_surf are the orientations in the rule (0 none, +1, -1), matched
 * as trigger_bimolecular does (rxn_utils.inl:58-84).  Volume reactants get
 * MCG_SPF_CAN_VOLSURF.
 *
 * Pathway k has num_products[k] (0..3) products in the rule's order:
 * species, orientation (+1, -1, or 0: none, drawn per reaction, rng_uint & 1)
 * and bit q of surf_products[k] set for a surface species
 * (outcome_products_random, diffuse_react_event.cpp:2446-2933).  A given
 * orientation is flipped when the surface reactant's orientation differs
 * from geom_surf (:2636-2658).  The first volume product of the volume
 * reactant's species keeps it: it goes on with its step, through the wall
 * when that product's orientation differs from geom_vol (RX_FLIP, :940-971)
 * and on its side otherwise; else the volume reactant is consumed.  Within
 * its step a kept molecule does not react again with the tile it emptied.
 * A surface product of S's species keeps S (its
 * orientation becomes the product's); at most one other surface product,
 * which takes S's tile (no neighbour-tile search on this path).  A volume
 * product starts at the hit point moved off the wall to the side its
 * orientation names (update_vol_mol_after_rxn_with_surf_mol, :2291-2320,
 * tiny_diffuse_3D) and does not react with that tile on its first hit. */
typedef struct mcg_volsurf_rxn {
  uint32_t vol_species, surf_species;
  int32_t geom_vol, geom_surf;
  uint32_t num_pathways;           /* 1..4 */
  uint32_t reserved;
  double cum_probs[4];
  uint32_t num_products[4];
  uint32_t products[4][3];
  int32_t product_orient[4][3];
  uint32_t surf_products[4];
} mcg_volsurf_rxn;

int mcg_set_volsurf_rxns(mcg_ctx* ctx, const mcg_volsurf_rxn* rxns, uint32_t num_rxns);

/* Surface molecules of one replica (Molecule's surface part: wall, tile,
 * orientation and position in the wall's uv frame, molecule.h:52-167).  A
 * wall gets its tile grid (Grid::initialize, wall.cpp:38-74: ceil(sqrt(area))^2
 * tiles) with its first molecule; a tile holds at most one molecule.  uv (2
 * doubles per molecule) may be NULL: the tiles' centres (grid2uv,
 * grid_utils.inl:233-250); diffusion_times may be NULL: the clock.  Species
 * flagged MCG_SPF_SURFACE with a space step diffuse on the mesh after each
 * iteration's volume moves (diffuse_surf_molecule, diffuse_react_event.cpp:
 * 1071-1246): every wall then gets a grid, and the walls are linked across
 * the edges two of them share (surface_net, geometry.cpp:256-353; an edge of
 * more than two walls is refused).  Their species' time steps follow the
 * volume molecules' rules: a step below one iteration is taken again in the
 * same iteration, a long step is shortened for a young molecule until it is
 * mature (get_max_time, diffuse_react_event.cpp:164-198); flags may be NULL
 * (none) or carry MCG_MOLF_MATURE only, birthdays (absolute, not after the
 * diffusion time) may be NULL or TIME_INVALID (-256): the clock.  Volume
 * molecules react with surface molecules through mcg_set_volsurf_rxns.  ids
 * share the replica's id space with the volume molecules.  Call after
 * mcg_set_geometry and mcg_set_species. */
int mcg_add_surface_molecules(mcg_ctx* ctx, uint32_t replica, const uint32_t* ids, const uint32_t* species,
                              const uint32_t* walls, const uint32_t* tiles, const int32_t* orients, const double* uv,
                              const double* diffusion_times, const uint32_t* flags, const double* birthdays,
                              uint64_t count);

/* The surface molecules back, in (replica, wall, tile) order, with their uv
 * positions, diffusion times, flags and birthdays (the birthdays of products
 * made while no surface species diffuses are not kept).  Any output pointer
 * may be NULL; *count gets the number. */
int mcg_get_surface_molecules(mcg_ctx* ctx, uint32_t* ids, uint32_t* species, uint32_t* replica, uint32_t* walls,
                              uint32_t* tiles, int32_t* orients, double* uv, double* diffusion_times,
                              uint32_t* flags, double* birthdays, uint64_t capacity, uint64_t* count);

/* Keeps up to `capacity` reaction events of each mcg_step call (0 = off). */
int mcg_set_rxn_log(mcg_ctx* ctx, uint64_t capacity);

/* Events of the last mcg_step call in (iteration, replica, mover id) order
 * within each settling phase; *count gets the number recorded (which may
 * exceed what the log kept). */
int mcg_get_rxn_log(mcg_ctx* ctx, mcg_rxn_event* events, uint64_t capacity, uint64_t* count);

/* Live molecules (volume and surface) per (replica, species):
 * counts[r * num_species + s] (MolOrRxnCountEvent's molecule counts over the
 * whole partition). */
int mcg_count_species(mcg_ctx* ctx, uint64_t* counts);

/* mcg_count_species without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The counters live in a device array of the library, at most
 * 65,536 of them (num_replicas * num_species; larger contexts get
 * MCG_E_LIMIT), 512 KiB that each GPU the library runs on holds until the
 * process ends.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (from several host threads or contexts) race.  Surface molecules
 * added since the last step are uploaded by this call only when their tile
 * buffers already exist (MCG_E_STATE otherwise: call mcg_step or
 * mcg_count_species once first). */
int mcg_count_species_noalloc(mcg_ctx* ctx, uint64_t* counts);

/* Volume molecules of one replica (Partition::add_volume_molecule semantics:
 * each position must lie in the partition).  ids are the molecule ids that
 * key their random streams; the replica's next product id becomes at least
 * max(id) + 1 (partition.h:467-469).
 *
 * The time fields of Molecule (molecule.h:130-141) are optional (NULL gives
 * the default): diffusion_times are absolute, in iterations, and lie in
 * [clock, clock + 1) where clock is the iteration the next mcg_step starts
 * at (default clock); unimol_times are absolute (TIME_INVALID = -256: no
 * lifetime drawn yet or no unimolecular class; default -256); flags are
 * MCG_MOLF_* (default MCG_MOLF_SCHEDULE_UNIMOL, a newly released molecule).
 * A diffusion time past clock + 1 belongs to a species whose time step
 * exceeds one iteration.  Molecules given a diffusion time other than the
 * clock make a context without reaction classes take the time-aware
 * reactive pipeline.  counted_volumes (Molecule::v.counted_volume_index,
 * molecule.h:76) may be NULL (0, outside every counted object); with
 * mcg_set_counted_volumes they follow the molecules through transparent
 * walls and into their products.  birthdays (Molecule::birthday, absolute,
 * not after the diffusion time) matter only for species whose time step
 * exceeds one iteration; NULL or TIME_INVALID entries give the clock (a
 * molecule released in this iteration, molecule.h:139-141). */
#define MCG_MOLF_SCHEDULE_UNIMOL 1u  /* MOLECULE_FLAG_SCHEDULE_UNIMOL_RXN: lifetime not drawn yet */
#define MCG_MOLF_MATURE 2u           /* MOLECULE_FLAG_MATURE: moves with its species' full time step */
int mcg_add_volume_molecules(mcg_ctx* ctx, uint32_t replica, const uint32_t* ids, const uint32_t* species,
                             const double* positions, const double* diffusion_times, const double* unimol_times,
                             const uint32_t* flags, const uint32_t* counted_volumes, const double* birthdays,
                             uint64_t count);

/* Counted volumes (GeometryObject::counted_volume_index_inside / _outside,
 * geometry.h:102-107): for each wall (mcg_set_geometry order) the index of
 * the counted volume inside and outside its object, or 0xFFFFFFFF for both
 * when the object is not a counted volume; indices < num_volumes.  A
 * molecule crossing a transparent wall of a counted object takes the index
 * on the other side (update_counted_volume_id_when_crossing_wall,
 * collision_utils.inl:1637-1694; objects whose volumes intersect are not on
 * this path); products take the volume the reaction happened in, a volume
 * product of a volume-surface reaction on a counted wall the side its
 * orientation names.  Set after the geometry and before the first step; a
 * context with counted volumes runs the reactive pipeline. */
int mcg_set_counted_volumes(mcg_ctx* ctx, uint32_t num_volumes, const uint32_t* wall_inside,
                            const uint32_t* wall_outside);

/* Live volume molecules per (replica, counted volume, species):
 * counts[(replica * num_volumes + volume) * num_species + species]. */
int mcg_count_species_by_volume(mcg_ctx* ctx, uint64_t* counts);

/* Reactions carried out since the first step (SimulationStats::
 * inc_rxn_occurred and Partition::inc_rxn_in_volume_occured_count,
 * diffuse_react_event.cpp:2509-2520), per replica, counted volume the mover
 * was in (one volume without counted volumes) and pathway slot: bimolecular
 * class i pathway k at i * 4 + k, then the unimolecular class of species s
 * at (num_bimol + s) * 4 + k, then volume-surface class i at
 * (num_bimol + num_species + i) * 4 + k.  dims (may be NULL) gets
 * {replicas, volumes, slots}; counts (may be NULL) gets
 * counts[(replica * volumes + volume) * slots + slot]. */
int mcg_get_rxn_counts(mcg_ctx* ctx, uint64_t* counts, uint32_t* dims);

/* Partition::next_molecule_id per replica (partition.h:420-470): the id the
 * replica's next product gets.  Get fills num_replicas values; set raises
 * each replica's counter to at least the given value (it never lowers it). */
int mcg_get_next_molecule_ids(mcg_ctx* ctx, uint32_t* next_ids);
int mcg_set_next_molecule_ids(mcg_ctx* ctx, const uint32_t* next_ids);

/* Runs num_iterations diffuse-react iterations starting at first_iteration
 * (DiffuseReactEvent::step with event_time = iteration, diffuse_react_event.cpp:
 * 53-64).  time_up_to_next_barrier is what Scheduler::handle_next_event hands
 * the event through set_barrier_time_for_next_execution before the call's
 * first iteration (scheduler.cpp:303-307): a whole number >= 1 of time units
 * from first_iteration to the next barrier.  Iteration first_iteration + k
 * clips every molecule's step at min(time_up_to_next_barrier - k, 100)
 * (DIFFUSION_TIME_UPPER_LIMIT, diffuse_react_event.h:37), so a call must not
 * run past its barrier: num_iterations <= time_up_to_next_barrier.
 * first_iteration must be the engine's clock (cfg.start_iteration, then the
 * iteration after the last call's).  stats may be NULL.
 *
 * Without reaction classes each molecule's step is the reference's step with
 * the molecule's own random stream.  With reaction classes, the molecules of
 * an iteration move at once against the positions at the iteration's start;
 * a molecule whose reaction test succeeds claims itself and its partner, the
 * proposal with the smaller mover id wins a contested molecule, and losers
 * repeat their step without the consumed molecules.  Products move for the
 * rest of the step afterwards, as the reference's new_diffuse_actions do. */
int mcg_step(mcg_ctx* ctx, uint32_t first_iteration, uint32_t num_iterations, double time_up_to_next_barrier,
             mcg_step_stats* stats);

/* Waits for all queued work. */
int mcg_sync(mcg_ctx* ctx);

/* Live volume molecules over all replicas. */
int mcg_num_molecules(mcg_ctx* ctx, uint64_t* count);

/* Copies the live volume molecules back (for Partition at a barrier).  Any
 * output pointer may be NULL.  Molecules come back in engine order (which
 * changes every iteration when reactions are on), with the time fields
 * mcg_add_volume_molecules takes: an upload of what this returns into a new
 * context (start_iteration = this clock, next molecule ids copied) continues
 * the run bit for bit.  Birthdays are TIME_INVALID where the context keeps
 * none (no species' time step exceeds one iteration). */
int mcg_get_volume_molecules(mcg_ctx* ctx, uint32_t* ids, uint32_t* species, uint32_t* replica, double* positions,
                             uint32_t* subparts, double* diffusion_times, double* unimol_times, uint32_t* flags,
                             uint32_t* counted_volumes, double* birthdays, uint64_t capacity, uint64_t* count);

/* Parity/inspection: wall constants (16 doubles per wall: normal, unit_u,
 * unit_v, distance_to_origin, uv_vert1_u, uv_vert2.u, uv_vert2.v, area, 0, 0)
 * and the per-subpartition wall lists as CSR (offsets has n^3 + 1 entries).
 * Any output pointer may be NULL; *num_entries gets the CSR length. */
int mcg_get_walls(mcg_ctx* ctx, double* wall_constants, uint32_t* offsets, uint32_t* indices, uint64_t* num_entries);

/* Parity: wall hits recorded in the last iteration when MCG_FLAG_TRACE_HITS
 * is set: for molecule k (engine order), counts[k] hits, walls[k*8 .. k*8+7]. */
int mcg_get_hit_trace(mcg_ctx* ctx, uint32_t* walls, uint32_t* counts, uint64_t capacity);

/* Average device time of the dominant (diffuse) kernel over the last
 * mcg_step call, measured with HIP events on the engine's stream. */
int mcg_last_kernel_time(mcg_ctx* ctx, double* avg_ms, uint32_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* MCG_H_ */
