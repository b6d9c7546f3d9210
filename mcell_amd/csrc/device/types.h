// Device-side data layout of the diffuse-react engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcg {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;

// Per-wall constants in FP64 (Wall::initialize_wall_constants,
// src4/wall.cpp:281-342) plus the three vertices, 24 doubles = 192 B.
struct WallRec64 {
  double nx, ny, nz, d;        // normal, distance_to_origin
  double ux, uy, uz, uv1u;     // unit_u, uv_vert1_u
  double vx, vy, vz, uv2u;     // unit_v, uv_vert2.u
  double p0x, p0y, p0z, uv2v;  // vertex 0, uv_vert2.v
  double p1x, p1y, p1z, area;  // vertex 1
  double p2x, p2y, p2z, pad;   // vertex 2
};

// The same in FP32 for the FP32 ray-triangle test, 16 floats = 64 B (one
// cache line).  Vertex 0 is stored relative to nothing: positions stay in
// partition units, which are bounded by the partition edge (<= 3e4 units).
struct WallRec32 {
  float nx, ny, nz, d;
  float ux, uy, uz, uv1u;
  float vx, vy, vz, uv2u;
  float p0x, p0y, p0z, uv2v;
};
static_assert(sizeof(WallRec32) == 64, "one cache line per FP32 wall");

// SimulationConfig / Partition geometry in internal units.
struct PartitionParams {
  double ox, oy, oz;     // origin_corner
  double qx, qy, qz;     // opposite_corner
  double edge;           // partition_edge_length
  double sp_len;         // subpart_edge_length
  double sp_rcp;         // subpart_edge_length_rcp
  double rxn_radius;     // rxn_radius_3d
  uint32_t n;            // num_subparts_per_partition_edge
  uint32_t n2;
  uint32_t use_expanded_list;
  uint32_t pad;
};

struct SpeciesRec {
  double space_step;
  double time_step;
  uint32_t flags;
  uint32_t pad;
};

enum : uint32_t {
  SPF_CAN_VOLVOL = 1u << 0,
  SPF_CAN_VOLWALL = 1u << 1,
  SPF_CAN_VOLSURF = 1u << 2,
  SPF_TARGET_ONLY = 1u << 3,
  SPF_SURFACE = 1u << 4,  // a surface species: its molecules sit on wall tiles
};

// molecule tag word: species (8 bits) | flags (8 bits) | replica (16 bits)
__device__ __host__ __forceinline__ uint32_t tag_species(uint32_t t) { return t & 0xFFu; }
__device__ __host__ __forceinline__ uint32_t tag_flags(uint32_t t) { return (t >> 8) & 0xFFu; }
__device__ __host__ __forceinline__ uint32_t tag_replica(uint32_t t) { return t >> 16; }
__device__ __host__ __forceinline__ uint32_t make_tag(uint32_t s, uint32_t f, uint32_t r) {
  return (s & 0xFFu) | ((f & 0xFFu) << 8) | (r << 16);
}
// Species parameters, one table for all replicas (stride 0) or one table of
// `stride` species per replica.
struct SpeciesTable {
  const SpeciesRec* rec;
  uint32_t stride;
  __device__ __forceinline__ const SpeciesRec& of(uint32_t tag) const {
    return rec[tag_replica(tag) * stride + tag_species(tag)];
  }
};

enum : uint32_t {
  MF_DEFUNCT = 1u << 0,
  MF_SCHEDULE_UNIMOL = 1u << 1,  // lifetime not drawn yet (MOLECULE_FLAG_SCHEDULE_UNIMOL_RXN)
  MF_TOFF = 1u << 2,             // reactive path: the diffusion-time offset array holds a non-zero value
  // reactive path, within one iteration only: the stream pass of the
  // molecule's next step (its n-th diffusion in the iteration reads pass n)
  // and whether that pass's first draw is already spent (the unimolecular
  // pathway draw of a reactant that stays)
  MF_PASS_SHIFT = 3,
  MF_PASS_MASK = 7u << 3,
  MF_SKIP_DRAW = 1u << 6,
  MF_STEP_FLAGS = MF_SCHEDULE_UNIMOL | MF_PASS_MASK | MF_SKIP_DRAW,  // cleared when a step ends
  // MOLECULE_FLAG_MATURE (molecule.h:34): a molecule of a species with a
  // time step above one iteration moves with its full step
  MF_MATURE = 1u << 7,
};
constexpr double kTimeInvalid = -256.0;   // TIME_INVALID (defines.h:144)
constexpr double kTimeForever = 1e20;     // TIME_FOREVER (defines.h:145)

// Surface classes: what a wall does to a species that hits it
// (collide_and_react_with_walls, diffuse_react_event.cpp:991-1067).
enum : uint32_t { SC_REFLECT = 0, SC_TRANSPARENT = 1, SC_ABSORB = 2 };
struct SurfaceClasses {
  const uint32_t* __restrict__ wall_class;  // per wall: 0 = none, else 1-based class
  const uint8_t* __restrict__ action;       // [(class - 1) * num_species + species]: front | back << 2
  uint32_t num_species;
  uint32_t on;                              // any class with a non-reflect action
  // hit_front: the hit was WALL_FRONT (the move ran against the normal)
  __device__ __forceinline__ uint32_t of(uint32_t wall, uint32_t species, bool hit_front) const {
    if (!on) return SC_REFLECT;
    const uint32_t c = wall_class[wall];
    if (c == 0) return SC_REFLECT;
    const uint32_t a = action[(c - 1) * num_species + species];
    return hit_front ? (a & 3u) : ((a >> 2) & 3u);
  }
};

// error bits reported through a device word
enum : uint32_t {
  ERR_ESCAPED = 1u << 0,     // a molecule left the partition (diffuse_react_event.cpp:592-612)
  ERR_SEGMENTS = 1u << 1,    // more than kMaxSegments reflections in one step
  ERR_REDO = 1u << 2,        // more than kMaxRedo edge re-draws for one wall search
  ERR_DDA = 1u << 3,         // subpartition walk did not terminate
  ERR_TRACE = 1u << 4,       // hit trace buffer overflow (trace only)
  ERR_CELLS = 1u << 5,       // more subpartitions around one move than the collision set holds
  ERR_GRID = 1u << 6,        // a reactive molecule outside the cell grid
  ERR_EXD = 1u << 7,         // exact_disk vertex pool exhausted
  ERR_QUEUE = 1u << 8,       // proposal / product capacity exceeded
  ERR_ROUND32 = 1u << 9,     // FP32 form: a point could not be rounded onto its side of every nearby plane
  ERR_COLLISIONS = 1u << 10, // more molecule collisions in one ray segment than the step's guard allows
  ERR_PASSES = 1u << 11,     // a molecule diffused more often in one iteration than its tag counts (8 calls)
};

// per-call statistics counters (SimulationStats::mol_wall_reflections and
// the transparent crossings), one wave-summed atomic per wave
enum : uint32_t { CTR_REFLECTIONS = 0, CTR_CROSSINGS = 1, CTR_WORDS = 2 };
__device__ __forceinline__ void wave_count(unsigned long long* ctr, uint32_t v) {
  if (ctr == nullptr) return;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if (__lane_id() == 0 && v) atomicAdd(ctr, (unsigned long long)v);
}

constexpr int kMaxSegments = 1024;
constexpr int kMaxRedo = 4096;
constexpr int kTraceHits = 8;

struct HitTrace {
  uint32_t* wall;     // [n * kTraceHits]
  uint32_t* count;    // [n]
  uint32_t enabled;
};

}  // namespace mcg
