// TEST INFRASTRUCTURE ONLY (CPU oracle; see oracle_core.h).
// RNG (ISAAC64, Philox4x32-10, Ziggurat), unit/partition derivation, wall
// constants and wall->subpartition binning.
#include <algorithm>
#include <stdexcept>

#include "oracle_core.h"
#include "ziggurat_tables.h"

namespace orc {

// ------------------------------------------------------------------- ISAAC64
// Bob Jenkins' ISAAC64 (public domain) as used by MCell: src/isaac64.c:84-155
// (init with a single 32-bit seed word in r[0]) and :60-82 (generate).
namespace {
inline void isaac_mix(uint64_t& a, uint64_t& b, uint64_t& c, uint64_t& d, uint64_t& e, uint64_t& f,
                      uint64_t& g, uint64_t& h) {
  a -= e; f ^= h >> 9;  h += a;
  b -= f; g ^= a << 9;  a += b;
  c -= g; h ^= b >> 23; b += c;
  d -= h; a ^= c << 15; c += d;
  e -= a; b ^= d >> 14; d += e;
  f -= b; c ^= e << 20; e += f;
  g -= c; d ^= f >> 17; f += g;
  h -= d; e ^= g << 14; g += h;
}
}  // namespace

void Isaac64::generate() {
  uint64_t a = aa;
  uint64_t b = bb + (++cc);
  // 256 steps; step i uses m[i] and m[(i + 128) % 256]; the mixing function
  // cycles through four shift patterns.
  for (int i = 0; i < 256; ++i) {
    uint64_t x = mm[i];
    uint64_t mix;
    switch (i & 3) {
      case 0: mix = ~(a ^ (a << 21)); break;
      case 1: mix = a ^ (a >> 5); break;
      case 2: mix = a ^ (a << 12); break;
      default: mix = a ^ (a >> 33); break;
    }
    a = mix + mm[(i + 128) & 255];
    uint64_t y = mm[(x >> 3) & 255] + a + b;
    mm[i] = y;
    b = mm[(y >> 11) & 255] + x;
    randrsl[i] = b;
  }
  bb = b;
  aa = a;
  ++rngblocks;
}

void Isaac64::init(uint32_t seed) {
  rngblocks = 0;
  aa = bb = cc = 0;
  uint64_t a, b, c, d, e, f, g, h;
  a = b = c = d = e = f = g = h = 0x9e3779b97f4a7c13ULL;
  for (int i = 0; i < 256; ++i) randrsl[i] = 0;
  randrsl[0] = seed;
  for (int i = 0; i < 4; ++i) isaac_mix(a, b, c, d, e, f, g, h);
  for (int pass = 0; pass < 2; ++pass) {
    const uint64_t* src = pass == 0 ? randrsl : mm;
    for (int i = 0; i < 256; i += 8) {
      a += src[i]; b += src[i + 1]; c += src[i + 2]; d += src[i + 3];
      e += src[i + 4]; f += src[i + 5]; g += src[i + 6]; h += src[i + 7];
      isaac_mix(a, b, c, d, e, f, g, h);
      mm[i] = a; mm[i + 1] = b; mm[i + 2] = c; mm[i + 3] = d;
      mm[i + 4] = e; mm[i + 5] = f; mm[i + 6] = g; mm[i + 7] = h;
    }
  }
  generate();
  randcnt = 512;
}

// ------------------------------------------------------------- Philox4x32-10
void Philox::block(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// ------------------------------------------------------------------ Ziggurat
// src/rng.c:173-218.  Bits 0-6 select the strip, bit 7 the sign, bits 8-31
// the position inside the strip; strip 0 samples the tail.
double rng_gauss(DrawSource& r) {
  static const double SCALE_FACTOR = 3.442619855899;
  static const double RECIP_SCALE_FACTOR = 1.0 / 3.442619855899;
  double x, y = 0;
  double sign = 1.0;
  do {
    uint64_t bits = r.u32();
    sign = (bits & 0x80) ? -1.0 : 1.0;
    uint64_t region = bits & 0x0000007f;
    uint64_t pos_within_region = bits & 0xffffff00;
    x = (double)pos_within_region * ZIG_WTAB[region];
    if (pos_within_region < ZIG_KTAB[region]) break;
    if (region != 0) {
      double yB = ZIG_YTAB[region];
      double yR = ZIG_YTAB[region - 1] - yB;
      y = yB + yR * r.dbl();
    } else {
      x = SCALE_FACTOR - std::log1p(-r.dbl()) * RECIP_SCALE_FACTOR;
      y = std::exp(-SCALE_FACTOR * (x - 0.5 * SCALE_FACTOR)) * r.dbl();
    }
  } while (y >= std::exp(-0.5 * x * x));
  return sign * x;
}

// ------------------------------------------------------------ configuration
static double floor_to_multiple_allow_negative(double val, double multiple) {  // defines.h:524-533
  if (val >= 0) return (double)((long long)((val + EPS) / multiple)) * multiple;
  return (double)((long long)((val + EPS - multiple) / multiple)) * multiple;
}
static double floor_to_multiple(double val, double multiple) {  // defines.h:518-522
  return (double)((long long)((val + EPS) / multiple)) * multiple;
}
static double ceil_to_multiple(double val, double multiple) {  // defines.h:543-550
  double res = floor_to_multiple(val, multiple);
  if (!cmp_eq(val, res)) res += multiple;
  return res;
}

Config derive_config(const UserConfig& u, const double bbox[6]) {
  Config c{};
  c.time_unit = u.time_step;                                   // converter:237
  c.grid_density = u.surface_grid_density;
  c.length_unit = 1 / std::sqrt(u.surface_grid_density);       // converter:246
  c.rcp_length_unit = 1 / c.length_unit;
  if (u.interaction_radius > 0)
    c.rxn_radius_3d = u.interaction_radius / c.length_unit;    // converter:252
  else                                                         // BNGConfig default, cf. converter:264
    c.rxn_radius_3d = (1.0 / std::sqrt(MY_PI * c.grid_density)) / c.length_unit;

  V3 llf{bbox[0], bbox[1], bbox[2]}, urb{bbox[3], bbox[4], bbox[5]};
  V3 llf_m = llf - v3(PARTITION_EDGE_EXTRA_MARGIN_UM);
  V3 urb_m = urb + v3(PARTITION_EDGE_EXTRA_MARGIN_UM);
  double auto_dim = max3(urb_m) - min3(llf_m);                // converter:289
  bool auto_origin = false;
  if (auto_dim > u.partition_dimension) {
    c.partition_edge_length = auto_dim / c.length_unit;
    c.origin = (llf - v3(PARTITION_EDGE_EXTRA_MARGIN_UM)) / v3(c.length_unit);
    auto_origin = true;
  } else {
    c.partition_edge_length = u.partition_dimension / c.length_unit;
  }
  if (!auto_origin) c.origin = -v3(c.partition_edge_length) / v3(2);

  double sp_len = u.subpartition_dimension / c.length_unit;   // converter:370
  uint32_t tentative = (uint32_t)(c.partition_edge_length / sp_len);
  if (tentative > MAX_SUBPARTS_PER_PARTITION) sp_len = c.partition_edge_length / MAX_SUBPARTS_PER_PARTITION;
  V3 orig = c.origin;
  c.origin = {floor_to_multiple_allow_negative(orig.x, sp_len), floor_to_multiple_allow_negative(orig.y, sp_len),
              floor_to_multiple_allow_negative(orig.z, sp_len)};
  V3 moved = orig - c.origin;
  double enlarged = c.partition_edge_length + max3(moved);
  c.partition_edge_length = ceil_to_multiple(enlarged, sp_len);
  c.n = (uint32_t)std::round(c.partition_edge_length / sp_len);
  c.use_expanded_list = u.has_bimol_vol_rxns ? 1 : 0;          // converter:84-87,387

  // simulation_config.cpp:45-70 (FP64 branch)
  c.subpart_edge_length = c.partition_edge_length / (double)c.n;
  c.subpart_edge_length_rcp = 1.0 / c.subpart_edge_length;
  c.partition_edge_length = c.subpart_edge_length * c.n;
  c.n2 = c.n * c.n;
  c.n3 = c.n2 * c.n;
  c.opposite = c.origin + v3(c.partition_edge_length);         // partition.cpp:57
  return c;
}

// -------------------------------------------------------------------- world
World::World(const Config& c) : cfg(c) {}

uint32_t World::add_vertex_um(V3 p_um) {
  verts.push_back(p_um * v3(cfg.rcp_length_unit));  // converter:921-923
  return (uint32_t)verts.size() - 1;
}

// src4/wall.cpp:281-342
void World::add_wall(uint32_t a, uint32_t b, uint32_t c, uint32_t object_id, uint32_t surf_class) {
  Wall w{};
  w.v[0] = a; w.v[1] = b; w.v[2] = c;
  w.object_id = object_id;
  w.surf_class = surf_class;
  V3 v0 = verts[a], v1 = verts[b], v2 = verts[c];
  V3 vA = v1 - v0, vB = v2 - v0;
  V3 vX = cross(vA, vB);
  w.area = 0.5 * std::sqrt(len3_squared(vX));
  if (!distinguishable(w.area, 0, EPS)) {
    w.normal = w.unit_u = w.unit_v = v3(0);
    walls.push_back(w);
    return;
  }
  V3 f1 = v1 - v0;
  double f1_len_squared = len3_squared(f1);
  double inv_f1_len = 1 / std::sqrt(f1_len_squared);
  w.unit_u = f1 * v3(inv_f1_len);
  V3 f2 = v2 - v0;
  w.normal = cross(w.unit_u, f2);
  double norm_len_squared = len3_squared(w.normal);
  double inv_norm_len = 1 / std::sqrt(norm_len_squared);
  w.normal = w.normal * v3(inv_norm_len);
  w.unit_v = cross(w.normal, w.unit_u);
  w.d = dot(v0, w.normal);
  w.uv_vert1_u = dot(f1, w.unit_u);
  w.uv2_u = dot(f2, w.unit_u);
  w.uv2_v = dot(f2, w.unit_v);
  walls.push_back(w);
}

// src4/wall_utils.inl:326-517
static int wall_in_box(const V3* vert, const Wall& w, V3 llf, V3 urb) {
  for (int i = 0; i < 3; i++) {  // point_in_box
    const V3& p = vert[i];
    if (p.x >= llf.x && p.x <= urb.x && p.y >= llf.y && p.y <= urb.y && p.z >= llf.z && p.z <= urb.z) return 1;
  }
  for (int i = 0; i < 3; i++) {
    double r, a3, a4;
    const V3* v2 = &vert[i];
    const V3* v1 = (i == 0) ? &vert[2] : &vert[i - 1];
    if ((v1->x <= llf.x && llf.x < v2->x) || (v1->x > llf.x && llf.x >= v2->x)) {
      r = (llf.x - v1->x) / (v2->x - v1->x);
      a3 = v1->y + r * (v2->y - v1->y);
      a4 = v1->z + r * (v2->z - v1->z);
      if (llf.y <= a3 && a3 <= urb.y && llf.z <= a4 && a4 <= urb.z) return 2;
    }
    if ((v1->x <= urb.x && urb.x < v2->x) || (v1->x > urb.x && urb.x >= v2->x)) {
      r = (urb.x - v1->x) / (v2->x - v1->x);
      a3 = v1->y + r * (v2->y - v1->y);
      a4 = v1->z + r * (v2->z - v1->z);
      if (llf.y <= a3 && a3 <= urb.y && llf.z <= a4 && a4 <= urb.z) return 3;
    }
    if ((v1->y <= llf.y && llf.y < v2->y) || (v1->y > llf.y && llf.y >= v2->y)) {
      r = (llf.y - v1->y) / (v2->y - v1->y);
      a3 = v1->x + r * (v2->x - v1->x);
      a4 = v1->z + r * (v2->z - v1->z);
      if (llf.x <= a3 && a3 <= urb.x && llf.z <= a4 && a4 <= urb.z) return 4;
    }
    if ((v1->y <= urb.y && urb.y < v2->y) || (v1->y > urb.y && urb.y >= v2->y)) {
      r = (urb.y - v1->y) / (v2->y - v1->y);
      a3 = v1->x + r * (v2->x - v1->x);
      a4 = v1->z + r * (v2->z - v1->z);
      if (llf.x <= a3 && a3 <= urb.x && llf.z <= a4 && a4 <= urb.z) return 5;
    }
    if ((v1->z <= llf.z && llf.z < v2->z) || (v1->z > llf.z && llf.z >= v2->z)) {
      r = (llf.z - v1->z) / (v2->z - v1->z);
      a3 = v1->y + r * (v2->y - v1->y);
      a4 = v1->x + r * (v2->x - v1->x);
      if (llf.y <= a3 && a3 <= urb.y && llf.x <= a4 && a4 <= urb.x) return 6;
    }
    if ((v1->z <= urb.z && urb.z < v2->z) || (v1->z > urb.z && urb.z >= v2->z)) {
      r = (urb.z - v1->z) / (v2->z - v1->z);
      a3 = v1->y + r * (v2->y - v1->y);
      a4 = v1->x + r * (v2->x - v1->x);
      if (llf.y <= a3 && a3 <= urb.y && llf.x <= a4 && a4 <= urb.x) return 7;
    }
  }
  // box edges against the triangle: the 12 cube edges visited as a walk over
  // the 8 corners (7 new corners, then 5 revisited edges)
  static const int wx1[12] = {0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 1};
  static const int wy1[12] = {0, 0, 1, 1, 1, 1, 0, 0, 0, 0, 1, 0};
  static const int wz1[12] = {0, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 0};
  static const int wx2[12] = {0, 0, 0, 1, 1, 1, 1, 0, 0, 1, 1, 0};
  static const int wy2[12] = {0, 1, 1, 1, 1, 0, 0, 0, 1, 0, 1, 1};
  static const int wz2[12] = {1, 1, 0, 0, 1, 1, 0, 0, 0, 1, 1, 0};
  static const int e1[12] = {0, 1, 3, 2, 6, 7, 5, 4, 0, 1, 3, 4};
  static const int e2[12] = {1, 3, 2, 6, 7, 5, 4, 0, 2, 5, 7, 2};
  V3 n = w.normal;
  double d = w.d;
  V3 u = vert[1] - vert[0];
  double r_u = 1 / std::sqrt(len3_squared(u));
  u = u * v3(r_u);
  V3 v = cross(n, u);
  double vu[3], vv[3];
  for (int j = 0; j < 3; j++) {
    vu[j] = dot(vert[j], u);
    vv[j] = dot(vert[j], v);
  }
  V3 bb = llf;
  double d_box[8];
  d_box[0] = dot(bb, n);
  for (int i = 0; i < 12; i++) {
    double a1, a2;
    V3 ba;
    if (i < 7) {
      ba = bb;
      bb.x = wx2[i] ? urb.x : llf.x;
      bb.y = wy2[i] ? urb.y : llf.y;
      bb.z = wz2[i] ? urb.z : llf.z;
      a2 = d_box[e2[i]] = dot(bb, n);
      a1 = d_box[e1[i]];
      if ((a1 - d < 0 && a2 - d < 0) || (a1 - d > 0 && a2 - d > 0)) continue;
    } else {
      a1 = d_box[e1[i]];
      a2 = d_box[e2[i]];
      if ((a1 - d < 0 && a2 - d < 0) || (a1 - d > 0 && a2 - d > 0)) continue;
      ba.x = wx1[i] ? urb.x : llf.x;
      ba.y = wy1[i] ? urb.y : llf.y;
      ba.z = wz1[i] ? urb.z : llf.z;
      bb.x = wx2[i] ? urb.x : llf.x;
      bb.y = wy2[i] ? urb.y : llf.y;
      bb.z = wz2[i] ? urb.z : llf.z;
    }
    double r = (d - a1) / (a2 - a1);
    V3 c = ba + v3(r) * (bb - ba);
    double cu = dot(c, u), cv = dot(c, v);
    int temp = 0;
    for (int j = 0; j < 3; j++) {
      int k = (j == 0) ? 2 : j - 1;
      if ((vu[k] < cu && cu <= vu[j]) || (vu[k] >= cu && cu > vu[j])) {
        r = (cu - vu[k]) / (vu[j] - vu[k]);
        if ((vv[k] + r * (vv[j] - vv[k])) > cv) temp++;
      }
    }
    if (temp & 1) return 8 + i;
  }
  return 0;
}

// src4/geometry_utils.inl:110-207 + partition.cpp:91-118
void World::finalize_walls() {
  std::vector<std::pair<uint32_t, uint32_t>> pairs;  // (subpart, wall)
  for (uint32_t wi = 0; wi < walls.size(); ++wi) {
    const Wall& w = walls[wi];
    V3 vert[3] = {verts[w.v[0]], verts[w.v[1]], verts[w.v[2]]};
    V3 llf{mind(mind(vert[0].x, vert[1].x), vert[2].x), mind(mind(vert[0].y, vert[1].y), vert[2].y),
           mind(mind(vert[0].z, vert[1].z), vert[2].z)};
    V3 urb{maxd(maxd(vert[0].x, vert[1].x), vert[2].x), maxd(maxd(vert[0].y, vert[1].y), vert[2].y),
           maxd(maxd(vert[0].z, vert[1].z), vert[2].z)};
    double leeway = 1;
    if (llf.x < -leeway) leeway = -llf.x;
    if (llf.y < -leeway) leeway = -llf.y;
    if (llf.z < -leeway) leeway = -llf.z;
    if (urb.x > leeway) leeway = urb.x;
    if (urb.y > leeway) leeway = urb.y;
    if (urb.z > leeway) leeway = urb.z;
    leeway = POS_EPS + leeway * POS_EPS;
    if (cfg.use_expanded_list) leeway += cfg.rxn_radius_3d;
    V3 lw = v3(leeway);
    llf = llf - lw;
    urb = urb + lw;
    int lo[3], hi[3];
    subpart_3d(llf, lo);
    subpart_3d(urb, hi);
    for (int x = lo[0]; x <= hi[0]; x++)
      for (int y = lo[1]; y <= hi[1]; y++)
        for (int z = lo[2]; z <= hi[2]; z++) {
          uint32_t s = subpart_from_3d(x, y, z);
          int i3[3];
          subpart_3d_from_index(s, i3);
          V3 sllf = cfg.origin + V3{(double)i3[0], (double)i3[1], (double)i3[2]} * v3(cfg.subpart_edge_length);
          V3 surb = sllf + v3(cfg.subpart_edge_length);
          sllf = sllf - lw;
          surb = surb + lw;
          if (wall_in_box(vert, w, sllf, surb) != 0) pairs.push_back({s, wi});
        }
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  sp_wall_off.assign((size_t)cfg.n3 + 1, 0);
  sp_wall_idx.resize(pairs.size());
  for (auto& p : pairs) sp_wall_off[p.first + 1]++;
  for (size_t i = 0; i < cfg.n3; ++i) sp_wall_off[i + 1] += sp_wall_off[i];
  for (size_t i = 0; i < pairs.size(); ++i) sp_wall_idx[i] = pairs[i].second;
}

// MCell3 closed forms for volume species (src/mcell_species.c:239-276):
// space_step = sqrt(4e8 * D * time_unit) / length_unit, time_step = 1; with a
// custom time step t (s), space_step = sqrt(4e8 * D * t) / length_unit and
// time_step = t / time_unit (:254-264).
uint32_t World::add_species(double D, bool is_vol, uint32_t extra_flags, double custom_time_step) {
  Species s{};
  s.D = D;
  s.is_vol = is_vol;
  s.flags = extra_flags;
  if (!distinguishable(D, 0, EPS)) {
    s.space_step = 0.0;
    s.time_step = 1.0;
  } else if (custom_time_step > 0) {
    s.space_step = std::sqrt(4.0 * 1.0e8 * D * custom_time_step) * cfg.rcp_length_unit;
    s.time_step = custom_time_step / cfg.time_unit;
  } else {
    s.space_step = std::sqrt(4.0 * 1.0e8 * D * cfg.time_unit) * cfg.rcp_length_unit;
    s.time_step = 1.0;
  }
  species.push_back(s);
  size_t n = species.size();
  std::vector<int32_t> np(n * n, -1);
  for (size_t i = 0; i + 1 < n; ++i)
    for (size_t j = 0; j + 1 < n; ++j) np[i * n + j] = pair_rxn[i * (n - 1) + j];
  pair_rxn.swap(np);
  reactants.resize(n);
  return (uint32_t)n - 1;
}

// vol-vol pb_factor, MCell3 src/react_util.c:163-181; rx_radius in um.
int World::add_bimol_rxn(uint32_t a, uint32_t b, double rate, const std::vector<uint32_t>& products) {
  BimolRxn r;
  r.a = a;
  r.b = b;
  r.rate = rate;
  double eff_vel_a = species[a].space_step / species[a].time_step;
  double eff_vel_b = species[b].space_step / species[b].time_step;
  double rx_radius_um = cfg.rxn_radius_3d * cfg.length_unit;
  double eff_vel = (eff_vel_a + eff_vel_b) * cfg.length_unit / cfg.time_unit;
  double pb = 1.0 / (2.0 * std::sqrt(MY_PI) * rx_radius_um * rx_radius_um * eff_vel);
  pb *= 1.0e15 / N_AV;
  r.pb_factor = pb;
  r.cum_probs.push_back(rate * pb);
  r.max_fixed_p = r.cum_probs.back();
  r.products.push_back(products);
  rxns.push_back(r);
  int idx = (int)rxns.size() - 1;
  size_t n = species.size();
  pair_rxn[a * n + b] = idx;
  pair_rxn[b * n + a] = idx;
  species[a].flags |= SPF_CAN_VOLVOL;
  species[b].flags |= SPF_CAN_VOLVOL;
  return idx;
}

}  // namespace orc

namespace orc {

void World::add_unimol_rxn(uint32_t sp, double rate, const std::vector<uint32_t>& products) {
  if (unimol.size() < species.size()) unimol.resize(species.size());
  UnimolRxn& u = unimol[sp];
  double prev = u.cum.empty() ? 0.0 : u.cum.back();
  u.cum.push_back(prev + rate * cfg.time_unit);
  u.products.push_back(products);
}

}  // namespace orc

namespace orc {

// Grid::initialize (src4/wall.cpp:38-74)
void World::init_grid(uint32_t wi) {
  if (grids.size() < walls.size()) grids.resize(walls.size());
  Grid& g = grids[wi];
  if (g.num_tiles) return;
  const Wall& w = walls[wi];
  g.n_axis = (uint32_t)std::ceil(std::sqrt(w.area));
  if (g.n_axis < 1) g.n_axis = 1;
  g.num_tiles = g.n_axis * g.n_axis;
  g.tile_mol.assign(g.num_tiles, INVALID_U32);
  g.strip_width_rcp = 1 / (w.uv2_v / ((double)g.n_axis));
  g.vert2_slope = w.uv2_u / w.uv2_v;
  g.fullslope = w.uv_vert1_u / w.uv2_v;
  g.binding_factor = ((double)g.num_tiles) / w.area;
  const V3 v0 = verts[w.v[0]];
  g.vert0_u = dot(v0, w.unit_u);
  g.vert0_v = dot(v0, w.unit_v);
}

// vol-surf pb_factor (MCell3 src/react_util.c:132-157): the volume reactant's
// D and time step; doubled when both reactants are oriented the same way.
int World::add_volsurf_rxn(uint32_t vol, uint32_t surf, int geom_vol, int geom_surf, double rate,
                           const std::vector<VolSurfRxn::Product>& products) {
  VolSurfRxn r;
  r.vol = vol;
  r.surf = surf;
  r.geom_vol = geom_vol;
  r.geom_surf = geom_surf;
  r.rate = rate;
  const Species& sv = species[vol];
  double t_step = sv.time_step * cfg.time_unit;
  double pb = (sv.D <= 0) ? 0.0 : 1.0e11 * cfg.grid_density / (2.0 * N_AV) * std::sqrt(MY_PI * t_step / sv.D);
  if ((geom_vol + geom_surf) * (geom_vol - geom_surf) == 0 && geom_vol * geom_surf != 0) pb *= 2.0;
  r.pb_factor = pb;
  r.cum_probs.push_back(rate * pb);
  r.max_fixed_p = r.cum_probs.back();
  r.products = products;
  vs_rxns.push_back(r);
  size_t n = species.size();
  if (pair_vs.size() != n * n) pair_vs.assign(n * n, -1);
  pair_vs[vol * n + surf] = (int32_t)(vs_rxns.size() - 1);
  species[vol].flags |= SPF_CAN_VOLSURF;
  return (int)vs_rxns.size() - 1;
}

uint32_t World::add_surface_molecule(uint32_t sp, uint32_t wi, uint32_t tile, int orient, const double* uv,
                                     const double* time) {
  init_grid(wi);
  Grid& g = grids[wi];
  if (tile >= g.num_tiles) throw std::runtime_error("tile index out of range");
  if (g.tile_mol[tile] != INVALID_U32) throw std::runtime_error("tile already occupied");
  Mol m{};
  m.id = (uint32_t)mols.size();
  m.species = sp;
  m.flags = MF_SURFACE;
  m.wall = wi;
  m.tile = tile;
  m.orient = orient;
  if (uv) {
    m.u = uv[0];
    m.v = uv[1];
  } else {
    grid2uv(wi, tile, m.u, m.v);
  }
  // a surface molecule that can diffuse is scheduled like a volume molecule
  // (GridUtils::place_single_molecule_onto_grid, grid_utils.inl:2097-2160);
  // one that cannot is never moved
  const bool moves = species[sp].space_step > 0;
  m.diffusion_time = moves ? (time ? *time : (double)iteration) : TIME_FOREVER;
  m.birthday = time ? *time : (double)iteration;
  m.unimol_rxn_time = TIME_INVALID;
  mols.push_back(m);
  g.tile_mol[tile] = m.id;
  if (moves) schedulable.push_back(m.id);
  return m.id;
}

}  // namespace orc
