// Host-side geometry set-up of the engine: wall constants and the
// per-subpartition wall lists, computed once when the geometry is handed over.
#include "geometry.h"

#include "device/wallbox.h"

#include <algorithm>
#include <cmath>

namespace mcg {

namespace {

struct V {
  double x, y, z;
};
inline V sub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V scale(V a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V a, V b) {
  double tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z;
  return (tx + ty) + tz;
}
// glm::cross component order (libs/glm/detail/func_geometric.inl:68-79)
inline V cross(V a, V b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
inline double len2(V a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double comp(const V& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

bool distinguishable(double a, double b, double eps) {
  double c = std::fabs(a - b);
  a = std::fabs(a);
  if (a < 1) a = 1;
  b = std::fabs(b);
  if (b < a) eps *= a;
  else eps *= b;
  return c > eps;
}

// Triangle-box overlap test of Partition::finalize_walls
// (WallUtils::wall_in_box, src4/wall_utils.inl:326-517): a vertex inside the
// box, a triangle edge through a box face, or a box edge through the triangle.
bool triangle_meets_box(const V vert[3], const WallRec64& w, V llf, V urb) {
  for (int i = 0; i < 3; ++i) {
    const V& p = vert[i];
    if (p.x >= llf.x && p.x <= urb.x && p.y >= llf.y && p.y <= urb.y && p.z >= llf.z && p.z <= urb.z) return true;
  }
  // triangle edges (v1 -> v2, v1 the previous vertex) against the six faces;
  // for the face normal to axis a the in-face coordinates are (b, c)
  static const int face_axes[3][3] = {{0, 1, 2}, {1, 0, 2}, {2, 1, 0}};
  for (int i = 0; i < 3; ++i) {
    const V& v2 = vert[i];
    const V& v1 = vert[i == 0 ? 2 : i - 1];
    for (int f = 0; f < 3; ++f) {
      int a = face_axes[f][0], b = face_axes[f][1], c = face_axes[f][2];
      double a1 = comp(v1, a), a2 = comp(v2, a);
      for (int side = 0; side < 2; ++side) {
        double plane = side == 0 ? comp(llf, a) : comp(urb, a);
        if ((a1 <= plane && plane < a2) || (a1 > plane && plane >= a2)) {
          double r = (plane - a1) / (a2 - a1);
          double pb = comp(v1, b) + r * (comp(v2, b) - comp(v1, b));
          double pc = comp(v1, c) + r * (comp(v2, c) - comp(v1, c));
          if (comp(llf, b) <= pb && pb <= comp(urb, b) && comp(llf, c) <= pc && pc <= comp(urb, c)) return true;
        }
      }
    }
  }
  // the twelve box edges (corner bits x|y<<1|z<<2, oriented as the reference
  // walks them) against the triangle
  static const int edge_from[12] = {0, 4, 6, 2, 3, 7, 5, 1, 0, 4, 6, 1};
  static const int edge_to[12] = {4, 6, 2, 3, 7, 5, 1, 0, 2, 5, 7, 2};
  V n = {w.nx, w.ny, w.nz};
  double d = w.d;
  V u = sub(vert[1], vert[0]);
  double r_u = 1 / std::sqrt(len2(u));
  u = scale(u, r_u);
  V v = cross(n, u);
  double vu[3], vv[3];
  for (int j = 0; j < 3; ++j) {
    vu[j] = dot(vert[j], u);
    vv[j] = dot(vert[j], v);
  }
  auto corner = [&](int bits) -> V {
    return {(bits & 1) ? urb.x : llf.x, (bits & 2) ? urb.y : llf.y, (bits & 4) ? urb.z : llf.z};
  };
  double dc[8];
  for (int k = 0; k < 8; ++k) dc[k] = dot(corner(k), n);
  for (int e = 0; e < 12; ++e) {
    double a1 = dc[edge_from[e]], a2 = dc[edge_to[e]];
    if ((a1 - d < 0 && a2 - d < 0) || (a1 - d > 0 && a2 - d > 0)) continue;
    V ba = corner(edge_from[e]), bb = corner(edge_to[e]);
    double r = (d - a1) / (a2 - a1);
    V dbe = sub(bb, ba);
    V c = {ba.x + r * dbe.x, ba.y + r * dbe.y, ba.z + r * dbe.z};
    double cu = dot(c, u), cv = dot(c, v);
    int crossings = 0;
    for (int j = 0; j < 3; ++j) {
      int k = (j == 0) ? 2 : j - 1;
      if ((vu[k] < cu && cu <= vu[j]) || (vu[k] >= cu && cu > vu[j])) {
        double rr = (cu - vu[k]) / (vu[j] - vu[k]);
        if ((vv[k] + rr * (vv[j] - vv[k])) > cv) ++crossings;
      }
    }
    if (crossings & 1) return true;
  }
  return false;
}

}  // namespace

WallRec64 wall_constants(const double* p0, const double* p1, const double* p2) {
  WallRec64 w{};
  V v0 = {p0[0], p0[1], p0[2]}, v1 = {p1[0], p1[1], p1[2]}, v2 = {p2[0], p2[1], p2[2]};
  w.p0x = v0.x; w.p0y = v0.y; w.p0z = v0.z;
  w.p1x = v1.x; w.p1y = v1.y; w.p1z = v1.z;
  w.p2x = v2.x; w.p2y = v2.y; w.p2z = v2.z;
  V f1 = sub(v1, v0), f2 = sub(v2, v0);
  w.area = 0.5 * std::sqrt(len2(cross(f1, f2)));
  if (!distinguishable(w.area, 0, 1e-12)) return w;  // degenerate: all constants 0
  V uu = scale(f1, 1 / std::sqrt(len2(f1)));
  V nn = cross(uu, f2);
  nn = scale(nn, 1 / std::sqrt(len2(nn)));
  V vv = cross(nn, uu);
  w.nx = nn.x; w.ny = nn.y; w.nz = nn.z;
  w.ux = uu.x; w.uy = uu.y; w.uz = uu.z;
  w.vx = vv.x; w.vy = vv.y; w.vz = vv.z;
  w.d = dot(v0, nn);
  w.uv1u = dot(f1, uu);
  w.uv2u = dot(f2, uu);
  w.uv2v = dot(f2, vv);
  return w;
}

WallRec32 wall_to_fp32(const WallRec64& w) {
  WallRec32 r;
  r.nx = (float)w.nx; r.ny = (float)w.ny; r.nz = (float)w.nz; r.d = (float)w.d;
  r.ux = (float)w.ux; r.uy = (float)w.uy; r.uz = (float)w.uz; r.uv1u = (float)w.uv1u;
  r.vx = (float)w.vx; r.vy = (float)w.vy; r.vz = (float)w.vz; r.uv2u = (float)w.uv2u;
  r.p0x = (float)w.p0x; r.p0y = (float)w.p0y; r.p0z = (float)w.p0z; r.uv2v = (float)w.uv2v;
  return r;
}

// Partition::finalize_walls + GeometryUtils::wall_subparts_collision_test
// (src4/partition.cpp:91-118, src4/geometry_utils.inl:110-207): every
// subpartition whose box, grown by leeway = POS_EPS*(1 + max|coord|)
// (+ rxn_radius_3d when the expanded list is on), meets the triangle.
void bin_walls(const PartitionParams& P, const std::vector<WallRec64>& walls, std::vector<uint32_t>& off,
               std::vector<uint32_t>& idx) {
  const double POS_EPS = 1e-12;
  std::vector<std::pair<uint32_t, uint32_t>> pairs;
  for (uint32_t wi = 0; wi < walls.size(); ++wi) {
    const WallRec64& w = walls[wi];
    V vert[3] = {{w.p0x, w.p0y, w.p0z}, {w.p1x, w.p1y, w.p1z}, {w.p2x, w.p2y, w.p2z}};
    V lo = vert[0], hi = vert[0];
    for (int k = 1; k < 3; ++k) {
      lo = {std::min(lo.x, vert[k].x), std::min(lo.y, vert[k].y), std::min(lo.z, vert[k].z)};
      hi = {std::max(hi.x, vert[k].x), std::max(hi.y, vert[k].y), std::max(hi.z, vert[k].z)};
    }
    double leeway = 1;
    if (lo.x < -leeway) leeway = -lo.x;
    if (lo.y < -leeway) leeway = -lo.y;
    if (lo.z < -leeway) leeway = -lo.z;
    if (hi.x > leeway) leeway = hi.x;
    if (hi.y > leeway) leeway = hi.y;
    if (hi.z > leeway) leeway = hi.z;
    leeway = POS_EPS + leeway * POS_EPS;
    if (P.use_expanded_list) leeway += P.rxn_radius;
    lo = {lo.x - leeway, lo.y - leeway, lo.z - leeway};
    hi = {hi.x + leeway, hi.y + leeway, hi.z + leeway};
    int i0[3] = {(int)((lo.x - P.ox) * P.sp_rcp), (int)((lo.y - P.oy) * P.sp_rcp), (int)((lo.z - P.oz) * P.sp_rcp)};
    int i1[3] = {(int)((hi.x - P.ox) * P.sp_rcp), (int)((hi.y - P.oy) * P.sp_rcp), (int)((hi.z - P.oz) * P.sp_rcp)};
    for (int x = i0[0]; x <= i1[0]; ++x)
      for (int y = i0[1]; y <= i1[1]; ++y)
        for (int z = i0[2]; z <= i1[2]; ++z) {
          uint32_t s = (uint32_t)x + (uint32_t)y * P.n + (uint32_t)z * P.n2;
          // llf from the index decomposition, as Partition::get_subpart_llf_point
          int ix = (int)(s % P.n), iy = (int)((s / P.n) % P.n), iz = (int)((s / P.n2) % P.n);
          V sl = {P.ox + (double)ix * P.sp_len, P.oy + (double)iy * P.sp_len, P.oz + (double)iz * P.sp_len};
          V su = {sl.x + P.sp_len, sl.y + P.sp_len, sl.z + P.sp_len};
          sl = {sl.x - leeway, sl.y - leeway, sl.z - leeway};
          su = {su.x + leeway, su.y + leeway, su.z + leeway};
          if (triangle_meets_box(vert, w, sl, su)) pairs.push_back({s, wi});
        }
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  size_t n3 = (size_t)P.n2 * P.n;
  off.assign(n3 + 1, 0);
  idx.resize(pairs.size());
  for (auto& p : pairs) off[p.first + 1]++;
  for (size_t i = 0; i < n3; ++i) off[i + 1] += off[i];
  for (size_t i = 0; i < pairs.size(); ++i) idx[i] = pairs[i].second;
}

void entry_boxes(const PartitionParams& P, const std::vector<WallRec64>& walls, const std::vector<uint32_t>& off,
                 const std::vector<uint32_t>& idx, std::vector<uint32_t>& box) {
  box.assign(std::max<size_t>(idx.size(), 1), kBoxAll);
  const size_t n3 = off.size() - 1;
  for (size_t s = 0; s < n3; ++s)
    for (uint32_t k = off[s]; k < off[s + 1]; ++k) {
      const WallRec64& w = walls[idx[k]];
      box[k] = region_box(P, (uint32_t)s, std::min({w.p0x, w.p1x, w.p2x}), std::min({w.p0y, w.p1y, w.p2y}),
                          std::min({w.p0z, w.p1z, w.p2z}), std::max({w.p0x, w.p1x, w.p2x}),
                          std::max({w.p0y, w.p1y, w.p2y}), std::max({w.p0z, w.p1z, w.p2z}));
    }
}

}  // namespace mcg
