// TEST INFRASTRUCTURE ONLY (CPU oracle; see oracle_core.h).
// Bimolecular volume reactions: collide_and_react_with_vol_mol, exact_disk,
// test_bimolecular and the product outcome (src4/diffuse_react_event.cpp:
// 786-829, 1833-1895, 2446-2933; src4/exact_disk_utils.inl; src4/rxn_utils.inl).
#include <algorithm>
#include <stdexcept>

#include "oracle_core.h"

namespace orc {

// ----------------------------------------------------------------- exact_disk
namespace {

enum class Role { UNDEFINED, HEAD, TAIL, CROSS, SPAN, OTHER };

struct ExdV {
  int next = -1;
  double u = 0, v = 0;  // (x, y)
  double r2 = NAN, zeta = NAN;
  int e = -1;
  int span = -1;
  Role role = Role::UNDEFINED;
};

// exact_disk_utils.inl:54-80
double exd_zetize(double y, double x) {
  if (y >= 0) {
    if (x >= 0) {
      if (x < y) return 1 - 0.5 * x / y;
      return 0.5 * y / x;
    }
    if (-x < y) return 1 - 0.5 * x / y;
    return 2 + 0.5 * y / x;
  }
  if (x <= 0) {
    if (y < x) return 3 - 0.5 * x / y;
    return 2 + 0.5 * y / x;
  }
  if (x < -y) return 3 - 0.5 * x / y;
  return 4 + 0.5 * y / x;
}

// exact_disk_utils.inl:94-145
void exd_coordize(V3 mv, V3& m, V3& u, V3& v) {
  double a = 1 / std::sqrt(dot(mv, mv));
  m = v3(a) * mv;
  double mx2 = m.x * m.x, my2 = m.y * m.y, mz2 = m.z * m.z;
  if (mx2 > my2) {
    if (mx2 > mz2) {
      if (my2 > mz2) {
        u = {m.y, -m.x, 0};
        a = 1 - mz2;
        v = {m.z * m.x, m.z * m.y, -a};
      } else {
        u = {m.z, 0, -m.x};
        a = 1 - my2;
        v = {-m.y * m.x, a, -m.y * m.z};
      }
    } else {
      u = {-m.z, 0, m.x};
      a = 1 - my2;
      v = {m.y * m.x, -a, m.y * m.z};
    }
  } else {
    if (my2 > mz2) {
      if (mx2 > mz2) {
        u = {-m.y, m.x, 0};
        a = 1 - mz2;
        v = {-m.z * m.x, -m.z * m.y, a};
      } else {
        u = {0, m.z, -m.y};
        a = 1 - mx2;
        v = {-a, m.x * m.y, m.x * m.z};
      }
    } else {
      u = {0, -m.z, m.y};
      a = 1 - mx2;
      v = {a, -m.x * m.y, -m.x * m.z};
    }
  }
  a = 1 / std::sqrt(a);
  u = u * a;
  v = v * a;
}

bool distinguishable_vec3(V3 a, V3 b, double eps) {  // src4/defines.h:766-807
  double c = std::fabs(a.x), d;
  d = std::fabs(a.y); if (d > c) c = d;
  d = std::fabs(a.z); if (d > c) c = d;
  d = std::fabs(b.x); if (d > c) c = d;
  d = std::fabs(b.y); if (d > c) c = d;
  d = std::fabs(b.z); if (d > c) c = d;
  double cc = std::fabs(a.x - b.x);
  d = std::fabs(a.y - b.y); if (d > cc) cc = d;
  d = std::fabs(a.z - b.z); if (d > cc) cc = d;
  if (c < eps) c = eps;
  return c * eps < cc;
}

enum class IRes { OCCLUDED, SKIP, CONTINUE };

// exact_disk_utils.inl:213-223
void intersect_w_m0(V3 p0, V3 p1, ExdV& out) {
  double t = p0.x / (p0.x - p1.x);
  out.u = p0.y + t * (p1.y - p0.y);
  out.v = p0.z + t * (p1.z - p0.z);
}

// exact_disk_utils.inl:226-272
IRes intersect_line_with_circle(const ExdV& pa, const ExdV& pb, const ExdV& sm, double R2, double& t, double& s) {
  t = 0;
  s = 1;
  if (pa.r2 > R2 || pb.r2 > R2) {
    double pa_pb = pa.u * pb.u + pa.v * pb.v;
    if (!distinguishable(pa.r2 + pb.r2, 2 * pa_pb, POS_EPS)) {
      if (sm.r2 < pa.r2 && sm.r2 < pb.r2 && distinguishable(sm.r2, pa.r2, POS_EPS) &&
          distinguishable(sm.r2, pa.r2, POS_EPS))
        return IRes::SKIP;
      if (!distinguishable(sm.u * pa.v, sm.v * pa.u, SQRT_EPS) || !distinguishable(sm.u * pb.v, sm.v * pb.u, SQRT_EPS))
        return IRes::OCCLUDED;
      return IRes::SKIP;
    }
    double a = 1 / (pa.r2 + pb.r2 - 2 * pa_pb);
    double b = (pa_pb - pa.r2) * a;
    double c = (R2 - pa.r2) * a;
    double d = b * b + c;
    if (d <= 0) return IRes::SKIP;
    d = std::sqrt(d);
    t = -b - d;
    if (t >= 1) return IRes::SKIP;
    if (t < 0) t = 0;
    s = -b + d;
    if (s <= 0) return IRes::SKIP;
    if (s > 1) s = 1;
  }
  return IRes::CONTINUE;
}

struct Pool {
  std::vector<ExdV> v;
  int make() {
    v.emplace_back();
    return (int)v.size() - 1;
  }
  ExdV& operator[](int i) { return v[i]; }
};

// exact_disk_utils.inl:275-307
void construct_final_endpoints(Pool& P, double ti, double si, const ExdV& pa, const ExdV& pb, int& ia, int& ib) {
  ia = P.make();
  ib = P.make();
  ExdV& A = P[ia];
  if (ti > 0) {
    A.u = pa.u + ti * (pb.u - pa.u);
    A.v = pa.v + ti * (pb.v - pa.v);
    A.r2 = A.u * A.u + A.v * A.v;
    A.zeta = exd_zetize(A.v, A.u);
  } else {
    A.u = pa.u; A.v = pa.v; A.r2 = pa.r2;
    A.zeta = exd_zetize(pa.v, pa.u);
  }
  ExdV& B = P[ib];
  if (si < 1) {
    B.u = pa.u + si * (pb.u - pa.u);
    B.v = pa.v + si * (pb.v - pa.v);
    B.r2 = B.u * B.u + B.v * B.v;
    B.zeta = exd_zetize(B.v, B.u);
  } else {
    B.u = pb.u; B.v = pb.v; B.r2 = pb.r2;
    B.zeta = exd_zetize(pb.v, pb.u);
  }
}

double exd_span(Pool& P, int v1, int v2, int p) {  // :505-508
  return (P[v1].u - P[p].u) * (P[v2].v - P[p].v) - (P[v2].u - P[p].u) * (P[v1].v - P[p].v);
}
double time_span(Pool& P, int v1, int v2, int p) {  // :511-514
  return (P[p].u * P[v1].v - P[p].v * P[v1].u) /
         (P[p].v * (P[v2].u - P[v1].u) - P[p].u * (P[v2].v - P[v1].v));
}

// exact_disk_utils.inl:517-817
double area_for_multiple_edges(Pool& P, int vertex_head, double R2) {
  // insertion sort by zeta
  int vp = P[vertex_head].next;
  int ppa = vertex_head, ppb = vertex_head;
  P[ppa].next = -1;
  P[ppa].span = -1;
  while (vp != -1) {
    P[vp].span = -1;
    int vq = P[vp].next;
    if (P[vp].zeta < P[ppa].zeta) {
      P[vp].next = ppa;
      ppa = vp;
    } else {
      int pqa;
      for (pqa = ppa; P[pqa].next != -1; pqa = P[pqa].next)
        if (P[vp].zeta < P[P[pqa].next].zeta) break;
      P[vp].next = P[pqa].next;
      P[pqa].next = vp;
      if (P[vp].next == -1) ppb = vp;
    }
    vp = vq;
  }
  vertex_head = ppa;
  P[ppb].next = ppa;

  // walk around the circle, inserting crossing points
  int pa = P.make(), pb = P.make();  // scratch displacement vectors / approval points
  ppb = -1;
  for (ppa = vertex_head; ppa != vertex_head || ppb == -1; ppa = P[ppa].next) {
    if (P[ppa].role != Role::HEAD) continue;
    ppb = P[ppa].e;
    int pqa;
    for (pqa = P[ppa].next; pqa != ppb; pqa = P[pqa].next) {
      if (P[pqa].role != Role::HEAD) continue;
      int pqb = P[pqa].e;
      P[pa].u = P[ppb].u - P[ppa].u;
      P[pa].v = P[ppb].v - P[ppa].v;
      P[pb].u = P[pqb].u - P[pqa].u;
      P[pb].v = P[pqb].v - P[pqa].v;
      double r = P[pb].u * P[pa].v - P[pa].u * P[pb].v;
      if (r * r < POS_EPS * (P[pa].u * P[pa].u + P[pa].v * P[pa].v) * (P[pb].u * P[pb].u + P[pb].v * P[pb].v)) {
        P[pqa].e = -1;
        P[pqa].role = Role::OTHER;
        double a = P[pqb].zeta - P[ppb].zeta;
        if (a < 0) a += 4;
        if (a > 2) {
          P[pqb].role = Role::OTHER;
        } else {
          P[ppa].e = pqb;
          P[ppb].role = Role::OTHER;
          ppb = pqb;
          pqa = ppa;
        }
        continue;
      }
      double s = (P[ppa].u - P[pqa].u) * P[pa].v - (P[ppa].v - P[pqa].v) * P[pa].u;
      if (s * r <= POS_EPS * R2 * R2) continue;
      double t = s / r;
      if (t >= 1 - POS_EPS) continue;
      if (P[pa].u * P[pa].u > P[pa].v * P[pa].v) {
        s = (P[pqa].u - P[ppa].u + t * P[pb].u) * P[pa].u;
        if (s <= POS_EPS * R2 || s >= P[pa].u * P[pa].u * (1 - POS_EPS)) continue;
      } else {
        s = (P[pqa].v - P[ppa].v + t * P[pb].v) * P[pa].v;
        if (s <= POS_EPS * R2 || s >= P[pa].v * P[pa].v * (1 - POS_EPS)) continue;
      }
      int vq = P.make();
      P[vq].u = P[pqa].u + t * P[pb].u;
      P[vq].v = P[pqa].v + t * P[pb].v;
      P[vq].r2 = P[vq].u * P[vq].u + P[vq].v * P[vq].v;
      P[vq].zeta = exd_zetize(P[vq].v, P[vq].u);
      P[vq].e = ppb;
      P[vq].span = -1;
      P[vq].role = Role::CROSS;
      int w;
      for (w = ppa; w != ppb; w = P[w].next) {
        double a = P[vq].zeta - P[P[w].next].zeta;
        if (a > 2) a -= 4;
        else if (a < -2) a += 4;
        if (a < 0) break;
      }
      P[vq].next = P[w].next;
      P[w].next = vq;
      if (P[vq].zeta < P[vertex_head].zeta) vertex_head = vq;
    }
  }

  // collapse nearby points in zeta and r2
  int vq = -1;
  for (vp = vertex_head, vq = -1; vq != vertex_head; vp = vq) {
    for (vq = P[vp].next; vq != vertex_head; vq = P[vq].next) {
      if (P[vq].zeta - P[vp].zeta < POS_EPS) {
        P[vq].zeta = P[vp].zeta;
        if (-POS_EPS < P[vq].r2 - P[vp].r2 && POS_EPS > P[vq].r2 - P[vp].r2) P[vq].r2 = P[vp].r2;
      } else {
        break;
      }
    }
  }

  // register spanning segments
  vq = -1;
  for (vp = vertex_head; vp != vertex_head || vq == -1; vp = P[vp].next) {
    if (P[vp].role != Role::HEAD) continue;
    for (vq = P[vp].next; vq != P[vp].e; vq = P[vq].next) {
      if (!distinguishable(P[vq].zeta, P[vp].zeta, POS_EPS)) continue;
      if (!distinguishable(P[vq].zeta, P[P[vp].e].zeta, POS_EPS)) break;
      if (P[vq].role == Role::OTHER) continue;
      int vr = P.make();
      P[vr].next = P[vq].span;
      P[vq].span = vr;
      P[vr].e = vp;
      P[vr].zeta = P[vq].zeta;
      P[vr].role = Role::SPAN;
    }
  }

  // walk and accumulate the area
  double A = 0, zeta = 0, last_zeta = -1;
  int vs = -1;
  for (vp = vertex_head; zeta < 4 - POS_EPS; vp = P[vp].next) {
    if (P[vp].role == Role::OTHER) continue;
    if (!distinguishable(P[vp].zeta, last_zeta, POS_EPS)) continue;
    last_zeta = P[vp].zeta;
    int vr = (vs == pa) ? pb : pa;
    P[vr].u = P[vp].u;
    P[vr].v = P[vp].v;
    P[vr].zeta = P[vp].zeta;
    if (P[vp].role == Role::TAIL) {
      P[vr].r2 = R2 * (1 + POS_EPS);
      P[vr].e = -1;
    } else {
      P[vr].r2 = P[vp].r2;
      P[vr].e = P[vp].e;
    }
    for (vq = P[vp].next; !distinguishable(P[vq].zeta, last_zeta, POS_EPS); vq = P[vq].next) {
      if (P[vq].role == Role::HEAD) {
        if (P[vq].r2 < P[vp].r2 || P[vr].e == -1) {
          P[vr].u = P[vq].u;
          P[vr].v = P[vq].v;
          P[vr].r2 = P[vq].r2;
          P[vr].e = P[vq].e;
        } else if (!distinguishable(P[vq].r2, P[vr].r2, POS_EPS)) {
          double b = exd_span(P, vr, P[vr].e, P[vq].e);
          if (b > 0) P[vr].e = P[vq].e;
        }
      }
    }
    for (vq = P[vp].span; vq != -1; vq = P[vq].next) {
      int qa = P[vq].e;
      int qb = P[qa].e;
      double b = exd_span(P, qa, qb, vr);
      double c = b * b;
      if (c < R2 * R2 * POS_EPS) {
        if (P[vr].e == -1) {
          P[vr].r2 = P[vr].u * P[vr].u + P[vr].v * P[vr].v;
          P[vr].e = qb;
        } else {
          b = exd_span(P, vr, P[vr].e, qb);
          if (b > 0) P[vr].e = qb;
        }
      } else if (b < 0 || P[vr].e == -1) {
        double t = time_span(P, qa, qb, vp);
        P[vr].u = P[qa].u + t * (P[qb].u - P[qa].u);
        P[vr].v = P[qa].v + t * (P[qb].v - P[qa].v);
        P[vr].r2 = P[vr].u * P[vr].u + P[vr].v * P[vr].v;
        P[vr].e = qb;
      }
    }
    if (vs == -1) {
      vs = vr;
    } else {
      double c = P[vr].zeta - P[vs].zeta;
      if (c < 0) c += 4;
      if (c > POS_EPS) {
        zeta += c;
        if (P[vs].e == -1 ||
            (P[P[vs].e].zeta - P[vs].zeta) * (P[P[vs].e].zeta - P[vs].zeta) < POS_EPS * POS_EPS) {
          if (c >= 2) {
            P[vs].u = -P[vs].u;
            P[vs].v = -P[vs].v;
            A += 0.5 * MY_PI * R2;
          }
          double a = P[vs].u * P[vr].u + P[vs].v * P[vr].v;
          double b = P[vs].u * P[vr].v - P[vs].v * P[vr].u;
          double s;
          if (a <= 0) s = std::atan(-a / b) + 0.5 * MY_PI;
          else s = std::atan(b / a);
          A += 0.5 * s * R2;
        } else {
          if (!distinguishable(P[P[vs].e].zeta, P[vr].zeta, POS_EPS)) {
            A += 0.5 * (P[vs].u * P[P[vs].e].v - P[vs].v * P[P[vs].e].u);
          } else {
            double t = time_span(P, vs, P[vs].e, vr);
            double b2 = P[vs].u + (P[P[vs].e].u - P[vs].u) * t;
            double c2 = P[vs].v + (P[P[vs].e].v - P[vs].v) * t;
            A += 0.5 * (P[vs].u * c2 - P[vs].v * b2);
          }
        }
        vs = vr;
      } else {
        if (P[vr].e != -1) vs = vr;
      }
    }
  }
  return A;
}

}  // namespace

// exact_disk_utils.inl:840-1145 with use_expanded_list = true (MCell4 always
// enables it when bimolecular volume reactions exist, converter:387), so
// subpartition boundaries never occlude the disk.  Walls of the collision
// subpartition that the mover could cross transparently are not yet modelled
// (no surface classes on this path).
double World::exact_disk(V3 loc, V3 mv, double R, const Mol& moving, const Mol& target) const {
  (void)moving;
  Pool P;
  int vertex_head = -1;
  int n_edges = 0;
  double R2 = R * R;
  double m2_i = 1 / dot(mv, mv);
  V3 m, u, v;
  exd_coordize(mv, m, u, v);
  V3 L{dot(loc, m), dot(loc, u), dot(loc, v)};
  ExdV sm;
  V3 tpos = snapshot ? snapshot_pos[target.id] : target.pos;
  if (!distinguishable_vec3(loc, tpos, POS_EPS)) {
    sm.u = sm.v = sm.r2 = sm.zeta = 0;
  } else {
    V3 td = tpos - loc;
    sm.u = dot(td, u);
    sm.v = dot(td, v);
    sm.r2 = sm.u * sm.u + sm.v * sm.v;
    sm.zeta = exd_zetize(sm.v, sm.u);
  }
  uint32_t csp = subpart_index(loc);
  for (uint32_t k = sp_wall_off[csp]; k < sp_wall_off[csp + 1]; ++k) {
    const Wall& w = walls[sp_wall_idx[k]];
    V3 wv[3] = {verts[w.v[0]], verts[w.v[1]], verts[w.v[2]]};
    double l_n = dot(loc, w.normal);
    double d = w.d - l_n;
    double m_n = dot(mv, w.normal);
    if (d * d >= R2 * (1 - m2_i * m_n * m_n)) continue;
    V3 llf = wv[0], urb = wv[0];
    if (wv[1].x < llf.x) llf.x = wv[1].x; else if (wv[1].x > urb.x) urb.x = wv[1].x;
    if (wv[2].x < llf.x) llf.x = wv[2].x; else if (wv[2].x > urb.x) urb.x = wv[2].x;
    if (wv[1].y < llf.y) llf.y = wv[1].y; else if (wv[1].y > urb.y) urb.y = wv[1].y;
    if (wv[2].y < llf.y) llf.y = wv[2].y; else if (wv[2].y > urb.y) urb.y = wv[2].y;
    if (wv[1].z < llf.z) llf.z = wv[1].z; else if (wv[1].z > urb.z) urb.z = wv[1].z;
    if (wv[2].z < llf.z) llf.z = wv[2].z; else if (wv[2].z > urb.z) urb.z = wv[2].z;
    double a, b;
    b = R2 * (1 - mv.x * mv.x * m2_i);
    a = llf.x - loc.x; if (a > 0 && a * a >= b) continue;
    a = loc.x - urb.x; if (a > 0 && a * a >= b) continue;
    b = R2 * (1 - mv.y * mv.y * m2_i);
    a = llf.y - loc.y; if (a > 0 && a * a >= b) continue;
    a = loc.y - urb.y; if (a > 0 && a * a >= b) continue;
    b = R2 * (1 - mv.z * mv.z * m2_i);
    a = llf.z - loc.z; if (a > 0 && a * a >= b) continue;
    a = loc.z - urb.z; if (a > 0 && a * a >= b) continue;
    V3 v0 = V3{dot(wv[0], m), dot(wv[0], u), dot(wv[0], v)} - L;
    V3 v1 = V3{dot(wv[1], m), dot(wv[1], u), dot(wv[1], v)} - L;
    V3 v2 = V3{dot(wv[2], m), dot(wv[2], u), dot(wv[2], v)} - L;
    ExdV pa, pb;
    if ((v0.x < 0) == (v1.x < 0)) {
      if ((v2.x < 0) == (v1.x < 0)) continue;
      intersect_w_m0(v0, v2, pa);
      intersect_w_m0(v1, v2, pb);
    } else if ((v0.x < 0) == (v2.x < 0)) {
      intersect_w_m0(v0, v1, pa);
      intersect_w_m0(v2, v1, pb);
    } else {
      intersect_w_m0(v1, v0, pa);
      intersect_w_m0(v2, v0, pb);
    }
    pa.r2 = pa.u * pa.u + pa.v * pa.v;
    pb.r2 = pb.u * pb.u + pb.v * pb.v;
    if (pa.r2 < POS_EPS * R2 || pb.r2 < POS_EPS * R2) return -1;
    if (!distinguishable(pa.u * pb.v, pb.u * pa.v, POS_EPS) && (pa.u * pb.u + pa.v * pb.v) < 0) return -1;
    double ti, si;
    IRes cr = intersect_line_with_circle(pa, pb, sm, R2, ti, si);
    if (cr == IRes::OCCLUDED) return -1;
    if (cr == IRes::SKIP) continue;
    int ia, ib;
    construct_final_endpoints(P, ti, si, pa, pb, ia, ib);
    a = P[ib].zeta - P[ia].zeta;
    if (a < 0) a += 4;
    if (a >= 2) {
      std::swap(ia, ib);
      a = 4 - a;
    }
    b = sm.zeta - P[ia].zeta;
    if (b < 0) b += 4;
    if (b < a) {
      double au = P[ia].u - sm.u, av = P[ia].v - sm.v;
      double bu = P[ib].u - sm.u, bv = P[ib].v - sm.v;
      double c = au * bv - av * bu;
      if (c < 0 || !distinguishable(au * bv, av * bu, POS_EPS)) return -1;
    }
    P[ia].role = Role::HEAD;
    P[ib].role = Role::TAIL;
    P[ia].e = ib;
    P[ib].e = -1;
    P[ib].next = vertex_head;
    P[ia].next = ib;
    vertex_head = ia;
    n_edges++;
  }
  if (n_edges == 0) return 1;
  if (n_edges == 1) {
    int ia = vertex_head, ib = P[ia].e;
    double ares = P[ia].u * P[ib].u + P[ia].v * P[ib].v;
    double bres = P[ia].u * P[ib].v - P[ia].v * P[ib].u;
    double sres;
    if (ares <= 0) sres = std::atan(-ares / bres) + 0.5 * MY_PI;
    else sres = std::atan(bres / ares);
    return (0.5 * bres + R2 * (MY_PI - 0.5 * sres)) / (MY_PI * R2);
  }
  double A = area_for_multiple_edges(P, vertex_head, R2);
  return A / (MY_PI * R2);
}

// ------------------------------------------------------------ reactions
// rxn_utils.inl:301-320
static int binary_search_double(const std::vector<double>& A, double match, int max_idx, double mult) {
  int min_idx = 0;
  while (max_idx - min_idx > 1) {
    int mid = (max_idx + min_idx) / 2;
    if (match > (A[mid] * mult)) min_idx = mid;
    else max_idx = mid;
  }
  if (match > A[min_idx] * mult) return max_idx;
  return min_idx;
}

// rxn_utils.inl:336-414 (local_prob_factor == 0).  The pathway is
// RxnClass::get_pathway_index_for_probability: a binary search over the
// cumulative pathway probabilities (MCell3 src/react_cond.c:120-173).
int World::test_bimolecular(const BimolRxn& rx, double scaling) {
  return test_bimolecular_p(rx.max_fixed_p, rx.cum_probs, scaling);
}

int World::test_bimolecular_p(double max_fixed_p, const std::vector<double>& cum_probs, double scaling) {
  return test_bimolecular_u(max_fixed_p, cum_probs, scaling, rng().dbl());
}

// The same with its one rng_dbl draw u given (every branch draws exactly once).
int test_bimolecular_u(double max_fixed_p, const std::vector<double>& cum_probs, double scaling, double u) {
  double prob;
  if (max_fixed_p < scaling) {
    prob = u * scaling;
    if (prob >= max_fixed_p) return -1;
  } else {
    float max_p = (float)max_fixed_p;  // rxn_utils.inl:367 keeps max_p as float
    if (max_p >= scaling) {
      prob = u * max_p;
    } else {
      prob = u * scaling;
      if (prob >= max_p) return -1;
    }
  }
  return binary_search_double(cum_probs, prob, (int)cum_probs.size() - 1, 1);
}

// diffuse_react_event.cpp:786-829
bool World::collide_and_react_with_vol_mol(const Collision& c, V3& disp, double t_steps, double r_rate_factor,
                                           double elapsed) {
  const Mol& colliding = mols[c.partner];
  const Mol& diffused = mols[c.mover];
  double factor = exact_disk(c.pos, disp, cfg.rxn_radius_3d, diffused, colliding);
  if (factor < 0) return false;
  const BimolRxn& rx = rxns[c.rxn];
  double abs_time = elapsed + t_steps * c.time;
  double scaling = factor * r_rate_factor;
  int i = test_bimolecular(rx, scaling);
  if (i < 0) return false;
  if (snapshot) {
    rxn_log.push_back({(uint32_t)iteration, c.mover, c.partner, c.rxn, i, abs_time, c.pos});
    return true;
  }
  outcome_bimolecular(c, i, abs_time);
  const Mol& mover = mols[c.mover];
  return (mover.flags & MF_DEFUNCT) != 0;
}

// outcome_bimolecular + outcome_products_random for two volume reactants
// (diffuse_react_event.cpp:1833-1895, 2446-2933): a reactant is kept when its
// species appears among the products; every other product is a new volume
// molecule at the collision point that diffuses for the rest of the step.
void World::outcome_bimolecular(const Collision& c, int pathway, double time) {
  const BimolRxn& rx = rxns[c.rxn];
  uint32_t reacA = c.mover, reacB = c.partner;
  if (mols[reacA].species != rx.a || mols[reacB].species != rx.b) std::swap(reacA, reacB);
  const std::vector<uint32_t>& prods = rx.products[pathway];
  bool keepA = std::find(prods.begin(), prods.end(), rx.a) != prods.end();
  bool keepB = std::find(prods.begin(), prods.end(), rx.b) != prods.end();
  if (rx.a == rx.b) keepB = std::count(prods.begin(), prods.end(), rx.a) > 1;
  bool usedA = false, usedB = false;
  for (uint32_t ps : prods) {
    if (keepA && !usedA && ps == rx.a) { usedA = true; continue; }
    if (keepB && !usedB && ps == rx.b) { usedB = true; continue; }
    uint32_t nid = add_volume_molecule(ps, c.pos, time);
    // the volume the collision point lies in: the mover's (it crossed no
    // wall between)
    mols[nid].cvol = mols[c.mover].cvol;
    if (cmp_lt(time, event_time + 1.0, EPS)) new_diffuse_actions.push_back({nid});
  }
  rxn_log.push_back({(uint32_t)iteration, c.mover, c.partner, c.rxn, pathway, time, c.pos});
  if (!keepA) {
    mols[reacA].flags |= MF_DEFUNCT;
    register_reactant(mols[reacA], false);
  }
  if (!keepB) {
    mols[reacB].flags |= MF_DEFUNCT;
    register_reactant(mols[reacB], false);
  }
}

// xyz2grid_tile_index (grid_utils.inl:48-118): the hit point to the
// strip / stripe / flip tile index; points on a vertex map to that vertex's
// corner tile.
uint32_t World::xyz2grid(V3 v, uint32_t wi) const {
  const Wall& w = walls[wi];
  const Grid& g = grids[wi];
  if (g.num_tiles == 1) return 0;
  uint32_t tile_idx_0 = 0;
  uint32_t tile_idx_mid = g.num_tiles - 2 * g.n_axis + 1;
  uint32_t tile_idx_last = g.num_tiles - 1;
  V3 v0 = verts[w.v[0]], v1 = verts[w.v[1]], v2 = verts[w.v[2]];
  if (!distinguishable_vec3(v, v0, POS_EPS)) return tile_idx_mid;
  if (!distinguishable_vec3(v, v1, POS_EPS)) return tile_idx_last;
  if (!distinguishable_vec3(v, v2, POS_EPS)) return tile_idx_0;
  double i = dot(v, w.unit_u) - g.vert0_u;
  double j = dot(v, w.unit_v) - g.vert0_v;
  double striploc = j * g.strip_width_rcp;
  int strip = (int)striploc;
  double striprem = striploc - strip;
  strip = (int)g.n_axis - strip - 1;
  double u0 = j * g.vert2_slope;
  double u1_u0 = w.uv_vert1_u - j * g.fullslope;
  double stripeloc = ((i - u0) / u1_u0) * (strip + (1 - striprem));
  int stripe = (int)stripeloc;
  double striperem = stripeloc - stripe;
  int flip = (striperem < 1 - striprem) ? 0 : 1;
  int idx = strip * strip + 2 * stripe + flip;
  if ((uint32_t)idx >= g.num_tiles) throw std::runtime_error("xyz2grid: tile index out of range");
  return (uint32_t)idx;
}

// collide_and_react_with_surf_mol (diffuse_react_event.cpp:845-975) for the
// classes of VolSurfRxn; 1 = the volume molecule was destroyed (or, in the
// snapshot mode, proposed a reaction), 0 = it reacted, stayed and passes
// through the wall (RX_FLIP), -1 = nothing happened or it reacted and stayed
// on its side.
int World::collide_and_react_with_surf_mol(const Collision& c, double t_steps, double r_rate_factor,
                                           double elapsed) {
  const Grid& g = grids[c.wall];
  uint32_t j = xyz2grid(c.pos, c.wall);
  uint32_t occ = g.tile_mol[j];
  if (occ == INVALID_U32) return -1;
  // a volume product created on this tile does not rebind on its first hit
  // (diffuse_react_event.cpp:876-885)
  if (born_wall == c.wall && born_tile == j) {
    born_wall = born_tile = INVALID_U32;
    return -1;
  }
  const Mol& sm = mols[occ];
  const Mol& vm = mols[c.mover];
  size_t n = species.size();
  if (pair_vs.size() != n * n) return -1;
  int32_t ri = pair_vs[vm.species * n + sm.species];
  if (ri < 0) return -1;
  const VolSurfRxn& rx = vs_rxns[ri];
  // trigger_bimolecular's orientation rule (rxn_utils.inl:58-84)
  int orient_vol = (c.type == CollType::WALL_FRONT) ? 1 : -1;
  int ga = rx.geom_vol, gb = rx.geom_surf;
  bool match = ga == 0 || gb == 0 || (ga + gb) * (ga - gb) != 0 || orient_vol * sm.orient * ga * gb > 0;
  if (!match) return -1;
  double scaling = r_rate_factor / g.binding_factor;
  double time = elapsed + t_steps * c.time;
  int pw = test_bimolecular_p(rx.max_fixed_p, rx.cum_probs, scaling);
  if (pw < 0) return -1;
  uint32_t code = 0x40000000u | (uint32_t)ri;
  if (snapshot) {
    rxn_log.push_back({(uint32_t)iteration, c.mover, occ, (int32_t)code, pw, time, c.pos});
    return 1;
  }
  // outcome_products_random (diffuse_react_event.cpp:2446-2933): products
  // without an orientation draw one each in the rule's order (:2620-2625); a
  // given orientation flips when the surface reactant is not oriented as in
  // the rule (:2636-2658).  The first surface product of the reactant's
  // species is the reactant itself (it takes the orientation); another
  // surface product takes the reactant's tile; volume products start next
  // to the wall (update_vol_mol_after_rxn_with_surf_mol, :2291-2320).
  const std::vector<VolSurfRxn::Product>& prods = rx.products;
  const uint32_t vol_species = vm.species;
  std::vector<int> orient(prods.size());
  for (size_t q = 0; q < prods.size(); ++q)
    if (prods[q].orient == 0) orient[q] = (rng().u32() & 1u) ? 1 : -1;
  const int flip = (rx.geom_surf != 0 && sm.orient != rx.geom_surf) ? -1 : 1;
  for (size_t q = 0; q < prods.size(); ++q)
    if (prods[q].orient != 0) orient[q] = prods[q].orient * flip;
  size_t keep = prods.size();
  for (size_t q = 0; q < prods.size(); ++q)
    if (prods[q].surf && prods[q].species == rx.surf) {
      keep = q;
      break;
    }
  // the first volume product of the mover's species is the mover itself
  // (is_simple_cplx_product_on_both_sides_of_rxn_w_identical_compartments)
  size_t keep_vol = prods.size();
  for (size_t q = 0; q < prods.size(); ++q)
    if (!prods[q].surf && prods[q].species == vol_species) {
      keep_vol = q;
      break;
    }
  const uint32_t wall = sm.wall, tile = sm.tile;
  if (keep == prods.size()) {
    mols[occ].flags |= MF_DEFUNCT;
    grids[wall].tile_mol[tile] = INVALID_U32;
  }
  for (size_t q = 0; q < prods.size(); ++q) {
    if (q == keep) {
      mols[occ].orient = orient[q];
      continue;
    }
    if (q == keep_vol) continue;
    if (prods[q].surf) {
      // the product takes the receptor's tile and position
      // (GridPos::make_with_mol_pos, diffuse_react_event.cpp:2040-2050) and,
      // when it can diffuse, moves for the rest of the iteration
      const double uv[2] = {mols[occ].u, mols[occ].v};
      const uint32_t nid = add_surface_molecule(prods[q].species, wall, tile, orient[q], uv, &time);
      if (species[prods[q].species].space_step > 0 && cmp_lt(time, event_time + 1.0, EPS))
        new_diffuse_actions.push_back({nid});
      continue;
    }
    const uint32_t nid = add_volume_molecule(prods[q].species, c.pos, time);
    Mol& pm = mols[nid];
    // update_counted_volume_id_when_crossing_wall with the product's
    // orientation (:2775-2781), else the volume the mover is in
    pm.cvol = (!cv_in.empty() && cv_in[wall] != INVALID_U32) ? (orient[q] > 0 ? cv_out[wall] : cv_in[wall])
                                                             : mols[c.mover].cvol;
    pm.pos = volume_product_position(nid, wall, orient[q]);
    register_reactant(pm, false);
    pm.subpart = subpart_index(pm.pos);
    pm.reactant_subpart = pm.subpart;
    register_reactant(pm, true);
    if (cmp_lt(time, event_time + 1.0, EPS)) new_diffuse_actions.push_back({nid, wall, tile});
  }
  ++volsurf_reactions;
  rxn_log.push_back({(uint32_t)iteration, c.mover, occ, (int32_t)code, pw, time, c.pos});
  if (keep_vol < prods.size()) {
    // the mover stays (:2690-2720): a product orientation other than the
    // rule's makes it pass through the wall (RX_FLIP, :940-971), else its
    // ray trace goes on as after no reaction
    return orient[keep_vol] != rx.geom_vol ? 0 : -1;
  }
  Mol& mv = mols[c.mover];
  mv.flags |= MF_DEFUNCT;
  register_reactant(mv, false);
  return 1;
}

}  // namespace orc
