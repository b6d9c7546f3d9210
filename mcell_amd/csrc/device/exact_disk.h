// exact_disk on the GPU (exact_disk_utils.inl:840-1145, use_expanded_list):
// the fraction of the interaction disk around a collision point that is not
// cut off by walls, or -1 when a wall lies between the mover and its target.
// Runs only for collisions with a wall of the collision subpartition within
// the reaction radius; the reference's linked lists become a fixed pool of
// vertex records addressed by index.  Host-callable too (the pool-bound
// test runs it on the CPU).
#pragma once
#include "path64.h"

namespace mcg {

namespace exd {

enum Role : int { UNDEFINED = 0, HEAD = 1, TAIL = 2, CROSS = 3, SPAN = 4, OTHER = 5 };

// 40-byte vertex record; links are pool indices
struct V {
  double u, v;
  double r2, zeta;
  int16_t next;
  int16_t e;
  int16_t span;
  int16_t role;
};
static_assert(sizeof(V) == 40, "vertex record");

// 2 vertices per wall edge plus crossings and spans.  A disk near a finely
// meshed surface (triangle edges below the reaction radius) is cut by tens
// of walls.  The pool lives in private memory, sized per kernel: the passes
// of the full step run with small pools and hand a disk that overflows one
// to a pass with the large pool (kPool), where running out is reported as an
// error.  A pool that runs out stops exact_disk with *full set.
constexpr int kPool = 256;

template <int N>
struct PoolT {
  static constexpr int kCap = N;
  V p[N];
  int n;
  bool full;
  // On overflow `full` is set and the last slot handed out again; every
  // caller checks `full` right after and stops before linking it anywhere.
  __host__ __device__ __forceinline__ int make() {
    if (n >= N) {
      full = true;
      return N - 1;
    }
    V& x = p[n];
    x.next = -1; x.u = 0; x.v = 0; x.r2 = 0; x.zeta = 0; x.e = -1; x.span = -1; x.role = UNDEFINED;
    return n++;
  }
};

// exact_disk_utils.inl:54-80
__host__ __device__ __forceinline__ double zetize(double y, double x) {
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
__host__ __device__ __forceinline__ void coordize(D3 mv, D3& m, D3& u, D3& v) {
  double a = 1 / sqrt(dot3(mv, mv));
  m = {a * mv.x, a * mv.y, a * mv.z};
  double mx2 = m.x * m.x, my2 = m.y * m.y, mz2 = m.z * m.z;
  if (mx2 > my2) {
    if (mx2 > mz2) {
      if (my2 > mz2) { u = {m.y, -m.x, 0}; a = 1 - mz2; v = {m.z * m.x, m.z * m.y, -a}; }
      else { u = {m.z, 0, -m.x}; a = 1 - my2; v = {-m.y * m.x, a, -m.y * m.z}; }
    } else { u = {-m.z, 0, m.x}; a = 1 - my2; v = {m.y * m.x, -a, m.y * m.z}; }
  } else {
    if (my2 > mz2) {
      if (mx2 > mz2) { u = {-m.y, m.x, 0}; a = 1 - mz2; v = {-m.z * m.x, -m.z * m.y, a}; }
      else { u = {0, m.z, -m.y}; a = 1 - mx2; v = {-a, m.x * m.y, m.x * m.z}; }
    } else { u = {0, -m.z, m.y}; a = 1 - mx2; v = {a, -m.x * m.y, -m.x * m.z}; }
  }
  a = 1 / sqrt(a);
  u = {u.x * a, u.y * a, u.z * a};
  v = {v.x * a, v.y * a, v.z * a};
}

// src4/defines.h:766-807
__host__ __device__ __forceinline__ bool distinguishable_vec3(D3 a, D3 b, double eps) {
  double c = fabs(a.x), d;
  d = fabs(a.y); if (d > c) c = d;
  d = fabs(a.z); if (d > c) c = d;
  d = fabs(b.x); if (d > c) c = d;
  d = fabs(b.y); if (d > c) c = d;
  d = fabs(b.z); if (d > c) c = d;
  double cc = fabs(a.x - b.x);
  d = fabs(a.y - b.y); if (d > cc) cc = d;
  d = fabs(a.z - b.z); if (d > cc) cc = d;
  if (c < eps) c = eps;
  return c * eps < cc;
}

template <class Pool>
__host__ __device__ __forceinline__ double span(const Pool& P, int v1, int v2, int p) {
  return (P.p[v1].u - P.p[p].u) * (P.p[v2].v - P.p[p].v) - (P.p[v2].u - P.p[p].u) * (P.p[v1].v - P.p[p].v);
}
template <class Pool>
__host__ __device__ __forceinline__ double time_span(const Pool& P, int v1, int v2, int p) {
  return (P.p[p].u * P.p[v1].v - P.p[p].v * P.p[v1].u) /
         (P.p[p].v * (P.p[v2].u - P.p[v1].u) - P.p[p].u * (P.p[v2].v - P.p[v1].v));
}

// exact_disk_utils.inl:517-817
template <class Pool>
__host__ __device__ double area_multiple(Pool& P, int head, double R2) {
  const double PI = 3.14159265358979323846;
  int vp = P.p[head].next;
  int ppa = head, ppb = head;
  P.p[ppa].next = -1;
  P.p[ppa].span = -1;
  while (vp != -1) {
    P.p[vp].span = -1;
    int vq = P.p[vp].next;
    if (P.p[vp].zeta < P.p[ppa].zeta) {
      P.p[vp].next = ppa;
      ppa = vp;
    } else {
      int pqa;
      for (pqa = ppa; P.p[pqa].next != -1; pqa = P.p[pqa].next)
        if (P.p[vp].zeta < P.p[P.p[pqa].next].zeta) break;
      P.p[vp].next = P.p[pqa].next;
      P.p[pqa].next = vp;
      if (P.p[vp].next == -1) ppb = vp;
    }
    vp = vq;
  }
  head = ppa;
  P.p[ppb].next = ppa;

  int pa = P.make(), pb = P.make();
  if (P.full) return 0;  // (a reused slot would link the lists into cycles)
  ppb = -1;
  int guard = 0;
  for (ppa = head; ppa != head || ppb == -1; ppa = P.p[ppa].next) {
    if (++guard > 4 * Pool::kCap) { P.full = true; return 0; }
    if (P.p[ppa].role != HEAD) continue;
    ppb = P.p[ppa].e;
    int pqa;
    for (pqa = P.p[ppa].next; pqa != ppb; pqa = P.p[pqa].next) {
      if (++guard > 16 * Pool::kCap) { P.full = true; return 0; }
      if (P.p[pqa].role != HEAD) continue;
      int pqb = P.p[pqa].e;
      P.p[pa].u = P.p[ppb].u - P.p[ppa].u;
      P.p[pa].v = P.p[ppb].v - P.p[ppa].v;
      P.p[pb].u = P.p[pqb].u - P.p[pqa].u;
      P.p[pb].v = P.p[pqb].v - P.p[pqa].v;
      double r = P.p[pb].u * P.p[pa].v - P.p[pa].u * P.p[pb].v;
      if (r * r < POS_EPS * (P.p[pa].u * P.p[pa].u + P.p[pa].v * P.p[pa].v) *
                      (P.p[pb].u * P.p[pb].u + P.p[pb].v * P.p[pb].v)) {
        P.p[pqa].e = -1;
        P.p[pqa].role = OTHER;
        double a = P.p[pqb].zeta - P.p[ppb].zeta;
        if (a < 0) a += 4;
        if (a > 2) {
          P.p[pqb].role = OTHER;
        } else {
          P.p[ppa].e = pqb;
          P.p[ppb].role = OTHER;
          ppb = pqb;
          pqa = ppa;
        }
        continue;
      }
      double s = (P.p[ppa].u - P.p[pqa].u) * P.p[pa].v - (P.p[ppa].v - P.p[pqa].v) * P.p[pa].u;
      if (s * r <= POS_EPS * R2 * R2) continue;
      double t = s / r;
      if (t >= 1 - POS_EPS) continue;
      if (P.p[pa].u * P.p[pa].u > P.p[pa].v * P.p[pa].v) {
        s = (P.p[pqa].u - P.p[ppa].u + t * P.p[pb].u) * P.p[pa].u;
        if (s <= POS_EPS * R2 || s >= P.p[pa].u * P.p[pa].u * (1 - POS_EPS)) continue;
      } else {
        s = (P.p[pqa].v - P.p[ppa].v + t * P.p[pb].v) * P.p[pa].v;
        if (s <= POS_EPS * R2 || s >= P.p[pa].v * P.p[pa].v * (1 - POS_EPS)) continue;
      }
      int vq = P.make();
      if (P.full) return 0;
      P.p[vq].u = P.p[pqa].u + t * P.p[pb].u;
      P.p[vq].v = P.p[pqa].v + t * P.p[pb].v;
      P.p[vq].r2 = P.p[vq].u * P.p[vq].u + P.p[vq].v * P.p[vq].v;
      P.p[vq].zeta = zetize(P.p[vq].v, P.p[vq].u);
      P.p[vq].e = ppb;
      P.p[vq].span = -1;
      P.p[vq].role = CROSS;
      int w;
      for (w = ppa; w != ppb; w = P.p[w].next) {
        double a = P.p[vq].zeta - P.p[P.p[w].next].zeta;
        if (a > 2) a -= 4;
        else if (a < -2) a += 4;
        if (a < 0) break;
      }
      P.p[vq].next = P.p[w].next;
      P.p[w].next = vq;
      if (P.p[vq].zeta < P.p[head].zeta) head = vq;
    }
  }

  int vq = -1;
  for (vp = head, vq = -1; vq != head; vp = vq) {
    for (vq = P.p[vp].next; vq != head; vq = P.p[vq].next) {
      if (P.p[vq].zeta - P.p[vp].zeta < POS_EPS) {
        P.p[vq].zeta = P.p[vp].zeta;
        if (-POS_EPS < P.p[vq].r2 - P.p[vp].r2 && POS_EPS > P.p[vq].r2 - P.p[vp].r2) P.p[vq].r2 = P.p[vp].r2;
      } else {
        break;
      }
    }
  }

  vq = -1;
  for (vp = head; vp != head || vq == -1; vp = P.p[vp].next) {
    if (P.p[vp].role != HEAD) continue;
    for (vq = P.p[vp].next; vq != P.p[vp].e; vq = P.p[vq].next) {
      if (!distinguishable(P.p[vq].zeta, P.p[vp].zeta, POS_EPS)) continue;
      if (!distinguishable(P.p[vq].zeta, P.p[P.p[vp].e].zeta, POS_EPS)) break;
      if (P.p[vq].role == OTHER) continue;
      int vr = P.make();
      if (P.full) return 0;
      P.p[vr].next = P.p[vq].span;
      P.p[vq].span = vr;
      P.p[vr].e = vp;
      P.p[vr].zeta = P.p[vq].zeta;
      P.p[vr].role = SPAN;
    }
  }

  double A = 0, zeta = 0, last_zeta = -1;
  int vs = -1;
  guard = 0;
  for (vp = head; zeta < 4 - POS_EPS; vp = P.p[vp].next) {
    if (++guard > 8 * Pool::kCap) { P.full = true; return 0; }
    if (P.p[vp].role == OTHER) continue;
    if (!distinguishable(P.p[vp].zeta, last_zeta, POS_EPS)) continue;
    last_zeta = P.p[vp].zeta;
    int vr = (vs == pa) ? pb : pa;
    P.p[vr].u = P.p[vp].u;
    P.p[vr].v = P.p[vp].v;
    P.p[vr].zeta = P.p[vp].zeta;
    if (P.p[vp].role == TAIL) {
      P.p[vr].r2 = R2 * (1 + POS_EPS);
      P.p[vr].e = -1;
    } else {
      P.p[vr].r2 = P.p[vp].r2;
      P.p[vr].e = P.p[vp].e;
    }
    for (vq = P.p[vp].next; !distinguishable(P.p[vq].zeta, last_zeta, POS_EPS); vq = P.p[vq].next) {
      if (P.p[vq].role == HEAD) {
        if (P.p[vq].r2 < P.p[vp].r2 || P.p[vr].e == -1) {
          P.p[vr].u = P.p[vq].u;
          P.p[vr].v = P.p[vq].v;
          P.p[vr].r2 = P.p[vq].r2;
          P.p[vr].e = P.p[vq].e;
        } else if (!distinguishable(P.p[vq].r2, P.p[vr].r2, POS_EPS)) {
          double b = span(P, vr, P.p[vr].e, P.p[vq].e);
          if (b > 0) P.p[vr].e = P.p[vq].e;
        }
      }
    }
    for (vq = P.p[vp].span; vq != -1; vq = P.p[vq].next) {
      int qa = P.p[vq].e;
      int qb = P.p[qa].e;
      double b = span(P, qa, qb, vr);
      double c = b * b;
      if (c < R2 * R2 * POS_EPS) {
        if (P.p[vr].e == -1) {
          P.p[vr].r2 = P.p[vr].u * P.p[vr].u + P.p[vr].v * P.p[vr].v;
          P.p[vr].e = qb;
        } else {
          b = span(P, vr, P.p[vr].e, qb);
          if (b > 0) P.p[vr].e = qb;
        }
      } else if (b < 0 || P.p[vr].e == -1) {
        double t = time_span(P, qa, qb, vp);
        P.p[vr].u = P.p[qa].u + t * (P.p[qb].u - P.p[qa].u);
        P.p[vr].v = P.p[qa].v + t * (P.p[qb].v - P.p[qa].v);
        P.p[vr].r2 = P.p[vr].u * P.p[vr].u + P.p[vr].v * P.p[vr].v;
        P.p[vr].e = qb;
      }
    }
    if (vs == -1) {
      vs = vr;
    } else {
      double c = P.p[vr].zeta - P.p[vs].zeta;
      if (c < 0) c += 4;
      if (c > POS_EPS) {
        zeta += c;
        if (P.p[vs].e == -1 || (P.p[P.p[vs].e].zeta - P.p[vs].zeta) * (P.p[P.p[vs].e].zeta - P.p[vs].zeta) <
                                   POS_EPS * POS_EPS) {
          if (c >= 2) {
            P.p[vs].u = -P.p[vs].u;
            P.p[vs].v = -P.p[vs].v;
            A += 0.5 * PI * R2;
          }
          double a = P.p[vs].u * P.p[vr].u + P.p[vs].v * P.p[vr].v;
          double b = P.p[vs].u * P.p[vr].v - P.p[vs].v * P.p[vr].u;
          double s;
          if (a <= 0) s = atan(-a / b) + 0.5 * PI;
          else s = atan(b / a);
          A += 0.5 * s * R2;
        } else {
          if (!distinguishable(P.p[P.p[vs].e].zeta, P.p[vr].zeta, POS_EPS)) {
            A += 0.5 * (P.p[vs].u * P.p[P.p[vs].e].v - P.p[vs].v * P.p[P.p[vs].e].u);
          } else {
            double t = time_span(P, vs, P.p[vs].e, vr);
            double b2 = P.p[vs].u + (P.p[P.p[vs].e].u - P.p[vs].u) * t;
            double c2 = P.p[vs].v + (P.p[P.p[vs].e].v - P.p[vs].v) * t;
            A += 0.5 * (P.p[vs].u * c2 - P.p[vs].v * b2);
          }
        }
        vs = vr;
      } else {
        if (P.p[vr].e != -1) vs = vr;
      }
    }
  }
  return A;
}

}  // namespace exd

// exact_disk_utils.inl:840-1145; `full` set when the pool ran out.
template <int N = exd::kPool>
__host__ __device__ __noinline__ double exact_disk_dev(const PartitionParams& PP, const WallRec64* __restrict__ walls,
                                              const uint32_t* __restrict__ off, const uint32_t* __restrict__ idx,
                                              const uint32_t* __restrict__ box, D3 loc, D3 mv, double R, D3 target,
                                              bool* full) {
  using namespace exd;
  const double PI = 3.14159265358979323846;
  PoolT<N> P;
  P.n = 0;
  P.full = false;
  *full = false;
  int head = -1;
  int n_edges = 0;
  double R2 = R * R;
  double m2_i = 1 / dot3(mv, mv);
  D3 m, u, v;
  coordize(mv, m, u, v);
  D3 L = {dot3(loc, m), dot3(loc, u), dot3(loc, v)};
  double sm_u, sm_v, sm_r2, sm_zeta;
  if (!distinguishable_vec3(loc, target, POS_EPS)) {
    sm_u = sm_v = sm_r2 = sm_zeta = 0;
  } else {
    D3 td = {target.x - loc.x, target.y - loc.y, target.z - loc.z};
    sm_u = dot3(td, u);
    sm_v = dot3(td, v);
    sm_r2 = sm_u * sm_u + sm_v * sm_v;
    sm_zeta = zetize(sm_v, sm_u);
  }
  const uint32_t csp = subpart_index(PP, loc);
  // the walls the reference keeps lie within R of loc on every axis (its
  // bounding-box test below), so the cube loc +- R bounds their entry boxes
  const uint32_t db = region_box(PP, csp, loc.x - R, loc.y - R, loc.z - R, loc.x + R, loc.y + R, loc.z + R);
  const uint32_t e1 = off[csp + 1];
  for (uint32_t k = next_overlap(box, off[csp], e1, db); k < e1; k = next_overlap(box, k + 1, e1, db)) {
    const WallRec64& w = walls[idx[k]];
    D3 wv[3] = {{w.p0x, w.p0y, w.p0z}, {w.p1x, w.p1y, w.p1z}, {w.p2x, w.p2y, w.p2z}};
    D3 n = {w.nx, w.ny, w.nz};
    double l_n = dot3(loc, n);
    double d = w.d - l_n;
    double m_n = dot3(mv, n);
    if (d * d >= R2 * (1 - m2_i * m_n * m_n)) continue;
    D3 llf = wv[0], urb = wv[0];
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
    D3 v0 = {dot3(wv[0], m) - L.x, dot3(wv[0], u) - L.y, dot3(wv[0], v) - L.z};
    D3 v1 = {dot3(wv[1], m) - L.x, dot3(wv[1], u) - L.y, dot3(wv[1], v) - L.z};
    D3 v2 = {dot3(wv[2], m) - L.x, dot3(wv[2], u) - L.y, dot3(wv[2], v) - L.z};
    double pa_u, pa_v, pb_u, pb_v;
    auto isect = [](D3 p0, D3 p1, double& ou, double& ov) {  // exact_disk_utils.inl:213-223
      double t = p0.x / (p0.x - p1.x);
      ou = p0.y + t * (p1.y - p0.y);
      ov = p0.z + t * (p1.z - p0.z);
    };
    if ((v0.x < 0) == (v1.x < 0)) {
      if ((v2.x < 0) == (v1.x < 0)) continue;
      isect(v0, v2, pa_u, pa_v);
      isect(v1, v2, pb_u, pb_v);
    } else if ((v0.x < 0) == (v2.x < 0)) {
      isect(v0, v1, pa_u, pa_v);
      isect(v2, v1, pb_u, pb_v);
    } else {
      isect(v1, v0, pa_u, pa_v);
      isect(v2, v0, pb_u, pb_v);
    }
    double pa_r2 = pa_u * pa_u + pa_v * pa_v;
    double pb_r2 = pb_u * pb_u + pb_v * pb_v;
    if (pa_r2 < POS_EPS * R2 || pb_r2 < POS_EPS * R2) return -1;
    if (!distinguishable(pa_u * pb_v, pb_u * pa_v, POS_EPS) && (pa_u * pb_u + pa_v * pb_v) < 0) return -1;
    // test_intersect_line_with_circle (exact_disk_utils.inl:226-272)
    double ti = 0, si = 1;
    if (pa_r2 > R2 || pb_r2 > R2) {
      double pa_pb = pa_u * pb_u + pa_v * pb_v;
      if (!distinguishable(pa_r2 + pb_r2, 2 * pa_pb, POS_EPS)) {
        if (sm_r2 < pa_r2 && sm_r2 < pb_r2 && distinguishable(sm_r2, pa_r2, POS_EPS) &&
            distinguishable(sm_r2, pa_r2, POS_EPS))
          continue;
        if (!distinguishable(sm_u * pa_v, sm_v * pa_u, 1e-6) || !distinguishable(sm_u * pb_v, sm_v * pb_u, 1e-6))
          return -1;
        continue;
      }
      double aa = 1 / (pa_r2 + pb_r2 - 2 * pa_pb);
      double bb = (pa_pb - pa_r2) * aa;
      double cc = (R2 - pa_r2) * aa;
      double dd = bb * bb + cc;
      if (dd <= 0) continue;
      dd = sqrt(dd);
      ti = -bb - dd;
      if (ti >= 1) continue;
      if (ti < 0) ti = 0;
      si = -bb + dd;
      if (si <= 0) continue;
      if (si > 1) si = 1;
    }
    // construct_final_endpoints (exact_disk_utils.inl:275-307)
    int ia = P.make(), ib = P.make();
    if (P.full) break;  // reported below; the lists are not walked
    if (ti > 0) {
      P.p[ia].u = pa_u + ti * (pb_u - pa_u);
      P.p[ia].v = pa_v + ti * (pb_v - pa_v);
      P.p[ia].r2 = P.p[ia].u * P.p[ia].u + P.p[ia].v * P.p[ia].v;
      P.p[ia].zeta = zetize(P.p[ia].v, P.p[ia].u);
    } else {
      P.p[ia].u = pa_u; P.p[ia].v = pa_v; P.p[ia].r2 = pa_r2;
      P.p[ia].zeta = zetize(pa_v, pa_u);
    }
    if (si < 1) {
      P.p[ib].u = pa_u + si * (pb_u - pa_u);
      P.p[ib].v = pa_v + si * (pb_v - pa_v);
      P.p[ib].r2 = P.p[ib].u * P.p[ib].u + P.p[ib].v * P.p[ib].v;
      P.p[ib].zeta = zetize(P.p[ib].v, P.p[ib].u);
    } else {
      P.p[ib].u = pb_u; P.p[ib].v = pb_v; P.p[ib].r2 = pb_r2;
      P.p[ib].zeta = zetize(pb_v, pb_u);
    }
    a = P.p[ib].zeta - P.p[ia].zeta;
    if (a < 0) a += 4;
    if (a >= 2) {
      int t = ia; ia = ib; ib = t;
      a = 4 - a;
    }
    b = sm_zeta - P.p[ia].zeta;
    if (b < 0) b += 4;
    if (b < a) {
      double au = P.p[ia].u - sm_u, av = P.p[ia].v - sm_v;
      double bu = P.p[ib].u - sm_u, bv = P.p[ib].v - sm_v;
      double c = au * bv - av * bu;
      if (c < 0 || !distinguishable(au * bv, av * bu, POS_EPS)) return -1;
    }
    P.p[ia].role = HEAD;
    P.p[ib].role = TAIL;
    P.p[ia].e = ib;
    P.p[ib].e = -1;
    P.p[ib].next = head;
    P.p[ia].next = ib;
    head = ia;
    n_edges++;
  }
  if (P.full) { *full = true; return 1; }
  if (n_edges == 0) return 1;
  if (n_edges == 1) {
    int ia = head, ib = P.p[ia].e;
    double ares = P.p[ia].u * P.p[ib].u + P.p[ia].v * P.p[ib].v;
    double bres = P.p[ia].u * P.p[ib].v - P.p[ia].v * P.p[ib].u;
    double sres;
    if (ares <= 0) sres = atan(-ares / bres) + 0.5 * PI;
    else sres = atan(bres / ares);
    return (0.5 * bres + R2 * (PI - 0.5 * sres)) / (PI * R2);
  }
  double A = area_multiple(P, head, R2);
  if (P.full) { *full = true; return 1; }
  return A / (PI * R2);
}

}  // namespace mcg
