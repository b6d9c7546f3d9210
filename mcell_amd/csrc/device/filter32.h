// FP32 ray-triangle filter with a rigorous error band.
//
// In the FP32+fallback form every wall test is first done in single
// precision.  The test returns "certain miss" only when the FP32 plane
// distances and edge functions clear the triangle by more than a bound on
// their rounding error; everything else (a hit, a near-edge or near-coplanar
// case, a nearly parallel move) is a candidate and is re-solved in FP64 with
// the reference's collide_wall.  Because no FP64 hit or re-draw (REDO) can be
// classified as a certain miss, the FP32+fallback form finds exactly the walls
// the FP64 form finds, given the same start point and displacement.
//
// The bounds use eps = 2^-23 (twice the unit roundoff) and the partition's
// largest coordinate magnitude S, and are widened by generous constant factors;
// tests/test_filter32.py checks them against FP64 on dense near-edge ray sets.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define MCG_HD __host__ __device__ __forceinline__
#else
#define MCG_HD inline
#endif

namespace mcg {

struct FilterWall {  // same layout as the first 16 floats of WallRec32
  float nx, ny, nz, d;
  float ux, uy, uz, uv1u;
  float vx, vy, vz, uv2u;
  float p0x, p0y, p0z, uv2v;
};

struct FilterParams {
  float eps;     // 2^-23
  float S;       // max |coordinate| in the partition
  float e_dd;    // error bound of n.p - d
  float mrel;    // relative error of the move components against the FP64 move (0: the FP64 move rounded)
};

MCG_HD FilterParams make_filter_params(float S, float move_ulps = 0.0f) {
  FilterParams f;
  f.eps = 1.1920928955078125e-07f;  // 2^-23
  f.S = S;
  f.e_dd = 64.0f * f.eps * S;
  f.mrel = move_ulps * f.eps;
  return f;
}

MCG_HD float fabs_f(float x) { return x < 0 ? -x : x; }
MCG_HD float fmax_f(float a, float b) { return a < b ? b : a; }

// Returns true when the move p -> p + m certainly misses the wall (no hit,
// no re-draw in the FP64 test); false means "re-solve in FP64".
// *end_dist gets the FP32 signed distance of the end point to the plane.
MCG_HD bool certain_miss32(const FilterWall& w, const FilterParams& F, float px, float py, float pz, float mx,
                           float my, float mz, float* end_dist) {
  float dd = ((w.nx * px + w.ny * py) + w.nz * pz) - w.d;
  float m1 = (fabs_f(mx) + fabs_f(my)) + fabs_f(mz);
  float dv = (w.nx * mx + w.ny * my) + w.nz * mz;
  // (a move computed in FP32 is within mrel of the FP64 move per component:
  // dv moves by at most mrel * m1, the hit point by t * mrel * max|m|)
  float e_dv = 16.0f * F.eps * m1 + F.mrel * m1;
  float M = 2.0f * (F.e_dd + e_dv);
  float de = dd + dv;
  *end_dist = de;
  if ((dd > M && de > M) || (dd < -M && de < -M)) return true;
  float adv = fabs_f(dv);
  if (adv < 4.0f * M) return false;  // nearly parallel or a tiny move: exact test
  float t = -dd / dv;
  float at = fabs_f(t);
  // bound of the error of t, then of the hit point (max norm)
  float et = 2.0f * (F.e_dd + (at + 1.0f) * e_dv) / adv + 8.0f * F.eps * (at + 1.0f);
  float mxa = fmax_f(fmax_f(fabs_f(mx), fabs_f(my)), fabs_f(mz));
  float eh = et * mxa + 8.0f * F.eps * (F.S + mxa) + at * F.mrel * mxa;
  float hx = px + t * mx, hy = py + t * my, hz = pz + t * mz;
  float lx = hx - w.p0x, ly = hy - w.p0y, lz = hz - w.p0z;
  float L = (fabs_f(lx) + fabs_f(ly)) + fabs_f(lz);
  float el = eh + 4.0f * F.eps * F.S;                 // local point error
  float ebc = 2.0f * el + 16.0f * F.eps * L;          // error of b and c
  float b = (lx * w.ux + ly * w.uy) + lz * w.uz;
  float c = (lx * w.vx + ly * w.vy) + lz * w.vz;
  float f;
  if (w.uv2v < 0) {
    c = -c;
    f = -w.uv2v;
  } else {
    f = w.uv2v;
  }
  float au = fabs_f(w.uv2u);
  // edge functions of the reference's inside test (collision_utils.inl:
  // 743-796): e0 = c, e1 = b*f - c*uv2u, e2 = uv1u*uv2v + c*uv2u - c*uv1u - b*f
  float e0 = c;
  if (e0 < -4.0f * ebc) return true;
  float bf = b * f, cu = c * w.uv2u;
  float e1 = bf - cu;
  float m_e1 = 4.0f * (ebc * (f + au) + 8.0f * F.eps * (fabs_f(bf) + fabs_f(cu)));
  if (e1 < -m_e1) return true;
  float k = w.uv1u * w.uv2v;
  float cv1 = c * w.uv1u;
  float e2 = ((k + cu) - cv1) - bf;
  float m_e2 = 4.0f * (ebc * (au + w.uv1u + f) +
                       8.0f * F.eps * (((fabs_f(k) + fabs_f(cu)) + fabs_f(cv1)) + fabs_f(bf)));
  if (e2 < -m_e2) return true;
  return false;
}

}  // namespace mcg
