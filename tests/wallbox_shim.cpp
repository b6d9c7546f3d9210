// Host build of the sub-cell box helpers (mcell_amd/csrc/device/wallbox.h)
// for tests/test_wallbox.py.
#include "../mcell_amd/csrc/device/wallbox.h"

using namespace mcg;

static PartitionParams params(const double* p) {
  PartitionParams P{};
  P.ox = p[0];
  P.oy = p[1];
  P.oz = p[2];
  P.sp_len = p[3];
  P.sp_rcp = 1.0 / p[3];
  P.n = (uint32_t)p[4];
  P.n2 = P.n * P.n;
  return P;
}

extern "C" {

// The entries next_overlap visits in [k0, k1) (written to out), their count.
uint64_t overlap_walk(const uint32_t* box, uint32_t k0, uint32_t k1, uint32_t sb, uint32_t* out) {
  uint64_t n = 0;
  for (uint32_t k = next_overlap(box, k0, k1, sb); k < k1; k = next_overlap(box, k + 1, k1, sb)) out[n++] = k;
  return n;
}

// The same walk with the light passes' group size.
uint64_t overlap_walk_light(const uint32_t* box, uint32_t k0, uint32_t k1, uint32_t sb, uint32_t* out) {
  uint64_t n = 0;
  for (uint32_t k = next_overlap<kLightOverlapGroup>(box, k0, k1, sb); k < k1;
       k = next_overlap<kLightOverlapGroup>(box, k + 1, k1, sb))
    out[n++] = k;
  return n;
}

// out[i] = region_box of the triangle tri[9i..9i+8] in subpartition s[i]
void box_triangles(const double* pp, uint64_t n, const uint32_t* s, const double* tri, uint32_t* out) {
  const PartitionParams P = params(pp);
  for (uint64_t i = 0; i < n; ++i) {
    const double* t = tri + 9 * i;
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(t[a], std::min(t[3 + a], t[6 + a]));
      hi[a] = std::max(t[a], std::max(t[3 + a], t[6 + a]));
    }
    out[i] = region_box(P, s[i], lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
  }
}

void box_segments(const double* pp, uint64_t n, const uint32_t* s, const double* pos, const double* mv,
                  uint32_t* out) {
  const PartitionParams P = params(pp);
  for (uint64_t i = 0; i < n; ++i)
    out[i] = segment_box(P, s[i], pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], mv[3 * i], mv[3 * i + 1], mv[3 * i + 2]);
}

void box_overlap(uint64_t n, const uint32_t* a, const uint32_t* b, uint8_t* out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = boxes_overlap(a[i], b[i]) ? 1 : 0;
}

}  // extern "C"
