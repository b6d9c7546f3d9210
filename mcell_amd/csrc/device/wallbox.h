// Sub-cell boxes: a cheap, order-preserving pre-test in front of every loop
// over a subpartition's wall list.
//
// Each subpartition is split into kSubCells^3 cells.  Every entry of a
// subpartition's wall list carries the cell range (per axis, 4 bits for the
// low and 4 for the high cell) covered by the wall's bounding box widened by
// kBoxMargin and clamped into the subpartition:
//
//   x_lo | x_hi << 4 | y_lo << 8 | y_hi << 12 | z_lo << 16 | z_hi << 20
//
// A move gets the same kind of box for its segment.  Clamping is monotone,
// so two boxes that share a point still share a cell after clamping: when the
// cell ranges of a wall entry and a segment are disjoint on some axis, the
// segment cannot come near the wall (no hit, no edge re-draw), and the loop
// skips the entry.  The entries that remain are visited in list order, so
// the first re-draw and the closest hit are the ones the full loop finds.
// The margin (1e-3 internal units) is far above every rounding difference
// between the host's boxes and the device's, and far below a cell.
#pragma once
#include <stdint.h>

#include "types.h"

#if defined(__HIPCC__) || defined(__HIP__)
#define MCG_BOX_HD __host__ __device__ __forceinline__
#else
#define MCG_BOX_HD inline
#endif

namespace mcg {

constexpr int kSubCells = 16;
constexpr double kBoxMargin = 1e-3;
constexpr uint32_t kBoxAll = 0x00F0F0F0u;  // every cell on every axis

// Cell range [lo, hi] of the interval [a, b] on one axis of a subpartition
// starting at `origin`, packed lo | hi << 4.
MCG_BOX_HD uint32_t cell_range(double a, double b, double origin, double scale) {
  double ta = (a - kBoxMargin - origin) * scale;
  double tb = (b + kBoxMargin - origin) * scale;
  ta = ta < 0.0 ? 0.0 : (ta > kSubCells - 1 ? (double)(kSubCells - 1) : ta);
  tb = tb < 0.0 ? 0.0 : (tb > kSubCells - 1 ? (double)(kSubCells - 1) : tb);
  return (uint32_t)(int)ta | ((uint32_t)(int)tb << 4);
}

// Box of the axis-aligned region [lo, hi] within subpartition s.
MCG_BOX_HD uint32_t region_box(const PartitionParams& P, uint32_t s, double lox, double loy, double loz, double hix,
                               double hiy, double hiz) {
  const uint32_t ix = s % P.n, iy = (s / P.n) % P.n, iz = s / P.n2;
  const double scale = (double)kSubCells * P.sp_rcp;
  return cell_range(lox, hix, P.ox + (double)ix * P.sp_len, scale) |
         (cell_range(loy, hiy, P.oy + (double)iy * P.sp_len, scale) << 8) |
         (cell_range(loz, hiz, P.oz + (double)iz * P.sp_len, scale) << 16);
}

// Box of the segment p -> p + m within subpartition s.
MCG_BOX_HD uint32_t segment_box(const PartitionParams& P, uint32_t s, double px, double py, double pz, double mx,
                                double my, double mz) {
  const double ex = px + mx, ey = py + my, ez = pz + mz;
  return region_box(P, s, px < ex ? px : ex, py < ey ? py : ey, pz < ez ? pz : ez, px < ex ? ex : px,
                    py < ey ? ey : py, pz < ez ? ez : pz);
}

// The same boxes computed in FP32 (the light passes of the FP32 form).  The
// coordinates' FP32 rounding (below 1e-4 internal units for partitions up to
// a few thousand units) is far inside kBoxMargin, so a segment box computed
// here still overlaps the entry box of every wall the segment comes within
// ~1e-3 of.
MCG_BOX_HD uint32_t cell_range32(float a, float b, float origin, float scale) {
  float ta = (a - (float)kBoxMargin - origin) * scale;
  float tb = (b + (float)kBoxMargin - origin) * scale;
  ta = ta < 0.0f ? 0.0f : (ta > kSubCells - 1 ? (float)(kSubCells - 1) : ta);
  tb = tb < 0.0f ? 0.0f : (tb > kSubCells - 1 ? (float)(kSubCells - 1) : tb);
  return (uint32_t)(int)ta | ((uint32_t)(int)tb << 4);
}

// Box of the segment p -> e (end point given) within the subpartition whose
// low corner is (cx, cy, cz); scale = kSubCells / subpartition edge.
MCG_BOX_HD uint32_t segment_box32(float cx, float cy, float cz, float scale, float px, float py, float pz, float ex,
                                  float ey, float ez) {
  return cell_range32(px < ex ? px : ex, px < ex ? ex : px, cx, scale) |
         (cell_range32(py < ey ? py : ey, py < ey ? ey : py, cy, scale) << 8) |
         (cell_range32(pz < ez ? pz : ez, pz < ez ? ez : pz, cz, scale) << 16);
}

// True when the two boxes share a cell on every axis.
MCG_BOX_HD bool boxes_overlap(uint32_t a, uint32_t b) {
  // per 4-bit field: a_lo <= b_hi and b_lo <= a_hi
  const uint32_t a_lo = a & 0x000F0F0Fu, a_hi = (a >> 4) & 0x000F0F0Fu;
  const uint32_t b_lo = b & 0x000F0F0Fu, b_hi = (b >> 4) & 0x000F0F0Fu;
  // (hi | 0x10) - lo keeps bit 4 of each byte set exactly when hi >= lo
  const uint32_t x = ((b_hi | 0x00101010u) - a_lo) & ((a_hi | 0x00101010u) - b_lo);
  return (x & 0x00101010u) == 0x00101010u;
}

// The first entry k' >= k of the list range [k, k1) whose box overlaps sb,
// or k1.  Boxes are read G at a time, so that a long run of skipped entries
// costs a few rounds of loads in flight together instead of one load round
// per entry; the caller visits the entries in list order by calling it again
// from k' + 1.  The full steps, which meet lists of hundreds of entries next
// to finely meshed surfaces, read kOverlapGroup at a time; the light passes,
// whose lists are mostly a few entries long and whose instruction count
// matters more, read kLightOverlapGroup.
constexpr uint32_t kOverlapGroup = 8;
constexpr uint32_t kLightOverlapGroup = 4;
template <uint32_t G = kOverlapGroup>
MCG_BOX_HD uint32_t next_overlap(const uint32_t* __restrict__ box, uint32_t k, uint32_t k1, uint32_t sb) {
  for (; k < k1; k += G) {
    uint32_t bx[G];
    for (uint32_t u = 0; u < G; ++u) bx[u] = box[k + u < k1 ? k + u : k1 - 1];
    for (uint32_t u = 0; u < G; ++u)
      if (k + u < k1 && boxes_overlap(bx[u], sb)) return k + u;
  }
  return k1;
}

}  // namespace mcg
