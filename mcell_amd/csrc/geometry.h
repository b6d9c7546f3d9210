// Host-side geometry set-up (wall constants, wall lists per subpartition).
#pragma once
#include <stdint.h>

#include <vector>

#include "device/types.h"

namespace mcg {

// Wall::initialize_wall_constants (src4/wall.cpp:281-342) for the triangle
// (p0, p1, p2), FP64, glm evaluation order.
WallRec64 wall_constants(const double* p0, const double* p1, const double* p2);

WallRec32 wall_to_fp32(const WallRec64& w);

// Per-subpartition wall lists in CSR form, wall indices ascending inside each
// list (Partition::walls_per_subpart holds ordered sets).
void bin_walls(const PartitionParams& P, const std::vector<WallRec64>& walls, std::vector<uint32_t>& off,
               std::vector<uint32_t>& idx);

// The sub-cell box (device/wallbox.h) of every list entry: the wall's
// bounding box within the entry's subpartition.
void entry_boxes(const PartitionParams& P, const std::vector<WallRec64>& walls, const std::vector<uint32_t>& off,
                 const std::vector<uint32_t>& idx, std::vector<uint32_t>& box);

}  // namespace mcg
