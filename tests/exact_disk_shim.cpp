// Test helper: runs the engine's exact_disk (device/exact_disk.h, compiled
// for the host) on the engine's own wall set-up (geometry.cpp: wall
// constants, wall lists, entry boxes) with vertex pools of several sizes, so
// that tests can bound the pool a mesh needs and check that a pool that runs
// out stops (reports full) instead of looping.
#include <stdint.h>

#include <vector>

#include "../mcell_amd/csrc/device/exact_disk.h"
#include "../mcell_amd/csrc/geometry.h"

namespace {
mcg::PartitionParams g_P;
std::vector<mcg::WallRec64> g_walls;
std::vector<uint32_t> g_off, g_idx, g_box;

template <int N>
double run(mcg::D3 loc, mcg::D3 mv, double R, mcg::D3 target, bool* full) {
  return mcg::exact_disk_dev<N>(g_P, g_walls.data(), g_off.data(), g_idx.data(), g_box.data(), loc, mv, R, target,
                                full);
}
}  // namespace

// part: origin x, y, z, partition edge, subpartition edge, its reciprocal,
// reaction radius (internal units); n subpartitions per edge.
extern "C" void exd_setup(const double* verts, uint64_t nv, const uint32_t* tris, uint64_t nw, const double* part,
                          uint32_t n) {
  (void)nv;
  mcg::PartitionParams& P = g_P;
  P.ox = part[0]; P.oy = part[1]; P.oz = part[2];
  P.edge = part[3];
  P.qx = P.ox + P.edge; P.qy = P.oy + P.edge; P.qz = P.oz + P.edge;
  P.sp_len = part[4];
  P.sp_rcp = part[5];
  P.rxn_radius = part[6];
  P.n = n;
  P.n2 = n * n;
  P.use_expanded_list = 1;
  P.pad = 0;
  g_walls.resize(nw);
  for (uint64_t i = 0; i < nw; ++i)
    g_walls[i] = mcg::wall_constants(verts + 3 * tris[3 * i], verts + 3 * tris[3 * i + 1], verts + 3 * tris[3 * i + 2]);
  mcg::bin_walls(P, g_walls, g_off, g_idx);
  mcg::entry_boxes(P, g_walls, g_off, g_idx, g_box);
  if (g_idx.empty()) g_idx.push_back(0);
}

// exact_disk for nq collisions (loc, mv, target: nq x 3) with a pool of
// `pool` vertex records (16, 24, 32, 48, 64, 96, 128 or 256); full[i] = 1
// when the pool ran out.  Returns -1 for an unsupported pool size.
extern "C" int exd_eval(uint64_t nq, const double* loc, const double* mv, const double* target, int pool,
                        double* factor, int32_t* full) {
  const double R = g_P.rxn_radius;
  for (uint64_t i = 0; i < nq; ++i) {
    const mcg::D3 l = {loc[3 * i], loc[3 * i + 1], loc[3 * i + 2]};
    const mcg::D3 m = {mv[3 * i], mv[3 * i + 1], mv[3 * i + 2]};
    const mcg::D3 t = {target[3 * i], target[3 * i + 1], target[3 * i + 2]};
    bool f = false;
    double r;
    switch (pool) {
      case 16: r = run<16>(l, m, R, t, &f); break;
      case 24: r = run<24>(l, m, R, t, &f); break;
      case 32: r = run<32>(l, m, R, t, &f); break;
      case 48: r = run<48>(l, m, R, t, &f); break;
      case 64: r = run<64>(l, m, R, t, &f); break;
      case 96: r = run<96>(l, m, R, t, &f); break;
      case 128: r = run<128>(l, m, R, t, &f); break;
      case 256: r = run<256>(l, m, R, t, &f); break;
      default: return -1;
    }
    factor[i] = r;
    full[i] = f ? 1 : 0;
  }
  return 0;
}
