"""The sub-cell box pre-test in front of the wall loops
(mcell_amd/csrc/device/wallbox.h) rejects only walls a move cannot reach.

* boxes_overlap (a packed, branch-free compare) equals the plain per-axis
  interval test on every kind of packed box.
* Conservative: for random small triangles and moves in and around a
  subpartition, every move whose segment passes through the triangle, or
  ends within 1e-6 of it (where the reference's collide_wall may re-draw),
  has a box overlapping the triangle's entry box; the test still rejects a
  good part of the other moves (it is useful).
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
P = np.array([-50.0, -50.0, -50.0, 5.0, 20.0])  # origin, subpartition edge, subpartitions per edge


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("wallbox") / "libwallbox.so"
    subprocess.check_call([hipcc, "-O2", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
                           "-ffp-contract=off", os.path.join(HERE, "wallbox_shim.cpp"), "-o", str(out)])
    lib = ctypes.CDLL(str(out))
    for f in (lib.box_triangles, lib.box_segments, lib.box_overlap):
        f.restype = None
    for f in (lib.overlap_walk, lib.overlap_walk_light):
        f.restype = ctypes.c_uint64
    return lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def test_packed_overlap_equals_interval_test(shim):
    rng = np.random.default_rng(3)
    n = 200000
    lo = rng.integers(0, 16, size=(2, n, 3))
    hi = np.minimum(lo + rng.integers(0, 16, size=(2, n, 3)), 15)
    pack = lambda l, h: sum((l[:, a].astype(np.uint32) | (h[:, a].astype(np.uint32) << 4)) << (8 * a) for a in range(3))
    a, b = pack(lo[0], hi[0]).astype(np.uint32), pack(lo[1], hi[1]).astype(np.uint32)
    out = np.zeros(n, np.uint8)
    shim.box_overlap(ctypes.c_uint64(n), _p(a), _p(b), _p(out))
    expect = np.all((lo[0] <= hi[1]) & (lo[1] <= hi[0]), axis=1)
    assert np.array_equal(out.astype(bool), expect)


def test_grouped_walk_visits_the_overlapping_entries_in_order(shim):
    # next_overlap reads the boxes in groups; the walk it drives must visit
    # exactly the entries a one-by-one loop keeps, in list order, for ranges
    # of every length modulo the group size (including empty ones)
    rng = np.random.default_rng(11)
    for trial in range(400):
        n = int(rng.integers(0, 40))
        lo = rng.integers(0, 16, size=(n + 1, 3))
        hi = np.minimum(lo + rng.integers(0, 6, size=(n + 1, 3)), 15)
        boxes = np.array([sum((int(lo[i, a]) | (int(hi[i, a]) << 4)) << (8 * a) for a in range(3))
                          for i in range(n + 1)], dtype=np.uint32)
        sb, box = boxes[n], np.ascontiguousarray(boxes[:n]) if n else np.zeros(1, np.uint32)
        k0 = int(rng.integers(0, n + 1))
        keep = [k for k in range(k0, n) if np.all((lo[k] <= hi[n]) & (lo[n] <= hi[k]))]
        for walk in (shim.overlap_walk, shim.overlap_walk_light):
            out = np.zeros(max(n, 1), np.uint32)
            cnt = walk(_p(box), ctypes.c_uint32(k0), ctypes.c_uint32(n), ctypes.c_uint32(int(sb)), _p(out))
            assert list(out[:cnt]) == keep, (trial, n, k0, walk)


def _segment_triangle_distance(p, m, t):
    """Smallest distance between each segment p + s m (s in [0, 1]) and
    triangle t (n x 3 x 3), by dense sampling of the segment and exact
    point-triangle distances (an upper bound tight to the sampling step)."""
    best = np.full(len(p), np.inf)
    for s in np.linspace(0.0, 1.0, 201):
        q = p + s * m
        best = np.minimum(best, _point_triangle_distance(q, t))
    return best


def _point_triangle_distance(q, t):
    a, b, c = t[:, 0], t[:, 1], t[:, 2]
    ab, ac = b - a, c - a
    n = np.cross(ab, ac)
    nn = np.einsum("ij,ij->i", n, n)
    # projection onto the plane and its barycentric coordinates
    w = q - a
    dist_plane = np.einsum("ij,ij->i", w, n) / np.sqrt(nn)
    proj = q - (dist_plane / np.sqrt(nn))[:, None] * n
    v2 = proj - a
    d00, d01, d11 = (np.einsum("ij,ij->i", ab, ab), np.einsum("ij,ij->i", ab, ac), np.einsum("ij,ij->i", ac, ac))
    d20, d21 = np.einsum("ij,ij->i", v2, ab), np.einsum("ij,ij->i", v2, ac)
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    u = (d00 * d21 - d01 * d20) / den
    inside = (v >= 0) & (u >= 0) & (v + u <= 1)
    out = np.abs(dist_plane)

    def seg_dist(x, y):
        d = y - x
        s = np.clip(np.einsum("ij,ij->i", q - x, d) / np.einsum("ij,ij->i", d, d), 0, 1)
        return np.linalg.norm(q - (x + s[:, None] * d), axis=1)

    edges = np.minimum(np.minimum(seg_dist(a, b), seg_dist(b, c)), seg_dist(c, a))
    return np.where(inside, out, edges)


def test_boxes_are_conservative_and_selective(shim):
    rng = np.random.default_rng(7)
    n = 60000
    s = np.full(n, 10 + 10 * 20 + 10 * 400, np.uint32)  # subpartition (10, 10, 10): [0, 5)^3
    # small triangles (edges ~0.4) around and across the subpartition, moves of ~2.5 units
    centre = rng.uniform(-1.0, 6.0, size=(n, 3))
    tri = centre[:, None, :] + rng.normal(scale=0.25, size=(n, 3, 3))
    pos = rng.uniform(-0.5, 5.5, size=(n, 3))
    # half of the moves aimed at the triangle (through a random point on it)
    aim = rng.uniform(size=n) < 0.5
    wts = rng.dirichlet([1, 1, 1], size=n)
    target = np.einsum("ij,ijk->ik", wts, tri)
    mv = rng.normal(scale=1.5, size=(n, 3))
    d = target - pos
    reach = rng.uniform(0.9, 2.0, size=n)
    mv[aim] = d[aim] * reach[aim, None]
    through = aim & (reach >= 1.0)  # the segment passes through a point of the triangle
    tb = np.zeros(n, np.uint32)
    sb = np.zeros(n, np.uint32)
    shim.box_triangles(_p(P), ctypes.c_uint64(n), _p(s), _p(np.ascontiguousarray(tri)), _p(tb))
    ov = np.zeros(n, np.uint8)
    # grazing moves: end points within 1e-6 of the triangle
    graze = ~aim & (rng.uniform(size=n) < 0.5)
    mv[graze] = (target[graze] + rng.normal(scale=1e-7, size=(graze.sum(), 3))) - pos[graze]
    shim.box_segments(_p(P), ctypes.c_uint64(n), _p(s), _p(pos), _p(mv), _p(sb))
    shim.box_overlap(ctypes.c_uint64(n), _p(tb), _p(sb), _p(ov))
    far = ~(through | graze)
    dist = _segment_triangle_distance(pos[far], mv[far], tri[far])
    print("through %d, grazing %d, rejected %.1f%% of the others" % (through.sum(), graze.sum(),
                                                                     100 * (1 - ov[far].mean())))
    assert through.sum() > 10000 and graze.sum() > 10000
    assert np.all(ov[through] == 1) and np.all(ov[graze] == 1)
    # of the other moves, those the sampling finds touching are kept too
    assert np.all(ov[far][dist <= 1e-9] == 1)
    assert (ov[far] == 0).mean() > 0.3
