"""exact_disk's fixed vertex pool on the crowded-spheres mesh (CPU).

The engine's exact_disk (device/exact_disk.h, compiled for the host by
exact_disk_shim.cpp on the engine's own wall set-up) is run on collisions
placed on and around the mesh -- at random points of walls, at edges and at
vertices, offset along the normal by up to the reaction radius -- with vertex
pools of 16 to 256 records.  Checked: every pool size returns (a pool that runs
out reports full instead of walking a list linked into a cycle), a pool that
does not run out gives exactly the 256-record result, the 256-record pool
never runs out here, and that result equals the oracle's exact_disk
(exact_disk_utils.inl:292-747 via oracle_react.cpp) bit for bit.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from mcell_amd import model as M
from mcell_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
POOLS = [16, 24, 32, 48, 64, 96, 128, 256]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    out = tmp_path_factory.mktemp("exd") / "libexd.so"
    subprocess.check_call(["hipcc", "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-fPIC", "-shared",
                           "-ffp-contract=off", os.path.join(HERE, "exact_disk_shim.cpp"), "-x", "hip",
                           os.path.join(ROOT, "mcell_amd", "csrc", "geometry.cpp"), "-o", str(out)],
                          stderr=subprocess.DEVNULL)
    lib = ctypes.CDLL(str(out))
    lib.exd_setup.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                              ctypes.c_uint32]
    lib.exd_eval.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                             ctypes.c_void_p, ctypes.c_void_p]
    lib.exd_eval.restype = ctypes.c_int
    return lib


@pytest.fixture(scope="module")
def mesh():
    # the full crowded mesh: 64 perturbed icospheres, 1,310,720 sphere walls
    return W.crowded_spheres(n_molecules=64, n_surface=0)


def _collisions(wl, nq, seed):
    """Collision points within R of walls (on faces, at edge midpoints and at
    vertices), moves of random direction and length, and targets within R of
    the move's line through the point, perpendicular to the move."""
    V, T, R = wl.vertices, wl.walls, wl.part.rxn_radius_3d
    rng = np.random.default_rng(seed)
    wi = rng.integers(0, len(T), nq)
    tri = V[T[wi]]
    bc = rng.dirichlet([1, 1, 1], nq)
    on_face = np.einsum("ij,ijk->ik", bc, tri)
    at_vertex = tri[np.arange(nq), rng.integers(0, 3, nq)]
    at_edge = 0.5 * (tri[:, 0] + tri[:, 1])
    mode = rng.integers(0, 3, nq)[:, None]
    base = np.where(mode == 0, on_face, np.where(mode == 1, at_vertex, at_edge))
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    n /= np.linalg.norm(n, axis=1)[:, None]
    loc = base + n * rng.uniform(-R, R, nq)[:, None] + rng.normal(0, 0.1 * R, (nq, 3))
    d = rng.normal(size=(nq, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    mv = d * rng.uniform(0.2, 4.0, nq)[:, None]
    q = rng.normal(size=(nq, 3))
    q -= (q * d).sum(1)[:, None] * d
    q /= np.linalg.norm(q, axis=1)[:, None]
    target = loc + q * (R * np.sqrt(rng.uniform(0, 1, nq)))[:, None]
    return (np.ascontiguousarray(loc), np.ascontiguousarray(mv), np.ascontiguousarray(target))


def _setup(shim, wl):
    p = wl.part
    part = np.array(list(p.origin) + [p.partition_edge_length, p.subpart_edge_length, p.subpart_edge_length_rcp,
                                      p.rxn_radius_3d])
    V = np.ascontiguousarray(wl.vertices, np.float64)
    T = np.ascontiguousarray(wl.walls, np.uint32)
    shim.exd_setup(_p(V), len(V), _p(T), len(T), _p(part), p.num_subparts_per_edge)


def _eval(shim, loc, mv, target, pool):
    f = np.zeros(len(loc))
    full = np.zeros(len(loc), np.int32)
    assert shim.exd_eval(len(loc), _p(loc), _p(mv), _p(target), pool, _p(f), _p(full)) == 0
    return f, full.astype(bool)


def test_pool_sizes_agree_and_stop_when_full(shim, mesh):
    _setup(shim, mesh)
    loc, mv, target = _collisions(mesh, 20000, seed=1)
    ref, ref_full = _eval(shim, loc, mv, target, 256)
    assert not ref_full.any()
    assert np.sum(ref != 1.0) > 10000 and np.sum(ref < 0) > 1000  # the disks do meet walls
    fulls = {}
    for pool in POOLS[:-1]:
        f, full = _eval(shim, loc, mv, target, pool)
        np.testing.assert_array_equal(f[~full], ref[~full])
        fulls[pool] = int(full.sum())
    # the small pools do run out on these disks, and fewer with more records
    assert fulls[16] > 100
    assert all(fulls[a] >= fulls[b] for a, b in zip(POOLS[:-2], POOLS[1:-1]))
    assert fulls[128] == 0


def test_pool_result_equals_oracle(shim, mesh):
    from oracle import pyoracle as po
    _setup(shim, mesh)
    loc, mv, target = _collisions(mesh, 3000, seed=2)
    ref, ref_full = _eval(shim, loc, mv, target, 256)
    assert not ref_full.any()
    w = po.OracleWorld(mesh.config, M.bounding_box(mesh.objects), True)
    for k, obj in enumerate(mesh.objects):
        w.add_object(obj, k)
    w.finalize_walls()
    w.add_species(1e-6)
    w.add_molecules(0, target)
    got = np.array([w.lib.orc_exact_disk(w.w, _p(loc[i]), _p(mv[i]), i) for i in range(len(loc))])
    w.close()
    np.testing.assert_array_equal(ref, got)
