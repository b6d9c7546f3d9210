"""The FP32 wall filter never rejects a wall that the FP64 collide_wall hits
or re-draws on (REDO).  Checked on dense ray sets aimed at triangle edges and
vertices, at grazing and in-plane moves, and at random, against the oracle's
FP64 collide_wall (collision_utils.inl:629-812)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from mcell_amd import model as M
from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    out = tmp_path_factory.mktemp("shim") / "libfilter32.so"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                           os.path.join(HERE, "filter32_shim.cpp"), "-o", str(out)])
    lib = ctypes.CDLL(str(out))
    lib.filter_batch.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_uint64, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.filter_batch.restype = None
    lib.filter_batch_perturbed.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_uint64,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.filter_batch_perturbed.restype = None
    return lib


def _world_with_triangles(tris_um):
    cfg = M.Config(subpartition_dimension=0.5)
    objs = [M.GeometryObject("t%d" % k, [tuple(v) for v in t], [(0, 1, 2)]) for k, t in enumerate(tris_um)]
    w = po.OracleWorld(cfg, M.bounding_box(objs), False)
    for k, o in enumerate(objs):
        w.add_object(o, k)
    w.finalize_walls()
    return w


def _check(shim, w, wall, S, pos, move, move_ulps=0.0):
    n = len(pos)
    pos = np.ascontiguousarray(pos, np.float64)
    move = np.ascontiguousarray(move, np.float64)
    types = np.zeros(n, np.int32)
    w.lib.orc_collide_wall_batch.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p]
    w.lib.orc_collide_wall_batch(w.w, wall, n, pos.ctypes.data, move.ctypes.data, types.ctypes.data)
    consts = w.walls()[wall]
    verts = w.vertices(3 * w.lib.orc_num_walls(w.w))[3 * wall:3 * wall + 3]
    f = np.array([consts[0], consts[1], consts[2], consts[9], consts[3], consts[4], consts[5], consts[10],
                  consts[6], consts[7], consts[8], consts[11], verts[0, 0], verts[0, 1], verts[0, 2], consts[12]],
                 np.float32)
    miss = np.zeros(n, np.int32)
    de = np.zeros(n, np.float32)
    shim.filter_batch_perturbed(f.ctypes.data, ctypes.c_float(S), ctypes.c_float(move_ulps), n, pos.ctypes.data,
                                move.ctypes.data, miss.ctypes.data, de.ctypes.data)
    exact_nonmiss = types != 0
    bad = exact_nonmiss & (miss == 1)
    return types, miss, bad


def _rays_at_points(rng, targets, normal, n_per, step):
    """Rays of length ~step that pass through the given target points from
    either side of the plane, at random angles (including grazing)."""
    k = len(targets)
    idx = np.repeat(np.arange(k), n_per)
    d = rng.normal(size=(len(idx), 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    # bias a third of them towards grazing the plane
    graze = rng.random(len(idx)) < 0.33
    d[graze] -= np.outer(d[graze] @ normal, normal) * (1 - 1e-7 * rng.random(graze.sum()))[:, None]
    frac = rng.random(len(idx))
    L = step * (0.2 + rng.random(len(idx)))
    start = targets[idx] - frac[:, None] * L[:, None] * d
    return start, L[:, None] * d


@pytest.mark.parametrize("seed,move_ulps", [(0, 0.0), (1, 0.0), (2, 0.0), (3, 4.0), (4, 4.0)])
def test_filter_is_conservative_near_edges_and_vertices(shim, seed, move_ulps):
    """move_ulps 4: the filter gets the move moved by up to 4 FP32 ulps per
    component (the light pass's FP32 displacement) and must still keep every
    FP64 hit and re-draw of the unmoved FP64 move."""
    rng = np.random.default_rng(seed)
    # triangles in um, well inside a 10 um partition (internal units x100)
    tris = []
    for _ in range(6):
        c = rng.uniform(-3, 3, size=3)
        t = c + rng.normal(scale=rng.choice([0.003, 0.05, 0.5]), size=(3, 3))
        tris.append(t)
    w = _world_with_triangles(tris)
    S = 500.0  # max |coordinate| of the partition in internal units
    total_nonmiss = 0
    for wi in range(len(tris)):
        consts = w.walls()[wi]
        normal = consts[0:3]
        V = w.vertices(3 * len(tris))[3 * wi:3 * wi + 3]
        # target points: on the edges (exactly and within tiny offsets), at
        # vertices, inside and just outside
        tt = rng.random((400, 1))
        edges = np.concatenate([V[0] + tt * (V[1] - V[0]), V[1] + tt * (V[2] - V[1]), V[2] + tt * (V[0] - V[2])])
        jitter = rng.normal(size=edges.shape) * np.abs(edges).max() * rng.choice([1e-15, 1e-12, 1e-9, 1e-7],
                                                                                   size=(len(edges), 1))
        targets = np.concatenate([edges, edges + jitter, np.repeat(V, 50, axis=0),
                                  V[0] + rng.dirichlet([1, 1, 1], 300) @ (V - V[0]) * 1.0])
        step = float(np.linalg.norm(V[1] - V[0]) + 2.0)
        pos, move = _rays_at_points(rng, targets, normal, 8, step)
        types, miss, bad = _check(shim, w, wi, S, pos, move, move_ulps)
        assert not bad.any(), "filter rejected %d FP64 hits/redos on wall %d" % (bad.sum(), wi)
        total_nonmiss += int((types != 0).sum())
        # the filter must also be useful: most far misses are rejected
    assert total_nonmiss > 1000


def test_filter_rejects_clear_misses(shim):
    rng = np.random.default_rng(7)
    tri = np.array([[0.0, 0.0, 0.0], [0.2, 0.0, 0.0], [0.0, 0.2, 0.0]])
    w = _world_with_triangles([tri])
    n = 20000
    pos = rng.uniform(-40, 40, size=(n, 3))
    move = rng.normal(scale=2.0, size=(n, 3))
    types, miss, bad = _check(shim, w, 0, 500.0, pos, move)
    assert not bad.any()
    exact_miss = types == 0
    # nearly every exact miss of random moves is rejected in FP32
    assert miss[exact_miss].mean() > 0.995


def test_filter_handles_in_plane_and_parallel_moves(shim):
    rng = np.random.default_rng(11)
    tri = np.array([[0.125, 0.125, 0.25], [0.375, 0.125, 0.25], [0.125, 0.375, 0.25]])  # plane z = 25 lu exactly
    w = _world_with_triangles([tri])
    n = 5000
    pos = np.column_stack([rng.uniform(5, 45, n), rng.uniform(5, 45, n), np.full(n, 25.0)])
    move = np.column_stack([rng.normal(size=n), rng.normal(size=n), np.zeros(n)])
    consts = w.walls()[0]
    assert consts[0] == 0 and consts[1] == 0 and consts[2] == 1 and consts[9] == 25.0
    types, miss, bad = _check(shim, w, 0, 500.0, pos, move)
    assert (types == 3).all()  # start in the plane, move parallel to it: the re-draw case
    assert not bad.any()
