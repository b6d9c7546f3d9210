"""Surface-molecule diffusion on the GPU against the oracle.

Surface molecules take their steps after the iteration's volume moves
(diffuse_surf_molecule, diffuse_react_event.cpp:1071-1246): a 2-D
displacement traced across the walls that share edges and reflected at an
open mesh's border (ray_trace_surf, :1578-1725), then the tile under the end
point if it is free (else a new draw, up to eleven).  The oracle moves them
one at a time in id order; where no two molecules contend for a tile in one
iteration the runs are equal.

* Whole runs equal the oracle's: sparse molecules on the closed icosphere
  and on the open sheet (walls and tiles exactly, uv positions to 1e-9: the
  polar method's log is the device's); ligands binding diffusing receptors
  (L + R' -> LR' on the sphere with R and LR diffusing), where the oracle's
  id order also moves every ligand before the receptors.
* At size: the sheet's 2-D mean squared displacement is 4 D t; at high
  density on the sphere every molecule keeps a tile of its own, the one
  under its position, and none is lost.
* Species time steps (diffuse_single_molecule's rescheduling and
  get_max_time's maturity ramp, diffuse_react_event.cpp:164-198, 287-318):
  steps of 1/2 and 1/3 iteration taken again in the same iteration, a step
  of 3 iterations lengthening from one iteration as the molecule ages,
  against the oracle (walls, tiles, uv, diffusion times, birthdays and the
  mature flag).
* A download and upload in the middle of a run continues it bit for bit,
  also with the maturity ramp under way.
"""
import math
from dataclasses import replace

import numpy as np
import pytest

from mcell_amd import _lib
from mcell_amd import model as M
from mcell_amd import workloads as W
from mcell_amd.engine import DiffuseReactEngine

pytestmark = pytest.mark.gpu


def _surface_by_id(eng):
    s = eng.surface_molecules()
    o = np.argsort(s["id"], kind="stable")
    return {k: v[o] for k, v in s.items()}


def _compare_surface(eng, w, label, first_id=0):
    g = _surface_by_id(eng)
    wa, ta, uv = w.surface_state()
    om = w.molecules()
    live = np.nonzero(((om["flags"] & 1) == 0) & (wa != 0xFFFFFFFF))[0]
    assert np.array_equal(g["id"], live), "%s: the same surface molecules" % label
    assert np.array_equal(g["species"], om["species"][live]), "%s: species" % label
    assert np.array_equal(g["wall"], wa[live]), "%s: walls (%d differ)" % (label, np.sum(g["wall"] != wa[live]))
    assert np.array_equal(g["tile"], ta[live]), "%s: tiles" % label
    d = np.abs(g["uv"] - uv[live]).max(initial=0)
    assert d <= 1e-9, "%s: max |duv| %g" % (label, d)
    return len(live)


@pytest.mark.parametrize("sheet", [False, True])
def test_sparse_surface_diffusion_matches_oracle(gpu_available, oracle_lib, sheet):
    from oracle import pyoracle as po
    seed = 51
    wl = W.surface_diffusion(n_s=400, iterations=100, sheet=sheet, n_sheet=200 if sheet else 0, d_surf=2e-7)
    eng = DiffuseReactEngine.from_workload(wl, seed=seed)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
    w.set_barrier_interval(wl.barrier_interval)
    for chunk in range(0, wl.iterations, 20):
        eng.step(chunk, 20)
        w.run(20)
        n = _compare_surface(eng, w, "iteration %d" % (chunk + 20))
        print("iteration %d: %d surface molecules, %s" % (chunk + 20, n, w.surface_stats()), flush=True)
    assert w.surface_stats()["moves_between_walls"] > 1000
    w.close()
    eng.close()


def test_receptors_that_diffuse_match_oracle(gpu_available, oracle_lib):
    from oracle import pyoracle as po
    seed = 53
    wl = W.with_diffusing_surface(W.receptor_sphere(n_l=20000, n_r=1500, iterations=60, rate=4e8), {1: 1e-8, 2: 1e-8})
    eng = DiffuseReactEngine.from_workload(wl, seed=seed)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
    w.set_barrier_interval(wl.barrier_interval)
    for chunk in range(0, wl.iterations, 10):
        eng.step(chunk, 10)
        w.run(10)
        g = eng.molecules()
        o = np.argsort(g["id"], kind="stable")
        om = w.molecules()
        live = np.nonzero(((om["flags"] & 1) == 0) & ((om["flags"] & 8) == 0))[0]
        assert np.array_equal(g["id"][o], live)
        assert np.abs(g["pos"][o] - om["pos"][live]).max(initial=0) <= 1e-10
        n = _compare_surface(eng, w, "iteration %d" % (chunk + 10))
        c = eng.count_species(1, 3)[0]
        print("iteration %d: L %d R %d LR %d" % (chunk + 10, c[0], c[1], c[2]), flush=True)
    assert c[1] + c[2] == 1500 and c[2] > 20
    w.close()
    eng.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sheet_msd_at_size(gpu_available, oracle_lib, precision):
    """600 molecules on a fifth of the tiles in the middle of the sheet:
    4 D t after 20 iterations (rms 4 length units, the border 10 away;
    positions and walls from the engine, frames from the mesh)."""
    from oracle import pyoracle as po
    n, iters, d = 600, 20, 2e-7
    wl = W.surface_diffusion(n_s=0, iterations=iters, sheet=True, n_sheet=0, d_surf=d, barrier_interval=50)
    nw = len(wl.walls)
    w = po.world_from_workload(wl)
    rng = np.random.default_rng(17)
    tri = wl.vertices[wl.walls[nw - 2]]
    f1, f2 = tri[1] - tri[0], tri[2] - tri[0]
    p1, p2 = np.array([np.linalg.norm(f1), 0.0]), None
    uvec = f1 / np.linalg.norm(f1)
    nrm = np.cross(f1, f2)
    vvec = np.cross(nrm / np.linalg.norm(nrm), uvec)
    p2 = np.array([f2 @ uvec, f2 @ vvec])
    used, tiles, uvs = set(), [], []
    while len(tiles) < n:
        uv = 0.5 * p2 + rng.uniform(-40, 40, size=2)
        a, b = np.linalg.solve(np.array([p1, p2]).T, uv)
        if not (a > 1e-3 and b > 1e-3 and a + b < 1 - 1e-3):
            continue
        t = w.uv2grid(nw - 2, uv[0], uv[1])
        if t in used:
            continue
        used.add(t)
        tiles.append(t)
        uvs.append(uv)
    w.close()
    uvs = np.array(uvs)
    sm = {"ids": np.arange(n, dtype=np.uint32), "species": np.zeros(n, np.uint32),
          "walls": np.full(n, nw - 2, np.uint32), "tiles": np.array(tiles, np.uint32),
          "orients": np.ones(n, np.int32), "replica": np.zeros(n, np.uint32), "uv": uvs}
    wl = replace(wl, surface_mols=sm)
    eng = DiffuseReactEngine.from_workload(wl, seed=23, precision=precision)
    eng.step(0, iters)
    g = _surface_by_id(eng)

    def xyz(wall, uv):
        p = wl.vertices[wl.walls[wall]]
        e1, e2 = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]
        uu = e1 / np.linalg.norm(e1, axis=1)[:, None]
        nn = np.cross(e1, e2)
        vv = np.cross(nn / np.linalg.norm(nn, axis=1)[:, None], uu)
        return p[:, 0] + uv[:, :1] * uu + uv[:, 1:] * vv

    d2 = ((xyz(g["wall"], g["uv"]) - xyz(np.full(n, nw - 2), uvs)) ** 2).sum(axis=1)
    expect = 4 * d * iters * wl.part.time_unit * 1e8 * wl.part.rcp_length_unit ** 2
    print("%s: msd %.4f expected %.4f, %d on the second triangle" % (precision, d2.mean(), expect,
                                                                     np.sum(g["wall"] == nw - 1)), flush=True)
    assert abs(d2.mean() - expect) < 5 * expect / math.sqrt(n)
    assert np.all(g["diffusion_time"] == iters)
    eng.close()


def test_dense_tiles_stay_consistent(gpu_available, oracle_lib):
    """Half of the sphere's tiles occupied: contended tiles every iteration;
    after 40 iterations every molecule is on a tile of its own, the tile
    uv2grid gives for its position, and none was lost."""
    from oracle import pyoracle as po
    n = 16000
    wl = W.surface_diffusion(n_s=n, iterations=40, d_surf=5e-7, sphere_subdivisions=4)
    eng = DiffuseReactEngine.from_workload(wl, seed=29)
    eng.step(0, wl.iterations)
    g = _surface_by_id(eng)
    assert len(g["id"]) == n and np.array_equal(g["id"], np.arange(n))
    assert len(set(zip(g["wall"].tolist(), g["tile"].tolist()))) == n
    w = po.world_from_workload(wl)
    for a, t, p in zip(g["wall"], g["tile"], g["uv"]):
        assert w.uv2grid(int(a), p[0], p[1]) == t
    w.close()
    eng.close()


def test_surface_round_trip_continues_exactly(gpu_available):
    wl = W.surface_diffusion(n_s=1500, iterations=30, sheet=True, n_sheet=300, d_surf=2e-7, barrier_interval=7)
    first, second = 11, 19
    a = DiffuseReactEngine.from_workload(wl, seed=31)
    a.step(0, first + second)
    whole = _surface_by_id(a)
    b = DiffuseReactEngine.from_workload(wl, seed=31)
    b.step(0, first)
    m = b.surface_molecules()
    nid = b.next_molecule_ids()
    b.close()
    c = DiffuseReactEngine.from_workload(wl, seed=31, add_molecules=False, start_iteration=first)
    c.add_surface_molecules(0, m["id"], m["species"], m["wall"], m["tile"], m["orient"], uv=m["uv"],
                            diffusion_times=m["diffusion_time"])
    c.set_next_molecule_ids(nid)
    c.step(first, second)
    cont = _surface_by_id(c)
    for k in ("id", "wall", "tile", "orient", "uv", "diffusion_time"):
        assert np.array_equal(whole[k], cont[k]), k
    a.close()
    c.close()


def _surface_time_steps(iterations, barrier_interval=7, n_s=600, n_sheet=200):
    """S, S2, S3 on the sphere and the sheet (ids round-robin over the
    species) with time steps of 1/2, 3 and 1/3 iterations."""
    wl = W.surface_diffusion(n_s=n_s, iterations=iterations, sheet=True, n_sheet=n_sheet, d_surf=2e-7,
                             barrier_interval=barrier_interval)
    sp = [wl.species[0], M.Species("S2", 3e-7, surface=True), M.Species("S3", 1e-7, surface=True)]
    steps = [(ss, ts, f | 16) for ss, ts, f in W._species_table(sp, wl.part, [])]
    sm = dict(wl.surface_mols)
    sm["species"] = (np.arange(len(sm["ids"])) % 3).astype(np.uint32)
    wl = replace(wl, species=sp, species_steps=steps, surface_mols=sm)
    return W.with_time_steps(wl, [0.5, 3.0, 1.0 / 3.0])


def test_surface_time_steps_match_oracle(gpu_available, oracle_lib):
    from oracle import pyoracle as po
    seed = 57
    wl = _surface_time_steps(iterations=42)
    eng = DiffuseReactEngine.from_workload(wl, seed=seed)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
    w.set_barrier_interval(wl.barrier_interval)
    for chunk in range(0, wl.iterations, 14):
        eng.step(chunk, 14)
        w.run(14)
        n = _compare_surface(eng, w, "iteration %d" % (chunk + 14))
        g = _surface_by_id(eng)
        om = w.molecules()
        ids = g["id"]
        dt = np.abs(g["diffusion_time"] - om["diffusion_time"][ids]).max(initial=0)
        assert dt <= 1e-12 * (chunk + 14), "diffusion times differ by %g" % dt
        assert np.array_equal(g["birthday"], om["birthday"][ids])
        mature = (om["flags"][ids] & 2) != 0
        assert np.array_equal((g["flags"] & _lib.MCG_MOLF_MATURE) != 0, mature)
        print("iteration %d: %d surface molecules, %d mature, %s" % (chunk + 14, n, mature.sum(), w.surface_stats()),
              flush=True)
    # the long-step species matured during the run (age 1 + 0.2 t > 3 from t = 10)
    assert mature[g["species"] == 1].all() and not mature[g["species"] != 1].any()
    # every molecule's diffusion time lies past the run's end, within its step
    assert np.all(g["diffusion_time"] >= wl.iterations - 1e-9)
    assert np.all(g["diffusion_time"][g["species"] != 1] <= wl.iterations + 1e-9)
    w.close()
    eng.close()


def test_surface_time_steps_round_trip(gpu_available):
    wl = _surface_time_steps(iterations=20, barrier_interval=50)
    first, second = 5, 15  # the long-step species still young at the split
    a = DiffuseReactEngine.from_workload(wl, seed=61)
    a.step(0, first + second)
    whole = _surface_by_id(a)
    b = DiffuseReactEngine.from_workload(wl, seed=61)
    b.step(0, first)
    m = b.surface_molecules()
    nid = b.next_molecule_ids()
    b.close()
    assert not (m["flags"] & _lib.MCG_MOLF_MATURE).any() and np.any(m["diffusion_time"] > first)
    c = DiffuseReactEngine.from_workload(wl, seed=61, add_molecules=False, start_iteration=first)
    c.add_surface_molecules(0, m["id"], m["species"], m["wall"], m["tile"], m["orient"], uv=m["uv"],
                            diffusion_times=m["diffusion_time"], flags=m["flags"], birthdays=m["birthday"])
    c.set_next_molecule_ids(nid)
    c.step(first, second)
    cont = _surface_by_id(c)
    for k in ("id", "wall", "tile", "orient", "uv", "diffusion_time", "flags", "birthday"):
        assert np.array_equal(whole[k], cont[k]), k
    assert (whole["flags"][whole["species"] == 1] & _lib.MCG_MOLF_MATURE).all()
    a.close()
    c.close()
