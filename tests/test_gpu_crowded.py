"""The crowded geometry on the GPU: perturbed icospheres with walls far
smaller than the reaction radius (exact_disk cut by tens of walls, tiles of
one per wall), A + B -> E, E -> A + B and A + S' -> AS'.

* Snapshot mode (no unimolecular class, where the oracle's snapshot has no
  counterpart): every proposal and every position equals the oracle's with
  the same Philox draws.
* Full runs in both forms keep the reaction invariants A + E + AS, B + E and
  S + AS.
"""
import dataclasses

import numpy as np
import pytest

from mcell_amd import model as M
from mcell_amd import workloads as W
from mcell_amd.engine import DiffuseReactEngine

pytestmark = pytest.mark.gpu


def _small(**kw):
    args = dict(n_molecules=200000, n_surface=20000, num_spheres=8, sphere_subdivisions=5, iterations=3)
    args.update(kw)
    return W.crowded_spheres(**args)


@pytest.mark.parametrize("drain", ["inline", "defer"])
def test_crowded_snapshot_proposals_match_oracle(gpu_available, oracle_lib, drain):
    from oracle import pyoracle as po
    seed = 13
    wl = _small(iterations=5)
    wl = dataclasses.replace(wl, unimol_rxns=[], rxns=[M.BimolReaction(0, 1, 1e9, [2])],
                             volsurf_rxns=[M.VolSurfReaction(0, 3, 1e9, 4, product_orient=1, geom_vol=0, geom_surf=1)])
    eng = DiffuseReactEngine.from_workload(wl, seed=seed, propose_only=True, log_capacity=1 << 16, drain=drain)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
    w.set_snapshot(True)
    w.set_barrier_interval(wl.barrier_interval)
    total = {"bimol": 0, "volsurf": 0}
    for it in range(wl.iterations):
        eng.step(it, 1)
        w.run(1)
        g = eng.rxn_log()
        u, rx, d = w.rxn_log()
        w.lib.orc_clear_rxn_log(w.w)
        g = g[np.argsort(g["mover_id"], kind="stable")]
        o = np.argsort(u[:, 1], kind="stable")
        assert np.array_equal(g["mover_id"], u[o, 1]), "iteration %d: movers" % it
        assert np.array_equal(g["partner_id"], u[o, 2]), "iteration %d: partners" % it
        assert np.array_equal(g["rxn"], rx[o].astype(np.uint32)), "iteration %d: classes" % it
        assert np.abs(g["time"] - d[o, 0]).max(initial=0) <= 1e-12
        vs = (g["rxn"] & 0x40000000) != 0
        total["volsurf"] += int(vs.sum())
        total["bimol"] += int((~vs).sum())
        m = eng.molecules()
        om = w.molecules()
        live = np.nonzero(((om["flags"] & 1) == 0) & (om["species"] < 3))[0]
        oo = np.argsort(m["id"], kind="stable")
        assert np.array_equal(m["id"][oo], live)
        assert np.abs(m["pos"][oo] - om["pos"][live]).max() <= 1e-12 * 1000, "iteration %d: positions" % it
    print("proposals: %d bimolecular, %d volume-surface" % (total["bimol"], total["volsurf"]), flush=True)
    assert total["bimol"] > 50 and total["volsurf"] > 20
    w.close()
    eng.close()


def test_full_crowded_mesh_snapshot_proposals_match_oracle(gpu_available, oracle_lib):
    """The crowded case's full mesh: 64 perturbed icospheres of 20,480
    triangles (1,310,732 walls with the cube), with 300,000 volume and
    50,000 surface molecules.  Every proposal and position of 4 snapshot
    iterations equals the oracle's."""
    from oracle import pyoracle as po
    seed = 17
    wl = W.crowded_spheres(n_molecules=300000, n_surface=50000, iterations=4)
    assert len(wl.walls) == 1310732
    wl = dataclasses.replace(wl, unimol_rxns=[], rxns=[M.BimolReaction(0, 1, 1e9, [2])],
                             volsurf_rxns=[M.VolSurfReaction(0, 3, 1e9, 4, product_orient=1, geom_vol=0, geom_surf=1)])
    eng = DiffuseReactEngine.from_workload(wl, seed=seed, propose_only=True, log_capacity=1 << 16)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
    w.set_snapshot(True)
    w.set_barrier_interval(wl.barrier_interval)
    total = {"bimol": 0, "volsurf": 0}
    for it in range(wl.iterations):
        eng.step(it, 1)
        w.run(1)
        g = eng.rxn_log()
        u, rx, d = w.rxn_log()
        w.lib.orc_clear_rxn_log(w.w)
        g = g[np.argsort(g["mover_id"], kind="stable")]
        o = np.argsort(u[:, 1], kind="stable")
        assert np.array_equal(g["mover_id"], u[o, 1]), "iteration %d: movers" % it
        assert np.array_equal(g["partner_id"], u[o, 2]), "iteration %d: partners" % it
        assert np.array_equal(g["rxn"], rx[o].astype(np.uint32)), "iteration %d: classes" % it
        vs = (g["rxn"] & 0x40000000) != 0
        total["volsurf"] += int(vs.sum())
        total["bimol"] += int((~vs).sum())
        m = eng.molecules()
        om = w.molecules()
        live = np.nonzero(((om["flags"] & 1) == 0) & (om["species"] < 3))[0]
        oo = np.argsort(m["id"], kind="stable")
        assert np.array_equal(m["id"][oo], live)
        assert np.abs(m["pos"][oo] - om["pos"][live]).max() <= 1e-12 * 1000, "iteration %d: positions" % it
        print("iteration %d: %s" % (it, total), flush=True)
    assert total["bimol"] > 20 and total["volsurf"] > 20
    w.close()
    eng.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_crowded_invariants(gpu_available, precision):
    wl = _small(n_molecules=400000, iterations=30)
    c0 = np.bincount(wl.mol_species, minlength=5).astype(np.int64)
    c0[3] = len(wl.surface_mols["ids"])
    eng = DiffuseReactEngine.from_workload(wl, precision=precision, seed=4)
    st = eng.step(0, wl.iterations)
    c = eng.count_species(1, 5)[0].astype(np.int64)
    print("%s: A %d B %d E %d S %d AS %d, %d reactions" % (precision, *c, st.reactions), flush=True)
    assert c[0] + c[2] + c[4] == c0[0] + c0[2]
    assert c[1] + c[2] == c0[1] + c0[2]
    assert c[3] + c[4] == c0[3]
    assert c[4] > 0 and st.reactions > 100
    assert eng.num_molecules() == c[0] + c[1] + c[2]
    eng.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_drain_forms_give_the_same_run(gpu_available, precision):
    """The drain's two forms (exact_disk in place with a small vertex pool, or
    every step whose disk meets walls deferred to a pass of its own) and the
    engine's timed choice between them give the same run, bit for bit:
    molecules, surface tiles and counts after 12 full iterations."""
    wl = _small(n_molecules=300000, iterations=12)
    runs = {}
    for drain in ("inline", "defer", "auto"):
        eng = DiffuseReactEngine.from_workload(wl, precision=precision, seed=6, drain=drain)
        st = eng.step(0, wl.iterations)
        if drain != "auto":
            assert st.drain_form == (1 if drain == "inline" else 2)
        m = eng.molecules()
        o = np.argsort(m["id"], kind="stable")
        sm = eng.surface_molecules()
        runs[drain] = (m["id"][o], m["species"][o], m["pos"][o], {k: np.asarray(v) for k, v in sm.items()},
                       eng.count_species(1, 5), st.reactions, st.wall_hits)
        print("%s %s: %d reactions, %d reflections, form %d" % (precision, drain, st.reactions, st.wall_hits,
                                                                st.drain_form), flush=True)
        eng.close()
    a = runs["inline"]
    assert a[5] > 100
    for other in ("defer", "auto"):
        b = runs[other]
        for k in (0, 1, 2, 4):
            assert np.array_equal(a[k], b[k]), (other, k)
        assert a[3].keys() == b[3].keys()
        for k in a[3]:
            assert np.array_equal(a[3][k], b[3][k]), (other, k)
        assert a[5:] == b[5:]
