"""GPU parity of volume-surface reactions (L' + R' -> LR' on the receptor
sphere) against the CPU oracle with the same per-molecule Philox draws.

A ligand reacts at most once and a receptor at most once per iteration; two
ligands on one tile settle as the oracle's molecule order does (the smaller
id first, the other finds the tile taken and reflects), so the whole run is
compared exactly: the surviving ligands and their positions, and which
receptors became LR with which ids.
"""
import numpy as np
import pytest

from mcell_amd import workloads as W
from mcell_amd.engine import DiffuseReactEngine

pytestmark = pytest.mark.gpu


def _oracle_surface(w, wl):
    """(id, species, wall, tile) of the oracle's live surface molecules."""
    om = w.molecules()
    walls = np.unique(wl.surface_mols["walls"])
    rows = []
    for wall in walls:
        t = w.tiles(int(wall))
        occ = np.nonzero(t != 0xFFFFFFFF)[0]
        for k in occ:
            rows.append((int(t[k]), int(om["species"][t[k]]), int(wall), int(k)))
    rows.sort()
    return np.array(rows, np.int64).reshape(-1, 4)


def _engine_surface(eng):
    s = eng.surface_molecules()
    rows = np.stack([s["id"], s["species"], s["wall"], s["tile"]], axis=1).astype(np.int64)
    return rows[np.argsort(rows[:, 0], kind="stable")]


@pytest.mark.parametrize("rate", [2e8, 4e8])
def test_volsurf_snapshot_proposals_match_oracle(gpu_available, oracle_lib, rate):
    """Proposals (mover, receptor, time, collision point) without carrying
    them out; rate 2e8 keeps max_fixed_p below the per-tile scaling, 4e8
    above it (the two branches of test_bimolecular)."""
    from oracle import pyoracle as po
    seed = 29
    wl = W.receptor_sphere(n_l=150000, n_r=30000, iterations=4, rate=rate)
    eng = DiffuseReactEngine.from_workload(wl, seed=seed, propose_only=True, log_capacity=1 << 16)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
    w.set_snapshot(True)
    w.set_barrier_interval(wl.barrier_interval)
    total = 0
    for it in range(wl.iterations):
        eng.step(it, 1)
        w.run(1)
        g = eng.rxn_log()
        u, rx, d = w.rxn_log()
        w.lib.orc_clear_rxn_log(w.w)
        g = g[np.argsort(g["mover_id"], kind="stable")]
        o = np.argsort(u[:, 1], kind="stable")
        assert np.array_equal(g["mover_id"], u[o, 1]), "iteration %d: movers" % it
        assert np.array_equal(g["partner_id"], u[o, 2]), "iteration %d: receptors" % it
        assert np.all(g["rxn"] == 0x40000000) and np.array_equal(g["rxn"], rx[o].astype(np.uint32))
        assert np.abs(g["time"] - d[o, 0]).max(initial=0) <= 1e-12
        assert np.abs(g["pos"] - d[o, 1:4]).max(initial=0) <= 1e-9
        total += len(g)
    print("rate %.0e: %d proposals" % (rate, total), flush=True)
    assert total > 50
    w.close()
    eng.close()


def test_volsurf_full_run_matches_oracle(gpu_available, oracle_lib):
    from oracle import pyoracle as po
    seed = 31
    wl = W.receptor_sphere(n_l=60000, n_r=30000, iterations=12, rate=4e8)
    eng = DiffuseReactEngine.from_workload(wl, seed=seed)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
    w.set_barrier_interval(wl.barrier_interval)
    for it in range(wl.iterations):
        eng.step(it, 1)
        w.run(1)
        g = eng.molecules()
        o = np.argsort(g["id"], kind="stable")
        om = w.molecules()
        live = np.nonzero(((om["flags"] & 1) == 0) & (om["species"] == 0))[0]
        assert np.array_equal(g["id"][o], live), "iteration %d: the same ligands" % it
        d = np.abs(g["pos"][o] - om["pos"][live]).max(initial=0)
        assert d <= 1e-12 * 100, "iteration %d: max |dpos| %g" % (it, d)
        assert np.array_equal(_engine_surface(eng), _oracle_surface(w, wl)), "iteration %d: surface" % it
    n_rx = w.volsurf_reactions()
    c = eng.count_species(1, 3)[0].astype(np.int64)
    print("%d reactions; counts L %d R %d LR %d" % (n_rx, c[0], c[1], c[2]), flush=True)
    assert n_rx > 100 and c[2] == n_rx and c[0] == wl.num_molecules - n_rx
    w.close()
    eng.close()


def test_volume_surface_case_matches_oracle(gpu_available, oracle_lib):
    """The volume-surface benchmark case itself (100,000 L outside the
    5,120-triangle sphere, 20,000 R' on it, 10,000 T crossing it through a
    transparent surface class, L' + R' -> LR' at 1e8 /M/s): 20 iterations
    of the whole run equal the oracle's, ligand for ligand and tile for
    tile, and so do the crossings of T."""
    from oracle import pyoracle as po
    seed = 37
    wl = W.volume_surface(iterations=20)
    eng = DiffuseReactEngine.from_workload(wl, seed=seed)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
    w.set_barrier_interval(wl.barrier_interval)
    crossings = 0
    for it in range(wl.iterations):
        st = eng.step(it, 1)
        crossings += st.transparent_crossings
        w.run(1)
        g = eng.molecules()
        o = np.argsort(g["id"], kind="stable")
        om = w.molecules()
        live = np.nonzero(((om["flags"] & 1) == 0) & ((om["species"] == 0) | (om["species"] == 3)))[0]
        assert np.array_equal(g["id"][o], live), "iteration %d: the same volume molecules" % it
        assert np.array_equal(g["species"][o], om["species"][live])
        d = np.abs(g["pos"][o] - om["pos"][live]).max(initial=0)
        assert d <= 1e-12 * 100, "iteration %d: max |dpos| %g" % (it, d)
        assert np.array_equal(_engine_surface(eng), _oracle_surface(w, wl)), "iteration %d: surface" % it
    n_rx = w.volsurf_reactions()
    print("%d reactions, %d transparent crossings (oracle %d)" % (n_rx, crossings,
                                                                  w.wall_stats()["transparent_crossings"]), flush=True)
    assert n_rx > 20 and crossings == w.wall_stats()["transparent_crossings"] > 100
    w.close()
    eng.close()


def test_volume_surface_case_fp32_takes_the_fp64_wall_hits(gpu_available):
    """The volume-surface case's molecules and geometry (transparent sphere
    for T) without its reaction classes: over 20 iterations the FP32+FP64-
    fallback form hits exactly the walls the FP64 form hits, molecule-step by
    molecule-step (the hit-index disagreement rate the benchmark reports)."""
    import bench
    r = bench.fp32_hit_agreement(W.volume_surface(iterations=20))
    print(r, flush=True)
    assert r["steps_with_hits"] > 10000
    assert r["hit_index_disagreements"] == 0
    assert r["position_rms_lu"] < 1e-3


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_volsurf_conservation_and_orientation(gpu_available, precision):
    wl = W.receptor_sphere(n_l=200000, n_r=40000, iterations=40, rate=4e8)
    sm = wl.surface_mols
    eng = DiffuseReactEngine.from_workload(wl, precision=precision, seed=3)
    eng.step(0, wl.iterations)
    c = eng.count_species(1, 3)[0].astype(np.int64)
    s = eng.surface_molecules()
    print("%s: L %d R %d LR %d" % (precision, c[0], c[1], c[2]), flush=True)
    assert c[2] > 1000
    assert c[0] + c[2] == wl.num_molecules and c[1] + c[2] == len(sm["ids"])
    assert eng.num_molecules() == c[0]
    # products sit on tiles of up receptors, oriented up; down receptors never react
    start = {(int(a), int(b)): int(o) for a, b, o in zip(sm["walls"], sm["tiles"], sm["orients"])}
    for sp, wall, tile, orient in zip(s["species"], s["wall"], s["tile"], s["orient"]):
        if sp == 2:
            assert start[(int(wall), int(tile))] == 1 and orient == 1
        else:
            assert orient == start[(int(wall), int(tile))]
    assert len(s["id"]) == len(sm["ids"]) and len(np.unique(s["id"])) == len(s["id"])
    eng.close()


@pytest.mark.parametrize("case", ["enzyme", "release"])
def test_volume_products_match_oracle(gpu_available, oracle_lib, case):
    """Volume products of volume-surface reactions, whole runs against the
    oracle: 'enzyme' L' + R' -> R' + P (the receptor stays, P's orientation
    drawn per reaction) with P + R' -> RP' (a new P does not rebind to its
    creation tile on its first hit there); 'release' L' + R' -> LR' + P, (P
    to the back side, LR' takes the tile).  Every volume molecule (L and P)
    and every tile equal the oracle's each iteration."""
    from oracle import pyoracle as po
    from mcell_amd import model as M
    seed = 41
    if case == "enzyme":
        wl = W.receptor_sphere(n_l=60000, n_r=30000, iterations=12, rate=4e8,
                               products=[(1, 1, True), (3, 0, False)], more_species=[("P", 1e-6), ("RP", 0.0)],
                               more_volsurf=[M.VolSurfReaction(3, 1, 4e8, 4, product_orient=1, geom_vol=0,
                                                               geom_surf=1)])
    else:
        wl = W.receptor_sphere(n_l=60000, n_r=30000, iterations=12, rate=4e8,
                               products=[(2, 1, True), (3, -1, False)], more_species=[("P", 1e-6)])
    eng = DiffuseReactEngine.from_workload(wl, seed=seed)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
    w.set_barrier_interval(wl.barrier_interval)
    for it in range(wl.iterations):
        eng.step(it, 1)
        w.run(1)
        g = eng.molecules()
        o = np.argsort(g["id"], kind="stable")
        om = w.molecules()
        live = np.nonzero(((om["flags"] & 1) == 0) & ((om["species"] == 0) | (om["species"] == 3)))[0]
        assert np.array_equal(g["id"][o], live), "iteration %d: the same volume molecules" % it
        assert np.array_equal(g["species"][o], om["species"][live])
        d = np.abs(g["pos"][o] - om["pos"][live]).max(initial=0)
        assert d <= 1e-12 * 100, "iteration %d: max |dpos| %g" % (it, d)
        assert np.array_equal(_engine_surface(eng), _oracle_surface(w, wl)), "iteration %d: surface" % it
    n_rx = w.volsurf_reactions()
    n_p = int(np.sum(om["species"][live] == 3))
    print("%s: %d reactions, %d P alive" % (case, n_rx, n_p), flush=True)
    assert n_rx > 100 and n_p > 50
    w.close()
    eng.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_counted_volumes(gpu_available, oracle_lib, precision):
    """The sphere as a counted volume (1 inside, 0 outside), T crossing it
    through a transparent class and L binding R' on it: every molecule's
    index equals the oracle's (FP64 form) and says on which side of the
    sphere it lies (both forms), and the per-volume counts are those of the
    molecules."""
    from oracle import pyoracle as po
    seed = 43
    wl = W.receptor_sphere(n_l=20000, n_r=20000, iterations=30, rate=4e8, n_t=40000, counted=True,
                           products=[(1, 1, True), (4, 0, False)], more_species=[("P", 1e-6)])
    eng = DiffuseReactEngine.from_workload(wl, seed=seed, precision=precision)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
    w.set_barrier_interval(wl.barrier_interval)
    nb = len(wl.objects[0].wall_list)
    for chunk in range(0, wl.iterations, 10):
        eng.step(chunk, 10)
        g = eng.molecules()
        o = np.argsort(g["id"], kind="stable")
        truth = W.inside_convex(g["pos"][o], wl.vertices, wl.walls[nb:]).astype(np.uint32)
        assert np.array_equal(g["counted_volume"][o], truth), "iteration %d: counted volumes" % (chunk + 10)
        if precision == "fp64":
            w.run(10)
            om = w.molecules()
            live = np.nonzero(((om["flags"] & 1) == 0) & ((om["species"] == 0) | (om["species"] >= 3)))[0]
            assert np.array_equal(g["id"][o], live)
            assert np.array_equal(g["counted_volume"][o], w.molecule_volumes()[live])
        c = eng.count_species_by_volume(1, 2, 5)[0].astype(np.int64)
        ref = np.zeros((2, 5), np.int64)
        np.add.at(ref, (g["counted_volume"], g["species"]), 1)
        assert np.array_equal(c, ref)
        print("%s iteration %d: T inside %d outside %d, P inside %d outside %d"
              % (precision, chunk + 10, c[1, 3], c[0, 3], c[1, 4], c[0, 4]), flush=True)
    assert c[1, 3] > 1000 and c[0, 4] + c[1, 4] > 50
    # every reaction happened outside the sphere (the ligands start there) and is counted there
    rc = eng.rxn_counts()
    n_vs = int(rc[0, 0, rc.shape[2] - 4:].sum())
    assert n_vs == int(rc.sum()) > 50 and int(rc[0, 1].sum()) == 0
    if precision == "fp64":
        assert n_vs == w.volsurf_reactions()
    w.close()
    eng.close()


@pytest.mark.parametrize("side", [1, -1])
def test_kept_ligand_matches_oracle(gpu_available, oracle_lib, side):
    """L' + R' -> L? + LR' (the ligand stays: outcome_products_random,
    diffuse_react_event.cpp:2690-2720): with L' it goes on outside, with L,
    it passes through the sphere (RX_FLIP, :940-971), its step repeated
    through the reaction.  Every L and every tile equal the oracle's each
    iteration, and the L inside the sphere are as many as the LR."""
    from oracle import pyoracle as po
    seed = 43
    wl = W.receptor_sphere(n_l=60000, n_r=30000, iterations=12, rate=4e8, down_fraction=0.0,
                           products=[(0, side, False), (2, 1, True)])
    eng = DiffuseReactEngine.from_workload(wl, seed=seed)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
    w.set_barrier_interval(wl.barrier_interval)
    for it in range(wl.iterations):
        eng.step(it, 1)
        w.run(1)
        g = eng.molecules()
        o = np.argsort(g["id"], kind="stable")
        om = w.molecules()
        live = np.nonzero(((om["flags"] & 1) == 0) & (om["species"] == 0))[0]
        assert np.array_equal(g["id"][o], live), "iteration %d: the same volume molecules" % it
        d = np.abs(g["pos"][o] - om["pos"][live]).max(initial=0)
        assert d <= 1e-12 * 100, "iteration %d: max |dpos| %g" % (it, d)
        assert np.array_equal(_engine_surface(eng), _oracle_surface(w, wl)), "iteration %d: surface" % it
    c = eng.count_species(1, 3)[0]
    inside = int(np.sum(W.inside_convex(g["pos"], wl.vertices, wl.walls[12:])))
    print("side %d: L %d (inside %d) R %d LR %d" % (side, c[0], inside, c[1], c[2]), flush=True)
    assert c[0] == 60000 and c[2] > 100 and inside == (0 if side == 1 else c[2])
    w.close()
    eng.close()
