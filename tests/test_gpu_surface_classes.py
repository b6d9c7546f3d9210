"""GPU parity of surface classes (transparent and absorptive walls) against
the CPU oracle with the same per-molecule Philox draws: the sphere-in-cube
mesh with class 1 on the sphere acting on species T only."""
import dataclasses

import numpy as np
import pytest

from mcell_amd import model as M
from mcell_amd import workloads as W
from mcell_amd.engine import DiffuseReactEngine

pytestmark = pytest.mark.gpu


def _by_id(eng):
    m = eng.molecules()
    o = np.argsort(m["id"], kind="stable")
    return {k: v[o] for k, v in m.items()}


@pytest.mark.parametrize("action", [1, 2])
def test_surface_class_fp64_matches_oracle(gpu_available, oracle_lib, action):
    from oracle import pyoracle as po
    seed = 41
    wl = W.sphere_in_cube(n_l=3000, n_t=3000, iterations=30, t_action=action)
    eng = DiffuseReactEngine.from_workload(wl, seed=seed)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
    for it in range(wl.iterations):
        eng.step(it, 1)
        w.run(1)
        g = _by_id(eng)
        om = w.molecules()
        live = (om["flags"] & 1) == 0
        assert np.array_equal(g["id"], np.nonzero(live)[0]), "iteration %d: the same molecules survive" % it
        d = np.abs(g["pos"] - om["pos"][live]).max()
        assert d <= 1e-12 * 100, "iteration %d: max |dpos| %g" % (it, d)
    ws = w.wall_stats()
    print("action %d: oracle transparent crossings %d, absorbed %d; engine live %d"
          % (action, ws["transparent_crossings"], ws["absorbed"], eng.num_molecules()), flush=True)
    assert ws["transparent_crossings"] + ws["absorbed"] > 50
    w.close()
    eng.close()


def test_transparent_fp32_keeps_the_fp64_wall_sequence(gpu_available):
    wl = W.sphere_in_cube(n_l=3000, n_t=3000, iterations=20, t_action=1)
    pos = {}
    for prec in ("fp64", "fp32"):
        eng = DiffuseReactEngine.from_workload(wl, precision=prec, seed=5)
        eng.step(0, wl.iterations)
        pos[prec] = _by_id(eng)["pos"]
        eng.close()
    rms = np.sqrt(np.mean((pos["fp64"] - pos["fp32"]) ** 2))
    r64 = np.linalg.norm(pos["fp64"], axis=1)
    r32 = np.linalg.norm(pos["fp32"], axis=1)
    # the same side of the sphere (well away from it) for every molecule
    far = np.abs(r64 - 50) > 1.0
    print("fp32 vs fp64 after %d iterations: position RMS %.3e lu" % (wl.iterations, rms), flush=True)
    assert rms < 1e-3
    assert np.array_equal(r64[far] < 50, r32[far] < 50)


def test_reactions_through_transparent_sphere_snapshot_parity(gpu_available, oracle_lib):
    """Reaction proposals of L + T (T crosses the sphere, L reflects) equal
    the oracle's in the snapshot mode."""
    from oracle import pyoracle as po
    seed = 19
    wl = W.sphere_in_cube(n_l=20000, n_t=20000, iterations=5, t_action=1, subpartition_dimension=0.1)
    part = M.derive_partition(wl.config, M.bounding_box(wl.objects), has_bimol_vol_rxns=True)
    wl = dataclasses.replace(wl, part=part, rxns=[M.BimolReaction(0, 1, 1.36e9, [])])
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
        assert np.array_equal(g["mover_id"], u[o, 1]) and np.array_equal(g["partner_id"], u[o, 2])
        total += len(g)
        s = _by_id(eng)
        assert np.abs(s["pos"] - w.molecules()["pos"]).max() <= 1e-12 * 100
    print("%d proposals" % total, flush=True)
    assert total > 20
    w.close()
    eng.close()
