"""GPU tests of the drop-in boundary (include/mcg.h): the scheduler's barrier
time, the molecule state a barrier round trip carries, the statistics, the
clip at the partition face, and what a failed step leaves behind.

The reference contract (src4/diffuse_react_event.h:118-150,
src4/scheduler.cpp:294-330): before every step the scheduler hands the event
the time up to the next barrier, and every molecule's step is clipped there.
"""
import ctypes

import numpy as np
import pytest

from mcell_amd import _lib
from mcell_amd import workloads as W
from mcell_amd.engine import DiffuseReactEngine
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _close(a, b, rel=1e-12):
    """FP64 values within 1e-12 relative (of max(|b|, 1)): the precision
    bar for positions and times.  Draw-derived values agree bit for bit
    except where the device's log (unimolecular lifetimes) rounds
    differently from the host's in the last bit."""
    return bool(np.all(np.abs(a - b) <= rel * np.maximum(np.abs(b), 1.0)))


def _by_id(m):
    o = np.argsort(m["id"], kind="stable")
    return {k: v[o] for k, v in m.items()}


def _oracle_live(w):
    om = w.molecules()
    live = np.nonzero((om["flags"] & 1) == 0)[0]
    return live, om


def test_partition_face_clip_matches_oracle(gpu_available):
    """Cube faces 2 units inside the partition faces: thousands of moves end
    outside the partition and are clipped at its face before the wall search
    (collision_utils.inl:48-114).  FP64 positions and every wall index equal
    the oracle's; the reflections the engine counts equal the oracle's."""
    wl = W.box_at_partition_faces(n=20000, iterations=20)
    n = wl.num_molecules
    eng = DiffuseReactEngine(wl.part, precision="fp64", seed=11, capacity=n, trace=True)
    eng.set_geometry(wl.vertices, wl.walls, wl.wall_class)
    eng.set_species(wl.species_steps)
    eng.add_volume_molecules(0, wl.ids, wl.mol_species, wl.positions)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=11)
    w.set_trace(True)
    reflections = 0
    for it in range(wl.iterations):
        st = eng.step(it, 1)
        w.clear_hits()
        w.run(1)
        reflections += st.wall_hits
        g = eng.molecules()
        o = w.molecules()
        assert np.array_equal(g["id"], np.arange(n))
        assert _close(g["pos"], o["pos"]), "iteration %d: positions" % it
        walls, counts = eng.hit_trace()
        u, _ = w.hits()
        seqs = [[] for _ in range(n)]
        for rec in u:
            seqs[rec[0]].append(int(rec[2]))
        for k in np.nonzero(counts)[0]:
            assert walls[k, :counts[k]].tolist() == seqs[k], "iteration %d molecule %d" % (it, k)
        assert int(counts.sum()) == sum(len(s) for s in seqs)
    stats = w.stats()
    print("partition clips %d, reflections %d (engine %d)" % (stats["partition_clips"],
                                                                stats["mol_wall_reflections"], reflections))
    assert stats["partition_clips"] > 1000  # the clip path was exercised
    assert reflections == stats["mol_wall_reflections"]
    half = 0.5 / wl.part.length_unit
    assert np.all(np.abs(eng.molecules()["pos"]) < half)
    eng.close()
    w.close()


@pytest.mark.parametrize("barriers", [[7, 10, 13, 31, 45], [1, 2, 3, 50]])
def test_irregular_barriers_match_oracle(gpu_available, barriers):
    """E -> A + B with products created inside iterations: a product's first
    step runs up to the next barrier, so where the barriers fall changes the
    trajectories.  The engine stepped by the scheduler's barrier times equals
    the oracle run with the same barrier list, molecule for molecule, and
    differs from the run with a barrier every 10 iterations."""
    seed = 31
    wl = W.decay_box(n_e=3000, iterations=60, rate=5e3)
    eng = DiffuseReactEngine.from_workload(wl, seed=seed, capacity=20000, barriers=barriers)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
    w.set_barriers(barriers)
    for chunk in range(0, wl.iterations, 12):
        eng.step(chunk, 12)
        w.run(12)
        g = _by_id(eng.molecules())
        live, om = _oracle_live(w)
        assert np.array_equal(g["id"], live), "iteration %d: the same molecules" % (chunk + 12)
        assert np.array_equal(g["species"], om["species"][live])
        assert _close(g["pos"], om["pos"][live]), "iteration %d: positions" % (chunk + 12)
        assert _close(g["diffusion_time"], om["diffusion_time"][live]), "iteration %d: diffusion times" % (chunk + 12)
    ref = DiffuseReactEngine.from_workload(wl, seed=seed, capacity=20000)  # a barrier every 10 iterations
    ref.step(0, wl.iterations)
    r = _by_id(ref.molecules())
    assert len(r["id"]) != len(g["id"]) or not np.array_equal(r["pos"], g["pos"])
    assert w.unimol_reactions() > 500
    ref.close()
    eng.close()
    w.close()


def _round_trip(wl, seed, first, second, capacity, precision="fp64"):
    """Runs first + second iterations in one engine, and first iterations,
    a download, an upload into a new engine and second more; returns both
    final states by id."""
    a = DiffuseReactEngine.from_workload(wl, seed=seed, capacity=capacity, precision=precision)
    a.step(0, first + second)
    whole = _by_id(a.molecules())
    b = DiffuseReactEngine.from_workload(wl, seed=seed, capacity=capacity, precision=precision)
    b.step(0, first)
    m = b.molecules()
    nid = b.next_molecule_ids()
    b.close()
    c = DiffuseReactEngine.from_workload(wl, seed=seed, capacity=capacity, precision=precision,
                                         add_molecules=False, start_iteration=first)
    for r in range(wl.num_replicas):
        sel = m["replica"] == r
        c.add_volume_molecules(r, m["id"][sel], m["species"][sel], m["pos"][sel], m["diffusion_time"][sel],
                               m["unimol_time"][sel], m["flags"][sel], birthdays=m["birthday"][sel])
    c.set_next_molecule_ids(nid)
    c.step(first, second)
    cont = _by_id(c.molecules())
    a.close()
    c.close()
    return whole, cont, m


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_round_trip_decay_continues_bit_identically(gpu_available, precision):
    """E -> A + B (lifetimes drawn, products with diffusion-time offsets
    inside barriers): downloading at iteration 23 and uploading into a fresh
    context continues the run bit for bit."""
    wl = W.decay_box(n_e=4000, iterations=50, rate=5e3)
    whole, cont, m = _round_trip(wl, 41, 23, 27, 30000, precision)
    assert np.any(m["diffusion_time"] != 23.0)        # offsets were carried
    assert np.any(m["unimol_time"] != _lib.TIME_INVALID)  # drawn lifetimes were carried
    for k in ("id", "species", "replica", "flags"):
        assert np.array_equal(whole[k], cont[k]), k
    for k in ("pos", "diffusion_time", "unimol_time"):
        assert np.array_equal(whole[k].view(np.uint64), cont[k].view(np.uint64)), k


def test_round_trip_bimol_continues_bit_identically(gpu_available):
    """A + B -> C with claims and products: the molecule ids the next
    products get travel with the state (mcg_get/set_next_molecule_ids)."""
    wl = W.bimol_box(n_a=3000, n_b=3000, iterations=40, seed=5, rate=1.36e9)
    whole, cont, m = _round_trip(wl, 43, 17, 23, 20000)
    assert int(m["id"].max()) >= 6000  # products exist
    for k in ("id", "species", "flags"):
        assert np.array_equal(whole[k], cont[k]), k
    for k in ("pos", "diffusion_time"):
        assert np.array_equal(whole[k].view(np.uint64), cont[k].view(np.uint64)), k


def test_step_contract_is_checked(gpu_available):
    wl = W.free_diffusion_box(scale=0.01)
    eng = DiffuseReactEngine.from_workload(wl)
    with pytest.raises(_lib.McgError) as e:
        eng.step_raw(0, 3, 2.0)  # past the barrier
    assert e.value.status == 1
    with pytest.raises(_lib.McgError) as e:
        eng.step_raw(0, 1, 1.5)  # not a whole number
    assert e.value.status == 1
    with pytest.raises(_lib.McgError) as e:
        eng.step_raw(5, 1, 1.0)  # not the engine's clock
    assert e.value.status == 1
    eng.step_raw(0, 2, 2.0)
    eng.step_raw(2, 1, 100.0)
    assert eng.molecules()["diffusion_time"][0] == 3.0
    eng.close()


def test_capacity_overflow_fails_the_context(gpu_available):
    """E -> A + B doubles molecules: a capacity barely above the start
    overflows.  The step reports MCG_E_CAPACITY and every later call on the
    context reports MCG_E_STATE instead of touching slots past the buffers."""
    wl = W.decay_box(n_e=1000, iterations=200, rate=5e4)
    eng = DiffuseReactEngine.from_workload(wl, seed=3, capacity=1100)
    with pytest.raises(_lib.McgError) as e:
        eng.step(0, 200)
    assert e.value.status == _lib_status("MCG_E_CAPACITY")
    for call in (lambda: eng.num_molecules(), lambda: eng.molecules(), lambda: eng.step(0, 1),
                 lambda: eng.count_species(1, 3), lambda: eng.next_molecule_ids()):
        with pytest.raises(_lib.McgError) as e:
            call()
        assert e.value.status == _lib_status("MCG_E_STATE")
        assert "earlier step" in str(e.value)
    eng.close()


def _lib_status(name):
    return {"MCG_E_CAPACITY": 3, "MCG_E_STATE": 8}[name]


def test_transparent_crossings_are_counted(gpu_available):
    """T crossing the transparent sphere: the engine's crossing count equals
    the oracle's (FP64, same draws)."""
    wl = W.sphere_in_cube(n_l=0, n_t=20000, iterations=30)
    eng = DiffuseReactEngine.from_workload(wl, seed=wl.config.seed)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX)
    st = eng.step(0, wl.iterations)
    w.run(wl.iterations)
    crossings = w.wall_stats()["transparent_crossings"]
    print("engine crossings %d (oracle %d), reflections %d" % (st.transparent_crossings, crossings, st.wall_hits))
    assert st.transparent_crossings > 100
    assert st.transparent_crossings == crossings
    assert st.wall_hits == w.stats()["mol_wall_reflections"]
    eng.close()
    w.close()
