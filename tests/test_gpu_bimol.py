"""GPU parity of the volume-volume reaction path against the CPU oracle.

* Snapshot mode (MCG_FLAG_PROPOSE_ONLY vs the oracle's snapshot mode, same
  Philox draws): every molecule moves against the iteration-start positions
  and successful reaction tests are logged.  Proposals (mover, partner,
  reaction, pathway) are compared bit-exactly, collision times and points
  within 1e-12 relative, and every molecule's end position bit-exactly.
* Full mode: conservation (A + C and B + C per replica), unique ids,
  containment, determinism, and the species counts over time against the
  oracle's 32-seed ensembles (tests/golden/bimol_ensemble.npz, 2,000 +
  2,000 molecules, and bimol_ensemble_full.npz at the benchmark case's full
  size): every sample within 3 sigma.
"""
import os

import numpy as np
import pytest

from mcell_amd import workloads as W
from mcell_amd.engine import DiffuseReactEngine

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _oracle_snapshot_world(wl, seed):
    from oracle import pyoracle as po
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed, replica=0)
    w.set_snapshot(True)
    w.set_barrier_interval(wl.barrier_interval)
    return w


def _gpu_state_by_id(eng):
    m = eng.molecules()
    order = np.argsort(m["id"], kind="stable")
    return {k: v[order] for k, v in m.items()}


def _near_walls(wl, depth, seed):
    """Moves every molecule to within `depth` internal units of a cube face
    (the collisions then need exact_disk)."""
    rng = np.random.default_rng(seed)
    n = len(wl.ids)
    pos = rng.uniform(-50 + depth, 50 - depth, size=(n, 3))
    axis = rng.integers(0, 3, n)
    side = rng.choice([-1.0, 1.0], n)
    pos[np.arange(n), axis] = side * (50 - rng.uniform(0, depth, n))
    wl.positions = pos
    return wl


@pytest.mark.parametrize("rate,near", [(1e8, False), (1.36e9, False), (1.36e9, True)])
def test_snapshot_proposals_and_positions_match_oracle(gpu_available, oracle_lib, rate, near):
    seed = 77
    wl = W.bimol_box(n_a=3000, n_b=3000, iterations=12, seed=5, rate=rate)
    if near:
        wl = _near_walls(wl, 1.0, 8)
    eng = DiffuseReactEngine.from_workload(wl, seed=seed, propose_only=True, log_capacity=1 << 16)
    w = _oracle_snapshot_world(wl, seed)
    total = 0
    for it in range(wl.iterations):
        eng.step(it, 1)
        g = eng.rxn_log()
        w.run(1)
        u, rx, d = w.rxn_log()
        w.lib.orc_clear_rxn_log(w.w)
        g = g[np.argsort(g["mover_id"], kind="stable")]
        o = np.argsort(u[:, 1], kind="stable")
        u, rx, d = u[o], rx[o], d[o]
        print("iteration %d: %d proposals (oracle %d)" % (it, len(g), len(u)), flush=True)
        assert len(g) == len(u)
        assert np.array_equal(g["mover_id"], u[:, 1])
        assert np.array_equal(g["partner_id"], u[:, 2])
        assert np.array_equal(g["pathway"], u[:, 3])
        assert np.array_equal(g["rxn"], rx.astype(np.uint32))
        assert np.all(g["iteration"] == it)
        np.testing.assert_allclose(g["time"], d[:, 0], rtol=1e-12, atol=0)
        np.testing.assert_allclose(g["pos"], d[:, 1:4], rtol=1e-12, atol=1e-12)
        total += len(g)
        # every molecule where the oracle has it
        s = _gpu_state_by_id(eng)
        om = w.molecules()
        assert np.array_equal(s["id"], np.arange(len(om["species"])))
        diff = np.abs(s["pos"] - om["pos"])
        assert diff.max() <= 1e-12 * np.abs(om["pos"]).max(), "iteration %d: max |dpos| %g" % (it, diff.max())
        assert np.array_equal(s["subpart"], om["subpart"])
    assert total > (5 if rate < 1e9 else 200)
    w.close()
    eng.close()


def _counts(eng, nrep):
    return eng.count_species(nrep, 3).astype(np.int64)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_full_mode_conservation_ids_and_containment(gpu_available, precision):
    nrep = 4
    wl = W.bimol_box(n_a=3000, n_b=2000, iterations=60, seed=11, num_replicas=nrep, rate=1.36e9)
    eng = DiffuseReactEngine.from_workload(wl, precision=precision, seed=5, log_capacity=1 << 18)
    c0 = _counts(eng, nrep)
    total_rx = 0
    for it in range(0, wl.iterations, 10):
        st = eng.step(it, 10)
        total_rx += st.reactions
        c = _counts(eng, nrep)
        assert np.array_equal(c[:, 0] + c[:, 2], c0[:, 0]), "A + C conserved"
        assert np.array_equal(c[:, 1] + c[:, 2], c0[:, 1]), "B + C conserved"
    c = _counts(eng, nrep)
    assert c[:, 2].sum() == total_rx > 100
    m = eng.molecules()
    assert len(m["id"]) == c.sum()
    for r in range(nrep):
        ids = m["id"][m["replica"] == r]
        assert len(np.unique(ids)) == len(ids), "ids unique within a replica"
        prod = ids[m["species"][m["replica"] == r] == 2]
        assert prod.min() >= 5000 and prod.max() < 5000 + c[r, 2], "product ids follow the initial ids"
    half = 50.0  # the 1 um cube is [-50, 50]^3 internal units
    assert np.all(np.abs(m["pos"]) < half + 1e-9)
    print("%s: %d reactions over %d replicas" % (precision, total_rx, nrep), flush=True)
    eng.close()


def test_full_mode_is_deterministic(gpu_available):
    wl = W.bimol_box(n_a=2000, n_b=2000, iterations=30, seed=3, num_replicas=2, rate=1.36e9)
    states = []
    for _ in range(2):
        eng = DiffuseReactEngine.from_workload(wl, seed=9, log_capacity=1 << 16)
        eng.step(0, wl.iterations)
        m = eng.molecules()
        key = np.lexsort((m["id"], m["replica"]))
        states.append({k: v[key] for k, v in m.items()})
        eng.close()
    for k in states[0]:
        assert np.array_equal(states[0][k], states[1][k]), k


def test_species_counts_within_3_sigma_of_cpu_ensemble(gpu_available):
    import importlib.util
    spec = importlib.util.spec_from_file_location("mk", os.path.join(HERE, "golden", "make_bimol_ensemble.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    ref = np.load(os.path.join(HERE, "golden", "bimol_ensemble.npz"))
    cpu = ref["counts"].astype(np.float64)  # [32, 101, 3]
    nrep = 128
    wl = mk.workload(num_replicas=nrep)
    eng = DiffuseReactEngine.from_workload(wl, seed=2024)
    gpu = np.zeros((nrep, cpu.shape[1], 3))
    gpu[:, 0] = _counts(eng, nrep)
    for k in range(1, cpu.shape[1]):
        eng.step((k - 1) * mk.SAMPLE, mk.SAMPLE)
        gpu[:, k] = _counts(eng, nrep)
    eng.close()
    _assert_within_3_sigma(gpu, cpu)
    assert np.all(gpu[:, 0] == cpu[0, 0])


def _assert_within_3_sigma(gpu, cpu):
    """Ensemble means of A, B and C at every sample: the GPU's within 3
    standard errors of the CPU ensemble's (both ensembles' spread)."""
    sig = np.sqrt(cpu.var(axis=0, ddof=1) / cpu.shape[0] + gpu.var(axis=0, ddof=1) / gpu.shape[0])
    z = (gpu.mean(axis=0) - cpu.mean(axis=0)) / np.where(sig > 0, sig, 1)
    print("final A: gpu %.1f +- %.1f, cpu %.1f +- %.1f; |z| max %.2f over %d samples"
          % (gpu[:, -1, 0].mean(), gpu[:, -1, 0].std(), cpu[:, -1, 0].mean(), cpu[:, -1, 0].std(),
             np.abs(z).max(), z.shape[0]), flush=True)
    assert np.abs(z).max() < 3.0, "a sample outside 3 sigma: %s" % (np.argwhere(np.abs(z) >= 3.0)[:5].tolist(),)


def test_full_size_species_counts_within_3_sigma_of_cpu_ensemble(gpu_available):
    """The bimolecular benchmark case at full size (50,000 A + 50,000 B in
    the 1 um cube, 0.05 um subpartitions, 2,000 iterations, counts every 10):
    32 GPU replicas against the oracle's 32-run ensemble
    (tests/golden/bimol_ensemble_full.npz, written by
    make_bimol_ensemble.py --full), every one of the 201 samples within 3
    sigma."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("mk", os.path.join(HERE, "golden", "make_bimol_ensemble.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    mk.use_full()
    ref = np.load(os.path.join(HERE, "golden", "bimol_ensemble_full.npz"))
    cpu = ref["counts"].astype(np.float64)  # [32, 201, 3]
    nrep = cpu.shape[0]
    wl = mk.workload(num_replicas=nrep)
    eng = DiffuseReactEngine.from_workload(wl, seed=2025)
    gpu = np.zeros((nrep, cpu.shape[1], 3))
    gpu[:, 0] = _counts(eng, nrep)
    for k in range(1, cpu.shape[1]):
        eng.step((k - 1) * mk.SAMPLE, mk.SAMPLE)
        gpu[:, k] = _counts(eng, nrep)
        if k % 50 == 0:
            print("iteration %d: A %.1f (cpu %.1f)" % (k * mk.SAMPLE, gpu[:, k, 0].mean(), cpu[:, k, 0].mean()),
                  flush=True)
    eng.close()
    assert np.all(gpu[:, 0] == cpu[:, 0])  # the same placements
    _assert_within_3_sigma(gpu, cpu)


def test_fp32_snapshot_proposals_agree_with_fp64(gpu_available):
    wl = W.bimol_box(n_a=3000, n_b=3000, iterations=20, seed=21, rate=1.36e9)
    logs = {}
    for prec in ("fp64", "fp32"):
        eng = DiffuseReactEngine.from_workload(wl, precision=prec, seed=31, propose_only=True, log_capacity=1 << 16)
        eng.step(0, wl.iterations)
        g = eng.rxn_log()
        logs[prec] = set(zip(g["iteration"].tolist(), g["mover_id"].tolist(), g["partner_id"].tolist()))
        eng.close()
    a, b = logs["fp64"], logs["fp32"]
    agree = len(a & b) / max(len(a | b), 1)
    print("fp32 vs fp64 proposals: %d / %d agree (%.4f)" % (len(a & b), len(a | b), agree), flush=True)
    assert len(a) > 100 and agree >= 0.98


def test_replica_params_sphere_mesh_snapshot_parity(gpu_available, oracle_lib):
    """Replicas with their own diffusion constants and reaction probabilities
    over the shared sphere-in-cube mesh: each replica's proposals and
    positions equal the oracle's for that instance."""
    from oracle import pyoracle as po
    seed = 13
    wl = W.replicas_sphere_in_cube(num_replicas=3, iterations=6, n_min=12000, n_max=20000, rate=1.36e9,
                                   subpartition_dimension=0.1)
    eng = DiffuseReactEngine.from_workload(wl, seed=seed, propose_only=True, log_capacity=1 << 16)
    worlds = []
    for r in range(wl.num_replicas):
        w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed, replica=r)
        w.set_snapshot(True)
        w.set_barrier_interval(wl.barrier_interval)
        worlds.append(w)
    total = 0
    for it in range(wl.iterations):
        eng.step(it, 1)
        g = eng.rxn_log()
        m = eng.molecules()
        for r, w in enumerate(worlds):
            w.run(1)
            u, rx, d = w.rxn_log()
            w.lib.orc_clear_rxn_log(w.w)
            gr = g[g["replica"] == r]
            gr = gr[np.argsort(gr["mover_id"], kind="stable")]
            o = np.argsort(u[:, 1], kind="stable")
            assert np.array_equal(gr["mover_id"], u[o, 1]) and np.array_equal(gr["partner_id"], u[o, 2])
            np.testing.assert_allclose(gr["time"], d[o, 0], rtol=1e-12, atol=0)
            total += len(gr)
            sel = m["replica"] == r
            ids = m["id"][sel]
            pos = m["pos"][sel][np.argsort(ids)]
            om = w.molecules()
            assert np.abs(pos - om["pos"]).max() <= 1e-12 * 100, (it, r)
        print("iteration %d: %d proposals so far" % (it, total), flush=True)
    assert total > 50
    for w in worlds:
        w.close()
    eng.close()
