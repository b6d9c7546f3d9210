"""GPU parity of the free-diffusion path (reflective 1 um cube, 10,000 A).

The engine (libmcg.so, FP64 form) and the CPU oracle consume the same
per-molecule Philox streams, so every molecule must take the same wall hits
in the same order, end in the same subpartition, and land on positions equal
to within 1e-12 relative (the FP64 forms use the same operation order and no
FMA, so in practice they are bit-identical).
"""
import numpy as np
import pytest

from mcell_amd import workloads as W
from mcell_amd.engine import DiffuseReactEngine
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _engine(wl, precision="fp64", trace=True, seed=None):
    eng = DiffuseReactEngine(wl.part, precision=precision, seed=wl.config.seed if seed is None else seed,
                             num_replicas=wl.num_replicas, capacity=max(wl.num_molecules, 1), trace=trace)
    eng.set_geometry(wl.vertices, wl.walls, wl.wall_class)
    eng.set_species(wl.species_steps)
    for r in range(wl.num_replicas):
        sel = wl.replica == r
        eng.add_volume_molecules(r, wl.ids[sel], wl.mol_species[sel], wl.positions[sel])
    return eng


def test_wall_constants_and_subpart_lists_match_oracle(gpu_available):
    wl = W.free_diffusion_box(scale=0.01)
    eng = _engine(wl, trace=False)
    orc = po.world_from_workload(wl)
    consts, off, idx = eng.wall_data(len(wl.walls))
    assert np.array_equal(consts.view(np.uint64), orc.walls().view(np.uint64))
    ooff, oidx = orc.subpart_walls(wl.part.num_subparts_per_edge ** 3)
    assert np.array_equal(off, ooff)
    assert np.array_equal(idx, oidx)
    eng.close()


def _oracle_hits_by_molecule(orc, n):
    u, _ = orc.hits()
    seqs = [[] for _ in range(n)]
    for rec in u:
        seqs[rec[0]].append(int(rec[2]))
    return seqs


def test_step_by_step_parity_fp64(gpu_available):
    wl = W.free_diffusion_box()
    n = wl.num_molecules
    eng = _engine(wl)
    orc = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX)
    orc.set_trace(True)
    total_hits = 0
    for it in range(25):
        eng.step(it, 1)
        orc.clear_hits()
        orc.run(1)
        g = eng.molecules()
        o = orc.molecules()
        assert np.array_equal(g["id"], np.arange(n))
        rel = np.abs(g["pos"] - o["pos"]) / np.maximum(np.abs(o["pos"]), 1.0)
        assert rel.max() <= 1e-12, "iteration %d: max rel diff %g" % (it, rel.max())
        assert np.array_equal(g["subpart"], o["subpart"]), "iteration %d: subpartition mismatch" % it
        walls, counts = eng.hit_trace()
        seqs = _oracle_hits_by_molecule(orc, n)
        for k in range(n):
            assert counts[k] == len(seqs[k]), "iteration %d molecule %d: hit count" % (it, k)
            assert walls[k, :counts[k]].tolist() == seqs[k], "iteration %d molecule %d: wall indices" % (it, k)
        total_hits += int(counts.sum())
    assert total_hits > 100  # the cube walls were actually exercised
    eng.close()


def test_full_run_fp64_matches_oracle_and_stays_inside(gpu_available):
    wl = W.free_diffusion_box(iterations=1000)
    eng = _engine(wl, trace=False)
    orc = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX)
    eng.step(0, wl.iterations)
    orc.run(wl.iterations)
    g = eng.molecules()
    o = orc.molecules()
    rel = np.abs(g["pos"] - o["pos"]) / np.maximum(np.abs(o["pos"]), 1.0)
    assert rel.max() <= 1e-12
    assert np.array_equal(g["subpart"], o["subpart"])
    half = 0.5 * wl.part.rcp_length_unit
    assert np.all(np.abs(g["pos"]) < half)
    eng.close()


def test_msd_matches_6Dt(gpu_available):
    """Free diffusion away from the walls: MSD after 10 steps = 6 D t."""
    wl = W.free_diffusion_box(scale=10.0)   # 100,000 molecules for a tight estimate
    eng = _engine(wl, trace=False)
    p0 = wl.positions.copy()
    eng.step(0, 10)
    p = eng.molecules()["pos"]
    inner = np.all(np.abs(p0) < 50.0 - 15.0, axis=1)  # >= 0.15 um from every face
    d2 = ((p - p0) ** 2).sum(axis=1)[inner]
    D = 1e-6
    expected = 6 * D * 10 * 1e-6 / (wl.part.length_unit * 1e-4) ** 2  # lu^2
    se = expected * np.sqrt(2.0 / 3.0) / np.sqrt(inner.sum())
    assert abs(d2.mean() - expected) < 5 * se
    eng.close()


def _fp32_start(wl):
    """Positions rounded to FP32 so that both forms start from the same point."""
    wl.positions = wl.positions.astype(np.float32).astype(np.float64)
    return wl


def test_fp32_fallback_takes_the_fp64_wall_hits(gpu_available):
    """From the same start, the FP32+fallback form hits exactly the walls the
    FP64 form hits, and its position is within a few FP32 ulps (of the step's
    start and end coordinates) of the FP64 one: the light pass adds an FP32
    displacement in FP32, the full step rounds the FP64 end point (moved by a
    few ulps where it sits on a wall plane)."""
    wl = _fp32_start(W.free_diffusion_box(scale=5.0))
    e64 = _engine(wl, "fp64")
    e32 = _engine(wl, "fp32")
    e64.step(0, 1)
    e32.step(0, 1)
    w64, c64 = e64.hit_trace()
    w32, c32 = e32.hit_trace()
    assert np.array_equal(c64, c32)
    assert c64.sum() > 50
    for k in np.nonzero(c64)[0]:
        assert np.array_equal(w64[k, :c64[k]], w32[k, :c32[k]])
    p64 = e64.molecules()["pos"]
    p32 = e32.molecules()["pos"]
    p0 = wl.positions
    scale = np.maximum(np.abs(p64), np.abs(p0)).astype(np.float32)
    ulp = np.spacing(scale).astype(np.float64)
    assert np.all(np.abs(p32 - p64) <= 4 * ulp)
    same = np.mean(p32 == p64.astype(np.float32).astype(np.float64))
    print("FP32 end points equal to the FP64 ones rounded: %.4f" % same)
    assert same > 0.5
    e64.close()
    e32.close()


def test_fp32_fallback_long_run_agreement_and_containment(gpu_available):
    """Over 200 iterations the FP32+fallback trajectories stay inside the cube
    and take the FP64 wall-hit sequence for nearly every molecule-step; the
    disagreement rate and the position RMS against FP64 are reported."""
    wl = _fp32_start(W.free_diffusion_box())
    e64 = _engine(wl, "fp64")
    e32 = _engine(wl, "fp32")
    hits = disagree = 0
    for it in range(200):
        e64.step(it, 1)
        e32.step(it, 1)
        w64, c64 = e64.hit_trace()
        w32, c32 = e32.hit_trace()
        same = (c64 == c32) & np.all((w64 == w32) | (np.arange(8)[None, :] >= c64[:, None]), axis=1)
        hits += int(c64.sum())
        disagree += int((~same & (c64 + c32 > 0)).sum())
    p64 = e64.molecules()["pos"]
    p32 = e32.molecules()["pos"]
    rms = float(np.sqrt(np.mean((p32 - p64) ** 2)))
    print("fp32 vs fp64 over 200 iterations: %d wall hits, %d molecule-steps with a different hit sequence "
          "(rate %.2e), position RMS %.3e lu" % (hits, disagree, disagree / max(hits, 1), rms))
    assert np.all(np.abs(p32) < 50.0)
    assert disagree / max(hits, 1) < 1e-3
    assert rms < 1e-3
    e64.close()
    e32.close()


def test_fp32_msd_matches_6Dt(gpu_available):
    wl = W.free_diffusion_box(scale=10.0)
    eng = _engine(wl, "fp32", trace=False)
    p0 = wl.positions.astype(np.float32).astype(np.float64)
    eng.step(0, 10)
    p = eng.molecules()["pos"]
    inner = np.all(np.abs(p0) < 35.0, axis=1)
    d2 = ((p - p0) ** 2).sum(axis=1)[inner]
    expected = 6 * 1e-6 * 10 * 1e-6 / (wl.part.length_unit * 1e-4) ** 2
    se = expected * np.sqrt(2.0 / 3.0) / np.sqrt(inner.sum())
    assert abs(d2.mean() - expected) < 5 * se
    eng.close()


def test_replicas_are_independent_streams(gpu_available):
    """Two replicas with identical molecules but different replica ids take
    different paths; each replica matches its own oracle run."""
    wl = W.free_diffusion_replicas(num_replicas=3, mols_per_replica=500, iterations=20)
    eng = _engine(wl, trace=False)
    eng.step(0, 20)
    g = eng.molecules()
    for r in range(3):
        orc = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, replica=r)
        orc.run(20)
        o = orc.molecules()
        sel = g["replica"] == r
        assert np.array_equal(g["id"][sel], np.arange(500))
        rel = np.abs(g["pos"][sel] - o["pos"]) / np.maximum(np.abs(o["pos"]), 1.0)
        assert rel.max() <= 1e-12
    eng.close()
