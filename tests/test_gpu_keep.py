"""Reaction classes that keep a reactant, on the GPU against the oracle.

A product of a reactant's species is that reactant (outcome_products_random,
diffuse_react_event.cpp:2551-2565, 2686-2727): a kept mover's ray trace goes
on after the reaction (collide_and_react_with_vol_mol returns false,
:823-827), a kept unimolecular reactant diffuses on from its reaction time
with a new lifetime (outcome_unimolecular, :2990-3003).

* Whole runs equal the oracle's bit for bit (molecule ids, species and
  positions) where the snapshot changes nothing: E -> E + A beside E -> B
  (no molecule collisions), and catalysis where the partners do not move
  (static B; the kept mover may react several times in one step).
* Kinetics at size in both forms: the kept species' count never changes and
  the other follows the pseudo-first-order law exp(-k [kept] t); E -> E + A
  produces A at the rate the lifetimes fix.
"""
import math

import numpy as np
import pytest

from mcell_amd import workloads as W
from mcell_amd.engine import DiffuseReactEngine

pytestmark = pytest.mark.gpu


def _compare(eng, w, label):
    g = eng.molecules()
    o = np.argsort(g["id"], kind="stable")
    om = w.molecules()
    live = np.nonzero((om["flags"] & 1) == 0)[0]
    assert np.array_equal(g["id"][o], live), "%s: the same molecules" % label
    assert np.array_equal(g["species"][o], om["species"][live]), "%s: species" % label
    d = np.abs(g["pos"][o] - om["pos"][live]).max(initial=0)
    assert d <= 1e-12 * 100, "%s: max |dpos| %g" % (label, d)
    return np.bincount(om["species"][live], minlength=3)


def test_unimolecular_keep_matches_oracle_exactly(gpu_available, oracle_lib):
    from oracle import pyoracle as po
    seed = 29
    wl = W.decay_box(n_e=3000, iterations=200, rate=2e3, products=(2, 0), more_pathways=[(1e3, (1,))])
    eng = DiffuseReactEngine.from_workload(wl, seed=seed, capacity=40000)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
    w.set_barrier_interval(wl.barrier_interval)
    for chunk in range(0, wl.iterations, 20):
        eng.step(chunk, 20)
        w.run(20)
        c = _compare(eng, w, "iteration %d" % (chunk + 20))
        print("iteration %d: A %d B %d E %d, %d reactions" % (chunk + 20, c[0], c[1], c[2], w.unimol_reactions()),
              flush=True)
    assert c[0] > 500 and c[1] > 200
    # reaction counts by pathway (E is species 2: slots 8 + pathway) equal the oracle's events
    u, rx, d = w.rxn_log()
    rc = eng.rxn_counts()
    assert rc.shape[:2] == (1, 1)
    assert int(rc[0, 0, 8]) == int(np.sum(u[:, 3] == 0)) and int(rc[0, 0, 9]) == int(np.sum(u[:, 3] == 1))
    assert int(rc.sum()) == w.unimol_reactions()
    w.close()
    eng.close()


@pytest.mark.parametrize("keep", ["mover", "partner"])
def test_catalysis_on_static_partners_matches_oracle_exactly(gpu_available, oracle_lib, keep):
    """50 A moving among 200,000 static B at 1e9 /M/s: A + B -> A + C (the
    mover stays, often reacting twice in one step) or A + B -> B + C (the
    partner stays and serves several movers)."""
    from oracle import pyoracle as po
    seed = 31
    products = (0, 2) if keep == "mover" else (1, 2)
    n_a = 50 if keep == "mover" else 3000
    wl = W.bimol_box(n_a=n_a, n_b=200000, iterations=60, rate=1e9, products=products, d_b=0.0,
                     barrier_interval=10)
    eng = DiffuseReactEngine.from_workload(wl, seed=seed)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
    w.set_barrier_interval(wl.barrier_interval)
    total = 0
    for chunk in range(0, wl.iterations, 10):
        st = eng.step(chunk, 10)
        w.run(10)
        total += st.reactions
        c = _compare(eng, w, "iteration %d" % (chunk + 10))
        print("%s iteration %d: A %d B %d C %d" % (keep, chunk + 10, c[0], c[1], c[2]), flush=True)
    assert c[2] == total and total > 300
    assert c[0] == n_a if keep == "mover" else c[1] == 200000
    w.close()
    eng.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("keep", ["a", "b"])
def test_catalysis_kinetics(gpu_available, precision, keep):
    n_cat, n_sub = 20000, 20000
    products = (0, 2) if keep == "a" else (1, 2)
    wl = W.bimol_box(n_a=n_cat if keep == "a" else n_sub, n_b=n_sub if keep == "a" else n_cat, iterations=600,
                     products=products, barrier_interval=100)
    eng = DiffuseReactEngine.from_workload(wl, precision=precision, seed=12)
    conc = n_cat / (6.02214076e23 * 1e-15)
    cat, sub = (0, 1) if keep == "a" else (1, 0)
    for k in range(1, 4):
        eng.step(200 * (k - 1), 200)
        c = eng.count_species(1, 3)[0].astype(np.int64)
        p = math.exp(-1e8 * conc * 1e-6 * 200 * k)
        sd = math.sqrt(n_sub * p * (1 - p))
        print("%s keep %s t=%d: catalyst %d substrate %d (%.0f +- %.0f) C %d"
              % (precision, keep, 200 * k, c[cat], c[sub], n_sub * p, sd, c[2]), flush=True)
        assert c[cat] == n_cat
        assert c[sub] + c[2] == n_sub
        assert abs(c[sub] - n_sub * p) < 4 * sd + 0.01 * n_sub
    eng.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_unimolecular_keep_kinetics(gpu_available, precision):
    n0, k1, k2 = 40000, 1e3, 500.0
    wl = W.decay_box(n_e=n0, iterations=400, rate=k1, products=(2, 0), more_pathways=[(k2, (1,))])
    eng = DiffuseReactEngine.from_workload(wl, precision=precision, seed=8, capacity=4 * n0)
    eng.step(0, 400)
    c = eng.count_species(1, 3)[0].astype(np.int64)
    t = 400e-6
    p = math.exp(-k2 * t)
    mean_life = (1 - p) / k2
    e_life2 = 2 * (1 - p - k2 * t * p) / k2 ** 2 + t * t * p
    mean_a = n0 * k1 * mean_life
    sd_a = math.sqrt(n0 * (k1 * mean_life + k1 * k1 * (e_life2 - mean_life ** 2)))
    sd_e = math.sqrt(n0 * p * (1 - p))
    print("%s: E %d (%.0f +- %.0f) A %d (%.0f +- %.0f) B %d" % (precision, c[2], n0 * p, sd_e, c[0], mean_a, sd_a,
                                                                c[1]), flush=True)
    assert abs(c[2] - n0 * p) < 4 * sd_e
    assert c[1] == n0 - c[2]
    assert abs(c[0] - mean_a) < 4 * sd_a
    eng.close()
