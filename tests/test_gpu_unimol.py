"""GPU parity of unimolecular reactions (E -> A + B) against the CPU oracle.

With the same per-molecule Philox draws the decay of E has no contention, so
the whole run is compared exactly: which molecules exist, their ids (products
numbered in (replica, reactant id) order, the oracle's creation order) and
their positions, bit for bit.  The decay itself is checked against the
exponential law E(t) ~ Binomial(E0, exp(-k t)).
"""
import math

import numpy as np
import pytest

from mcell_amd import workloads as W
from mcell_amd.engine import DiffuseReactEngine

pytestmark = pytest.mark.gpu


def test_decay_matches_oracle_exactly(gpu_available, oracle_lib):
    from oracle import pyoracle as po
    seed = 23
    wl = W.decay_box(n_e=3000, iterations=200, rate=5e3)
    eng = DiffuseReactEngine.from_workload(wl, seed=seed, capacity=20000)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
    w.set_barrier_interval(wl.barrier_interval)
    for chunk in range(0, wl.iterations, 20):
        eng.step(chunk, 20)
        w.run(20)
        g = eng.molecules()
        o = np.argsort(g["id"], kind="stable")
        om = w.molecules()
        live = np.nonzero((om["flags"] & 1) == 0)[0]
        assert np.array_equal(g["id"][o], live), "iteration %d: the same molecules" % (chunk + 20)
        assert np.array_equal(g["species"][o], om["species"][live])
        d = np.abs(g["pos"][o] - om["pos"][live]).max()
        assert d <= 1e-12 * 100, "iteration %d: max |dpos| %g" % (chunk + 20, d)
        print("iteration %d: %d molecules, %d decays" % (chunk + 20, len(live), w.unimol_reactions()), flush=True)
    assert w.unimol_reactions() > 1000
    w.close()
    eng.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_decay_is_exponential(gpu_available, precision):
    n0, k = 50000, 1e3
    wl = W.decay_box(n_e=n0, iterations=900, rate=k)
    eng = DiffuseReactEngine.from_workload(wl, precision=precision, seed=3, capacity=3 * n0)
    for step in range(3):
        eng.step(300 * step, 300)
        c = eng.count_species(1, 3)[0].astype(np.int64)
        p = math.exp(-k * 1e-6 * 300 * (step + 1))
        sd = math.sqrt(n0 * p * (1 - p))
        print("%s t=%d: E %d (expected %.0f +- %.0f), A %d, B %d"
              % (precision, 300 * (step + 1), c[2], n0 * p, sd, c[0], c[1]), flush=True)
        assert abs(c[2] - n0 * p) < 4 * sd
        assert c[0] == c[1] == n0 - c[2]
    eng.close()
