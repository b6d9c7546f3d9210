"""Writes tests/golden/bimol_ensemble.npz: the CPU oracle's 32-seed ensemble
of A + B -> C in the 1 um reflective cube.  With --full it writes
tests/golden/bimol_ensemble_full.npz instead: the same at the full size of
the bimolecular benchmark case (50,000 A + 50,000 B, 2,000 iterations;
counts [32, 201, 3]).

The oracle follows the reference's sequential diffuse-react loop with the
reference's ISAAC64 + Ziggurat streams (one stream per run, seeded 1..32).
Workload: 2,000 A + 2,000 B (D = 1e-6 cm^2/s), C with D = 5e-7,
A + B -> C at 1e8 /M/s, 0.05 um subpartitions, counts every 10 iterations,
1000 iterations.  Run r (0-based) places its molecules with seed
PLACEMENT_BASE + r, the placement the engine's replica r uses in the parity
test, and draws its dynamics from ISAAC64 seed r + 1.

Contents:
  counts     uint32 [32, 101, 3]  A, B, C after iterations 0, 10, ..., 1000
  iterations uint32 [101]
"""
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

N_A = 2000
N_B = 2000
ITERATIONS = 1000
SAMPLE = 10
RUNS = 32
PLACEMENT_BASE = 1000
# the full-size ensemble (--full)
FULL = dict(N_A=50000, N_B=50000, ITERATIONS=2000, OUT="bimol_ensemble_full.npz")


def workload(num_replicas=1, base=PLACEMENT_BASE):
    from mcell_amd import workloads as W
    return W.bimol_box(n_a=N_A, n_b=N_B, iterations=ITERATIONS, seed=base, num_replicas=num_replicas,
                       barrier_interval=SAMPLE)


def run_one(r, iterations=None):
    """Run r's counts; iterations < ITERATIONS gives the run's first samples."""
    from oracle import pyoracle as po
    wl = workload(base=PLACEMENT_BASE + r)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_ISAAC, seed=r + 1)
    w.set_barrier_interval(SAMPLE)
    its = ITERATIONS if iterations is None else iterations
    out = np.zeros((its // SAMPLE + 1, 3), np.uint32)

    def count():
        m = w.molecules()
        live = (m["flags"] & 1) == 0
        return np.bincount(m["species"][live], minlength=3)[:3]

    out[0] = count()
    for k in range(its // SAMPLE):
        w.run(SAMPLE)
        out[k + 1] = count()
    w.close()
    return out


def use_full():
    """Switches the module to the full-size ensemble (before workload/run_one)."""
    global N_A, N_B, ITERATIONS
    N_A, N_B, ITERATIONS = FULL["N_A"], FULL["N_B"], FULL["ITERATIONS"]


def _init(full):
    if full:
        use_full()


def main():
    from oracle import pyoracle
    full = "--full" in sys.argv
    if full:
        use_full()
    pyoracle.build()
    with Pool(min(8, os.cpu_count() or 1), initializer=_init, initargs=(full,)) as p:
        counts = np.stack(p.map(run_one, range(RUNS)))
    its = np.arange(0, ITERATIONS + 1, SAMPLE, dtype=np.uint32)
    out = FULL["OUT"] if full else "bimol_ensemble.npz"
    np.savez_compressed(os.path.join(HERE, out), counts=counts, iterations=its)
    print("final A mean %.1f sd %.1f" % (counts[:, -1, 0].mean(), counts[:, -1, 0].std()))


if __name__ == "__main__":
    main()
