"""Count output from a GPU run: the count event every 5 iterations writes
L and LR as .dat files and L, LR and the binding reactions as one .gdat; the
files hold the engine's counters at those times (ligands and complexes
conserved: L + LR constant, LR = reactions)."""
import numpy as np
import pytest

from mcell_amd import counts as C
from mcell_amd import workloads as W
from mcell_amd.engine import DiffuseReactEngine

pytestmark = pytest.mark.gpu


def test_count_files_from_a_run(gpu_available, tmp_path):
    wl = W.receptor_sphere(n_l=40000, n_r=8000, iterations=20, rate=4e8)
    eng = DiffuseReactEngine.from_workload(wl, seed=7)
    dl = C.CountBuffer(str(tmp_path / "L.dat"), ["L"])
    dlr = C.CountBuffer(str(tmp_path / "LR.dat"), ["LR"])
    g = C.CountBuffer(str(tmp_path / "all.gdat"), ["L", "LR", "bind"], fmt="gdat")
    slot = (0 + 3 + 0) * 4 + 0  # volume-surface class 0, pathway 0 (no bimolecular classes, 3 species)
    items = [C.CountItem(dl, 0, [C.CountTerm("species", 0)]), C.CountItem(dlr, 0, [C.CountTerm("species", 2)]),
             C.CountItem(g, 0, [C.CountTerm("species", 0)]), C.CountItem(g, 1, [C.CountTerm("species", 2)]),
             C.CountItem(g, 2, [C.CountTerm("rxn", slot)])]
    ev = C.MolOrRxnCountEvent(eng, items, num_replicas=1, num_species=3, time_unit=wl.part.time_unit)
    expect = []
    for it in range(0, wl.iterations + 1, 5):
        if it:
            eng.step(it - 5, 5)
        ev.step(it)
        c = eng.count_species(1, 3)[0]
        expect.append((it * wl.part.time_unit, int(c[0]), int(c[2]), int(eng.rxn_counts()[0, :, slot].sum())))
    for b in (dl, dlr, g):
        b.flush_and_close()
    L = np.loadtxt(tmp_path / "L.dat")
    LR = np.loadtxt(tmp_path / "LR.dat")
    G = np.loadtxt(tmp_path / "all.gdat", comments="#")
    e = np.array(expect, dtype=float)
    assert np.allclose(L[:, 0], e[:, 0], rtol=1e-5) and np.array_equal(L[:, 1], e[:, 1])
    assert np.array_equal(LR[:, 1], e[:, 2])
    assert np.allclose(G[:, 0], e[:, 0], rtol=1e-8) and np.array_equal(G[:, 1:], e[:, 1:])
    assert np.all(G[:, 1] + G[:, 2] == 40000) and np.array_equal(G[:, 2], G[:, 3]) and G[-1, 2] > 100
    print(open(tmp_path / "all.gdat").read(), flush=True)
    eng.close()
