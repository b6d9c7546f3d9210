"""Edge cases of the GPU step.

* Empty contexts: a context with geometry, species and reaction classes but
  no molecules steps without error and stays empty (both forms, with and
  without reaction classes).
* Ragged batches: replicas are independent instances (their streams are
  keyed by replica and molecule id, partners and claims never cross
  replicas), so a replica's whole state after a run is bit-identical whether
  the other replicas of the batch are empty or full.
"""
import dataclasses

import numpy as np
import pytest

from mcell_amd import workloads as W
from mcell_amd.engine import DiffuseReactEngine

pytestmark = pytest.mark.gpu


def _keep(wl, keep):
    return dataclasses.replace(wl, ids=wl.ids[keep], mol_species=wl.mol_species[keep],
                               positions=wl.positions[keep], replica=wl.replica[keep])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["bimol", "free"])
def test_empty_context_steps(gpu_available, precision, kind):
    wl = W.bimol_box(n_a=100, n_b=100, iterations=5, seed=2) if kind == "bimol" else \
        W.free_diffusion_box(scale=0.01, iterations=5, seed=1)
    wl = _keep(wl, np.zeros(len(wl.ids), bool))
    eng = DiffuseReactEngine.from_workload(wl, precision=precision, capacity=1024)
    eng.step(0, 5)
    eng.sync()
    assert eng.num_molecules() == 0
    m = eng.molecules()
    assert all(len(v) == 0 for v in m.values())
    counts = eng.count_species(1, len(wl.species))
    assert not np.any(counts)
    eng.close()


def _state(wl, precision, iterations):
    eng = DiffuseReactEngine.from_workload(wl, precision=precision, seed=21, capacity=2 * max(len(wl.ids), 1))
    eng.step(0, iterations)
    eng.sync()
    m = eng.molecules()
    eng.close()
    return m


def _replica(m, r):
    sel = m["replica"] == r
    order = np.argsort(m["id"][sel], kind="stable")
    return {k: v[sel][order] for k, v in m.items()}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_ragged_batch_replicas_are_independent(gpu_available, precision):
    wl = W.bimol_box(n_a=3000, n_b=2000, iterations=12, seed=11, num_replicas=3, rate=1.36e9)
    full = _state(_keep(wl, wl.replica != 1), precision, 12)          # replica 1 empty
    only0 = _state(_keep(wl, wl.replica == 0), precision, 12)
    only2 = _state(_keep(wl, wl.replica == 2), precision, 12)
    assert not np.any(full["replica"] == 1)
    reacted = 0
    for r, alone in ((0, only0), (2, only2)):
        a, b = _replica(full, r), _replica(alone, r)
        assert len(a["id"]) == len(b["id"]) > 0
        for k in a:
            assert np.array_equal(a[k], b[k]), (r, k)
        reacted += int(np.sum(a["species"] == 2))
    assert reacted > 0, "the run must contain reactions to test their independence"
    print("%s: %d products in replicas 0 and 2" % (precision, reacted), flush=True)
