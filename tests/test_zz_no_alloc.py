"""The allocation-free species count (mcg_count_species_noalloc, include/
mcg.h) against mcg_count_species, both through the C ABI.

Two calls in a row share the library's device counter array: the first on
a larger context (8 replicas x 3 species of A + B -> C after some
iterations), the second on a smaller one (receptor sphere: 1 replica with
volume and surface species), so a stale count left from the first call
would show in the second.  A context with more counters than the array
holds is refused with MCG_E_LIMIT, and one whose surface molecules still
need tile buffers with MCG_E_STATE."""
import ctypes

import numpy as np
import pytest

from mcell_amd import _lib
from mcell_amd import workloads as W
from mcell_amd.engine import DiffuseReactEngine

pytestmark = pytest.mark.gpu


def _noalloc(eng, num_replicas, num_species):
    out = np.full((num_replicas, num_species), 0xDEAD, np.uint64)
    _lib.check(eng.lib.mcg_count_species_noalloc(eng.ctx, ctypes.c_void_p(out.ctypes.data)))
    return out


def test_noalloc_counts_equal_allocating_counts(gpu_available):
    big = W.bimol_box(n_a=3000, n_b=2000, iterations=6, seed=5, num_replicas=8, rate=1.36e9)
    a = DiffuseReactEngine.from_workload(big, seed=9)
    a.step(0, big.iterations)
    ref = a.count_species(8, 3)
    got = _noalloc(a, 8, 3)
    print("8 replicas x 3 species:", got.sum(axis=0).tolist(), flush=True)
    assert ref[:, 2].sum() > 0  # products exist
    assert np.array_equal(got, ref)
    a.close()

    small = W.receptor_sphere(n_l=4000, n_r=1500, iterations=4)
    b = DiffuseReactEngine.from_workload(small, seed=4)
    ns = len(small.species_steps)
    # before any step the surface molecules' tile buffers do not exist yet,
    # and this call allocates none: refused, and the context stays usable
    out = np.zeros(ns, np.uint64)
    assert b.lib.mcg_count_species_noalloc(b.ctx, ctypes.c_void_p(out.ctypes.data)) == 8  # MCG_E_STATE
    assert "tile buffers" in b.lib.mcg_last_error().decode()
    ref = b.count_species(1, ns)  # allocates the tile buffers
    assert np.array_equal(_noalloc(b, 1, ns), ref)
    b.step(0, small.iterations)
    ref = b.count_species(1, ns)
    got = _noalloc(b, 1, ns)
    print("receptor sphere:", got.tolist(), flush=True)
    assert np.array_equal(got, ref)
    assert got.sum() == ref.sum() > 0
    b.close()


def test_noalloc_refuses_more_counters_than_it_holds(gpu_available):
    wl = W.free_diffusion_replicas(num_replicas=40000, mols_per_replica=1, iterations=1)
    eng = DiffuseReactEngine.from_workload(wl)
    eng.set_species(list(wl.species_steps) * 2)  # 2 species x 40,000 replicas > 65,536
    out = np.zeros(80000, np.uint64)
    st = eng.lib.mcg_count_species_noalloc(eng.ctx, ctypes.c_void_p(out.ctypes.data))
    assert st == 7, "MCG_E_LIMIT expected"
    assert "exceeds" in eng.lib.mcg_last_error().decode()
    eng.close()
