"""The synthetic workload generators build what their docstrings state:
the crowded-sphere geometry (non-overlapping spheres inside the cube,
molecules outside every sphere, distinct surface tiles, species ratios) and
the receptor sphere (receptors on distinct tiles of the sphere's walls)."""
import numpy as np

from mcell_amd import model as M
from mcell_amd import workloads as W


def test_crowded_spheres_geometry_and_placement():
    wl = W.crowded_spheres(n_molecules=50000, n_surface=5000, num_spheres=12, sphere_subdivisions=4)
    assert len(wl.objects) == 13 and len(wl.walls) == 12 + 12 * 1280
    centres, radii = wl.spheres[:, :3], wl.spheres[:, 3]
    assert np.all((radii >= 0.1) & (radii <= 0.3))
    for o, c, r0 in zip(wl.objects[1:], centres, radii):
        r = np.linalg.norm(o.vertices() - c, axis=1)
        # every vertex moved radially by a factor within 1 +- 2 %
        assert np.all((r >= 0.98 * r0 - 1e-12) & (r <= 1.02 * r0 + 1e-12))
        assert np.all(np.abs(o.vertices()) <= 2.0)
    for i in range(len(radii)):
        for j in range(i):
            assert np.linalg.norm(centres[i] - centres[j]) > 1.02 * (radii[i] + radii[j])
    pos_um = wl.positions * wl.part.length_unit
    for c, r in zip(centres, radii):
        assert np.all(np.linalg.norm(pos_um - c, axis=1) > 1.02 * r - 1e-9)
    assert np.all(np.abs(pos_um) <= 2.0)
    assert np.array_equal(np.bincount(wl.mol_species, minlength=3), [25000, 20000, 5000])
    sm = wl.surface_mols
    keys = sm["walls"].astype(np.int64) * 1024 + sm["tiles"]
    assert len(np.unique(keys)) == len(keys) == 5000
    assert np.all(sm["walls"] >= 12) and np.all(sm["orients"] == 1)
    nt = M.wall_num_tiles(wl.vertices, wl.walls[sm["walls"]])
    assert np.all(sm["tiles"] < nt)
    assert np.array_equal(sm["ids"], 50000 + np.arange(5000))


def test_receptor_sphere_tiles():
    wl = W.receptor_sphere(n_l=1000, n_r=3000, iterations=1, n_t=500)
    sm = wl.surface_mols
    assert np.all(sm["walls"] >= 12)
    keys = sm["walls"].astype(np.int64) * 1024 + sm["tiles"]
    assert len(np.unique(keys)) == 3000
    nt = M.wall_num_tiles(wl.vertices, wl.walls[sm["walls"]])
    assert np.all(sm["tiles"] < nt)
    assert np.array_equal(np.bincount(wl.mol_species, minlength=4), [1000, 0, 0, 500])
    # L outside the sphere
    r = np.linalg.norm(wl.positions[:1000] * wl.part.length_unit, axis=1)
    assert np.all(r > 0.5)
