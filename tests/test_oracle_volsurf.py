"""The oracle's volume-surface pieces against what the mathematics fixes.

* pb_factor (MCell3 src/react_util.c:132-157): 1e11 sigma / (2 N_AV)
  sqrt(pi dt / D) = 1.47161e-9 per collision and /M/s for D = 1e-6 cm^2/s,
  dt = 1 us and sigma = 1e4 /um^2, doubled for a rule with both orientations
  set; the host (model.py) and the oracle agree.
* Grid (Grid::initialize, wall.cpp:38-74): ceil(sqrt(area))^2 tiles, binding
  factor tiles / area.
* xyz2grid (grid_utils.inl:48-118): the tiles split the triangle into equal
  areas (uniform points fall uniformly on the tiles), strip k holds tiles
  k^2 .. k^2 + 2k from the apex at vertex 2 (tile 0) down to the edge from
  vertex 0 (first tile of the last strip) to vertex 1 (last tile), and the
  vertices themselves map to those three tiles.
* Reactions: only receptors whose orientation matches the rule react
  (trigger_bimolecular, rxn_utils.inl:58-84), molecules are conserved
  (L + LR, R + LR), and the first iteration's reactions from a uniform
  ligand follow the mean-field count k [L] N_up dt.
"""
import math

import numpy as np
import pytest

from mcell_amd import model as M
from mcell_amd import workloads as W


def _sphere_world(subdivisions):
    from oracle import pyoracle as po
    cfg = M.Config(time_step=1e-6, subpartition_dimension=0.1, seed=1)
    ball = M.create_icosphere("sphere", 0.5, subdivisions)
    w = po.OracleWorld(cfg, M.bounding_box([ball]), False)
    w.add_object(ball, 0)
    w.finalize_walls()
    w.add_species(1e-6)
    w.add_species(0.0, is_vol=False)
    return w, len(ball.wall_list)


def test_pb_factor_closed_form_and_oracle_agree(oracle_lib):
    wl = W.receptor_sphere(n_l=10, n_r=10, iterations=1)
    from oracle import pyoracle as po
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=1)
    pb_oracle, maxp = w.volsurf_rxn(0)
    pb_host = wl.volsurf_rxns[0].pb_factor(wl.part, wl.species[0])
    # 1e11 * 1e4 / (2 * 6.0221417930e23) * sqrt(pi), doubled for L' + R'
    assert abs(pb_oracle - 2 * 1.47161418e-9) < 1e-17
    assert pb_host == pytest.approx(pb_oracle, rel=1e-15)
    assert maxp == pytest.approx(wl.volsurf_rxns[0].rate * pb_oracle, rel=1e-15)
    unoriented = M.VolSurfReaction(0, 1, 1e8, 2, geom_vol=0, geom_surf=1)
    assert unoriented.pb_factor(wl.part, wl.species[0]) == pytest.approx(pb_oracle / 2, rel=1e-15)
    w.close()


def test_grid_tiles_and_binding_factor(oracle_lib):
    w, nw = _sphere_world(2)
    walls = w.walls()
    for k in range(0, nw, 37):
        n_axis, n_tiles, _, _, _, bf, _, _ = w.grid_info(k)
        area = walls[k, 13]
        assert n_axis == math.ceil(math.sqrt(area))
        assert n_tiles == n_axis * n_axis
        assert bf == pytest.approx(n_tiles / area, rel=1e-15)
    w.close()


def _triangle(w, wall):
    from oracle import pyoracle as po
    cnt = w.lib.orc_num_walls(w.w)
    tri = np.zeros((cnt, 3), np.uint32)
    w.lib.orc_get_wall_vertices(w.w, po._p(tri))
    verts = w.vertices(int(tri.max()) + 1)
    return verts[tri[wall]]


def test_xyz2grid_splits_the_wall_into_equal_tiles(oracle_lib):
    w, nw = _sphere_world(2)
    wall = 5
    n_axis, n_tiles = (int(x) for x in w.grid_info(wall)[:2])
    assert n_tiles >= 64
    v = _triangle(w, wall)
    rng = np.random.default_rng(1)
    n = 60 * n_tiles
    a, b = rng.uniform(size=(2, n))
    flip = a + b > 1
    a[flip], b[flip] = 1 - a[flip], 1 - b[flip]
    pts = v[0] + a[:, None] * (v[1] - v[0]) + b[:, None] * (v[2] - v[0])
    idx = np.array([w.xyz2grid(wall, p) for p in pts])
    assert idx.min() >= 0 and idx.max() < n_tiles
    counts = np.bincount(idx, minlength=n_tiles)
    expected = n / n_tiles
    chi2 = ((counts - expected) ** 2 / expected).sum()
    # chi-square with n_tiles - 1 degrees of freedom: mean dof, sd sqrt(2 dof)
    assert chi2 < (n_tiles - 1) + 5 * math.sqrt(2 * (n_tiles - 1)), chi2
    # the strip layout: apex tile at vertex 2, the last strip from vertex 0 to vertex 1
    mid, last = (n_axis - 1) ** 2, n_tiles - 1
    for vert, tile in ((2, 0), (0, mid), (1, last)):
        assert w.xyz2grid(wall, v[vert]) == tile
        c = v.mean(axis=0)
        assert w.xyz2grid(wall, v[vert] + 1e-3 * (c - v[vert])) == tile
    # points along the edge v0 -> v1 walk the last strip's upright tiles in order
    ts = [w.xyz2grid(wall, v[0] + f * (v[1] - v[0]) + 1e-6 * (v[2] - v[0]))
          for f in (np.arange(n_axis) + 0.5) / n_axis]
    assert ts == [mid + 2 * k for k in range(n_axis)]
    w.close()


def test_only_matching_orientations_react_and_molecules_are_conserved(oracle_lib):
    from oracle import pyoracle as po
    wl = W.receptor_sphere(n_l=40000, n_r=20000, iterations=4)
    sm = wl.surface_mols
    up = set(sm["ids"][sm["orients"] > 0].tolist())
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=7)
    w.set_snapshot(True)
    w.run(2)
    u, rx, _ = w.rxn_log()
    assert len(u) > 10
    assert np.all(rx == 0x40000000)
    assert set(u[:, 2].tolist()) <= up
    w.close()
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=7)
    w.run(wl.iterations)
    m = w.molecules()
    live = (m["flags"] & 1) == 0
    c = np.bincount(m["species"][live], minlength=3)
    n_rx = w.volsurf_reactions()
    assert n_rx > 10
    assert c[0] == wl.num_molecules - n_rx and c[2] == n_rx and c[1] == len(sm["ids"]) - n_rx
    w.close()


def test_first_iteration_follows_mean_field(oracle_lib):
    """From a uniform ligand the first iteration's reactions average
    k [L] N_up dt (the rate the pb_factor is calibrated to)."""
    from oracle import pyoracle as po
    n_l, n_r, rate = 100000, 30000, 1e8
    wl = W.receptor_sphere(n_l=n_l, n_r=n_r, iterations=1, rate=rate)
    n_up = int((wl.surface_mols["orients"] > 0).sum())
    vol_l = (2.0 ** 3 - 4.0 / 3.0 * math.pi * 0.5 ** 3) * 1e-15  # liters outside the sphere
    conc = n_l / (vol_l * M.N_AV)
    expected = rate * conc * n_up * 1e-6
    total = 0
    seeds = range(8)
    for seed in seeds:
        w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=seed)
        w.run(1)
        total += w.volsurf_reactions()
        w.close()
    mean = total / len(seeds)
    sd = math.sqrt(expected / len(seeds))
    print("first-iteration reactions: %.1f per run, mean field %.1f" % (mean, expected))
    # Poisson noise plus 10% for the sphere's curvature and its 5120 facets
    assert abs(mean - expected) < 4 * sd + 0.1 * expected


@pytest.mark.parametrize("orient", [1, -1, 0])
def test_volume_product_starts_next_to_the_wall(oracle_lib, orient):
    """L' + R' -> R' + P (the receptor stays, P a volume product that does
    not move): P starts 2 * 16 POS_EPS from the hit point along the wall
    normal, on the side its orientation names (outside the sphere for +1,
    inside for -1, each side about half the time when drawn), R is never
    consumed and every reaction leaves exactly one P
    (update_vol_mol_after_rxn_with_surf_mol, diffuse_react_event.cpp:2291-
    2320; outcome_products_random, :2620-2625)."""
    from oracle import pyoracle as po
    wl = W.receptor_sphere(n_l=60000, n_r=20000, iterations=6, rate=4e8, down_fraction=0.0,
                           products=[(1, 1, True), (3, orient, False)], more_species=[("P", 0.0)])
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=7)
    w.set_barrier_interval(wl.barrier_interval)
    w.run(wl.iterations)
    u, rx, d = w.rxn_log()
    om = w.molecules()
    p_ids = np.nonzero(om["species"] == 3)[0]
    assert len(p_ids) == len(u) > 100
    assert np.sum(om["species"] == 1) == 20000
    # products are numbered in reaction order; P i belongs to reaction i
    hit = d[:, 1:4]
    off = om["pos"][p_ids] - hit
    dist = np.linalg.norm(off, axis=1)
    assert np.abs(dist - 32e-12).max() < 1e-13
    centre = (np.asarray(wl.objects[1].vertices()).mean(axis=0)) / wl.part.length_unit
    radial = np.einsum("ij,ij->i", off, hit - centre)
    side = np.sign(radial)
    if orient:
        assert np.all(side == orient)
    else:
        n_out = int(np.sum(side > 0))
        assert abs(n_out - len(side) / 2) < 4 * math.sqrt(len(side) / 4), (n_out, len(side))
    w.close()


def test_counted_volumes_follow_the_geometry(oracle_lib):
    """The sphere as a counted volume (1 inside, 0 outside) with T crossing
    it through a transparent class: after 60 iterations every live volume
    molecule's index says on which side of the sphere it is
    (update_counted_volume_id_when_crossing_wall, collision_utils.inl:
    1637-1694), many T having crossed."""
    from oracle import pyoracle as po
    wl = W.receptor_sphere(n_l=2000, n_r=500, iterations=60, n_t=20000, counted=True)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=5)
    w.set_barrier_interval(wl.barrier_interval)
    cv0 = w.molecule_volumes()
    w.run(wl.iterations)
    om = w.molecules()
    cv = w.molecule_volumes()
    live = np.nonzero(((om["flags"] & 1) == 0) & ((om["species"] == 0) | (om["species"] == 3)))[0]
    nb = len(wl.objects[0].wall_list)
    truth = W.inside_convex(om["pos"][live], wl.vertices, wl.walls[nb:]).astype(np.uint32)
    assert np.array_equal(cv[live], truth)
    changed = int(np.sum(cv[live] != cv0[live]))
    print("%d of %d volume molecules changed volume" % (changed, len(live)))
    assert changed > 100 and w.wall_stats()["transparent_crossings"] >= changed
    w.close()
