"""The oracle's reaction pieces against what the mathematics fixes.

* exact_disk (exact_disk_utils.inl:840-1145): 1 with no wall near the disk;
  one wall crossing the disk along a chord at distance h leaves
  1 - (R^2 acos(h/R) - h sqrt(R^2 - h^2)) / (pi R^2); two walls at a box edge
  leave the disk area inside the quadrant, checked against a quadrature; a
  target behind the wall is occluded (-1).
* pb_factor (MCell3 src/react_util.c:163-181): 0.0367904 per collision for
  A + B at 1e8 /M/s with D = 1e-6 cm^2/s, equal in the oracle and the host.
* Kinetics: the 32-seed ensemble in tests/golden/bimol_ensemble.npz follows
  the mean-field law A(t) = A0 / (1 + k A0 t) for A0 = B0, and its first run
  is what the oracle produces today.
"""
import ctypes
import math
import os

import numpy as np
import pytest

from mcell_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))


def _world(oracle_lib):
    from oracle import pyoracle as po
    wl = W.bimol_box(n_a=1, n_b=1, iterations=1)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=1)
    return wl, w


def _box_world(wl, target_pos):
    """The 1 um cube with the A + B -> C classes and one B molecule (id 0) at
    target_pos."""
    from oracle import pyoracle as po
    from mcell_amd import model as M
    w = po.OracleWorld(wl.config, M.bounding_box(wl.objects), True)
    for k, obj in enumerate(wl.objects):
        w.add_object(obj, k)
    w.finalize_walls()
    for s in wl.species:
        w.add_species(s.diffusion_constant_3d)
    w.add_bimol_rxn(0, 1, 1e8, [2])
    w.add_molecules(1, np.asarray(target_pos, np.float64)[None, :])
    return w


def _exact_disk(w, loc, mv, target):
    w.lib.orc_exact_disk.restype = ctypes.c_double
    loc = np.asarray(loc, np.float64)
    mv = np.asarray(mv, np.float64)
    return w.lib.orc_exact_disk(w.w, loc.ctypes.data, mv.ctypes.data, target)


def _segment_fraction(h, R):
    seg = R * R * math.acos(h / R) - h * math.sqrt(R * R - h * h)
    return 1.0 - seg / (math.pi * R * R)


def test_exact_disk_without_walls_is_one(oracle_lib):
    wl, w = _world(oracle_lib)
    # molecule 0 is the target; its own position is irrelevant when the
    # collision point has no wall within the radius
    v = _exact_disk(w, [3.0, -7.0, 11.0], [0.3, 1.2, -0.4], 0)
    assert v == 1.0
    w.close()


@pytest.mark.parametrize("h", [0.05, 0.2, 0.35, 0.5])
def test_exact_disk_single_wall_chord_closed_form(oracle_lib, h):
    wl = W.bimol_box(n_a=1, n_b=1, iterations=1)
    R = wl.part.rxn_radius_3d
    loc = np.array([50.0 - h, -20.0, 20.0])  # face x = 50, away from its diagonal y = z
    w = _box_world(wl, loc)  # the target sits at the collision point
    for mv in ([0.0, 1.3, 0.0], [0.0, 0.0, -0.9], [0.0, 0.6, 0.8]):  # moves parallel to the wall
        v = _exact_disk(w, loc, mv, 0)
        assert abs(v - _segment_fraction(h, R)) < 1e-12, (mv, v)
    w.close()


def _quadrant_fraction(hx, hz, R, n=4001):
    # disk in the (x, z) plane around (0, 0); keep x < hx and z < hz
    xs = (np.arange(n) + 0.5) / n * 2 * R - R
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    inside = X * X + Z * Z <= R * R
    keep = inside & (X < hx) & (Z < hz)
    return keep.sum() / inside.sum()


@pytest.mark.parametrize("hx,hz", [(0.1, 0.2), (0.3, 0.05), (0.25, 0.25)])
def test_exact_disk_box_edge_matches_quadrature(oracle_lib, hx, hz):
    wl = W.bimol_box(n_a=1, n_b=1, iterations=1)
    R = wl.part.rxn_radius_3d
    loc = np.array([50.0 - hx, -20.0, 50.0 - hz])  # edge x = z = 50
    w = _box_world(wl, loc)
    v = _exact_disk(w, loc, [0.0, 1.0, 0.0], 0)
    ref = _quadrant_fraction(hx, hz, R)
    assert abs(v - ref) < 2e-3, (v, ref)
    w.close()


def test_exact_disk_target_behind_wall_is_occluded(oracle_lib):
    wl = W.bimol_box(n_a=1, n_b=1, iterations=1)
    loc = np.array([49.8, -20.0, 20.0])
    # a target inside the disk but beyond the wall plane x = 50 (still in the partition)
    w = _box_world(wl, [50.2, -20.0, 20.0])
    assert _exact_disk(w, loc, [0.0, 1.0, 0.0], 0) == -1.0
    w.close()


def test_pb_factor_closed_form_and_oracle_agree(oracle_lib):
    from oracle import pyoracle as po
    wl = W.bimol_box(n_a=1, n_b=1, iterations=1)
    r = wl.rxns[0]
    pb = r.pb_factor(wl.part, wl.species[0], wl.species[1])
    # 1e15/N_AV / (2 sqrt(pi) R^2 v_eff), R = 1/sqrt(pi 1e4) um, v_eff = 2 ss/dt (ss = 0.02 um)
    R = 1 / math.sqrt(math.pi * 1e4)
    v_eff = 2 * 0.02 / 1e-6
    closed = 1e15 / 6.0221417930e23 / (2 * math.sqrt(math.pi) * R * R * v_eff)
    assert abs(pb - closed) < 1e-12 * closed
    assert abs(1e8 * pb - 0.0367903545) < 1e-9
    w = po.world_from_workload(wl, seed=1)
    o = w.rxn(0)
    assert abs(o[0] - pb) <= 1e-15 * pb and abs(o[1] - 1e8 * pb) <= 1e-15
    w.close()


def test_ensemble_follows_mean_field_kinetics():
    ref = np.load(os.path.join(HERE, "golden", "bimol_ensemble.npz"))
    counts = ref["counts"].astype(np.float64)
    its = ref["iterations"].astype(np.float64)
    A0 = counts[0, 0, 0]
    V_l = 1e-15  # 1 um^3 in litres
    kA0 = 1e8 * A0 / (6.0221417930e23 * V_l)  # s^-1
    t = its * 1e-6
    mf = A0 / (1 + kA0 * t)
    mean = counts[:, :, 0].mean(axis=0)
    assert np.all(np.abs(mean - mf) / mf < 0.02), np.abs(mean - mf).max()
    # conservation in every run
    assert np.all(counts[:, :, 0] + counts[:, :, 2] == A0)
    assert np.all(counts[:, :, 1] + counts[:, :, 2] == counts[0, 0, 1])


def test_ensemble_fixture_is_reproduced_by_the_oracle(oracle_lib):
    import importlib.util
    spec = importlib.util.spec_from_file_location("mk", os.path.join(HERE, "golden", "make_bimol_ensemble.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    ref = np.load(os.path.join(HERE, "golden", "bimol_ensemble.npz"))
    run0 = mk.run_one(0)
    assert np.array_equal(run0, ref["counts"][0])


def test_full_size_ensemble_kinetics_and_reproduction(oracle_lib):
    """bimol_ensemble_full.npz (50,000 + 50,000, 2,000 iterations): A + C
    and B + C conserved in every run, the mean within 3 % of the mean-field
    law A(t) = A0 / (1 + k A0 t) (the finite-radius collision rate runs a
    little above mean field), and run 0's first 200 iterations re-run by the
    oracle give the fixture's first 21 samples exactly."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("mk", os.path.join(HERE, "golden", "make_bimol_ensemble.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    mk.use_full()
    ref = np.load(os.path.join(HERE, "golden", "bimol_ensemble_full.npz"))
    counts = ref["counts"].astype(np.float64)
    assert counts.shape == (32, 201, 3)
    A0 = counts[0, 0, 0]
    assert np.all(counts[:, :, 0] + counts[:, :, 2] == A0)
    assert np.all(counts[:, :, 1] + counts[:, :, 2] == counts[:, 0, 1:2])
    kA0 = 1e8 * A0 / (6.0221417930e23 * 1e-15)
    mf = A0 / (1 + kA0 * ref["iterations"].astype(np.float64) * 1e-6)
    mean = counts[:, :, 0].mean(axis=0)
    print("full ensemble: final A %.1f, mean field %.1f" % (mean[-1], mf[-1]))
    assert np.all(np.abs(mean - mf) / mf < 0.03), (np.abs(mean - mf) / mf).max()
    run0 = mk.run_one(0, iterations=200)
    assert np.array_equal(run0, ref["counts"][0, :21])


def test_unimolecular_decay_is_exponential(oracle_lib):
    """E -> A + B at 1e3 /s: E(t) ~ Binomial(E0, exp(-k t)) with the
    lifetimes -ln(u)/k_tot of rxn_utils.inl:720-735; every decay leaves one A
    and one B at the E's position."""
    from oracle import pyoracle as po
    wl = W.decay_box(n_e=20000, iterations=600)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_ISAAC, seed=3)
    w.set_barrier_interval(wl.barrier_interval)
    n0 = 20000
    for k in range(1, 4):
        w.run(200)
        m = w.molecules()
        live = (m["flags"] & 1) == 0
        c = np.bincount(m["species"][live], minlength=3)
        p = math.exp(-1e-3 * 200 * k)
        sd = math.sqrt(n0 * p * (1 - p))
        assert abs(c[2] - n0 * p) < 4 * sd, (k, c, n0 * p)
        assert c[0] == c[1] == n0 - c[2] == w.unimol_reactions()
    w.close()


def test_unimolecular_class_that_keeps_its_reactant(oracle_lib):
    """E -> E + A at k1 = 1e3 /s beside E -> B at k2 = 500 /s (the reactant
    stays in the first pathway: outcome_unimolecular,
    diffuse_react_event.cpp:2990-3003).  Each E is a Poisson source of A
    while it lives and dies at rate k2, so E(t) ~ Binomial(E0, exp(-k2 t)),
    B = E0 - E, and the expected A count is E0 k1 (1 - exp(-k2 t)) / k2 with
    variance E0 k1 / k2^2 (...) bounded by its Poisson-mixture form."""
    from oracle import pyoracle as po
    n0, k1, k2 = 4000, 1e3, 500.0
    wl = W.decay_box(n_e=n0, iterations=400, products=(2, 0), rate=k1, more_pathways=[(k2, (1,))])
    w = po.world_from_workload(wl, draw_mode=po.DRAW_ISAAC, seed=5)
    w.set_barrier_interval(wl.barrier_interval)
    for k in range(1, 3):
        w.run(200)
        t = 1e-6 * 200 * k
        m = w.molecules()
        live = (m["flags"] & 1) == 0
        c = np.bincount(m["species"][live], minlength=3)
        p = math.exp(-k2 * t)
        sd_e = math.sqrt(n0 * p * (1 - p))
        assert abs(c[2] - n0 * p) < 4 * sd_e, (k, c)
        assert c[1] == n0 - c[2]
        # A per E: Poisson(k1 * lifetime within t), lifetime ~ min(Exp(k2), t)
        mean_life = (1 - p) / k2
        e_life2 = 2 * (1 - p - k2 * t * p) / k2 ** 2 + t * t * p  # E[min(L, t)^2]
        mean_a = n0 * k1 * mean_life
        var_a = n0 * (k1 * mean_life + k1 * k1 * (e_life2 - mean_life ** 2))
        print("t=%g: E %d (%.0f), A %d (%.0f +- %.0f), B %d" % (t, c[2], n0 * p, c[0], mean_a, math.sqrt(var_a),
                                                                 c[1]))
        assert abs(c[0] - mean_a) < 4 * math.sqrt(var_a)
        assert w.unimol_reactions() == c[0] + c[1]
    w.close()


@pytest.mark.parametrize("keep", ["a", "b"])
def test_catalysis_keeps_the_catalyst_and_is_pseudo_first_order(oracle_lib, keep):
    """A + B -> A + C (keep a) or A + B -> B + C (keep b) in the 1 um cube:
    the kept species' count never changes and the other decays as
    exp(-k [kept] t), [kept] = N / (N_A V) (k = 1e8 /M/s, the mean-field law
    the unkept A + B -> C ensemble follows)."""
    from oracle import pyoracle as po
    n_cat, n_sub = 2000, 2000
    products = (0, 2) if keep == "a" else (1, 2)
    wl = W.bimol_box(n_a=n_cat if keep == "a" else n_sub, n_b=n_sub if keep == "a" else n_cat, iterations=500,
                     products=products, barrier_interval=100)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_ISAAC, seed=9)
    w.set_barrier_interval(wl.barrier_interval)
    conc = n_cat / (6.02214076e23 * 1e-15)  # molar, 1 um^3 = 1e-15 l
    cat, sub = (0, 1) if keep == "a" else (1, 0)
    for k in range(1, 3):
        w.run(250)
        m = w.molecules()
        live = (m["flags"] & 1) == 0
        c = np.bincount(m["species"][live], minlength=3)
        p = math.exp(-1e8 * conc * 1e-6 * 250 * k)
        sd = math.sqrt(n_sub * p * (1 - p))
        print("keep %s t=%d: catalyst %d, substrate %d (%.0f +- %.0f), C %d"
              % (keep, 250 * k, c[cat], c[sub], n_sub * p, sd, c[2]))
        assert c[cat] == n_cat
        assert c[sub] + c[2] == n_sub
        assert abs(c[sub] - n_sub * p) < 4 * sd + 0.01 * n_sub
    w.close()
