"""Volume-surface classes that keep the volume reactant, in the oracle.

outcome_products_random (diffuse_react_event.cpp:2690-2720) keeps the
volume molecule when a product is of its species; collide_and_react_with_
surf_mol (:940-971) then lets it pass through the wall when that product's
orientation differs from the rule's volume reactant (RX_FLIP), else its ray
trace goes on as after no reaction.  On the receptor sphere (every receptor
up, L' + R' -> L? + LR', L hitting from outside) this is pinned by counts:

* L' + R' -> L' + LR': L is conserved and never enters the sphere; each
  reaction turns one R into LR.
* L' + R' -> L, + LR': each reaction moves one L inside, where it cannot
  react again (a back hit on an up receptor fails the orientation rule), so
  the number of L inside equals the number of LR, and L is conserved.
"""
import numpy as np
import pytest

from mcell_amd import workloads as W


@pytest.mark.parametrize("side", [1, -1])
def test_kept_ligand_counts(oracle_lib, side):
    from oracle import pyoracle as po
    wl = W.receptor_sphere(n_l=20000, n_r=6000, iterations=20, rate=4e8, down_fraction=0.0,
                           products=[(0, side, False), (2, 1, True)])
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=5)
    w.set_barrier_interval(wl.barrier_interval)
    nb = 12
    for it in range(0, wl.iterations, 5):
        w.run(5)
        m = w.molecules()
        live = (m["flags"] & 1) == 0
        sp = m["species"][live]
        n_l = int(np.sum(sp == 0))
        n_r, n_lr = int(np.sum(sp == 1)), int(np.sum(sp == 2))
        inside = int(np.sum(W.inside_convex(m["pos"][live][sp == 0], wl.vertices, wl.walls[nb:])))
        print("side %d iteration %d: L %d (inside %d) R %d LR %d" % (side, it + 5, n_l, inside, n_r, n_lr))
        assert n_l == 20000 and n_r + n_lr == 6000
        assert n_lr == w.volsurf_reactions()
        assert inside == (0 if side == 1 else n_lr)
    assert n_lr > 50
    w.close()
