"""Species time steps other than one iteration in the oracle, pinned by
closed forms and by MCell3's maturity rule.

* Custom time step t (s): space_step = sqrt(4e8 D t) / length_unit and
  time_step = t / time_unit (MCell3 src/mcell_species.c:254-264), so the
  per-axis variance of one step is proportional to t.
* A step shorter than one iteration diffuses the molecule again in the same
  iteration (diffuse_single_molecule, diffuse_react_event.cpp:294-319): the
  variance per iteration is 2 D dt whatever the step, and the number of
  diffusion calls is the iteration divided into steps, clipped at barriers.
* A step longer than one iteration starts at one iteration and lengthens
  with age, 1 + 0.2 (t - birthday), until that exceeds the species' step
  (MCell3 set_inertness_and_maxtime, src/diffuse.c:5239-5256; get_max_time,
  diffuse_react_event.cpp:164-198); a partial step scales the displacement
  by sqrt(steps) (compute_vol_displacement, diffusion_utils.inl:366-432), so
  the variance accumulated up to the molecule's diffusion time is again
  2 D t.
"""
import math
from dataclasses import replace

import numpy as np
import pytest

from mcell_amd import model as M
from mcell_amd import workloads as W

SQRT_EPS = 1e-6


def _centered(n, iterations, factors, d=1e-6, seed=5):
    """n molecules of species 0 near the centre of the 1 um reflective cube
    (no wall within reach of the run), species time steps from factors."""
    wl = W.free_diffusion_box(scale=0.1, iterations=iterations)
    wl = replace(wl, species=[M.Species("A", d)])
    wl = W.with_time_steps(wl, factors)
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-2.0, 2.0, size=(n, 3))
    return replace(wl, ids=np.arange(n, dtype=np.uint32), mol_species=np.zeros(n, np.uint32), positions=pos,
                   replica=np.zeros(n, np.uint32))


def _msd_per_axis(w, pos0):
    d = w.molecules()["pos"] - pos0
    return (d ** 2).mean(axis=0)


def test_custom_step_constants(oracle_lib):
    from oracle import pyoracle as po
    wl = _centered(10, 1, [1])
    w = po.world_from_workload(wl)
    for f in (0.25, 0.5, 3.0, 10.0):
        t = f * wl.part.time_unit
        s = w.add_species(1e-6, custom_time_step=t)
        D, ss, ts, _ = w.species(s)
        glob = w.species(0)[1]
        assert ts == pytest.approx(f, rel=1e-15)
        assert ss == pytest.approx(glob * math.sqrt(f), rel=1e-14)
        # the engine's species table is built by the same rule
        ss2, ts2 = M.Species("A", 1e-6, custom_time_step=t).steps(wl.part)
        assert (ss2, ts2) == (ss, ts)
    w.close()


@pytest.mark.parametrize("factor,barrier", [(0.25, 1), (0.5, 10), (0.3, 1), (0.3, 10)])
def test_short_steps_variance_and_calls(oracle_lib, factor, barrier):
    from oracle import pyoracle as po
    n, iters = 3000, 10
    wl = _centered(n, iters, [factor])
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=11)
    w.set_barrier_interval(barrier)
    w.run(iters)
    st = w.stats()
    assert st["mol_wall_reflections"] == 0
    # calls: each barrier interval of B iterations is covered by ceil(B / f)
    # steps, the last one clipped to end at the barrier
    per_interval = math.ceil(barrier / factor - 1e-9)
    assert st["diffuse_3d_calls"] == n * per_interval * (iters // barrier)
    msd = _msd_per_axis(w, wl.positions)
    var = 2 * 1e-6 * iters * wl.part.time_unit * 1e8 * wl.part.rcp_length_unit ** 2  # 2 D t per axis, in lu^2
    se = var * math.sqrt(2.0 / n)
    assert np.all(np.abs(msd - var) < 5 * se), (msd, var)
    m = w.molecules()
    assert np.all(m["diffusion_time"] == iters)
    w.close()


def _ramp_schedule(iterations, factor, barrier, birthday=0.0):
    """MCell3's schedule of one molecule of a species with time step
    `factor` (> 1) from its birthday: per iteration it, while the diffusion
    time t lies before it + 1, a step of min(factor, time to the barrier,
    1 + 0.2 (t - birthday) while not mature)."""
    t, mature, calls = birthday, False, 0
    for it in range(iterations):
        while t < it + 1 and abs(t - (it + 1)) >= 1e-12:
            next_barrier = (it // barrier + 1) * barrier
            max_time = min(next_barrier - it, 100) - (t - it)
            if not mature:
                f = 1.0 + 0.2 * (t - birthday)
                max_time = min(max_time, f)
                if f > factor:
                    mature = True
            t += min(factor, max_time)
            calls += 1
            if not (t < it + 1 and abs(t - (it + 1)) >= 1e-12):
                r = round(t)
                if abs(t - r) < SQRT_EPS:
                    t = float(r)
    return t, mature, calls


@pytest.mark.parametrize("barrier", [100, 7, 1])
def test_long_steps_maturity_ramp(oracle_lib, barrier):
    from oracle import pyoracle as po
    n, iters, factor = 2000, 70, 10.0
    wl = _centered(n, iters, [factor], d=1e-7)
    w = po.world_from_workload(wl, draw_mode=po.DRAW_PHILOX, seed=12)
    w.set_barrier_interval(barrier)
    w.run(iters)
    t, mature, calls = _ramp_schedule(iters, factor, barrier)
    m = w.molecules()
    assert np.all(m["diffusion_time"] == t), (m["diffusion_time"][:3], t)
    assert np.all(((m["flags"] & 2) != 0) == mature)
    assert np.all(m["birthday"] == 0.0)
    assert w.stats()["diffuse_3d_calls"] == n * calls
    if barrier == 100:
        assert mature and calls < 20  # the ramp, then steps of ten iterations
    if barrier == 1:
        # every step clipped at the next barrier; the age still matures it
        assert mature and calls == iters
    # variance up to the molecules' own diffusion time: 2 D t per axis
    msd = _msd_per_axis(w, wl.positions)
    var = 2 * 1e-7 * t * wl.part.time_unit * 1e8 * wl.part.rcp_length_unit ** 2
    assert np.all(np.abs(msd - var) < 5 * var * math.sqrt(2.0 / n)), (msd, var)
    w.close()
