"""Chooses the subpartition edge of a workload by measurement.

The subpartition grid sets how many walls a molecule tests per step, how many
grid cells a move crosses and how many partner candidates each reaction scan
reads.  Small cells mean short wall and partner lists but more crossings and a
larger cell index; the best edge depends on the mesh, the densities and the
step lengths, so it is measured: each candidate edge runs a few warm-up
iterations and a short timed burst of the workload on the GPU, and the edge
with the highest molecule-timesteps/s wins.  MCell4 takes the edge from the
model (Config.subpartition_dimension); a result here is a performance choice,
not a change of the model.
"""
from __future__ import annotations

import time
from typing import Dict, Sequence, Tuple

from . import workloads as W
from .engine import DiffuseReactEngine

DEFAULT_EDGES_UM = (0.05, 0.1, 0.2)


def measure(wl: W.Workload, precision: str, warmup: int = 9, steps: int = 5) -> float:
    """Molecule-timesteps/s of `steps` iterations after `warmup` ones (the
    warm-up covers the engine's timed choice of its drain form, 8
    iterations)."""
    eng = DiffuseReactEngine.from_workload(wl, precision=precision,
                                           capacity=wl.num_molecules * (2 if wl.rxns else 1))
    try:
        eng.step(0, warmup)
        eng.sync()
        t0 = time.perf_counter()
        st = eng.step(warmup, steps)
        eng.sync()
        secs = time.perf_counter() - t0
        done = st.molecule_steps if wl.rxns else wl.num_molecules * steps
        return done / secs
    finally:
        eng.close()


def choose_subpartition(wl: W.Workload, precision: str = "fp64", edges_um: Sequence[float] = DEFAULT_EDGES_UM,
                        warmup: int = 9, steps: int = 5, log=None) -> Tuple[float, Dict[float, float]]:
    """Returns (best edge in um, {edge: molecule-timesteps/s})."""
    rates = {}
    for e in edges_um:
        cand = W.with_subpartition(wl, e)
        rates[e] = measure(cand, precision, warmup, steps)
        if log:
            log("autotune: subpartition %.3f um (n = %d): %.3e molecule-timesteps/s"
                % (e, cand.part.num_subparts_per_edge, rates[e]))
    best = max(rates, key=rates.get)
    return best, rates
