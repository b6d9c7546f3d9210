"""The benchmark/parity workloads, built with the reference's geometry and
unit rules (see ``model``).  Each builder returns a ``Workload`` holding
exactly the flattened inputs the engine and the CPU oracle both receive.

Sizes follow the five workloads of the MCell4 diffuse-react benchmark:

* free_diffusion_box: 1 um reflective cube, 10,000 A (D = 1e-6 cm^2/s),
  0.1 um subpartitions, 1000 iterations (no reactions).
* bimol_box: same cube, 0.05 um subpartitions, 50,000 A + 50,000 B,
  A + B -> C at 1e8 /M/s, C with D = 5e-7, 2000 iterations.
* replicas_sphere_in_cube / crowded_spheres: larger shapes built on the same
  pieces (icosphere meshes, replicas).

``scale`` shrinks molecule counts for quick tests without changing geometry.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import model as M


@dataclass
class Workload:
    name: str
    part: M.PartitionConfig
    vertices: np.ndarray          # internal units, (nv, 3)
    walls: np.ndarray             # (nw, 3) uint32
    wall_class: np.ndarray        # (nw,) uint32
    species: List[M.Species]
    species_steps: List[tuple]    # (space_step, time_step, flags)
    ids: np.ndarray               # uint32
    mol_species: np.ndarray       # uint32
    positions: np.ndarray         # (n, 3) internal units
    replica: np.ndarray           # uint32
    iterations: int
    rxns: List[M.BimolReaction] = field(default_factory=list)
    num_replicas: int = 1
    config: M.Config = None
    objects: List[M.GeometryObject] = field(default_factory=list)
    barrier_interval: int = 1     # iterations between count barriers

    @property
    def num_molecules(self):
        return len(self.ids)


def _species_table(species, part, rxns):
    flags = [0] * len(species)
    for r in rxns:
        flags[r.reactant_a] |= 1
        flags[r.reactant_b] |= 1
    out = []
    for s, f in zip(species, flags):
        ss, ts = s.steps(part)
        out.append((ss, ts, f))
    return out


def rxn_table(wl: Workload):
    """The reaction classes as the engine takes them: (a, b, cum_probs,
    products per pathway), with P = rate * pb_factor (one pathway)."""
    out = []
    for r in wl.rxns:
        pb = r.pb_factor(wl.part, wl.species[r.reactant_a], wl.species[r.reactant_b])
        out.append((r.reactant_a, r.reactant_b, [r.rate * pb], [list(r.products)]))
    return out


def free_diffusion_box(scale: float = 1.0, iterations: int = 1000, seed: int = 1) -> Workload:
    cfg = M.Config(time_step=1e-6, subpartition_dimension=0.1, seed=seed)
    box = M.create_box("box", edge_dimension=1.0)
    part = M.derive_partition(cfg, M.bounding_box([box]), has_bimol_vol_rxns=False)
    verts, walls, classes = M.vertices_internal([box], part)
    species = [M.Species("A", 1e-6)]
    n = max(1, int(round(10000 * scale)))
    pos = M.place_uniform_in_box(n, 0.5, seed, part)
    return Workload("free_diffusion_box", part, verts, walls, classes, species, _species_table(species, part, []),
                    np.arange(n, dtype=np.uint32), np.zeros(n, np.uint32), pos, np.zeros(n, np.uint32), iterations,
                    config=cfg, objects=[box])


def free_diffusion_replicas(num_replicas: int, mols_per_replica: int, iterations: int = 100, seed: int = 4,
                            subpartition_dimension: float = 0.1, box_um: float = 1.0) -> Workload:
    """Many independent free-diffusion instances in one shared reflective cube
    (every replica has its own molecules and streams)."""
    cfg = M.Config(time_step=1e-6, subpartition_dimension=subpartition_dimension, seed=seed)
    box = M.create_box("box", edge_dimension=box_um)
    part = M.derive_partition(cfg, M.bounding_box([box]), has_bimol_vol_rxns=False)
    verts, walls, classes = M.vertices_internal([box], part)
    species = [M.Species("A", 1e-6)]
    n = num_replicas * mols_per_replica
    pos = M.place_uniform_in_box(n, box_um / 2, seed, part)
    rep = np.repeat(np.arange(num_replicas, dtype=np.uint32), mols_per_replica)
    ids = np.tile(np.arange(mols_per_replica, dtype=np.uint32), num_replicas)
    return Workload("free_diffusion_replicas", part, verts, walls, classes, species,
                    _species_table(species, part, []), ids, np.zeros(n, np.uint32), pos, rep, iterations,
                    num_replicas=num_replicas, config=cfg, objects=[box])


def bimol_box(n_a: int = 50000, n_b: int = 50000, iterations: int = 2000, seed: int = 2,
              num_replicas: int = 1, rate: float = 1e8, subpartition_dimension: float = 0.05,
              barrier_interval: int = 10) -> Workload:
    """A + B -> C in the 1 um reflective cube: A and B with D = 1e-6 cm^2/s,
    C with D = 5e-7, uniform placement, counts every ``barrier_interval``
    iterations.  Replica r is placed with seed ``seed + r`` (replica 0 is the
    single-instance workload); ids are 0..n_a-1 for A, then B."""
    cfg = M.Config(time_step=1e-6, subpartition_dimension=subpartition_dimension, seed=seed)
    box = M.create_box("box", edge_dimension=1.0)
    part = M.derive_partition(cfg, M.bounding_box([box]), has_bimol_vol_rxns=True)
    verts, walls, classes = M.vertices_internal([box], part)
    species = [M.Species("A", 1e-6), M.Species("B", 1e-6), M.Species("C", 5e-7)]
    rxns = [M.BimolReaction(0, 1, rate, [2])]
    n = n_a + n_b
    pos = np.concatenate([M.place_uniform_in_box(n, 0.5, seed + r, part) for r in range(num_replicas)])
    sp1 = np.concatenate([np.zeros(n_a, np.uint32), np.ones(n_b, np.uint32)])
    return Workload("bimol_box", part, verts, walls, classes, species, _species_table(species, part, rxns),
                    np.tile(np.arange(n, dtype=np.uint32), num_replicas), np.tile(sp1, num_replicas), pos,
                    np.repeat(np.arange(num_replicas, dtype=np.uint32), n), iterations, rxns=rxns,
                    num_replicas=num_replicas, config=cfg, objects=[box], barrier_interval=barrier_interval)
