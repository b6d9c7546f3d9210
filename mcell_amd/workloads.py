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

from dataclasses import dataclass, field, replace
from typing import List, Optional, Sequence

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
    replica_D: Optional[np.ndarray] = None  # (num_replicas, num_species) cm^2/s overrides, or None
    surface_classes: Optional[np.ndarray] = None  # uint8 (num_classes, num_species): front | back << 2
    unimol_rxns: List[M.UnimolReaction] = field(default_factory=list)
    volsurf_rxns: List[M.VolSurfReaction] = field(default_factory=list)
    # surface molecules: dict of equal-length arrays ids, species, walls,
    # tiles, orients, replica (or None)
    surface_mols: Optional[dict] = None
    spheres: Optional[np.ndarray] = None  # (k, 4) nominal centres and radii (um) of generated spheres
    # counted volumes: {"num": n, "wall_inside": (nw,), "wall_outside": (nw,),
    # "molecules": (num_molecules,) index of every volume molecule}, or None
    counted_volumes: Optional[dict] = None

    def species_for(self, replica: int) -> List[M.Species]:
        if self.replica_D is None:
            return self.species
        return [M.Species(s.name, float(d), s.target_only, s.custom_time_step, s.surface)
                for s, d in zip(self.species, self.replica_D[replica])]

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


def rxn_table(wl: Workload, replica: Optional[int] = None):
    """The reaction classes as the engine takes them: (a, b, cum_probs,
    products per pathway), with P = rate * pb_factor (one pathway), for the
    base species or for one replica's species."""
    species = wl.species if replica is None else wl.species_for(replica)
    out = []
    for r in wl.rxns:
        pb = r.pb_factor(wl.part, species[r.reactant_a], species[r.reactant_b])
        out.append((r.reactant_a, r.reactant_b, [r.rate * pb], [list(r.products)]))
    return out


def with_time_steps(wl: Workload, factors: Sequence[float]) -> Workload:
    """The workload with species i's time step set to factors[i] global
    steps (a custom time step, mcell_species.c:254-264; 0 or 1: the global
    step).  Space steps and reaction probabilities follow."""
    sp = [M.Species(s.name, s.diffusion_constant_3d, s.target_only,
                    f * wl.part.time_unit if f not in (0, 1) else 0.0, s.surface) for s, f in zip(wl.species, factors)]
    steps = _species_table(sp, wl.part, wl.rxns)
    if len(wl.species_steps) == len(steps):  # keep the builder's flags
        steps = [(ss, ts, old[2]) for (ss, ts, _), old in zip(steps, wl.species_steps)]
    return replace(wl, species=sp, species_steps=steps, name=wl.name + "_timesteps")


def with_diffusing_surface(wl: Workload, diffusion: dict) -> Workload:
    """The workload with the surface species {index: D (cm^2/s)} diffusing
    on the mesh (space step as for volume species, MCell3
    src/mcell_species.c:270-276)."""
    sp = list(wl.species)
    steps = list(wl.species_steps)
    for k, d in diffusion.items():
        sp[k] = M.Species(sp[k].name, d, surface=True)
        ss, ts = sp[k].steps(wl.part)
        steps[k] = (ss, ts, steps[k][2] | 16)
    return replace(wl, species=sp, species_steps=steps, name=wl.name + "_surface_diffusion")


def species_steps(wl: Workload, replica: int):
    """(space_step, time_step, flags) of one replica's species."""
    return _species_table(wl.species_for(replica), wl.part, wl.rxns)


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


def box_at_partition_faces(n: int = 20000, iterations: int = 50, seed: int = 7) -> Workload:
    """Free diffusion in the reflective 1 um cube with the partition sized to
    the cube (partition_dimension below the cube's extent: the converter
    takes the bounding box plus its 0.01 um margin, mcell4_converter.cpp:
    280-310) and 0.02 um subpartitions, so that every face of the cube lies
    2 internal units inside a partition face.  The molecules start within
    1.5 units of the walls: a move towards a wall often ends outside the
    partition, and ray_trace_vol clips it at the partition face before the
    wall search (get_displacement_up_to_partition_boundary,
    collision_utils.inl:48-114)."""
    cfg = M.Config(time_step=1e-6, subpartition_dimension=0.02, partition_dimension=0.5, seed=seed)
    box = M.create_box("box", edge_dimension=1.0)
    part = M.derive_partition(cfg, M.bounding_box([box]), has_bimol_vol_rxns=False)
    verts, walls, classes = M.vertices_internal([box], part)
    species = [M.Species("A", 1e-6)]
    rng = np.random.default_rng(seed)
    half = 0.5 / part.length_unit
    pos = rng.uniform(-half, half, size=(n, 3))
    # push one coordinate of every molecule to within 1.5 units of a face
    axis = rng.integers(0, 3, n)
    side = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    pos[np.arange(n), axis] = side * (half - rng.uniform(0.01, 1.5, n))
    return Workload("box_at_partition_faces", part, verts, walls, classes, species,
                    _species_table(species, part, []), np.arange(n, dtype=np.uint32), np.zeros(n, np.uint32), pos,
                    np.zeros(n, np.uint32), iterations, config=cfg, objects=[box])


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
              barrier_interval: int = 10, products=(2,), d_a: float = 1e-6, d_b: float = 1e-6) -> Workload:
    """A + B -> C in the 1 um reflective cube: A and B with D = 1e-6 cm^2/s
    (d_a, d_b), C with D = 5e-7, uniform placement, counts every
    ``barrier_interval`` iterations.  ``products`` replaces C (a product of a
    reactant's species keeps that reactant, e.g. (0, 2): A + B -> A + C).
    Replica r is placed with seed ``seed + r`` (replica 0 is the
    single-instance workload); ids are 0..n_a-1 for A, then B."""
    cfg = M.Config(time_step=1e-6, subpartition_dimension=subpartition_dimension, seed=seed)
    box = M.create_box("box", edge_dimension=1.0)
    part = M.derive_partition(cfg, M.bounding_box([box]), has_bimol_vol_rxns=True)
    verts, walls, classes = M.vertices_internal([box], part)
    species = [M.Species("A", d_a), M.Species("B", d_b), M.Species("C", 5e-7)]
    rxns = [M.BimolReaction(0, 1, rate, list(products))]
    n = n_a + n_b
    pos = np.concatenate([M.place_uniform_in_box(n, 0.5, seed + r, part) for r in range(num_replicas)])
    sp1 = np.concatenate([np.zeros(n_a, np.uint32), np.ones(n_b, np.uint32)])
    return Workload("bimol_box", part, verts, walls, classes, species, _species_table(species, part, rxns),
                    np.tile(np.arange(n, dtype=np.uint32), num_replicas), np.tile(sp1, num_replicas), pos,
                    np.repeat(np.arange(num_replicas, dtype=np.uint32), n), iterations, rxns=rxns,
                    num_replicas=num_replicas, config=cfg, objects=[box], barrier_interval=barrier_interval)


def place_outside_sphere(n: int, half_edge_um: float, radius_um: float, seed: int, part: M.PartitionConfig):
    """Uniform positions in the cube [-h, h]^3 outside the centred sphere of
    the given radius (rejection sampling, numpy PCG64), internal units."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.empty((0, 3))
    r2 = radius_um * radius_um
    while len(out) < n:
        p = rng.uniform(-half_edge_um, half_edge_um, size=(2 * (n - len(out)) + 16, 3))
        p = p[(p * p).sum(axis=1) > r2]
        out = np.concatenate([out, p])
    return out[:n] * part.rcp_length_unit


def replicas_sphere_in_cube(num_replicas: int = 256, iterations: int = 500, seed: int = 4,
                            subpartition_dimension: float = 0.05, n_min: int = 1000, n_max: int = 200000,
                            d_min: float = 1e-7, d_max: float = 5e-6, rate: float = 1e8,
                            sphere_subdivisions: int = 5) -> Workload:
    """Batched replicas of A + B -> C over one mesh: a reflective 2 um cube
    holding a reflective icosphere of radius 0.5 um (API subdivisions 5: 5,120
    triangles).  Instance i draws N_A ~ U[n_min, n_max] (N_B = N_A / 2) and one
    D ~ U[d_min, d_max] cm^2/s for A and B from PCG64(seed); C diffuses with
    D = 5e-7.  Instance i places its molecules outside the sphere with seed
    seed + i; its ids are 0..N_A-1 (A) then B."""
    cfg = M.Config(time_step=1e-6, subpartition_dimension=subpartition_dimension, seed=seed)
    box = M.create_box("box", edge_dimension=2.0)
    ball = M.create_icosphere("sphere", 0.5, sphere_subdivisions)
    objects = [box, ball]
    part = M.derive_partition(cfg, M.bounding_box(objects), has_bimol_vol_rxns=True)
    verts, walls, classes = M.vertices_internal(objects, part)
    species = [M.Species("A", 1e-6), M.Species("B", 1e-6), M.Species("C", 5e-7)]
    rxns = [M.BimolReaction(0, 1, rate, [2])]
    rng = np.random.Generator(np.random.PCG64(seed))
    n_a = rng.integers(n_min, n_max + 1, size=num_replicas)
    d = rng.uniform(d_min, d_max, size=num_replicas)
    replica_D = np.stack([d, d, np.full(num_replicas, 5e-7)], axis=1)
    pos, ids, sp, rep = [], [], [], []
    for i in range(num_replicas):
        na, nb = int(n_a[i]), int(n_a[i]) // 2
        pos.append(place_outside_sphere(na + nb, 1.0, 0.5, seed + i, part))
        ids.append(np.arange(na + nb, dtype=np.uint32))
        sp.append(np.concatenate([np.zeros(na, np.uint32), np.ones(nb, np.uint32)]))
        rep.append(np.full(na + nb, i, np.uint32))
    wl = Workload("replicas_sphere_in_cube", part, verts, walls, classes, species,
                  _species_table(species, part, rxns), np.concatenate(ids), np.concatenate(sp),
                  np.concatenate(pos), np.concatenate(rep), iterations, rxns=rxns, num_replicas=num_replicas,
                  config=cfg, objects=objects, barrier_interval=10, replica_D=replica_D)
    return wl


def with_subpartition(wl: Workload, subpartition_dimension: float) -> Workload:
    """The same workload on a partition with another subpartition edge (um):
    molecules and mesh are unchanged in internal units, only the grid the
    engine bins walls and molecules into differs."""
    import dataclasses
    cfg = dataclasses.replace(wl.config, subpartition_dimension=subpartition_dimension)
    part = M.derive_partition(cfg, M.bounding_box(wl.objects), has_bimol_vol_rxns=bool(wl.rxns))
    return dataclasses.replace(wl, part=part, config=cfg, species_steps=_species_table(wl.species, part, wl.rxns))


def sphere_in_cube(n_l: int = 100000, n_t: int = 10000, iterations: int = 1000, seed: int = 3,
                   subpartition_dimension: float = 0.5, sphere_subdivisions: int = 5,
                   t_action: int = 1) -> Workload:
    """The volume part of the sphere-in-cube workload: a reflective 2 um cube
    holding an icosphere of radius 0.5 um (API subdivisions 5: 5,120
    triangles) with surface class 1 on the sphere.  L (D = 1e-6 cm^2/s) starts
    outside the sphere and is reflected by it; T (D = 1e-6) starts anywhere in
    the cube and the class acts on it with `t_action` from both sides
    (1 = transparent, 2 = absorptive).  Placement seed `seed` (L), seed + 1 (T);
    ids are 0..n_l-1 (L) then T."""
    cfg = M.Config(time_step=1e-6, subpartition_dimension=subpartition_dimension, seed=seed)
    box = M.create_box("box", edge_dimension=2.0)
    ball = M.create_icosphere("sphere", 0.5, sphere_subdivisions)
    ball.surface_class = 1
    objects = [box, ball]
    part = M.derive_partition(cfg, M.bounding_box(objects), has_bimol_vol_rxns=False)
    verts, walls, classes = M.vertices_internal(objects, part)
    species = [M.Species("L", 1e-6), M.Species("T", 1e-6)]
    sc = np.array([[0, t_action | (t_action << 2)]], np.uint8)
    pos = np.concatenate([place_outside_sphere(n_l, 1.0, 0.5, seed, part),
                          M.place_uniform_in_box(n_t, 1.0, seed + 1, part)])
    sp = np.concatenate([np.zeros(n_l, np.uint32), np.ones(n_t, np.uint32)])
    n = n_l + n_t
    steps = _species_table(species, part, [])
    steps = [(ss, ts, f | (2 if sc[0, k] else 0)) for k, (ss, ts, f) in enumerate(steps)]
    return Workload("sphere_in_cube", part, verts, walls, classes, species, steps, np.arange(n, dtype=np.uint32),
                    sp, pos, np.zeros(n, np.uint32), iterations, config=cfg, objects=objects,
                    surface_classes=sc)


def unimol_table(wl: Workload):
    """Unimolecular classes as the engine takes them: (species, cumulative
    rates per internal time unit, products per pathway), one class per species."""
    by = {}
    for u in wl.unimol_rxns:
        by.setdefault(u.reactant, []).append(u)
    out = []
    for sp, us in sorted(by.items()):
        cum, acc = [], 0.0
        for u in us:
            acc += u.rate * wl.part.time_unit
            cum.append(acc)
        out.append((sp, cum, [list(u.products) for u in us]))
    return out


def volsurf_table(wl: Workload):
    """Volume-surface classes as the engine takes them: (vol, surf, geom_vol,
    geom_surf, cum_probs, products per pathway as (species, orientation, is
    surface) lists), P = rate * pb_factor."""
    out = []
    for v in wl.volsurf_rxns:
        pb = v.pb_factor(wl.part, wl.species[v.vol])
        out.append((v.vol, v.surf, v.geom_vol, v.geom_surf, [v.rate * pb], [v.product_list()]))
    return out


def place_on_walls(n: int, wall_ids: np.ndarray, vertices: np.ndarray, walls: np.ndarray, seed: int):
    """n distinct tiles of the given walls drawn uniformly by area (each tile
    of a wall covers area / tiles of it; weighted sampling without replacement
    by exponential keys, numpy PCG64): (walls, tiles)."""
    v = vertices[walls[wall_ids]]
    area = 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
    nt = M.wall_num_tiles(vertices, walls[wall_ids])
    base = np.concatenate([[0], np.cumsum(nt)])
    if n > base[-1]:
        raise ValueError("more surface molecules than tiles")
    weight = np.repeat(area / nt, nt)
    rng = np.random.Generator(np.random.PCG64(seed))
    key = rng.exponential(size=len(weight)) / weight
    flat = np.sort(np.argpartition(key, n - 1)[:n]) if n < len(key) else np.arange(n)
    k = np.searchsorted(base, flat, side="right") - 1
    return wall_ids[k].astype(np.uint32), (flat - base[k]).astype(np.uint32)


def surface_diffusion(n_s: int = 2000, iterations: int = 100, seed: int = 9, d_surf: float = 1e-7,
                      sphere_subdivisions: int = 4, sheet: bool = False, n_sheet: int = 0,
                      barrier_interval: int = 10) -> Workload:
    """Surface molecules S diffusing on the icosphere of radius 0.5 um (a
    closed surface of 20 * 4^s triangles) inside the reflective 2 um cube,
    n_s of them on distinct tiles drawn uniformly by area, at the tiles'
    centres, half oriented up; with ``sheet`` a flat square sheet of side
    1 um (two triangles: an open mesh whose outer edges reflect) at z =
    -0.8 um holds n_sheet more.  D_surf = d_surf cm^2/s (space step
    sqrt(4e8 D dt) / length unit, as for volume species).  Species S only;
    ids follow the placement order (sphere, then sheet)."""
    cfg = M.Config(time_step=1e-6, subpartition_dimension=0.1, seed=seed)
    box = M.create_box("box", edge_dimension=2.0)
    ball = M.create_icosphere("sphere", 0.5, sphere_subdivisions)
    objects = [box, ball]
    if sheet:
        z = -0.8
        objects.append(M.GeometryObject("sheet", [(-0.5, -0.5, z), (0.5, -0.5, z), (0.5, 0.5, z), (-0.5, 0.5, z)],
                                        [(0, 1, 2), (0, 2, 3)]))
    part = M.derive_partition(cfg, M.bounding_box(objects), has_bimol_vol_rxns=False)
    verts, walls, classes = M.vertices_internal(objects, part)
    species = [M.Species("S", d_surf, surface=True)]
    nb, nbl = len(box.wall_list), len(ball.wall_list)
    sw, st = place_on_walls(n_s, np.arange(nb, nb + nbl, dtype=np.uint32), verts, walls, seed)
    if sheet and n_sheet:
        w2, t2 = place_on_walls(n_sheet, np.arange(nb + nbl, nb + nbl + 2, dtype=np.uint32), verts, walls, seed + 1)
        sw, st = np.concatenate([sw, w2]), np.concatenate([st, t2])
    n = len(sw)
    rng = np.random.Generator(np.random.PCG64(seed + 2))
    orients = np.where(rng.uniform(size=n) < 0.5, -1, 1).astype(np.int32)
    surf = {"ids": np.arange(n, dtype=np.uint32), "species": np.zeros(n, np.uint32), "walls": sw, "tiles": st,
            "orients": orients, "replica": np.zeros(n, np.uint32)}
    steps = [(s, t, f | 16) for s, t, f in _species_table(species, part, [])]
    e = np.zeros(0, np.uint32)
    return Workload("surface_diffusion", part, verts, walls, classes, species, steps, e, e,
                    np.zeros((0, 3)), e, iterations, config=cfg, objects=objects, surface_mols=surf,
                    barrier_interval=barrier_interval)


def receptor_sphere(n_l: int = 100000, n_r: int = 10000, iterations: int = 10, seed: int = 3,
                    subpartition_dimension: float = 0.1, sphere_subdivisions: int = 5, rate: float = 1e8,
                    down_fraction: float = 0.5, n_t: int = 0, t_action: int = 1, products=None,
                    more_species=(), more_volsurf=(), counted: bool = False) -> Workload:
    """Ligand-receptor binding on a sphere: the reflective 2 um cube holding
    the icosphere of radius 0.5 um; L (D = 1e-6 cm^2/s) starts outside the
    sphere and is reflected by it, the surface receptors R sit on distinct
    tiles of the sphere drawn uniformly by area, a fraction `down_fraction`
    of them oriented down (-1), the rest up (+1).  L' + R' -> LR' at `rate`
    /M/s: only up receptors hit from the outside (front) react, and the
    product LR takes the receptor's tile.  With n_t > 0 a fourth species T
    (D = 1e-6) starts anywhere in the cube and surface class 1 on the sphere
    acts on it with `t_action` from both sides (1 = transparent, 2 =
    absorptive).  Species L, R, LR, T; placement seeds: seed (L), seed + 1
    (T), seed + 2 (receptor tiles), seed + 3 (orientations); ids 0..n_l-1 (L),
    then T, then the receptors.  ``products`` replaces LR' by a product list
    of the class ((species, orientation, is surface), model.VolSurfReaction);
    ``more_species`` ((name, D), ...) follow the others and ``more_volsurf``
    adds classes.  counted: the sphere is a counted volume (index 1 inside,
    0 outside), every volume molecule starting with the index of where it
    lies."""
    cfg = M.Config(time_step=1e-6, subpartition_dimension=subpartition_dimension, seed=seed)
    box = M.create_box("box", edge_dimension=2.0)
    ball = M.create_icosphere("sphere", 0.5, sphere_subdivisions)
    if n_t:
        ball.surface_class = 1
    objects = [box, ball]
    part = M.derive_partition(cfg, M.bounding_box(objects), has_bimol_vol_rxns=False)
    verts, walls, classes = M.vertices_internal(objects, part)
    species = [M.Species("L", 1e-6), M.Species("R", 0.0), M.Species("LR", 0.0)]
    if n_t:
        species.append(M.Species("T", 1e-6))
    species += [M.Species(name, d) for name, d in more_species]
    rx = [M.VolSurfReaction(0, 1, rate, 2, product_orient=1, geom_vol=1, geom_surf=1, products=products)]
    rx += list(more_volsurf)
    sphere_walls = np.arange(len(box.wall_list), len(walls), dtype=np.uint32)
    sw, st = place_on_walls(n_r, sphere_walls, verts, walls, seed + 2)
    rng = np.random.Generator(np.random.PCG64(seed + 3))
    orients = np.where(rng.uniform(size=n_r) < down_fraction, -1, 1).astype(np.int32)
    pos = place_outside_sphere(n_l, 1.0, 0.5, seed, part)
    sp = np.zeros(n_l, np.uint32)
    sc = None
    if n_t:
        pos = np.concatenate([pos, M.place_uniform_in_box(n_t, 1.0, seed + 1, part)])
        sp = np.concatenate([sp, np.full(n_t, 3, np.uint32)])
        sc = np.zeros((1, len(species)), np.uint8)
        sc[0, 3] = t_action | (t_action << 2)
    n_vol = len(sp)
    surf = {"ids": np.arange(n_vol, n_vol + n_r, dtype=np.uint32), "species": np.ones(n_r, np.uint32),
            "walls": sw, "tiles": st, "orients": orients, "replica": np.zeros(n_r, np.uint32)}
    steps = _species_table(species, part, [])
    for v in rx:
        steps[v.vol] = (steps[v.vol][0], steps[v.vol][1], steps[v.vol][2] | 4)
    if n_t:
        steps[3] = (steps[3][0], steps[3][1], steps[3][2] | 2)
    cv = None
    if counted:
        nb = len(box.wall_list)
        wi = np.full(len(walls), 0xFFFFFFFF, np.uint32)
        wo = wi.copy()
        wi[nb:] = 1
        wo[nb:] = 0
        cv = {"num": 2, "wall_inside": wi, "wall_outside": wo,
              "molecules": inside_convex(pos, verts, walls[nb:]).astype(np.uint32)}
    return Workload("receptor_sphere", part, verts, walls, classes, species, steps,
                    np.arange(n_vol, dtype=np.uint32), sp, pos, np.zeros(n_vol, np.uint32),
                    iterations, config=cfg, objects=objects, volsurf_rxns=rx, surface_mols=surf,
                    surface_classes=sc, counted_volumes=cv)


def inside_convex(points: np.ndarray, vertices: np.ndarray, walls: np.ndarray) -> np.ndarray:
    """Whether each point lies inside the convex mesh whose walls face
    outward (behind every wall plane)."""
    tri = vertices[walls]
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    inside = np.ones(len(points), bool)
    for k in range(0, len(points), 4096):
        p = points[k:k + 4096]
        inside[k:k + 4096] = np.all(np.einsum("pj,wj->pw", p, n) - np.einsum("wj,wj->w", tri[:, 0], n) < 0, axis=1)
    return inside


def volume_surface(iterations: int = 1000, seed: int = 3, subpartition_dimension: float = 0.1) -> Workload:
    """The volume-surface case: the reflective 2 um cube with the 0.5 um
    icosphere (5,120 triangles); 20,000 receptors R' on the sphere drawn
    uniformly by area (seed 3), 100,000 L (D = 1e-6 cm^2/s) outside the
    sphere, L' + R' -> LR' at 1e8 /M/s, and 10,000 T that cross the sphere
    through a transparent surface class; dt 1e-6 s."""
    return receptor_sphere(n_l=100000, n_r=20000, iterations=iterations, seed=seed,
                           subpartition_dimension=subpartition_dimension, rate=1e8, down_fraction=0.0,
                           n_t=10000, t_action=1)


def crowded_spheres(n_molecules: int = 10_000_000, n_surface: int = 200_000, iterations: int = 200, seed: int = 5,
                    num_spheres: int = 64, sphere_subdivisions: int = 6, r_min: float = 0.1, r_max: float = 0.3,
                    perturb: float = 0.02, cube_um: float = 4.0, subpartition_dimension: float = 0.05) -> Workload:
    """Large crowded geometry: `num_spheres` icospheres (API subdivisions 6:
    20,480 triangles each) with radii ~ U[r_min, r_max] um, centres
    rejection-sampled so that no two overlap and all lie inside the
    reflective `cube_um` cube, every vertex moved radially by a factor
    1 + U[-perturb, perturb] (non-convex surfaces; one PCG64(seed) stream for
    centres, radii and perturbations).  n_molecules volume molecules A:B:E =
    5:4:1 (D 1e-6, 1e-6, 2e-7 cm^2/s) uniform in the cube outside the spheres
    (seed + 1); A + B -> E at 1e8 /M/s, E -> A + B at 1e3 /s; n_surface
    surface molecules S' on the spheres drawn uniformly by area (seed + 2) and
    A + S' -> AS' at 5e7 /M/s.  Species A, B, E, S, AS; ids: volume molecules
    in species order, then S."""
    cfg = M.Config(time_step=1e-6, subpartition_dimension=subpartition_dimension, seed=seed)
    h = cube_um / 2
    box = M.create_box("box", edge_dimension=cube_um)
    unit = M.create_icosphere("unit", 1.0, sphere_subdivisions)
    uv, uw = unit.vertices(), unit.walls()
    rng = np.random.Generator(np.random.PCG64(seed))
    centres, radii = [], []
    margin = 1.0 + perturb
    while len(radii) < num_spheres:
        r = rng.uniform(r_min, r_max)
        c = rng.uniform(-h + r * margin, h - r * margin, size=3)
        if all(np.linalg.norm(c - c2) > (r + r2) * margin for c2, r2 in zip(centres, radii)):
            centres.append(c)
            radii.append(r)
    objects = [box]
    for i, (c, r) in enumerate(zip(centres, radii)):
        scale = r * (1.0 + rng.uniform(-perturb, perturb, size=len(uv)))
        objects.append(M.GeometryObject("sphere%d" % i, c + uv * scale[:, None], uw))
    part = M.derive_partition(cfg, M.bounding_box(objects), has_bimol_vol_rxns=True)
    verts, walls, classes = M.vertices_internal(objects, part)
    species = [M.Species("A", 1e-6), M.Species("B", 1e-6), M.Species("E", 2e-7), M.Species("S", 0.0),
               M.Species("AS", 0.0)]
    rxns = [M.BimolReaction(0, 1, 1e8, [2])]
    unimol = [M.UnimolReaction(2, 1e3, [0, 1])]
    vs = [M.VolSurfReaction(0, 3, 5e7, 4, product_orient=1, geom_vol=0, geom_surf=1)]
    # molecules: uniform in the cube, outside every sphere's perturbed extent
    prng = np.random.Generator(np.random.PCG64(seed + 1))
    cs, rs = np.array(centres), np.array(radii) * margin
    pos = np.empty((0, 3))
    while len(pos) < n_molecules:
        p = prng.uniform(-h, h, size=(min(4_000_000, 2 * (n_molecules - len(pos)) + 16), 3))
        inside = np.zeros(len(p), bool)
        order = np.argsort(p[:, 0], kind="stable")
        xs = p[order, 0]
        for c, r in zip(cs, rs):
            lo, hi = np.searchsorted(xs, [c[0] - r, c[0] + r])
            idx = order[lo:hi]
            inside[idx] |= ((p[idx] - c) ** 2).sum(axis=1) <= r * r
        pos = np.concatenate([pos, p[~inside]])
    pos = pos[:n_molecules] * part.rcp_length_unit
    n_a, n_b = n_molecules * 5 // 10, n_molecules * 4 // 10
    n_e = n_molecules - n_a - n_b
    sp = np.concatenate([np.zeros(n_a, np.uint32), np.ones(n_b, np.uint32), np.full(n_e, 2, np.uint32)])
    sphere_walls = np.arange(len(box.wall_list), len(walls), dtype=np.uint32)
    sw, st = place_on_walls(n_surface, sphere_walls, verts, walls, seed + 2)
    surf = {"ids": np.arange(n_molecules, n_molecules + n_surface, dtype=np.uint32),
            "species": np.full(n_surface, 3, np.uint32), "walls": sw, "tiles": st,
            "orients": np.ones(n_surface, np.int32), "replica": np.zeros(n_surface, np.uint32)}
    steps = _species_table(species, part, rxns)
    steps[0] = (steps[0][0], steps[0][1], steps[0][2] | 4)
    return Workload("crowded_spheres", part, verts, walls, classes, species, steps,
                    np.arange(n_molecules, dtype=np.uint32), sp, pos, np.zeros(n_molecules, np.uint32), iterations,
                    rxns=rxns, config=cfg, objects=objects, barrier_interval=10, unimol_rxns=unimol,
                    volsurf_rxns=vs, surface_mols=surf,
                    spheres=np.concatenate([np.array(centres), np.array(radii)[:, None]], axis=1))


def decay_box(n_e: int = 20000, iterations: int = 2000, seed: int = 6, rate: float = 1e3,
              d_e: float = 2e-7, back_rate: float = 0.0, products=(0, 1), more_pathways=()) -> Workload:
    """E -> A + B at `rate` /s in the reflective 1 um cube (E with D = d_e
    cm^2/s, A and B with 1e-6), optionally A + B -> E at `back_rate` /M/s;
    E uniform (seed), counts every 10 iterations.  ``products`` replaces
    A + B (E among them keeps E: E -> E + A); ``more_pathways`` adds
    (rate, products) pathways of E."""
    cfg = M.Config(time_step=1e-6, subpartition_dimension=0.05, seed=seed)
    box = M.create_box("box", edge_dimension=1.0)
    rxns = [M.BimolReaction(0, 1, back_rate, [2])] if back_rate > 0 else []
    part = M.derive_partition(cfg, M.bounding_box([box]), has_bimol_vol_rxns=bool(rxns))
    verts, walls, classes = M.vertices_internal([box], part)
    species = [M.Species("A", 1e-6), M.Species("B", 1e-6), M.Species("E", d_e)]
    pos = M.place_uniform_in_box(n_e, 0.5, seed, part)
    return Workload("decay_box", part, verts, walls, classes, species, _species_table(species, part, rxns),
                    np.arange(n_e, dtype=np.uint32), np.full(n_e, 2, np.uint32), pos, np.zeros(n_e, np.uint32),
                    iterations, rxns=rxns, config=cfg, objects=[box], barrier_interval=10,
                    unimol_rxns=[M.UnimolReaction(2, rate, list(products))] +
                    [M.UnimolReaction(2, r, list(pr)) for r, pr in more_pathways])
