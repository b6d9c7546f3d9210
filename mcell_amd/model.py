"""Host-side model description: the pieces of MCell4's libmcell API that feed
the diffuse-react path.

This mirrors, in plain Python floats (IEEE double, same operation order), what
the reference computes before the first DiffuseReactEvent runs:

* geometry builders ``create_box`` / ``create_icosphere`` with the reference's
  vertex and face order (libmcell/api/geometry_utils.cpp:32-266),
* unit conversion and partition sizing of ``MCell4Converter``
  (libmcell/api/mcell4_converter.cpp:237-393) and
  ``SimulationConfig::init_subpartition_edge_length``
  (src4/simulation_config.cpp:45-70),
* species space/time steps (MCell3 closed form, src/mcell_species.c:237-300),
* the vol-vol reaction probability factor (src/react_util.c:163-181).

The results are handed to the engine's C ABI (include/mcg.h) in the
reference's internal units.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

MY_PI = 3.14159265358979323846
N_AV = 6.0221417930e23
EPS = 1e-12
PARTITION_EDGE_EXTRA_MARGIN_UM = 0.01
MAX_SUBPARTS_PER_PARTITION = 300
POS_SQRT2 = 1.41421356238  # src4/defines.h:148


# --------------------------------------------------------------- geometry ---
@dataclass
class GeometryObject:
    """Vertices in um and triangles (vertex index triples), reference order."""

    name: str
    vertex_list: List[Tuple[float, float, float]]
    wall_list: List[Tuple[int, int, int]]
    surface_class: int = 0  # 0 = reflective

    def vertices(self) -> np.ndarray:
        return np.asarray(self.vertex_list, dtype=np.float64).reshape(-1, 3)

    def walls(self) -> np.ndarray:
        return np.asarray(self.wall_list, dtype=np.uint32).reshape(-1, 3)


def create_box(name: str, edge_dimension: Optional[float] = None,
               xyz_dimensions: Optional[Sequence[float]] = None) -> GeometryObject:
    """geometry_utils::create_box (libmcell/api/geometry_utils.cpp:32-94)."""
    if (edge_dimension is None) == (xyz_dimensions is None):
        raise ValueError("Exactly one edge_dimension or individual x, y, z dimensions can be set.")
    if edge_dimension is not None:
        x = y = z = edge_dimension / 2
    else:
        if len(xyz_dimensions) != 3:
            raise ValueError("xyz_dimensions must have 3 entries")
        x, y, z = (xyz_dimensions[0] / 2, xyz_dimensions[1] / 2, xyz_dimensions[2] / 2)
    verts = [(-x, -y, -z), (-x, -y, z), (-x, y, -z), (-x, y, z),
             (x, -y, -z), (x, -y, z), (x, y, -z), (x, y, z)]
    # CellBlender face order; surface releases depend on it
    faces = [(1, 2, 0), (3, 6, 2), (7, 4, 6), (5, 0, 4), (6, 0, 2), (3, 5, 7),
             (1, 3, 2), (3, 7, 6), (7, 5, 4), (5, 1, 0), (6, 4, 0), (3, 1, 5)]
    return GeometryObject(name, verts, faces)


def _normalize(v):
    # geometry_utils.cpp:120-124: lrcp = 1/sqrt(dot(a, a)), then a * lrcp
    d = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]
    lrcp = 1.0 / math.sqrt(d)
    return (v[0] * lrcp, v[1] * lrcp, v[2] * lrcp)


def create_icosphere(name: str, radius: float, subdivisions: int) -> GeometryObject:
    """geometry_utils::create_icosphere (geometry_utils.cpp:127-266).

    ``subdivisions`` is the API parameter: faces = 20 * 4**(subdivisions - 1).
    """
    if subdivisions < 1 or subdivisions > 8:
        raise ValueError("subdivisions must be in 1..8")
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [_normalize(v) for v in [
        (-1.0, t, 0.0), (1.0, t, 0.0), (-1.0, -t, 0.0), (1.0, -t, 0.0),
        (0.0, -1.0, t), (0.0, 1.0, t), (0.0, -1.0, -t), (0.0, 1.0, -t),
        (t, 0.0, -1.0), (t, 0.0, 1.0), (-t, 0.0, -1.0), (-t, 0.0, 1.0)]]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
             (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
             (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
             (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions - 1):
        out_v = list(verts)
        divisions = {}

        def edge_vertex(f0, f1, v0, v1):
            key = (f0, f1) if f0 < f1 else (f1, f0)
            if key in divisions:
                return divisions[key]
            # normalize(Vec3(0.5) * (v0 + v1))
            m = (0.5 * (v0[0] + v1[0]), 0.5 * (v0[1] + v1[1]), 0.5 * (v0[2] + v1[2]))
            out_v.append(_normalize(m))
            divisions[key] = len(out_v) - 1
            return divisions[key]

        out_f = []
        for (f0, f1, f2) in faces:
            v0, v1, v2 = verts[f0], verts[f1], verts[f2]
            f3 = edge_vertex(f0, f1, v0, v1)
            f4 = edge_vertex(f1, f2, v1, v2)
            f5 = edge_vertex(f2, f0, v2, v0)
            out_f += [(f0, f3, f5), (f3, f1, f4), (f4, f2, f5), (f3, f4, f5)]
        verts, faces = out_v, out_f
    verts = [(v[0] * radius, v[1] * radius, v[2] * radius) for v in verts]
    return GeometryObject(name, verts, faces)


def translate(obj: GeometryObject, offset: Sequence[float]) -> GeometryObject:
    """GeometryObject.translate: adds the offset to every vertex (um)."""
    verts = [(v[0] + offset[0], v[1] + offset[1], v[2] + offset[2]) for v in obj.vertex_list]
    return GeometryObject(obj.name, verts, list(obj.wall_list), obj.surface_class)


# ----------------------------------------------------------------- config ---
@dataclass
class Config:
    """libmcell Config fields on this path (simulation_setup.yaml defaults)."""

    time_step: float = 1e-6
    surface_grid_density: float = 10000.0
    interaction_radius: Optional[float] = None
    partition_dimension: float = 10.0
    subpartition_dimension: float = 0.5
    seed: int = 1


@dataclass
class PartitionConfig:
    """SimulationConfig + Partition corner values in internal units."""

    time_unit: float
    length_unit: float
    rcp_length_unit: float
    grid_density: float
    rxn_radius_3d: float
    partition_edge_length: float
    origin: Tuple[float, float, float]
    subpart_edge_length: float
    subpart_edge_length_rcp: float
    num_subparts_per_edge: int
    use_expanded_list: int

    @property
    def opposite(self):
        return tuple(o + self.partition_edge_length for o in self.origin)


def _floor_to_multiple_allow_negative(val, multiple):
    if val >= 0:
        return float(int((val + EPS) / multiple)) * multiple
    return float(int((val + EPS - multiple) / multiple)) * multiple


def _ceil_to_multiple(val, multiple):
    res = float(int((val + EPS) / multiple)) * multiple
    if not abs(val - res) < EPS:
        res += multiple
    return res


def _c_round(x):
    return float(math.floor(x + 0.5)) if x >= 0 else -float(math.floor(-x + 0.5))


def derive_partition(cfg: Config, bbox_um, has_bimol_vol_rxns: bool) -> PartitionConfig:
    """MCell4Converter::convert_simulation_setup (mcell4_converter.cpp:237-393)
    and SimulationConfig::init (simulation_config.cpp:45-70).

    ``bbox_um`` = (llf, urb) over all geometry vertices in um.
    """
    time_unit = cfg.time_step
    grid_density = cfg.surface_grid_density
    length_unit = 1 / math.sqrt(grid_density)
    rcp_length_unit = 1 / length_unit
    if cfg.interaction_radius is not None:
        rxn_radius_3d = cfg.interaction_radius / length_unit
    else:
        rxn_radius_3d = (1.0 / math.sqrt(MY_PI * grid_density)) / length_unit
    llf, urb = bbox_um
    llf_m = [v - PARTITION_EDGE_EXTRA_MARGIN_UM for v in llf]
    urb_m = [v + PARTITION_EDGE_EXTRA_MARGIN_UM for v in urb]
    max3 = lambda v: max(max(v[0], v[1]), v[2])
    min3 = lambda v: min(min(v[0], v[1]), v[2])
    auto_dim = max3(urb_m) - min3(llf_m)
    auto_origin = False
    if auto_dim > cfg.partition_dimension:
        pel = auto_dim / length_unit
        origin = [(v - PARTITION_EDGE_EXTRA_MARGIN_UM) / length_unit for v in llf]
        auto_origin = True
    else:
        pel = cfg.partition_dimension / length_unit
    if not auto_origin:
        origin = [(-pel) / 2] * 3
    sp_len = cfg.subpartition_dimension / length_unit
    if int(pel / sp_len) > MAX_SUBPARTS_PER_PARTITION:
        sp_len = pel / MAX_SUBPARTS_PER_PARTITION
    orig = list(origin)
    origin = [_floor_to_multiple_allow_negative(v, sp_len) for v in orig]
    moved = [a - b for a, b in zip(orig, origin)]
    pel = _ceil_to_multiple(pel + max3(moved), sp_len)
    n = int(_c_round(pel / sp_len))
    sp_edge = pel / float(n)
    rcp = 1.0 / sp_edge
    pel = sp_edge * n
    # MCell4Converter's check (mcell4_converter.cpp:121-127)
    if not rxn_radius_3d * POS_SQRT2 < sp_edge / 2:
        raise ValueError("Reaction radius multiplied by sqrt(2) %g must be less than half of subpartition edge "
                         "length %g" % (rxn_radius_3d * length_unit * POS_SQRT2, sp_edge * length_unit / 2))
    return PartitionConfig(time_unit, length_unit, rcp_length_unit, grid_density, rxn_radius_3d, pel,
                           tuple(origin), sp_edge, rcp, n, 1 if has_bimol_vol_rxns else 0)


def bounding_box(objects: Sequence[GeometryObject]):
    """MCell4Converter::get_geometry_bounding_box (mcell4_converter.cpp:176-208)."""
    llf = [1e140] * 3
    urb = [-1e140] * 3
    for o in objects:
        for v in o.vertex_list:
            for k in range(3):
                if v[k] < llf[k]:
                    llf[k] = v[k]
                if v[k] > urb[k]:
                    urb[k] = v[k]
    return llf, urb


# ---------------------------------------------------------------- species ---
@dataclass
class Species:
    name: str
    diffusion_constant_3d: float  # cm^2/s
    target_only: bool = False
    custom_time_step: float = 0.0  # s (0: the global time step)
    surface: bool = False          # a surface species (its molecules sit on wall tiles)

    def steps(self, part: PartitionConfig):
        """(space_step, time_step) in internal units (mcell_species.c:239-276;
        a custom time step t gives sqrt(4e8 D t) / length_unit and
        t / time_unit, :254-264)."""
        D = self.diffusion_constant_3d
        if abs(D - 0) <= EPS * max(1.0, abs(D)):
            return 0.0, 1.0
        if self.custom_time_step > 0:
            t = self.custom_time_step
            return math.sqrt(4.0 * 1.0e8 * D * t) * part.rcp_length_unit, t / part.time_unit
        return math.sqrt(4.0 * 1.0e8 * D * part.time_unit) * part.rcp_length_unit, 1.0


@dataclass
class BimolReaction:
    """A + B -> products for two volume species (rate in M^-1 s^-1)."""

    reactant_a: int
    reactant_b: int
    rate: float
    products: List[int] = field(default_factory=list)

    def pb_factor(self, part: PartitionConfig, sp_a: Species, sp_b: Species) -> float:
        """Vol-vol pb_factor (src/react_util.c:163-181), radius in um."""
        ss_a, ts_a = sp_a.steps(part)
        ss_b, ts_b = sp_b.steps(part)
        eff_vel = (ss_a / ts_a + ss_b / ts_b) * part.length_unit / part.time_unit
        r_um = part.rxn_radius_3d * part.length_unit
        pb = 1.0 / (2.0 * math.sqrt(MY_PI) * r_um * r_um * eff_vel)
        pb *= 1.0e15 / N_AV
        return pb


@dataclass
class UnimolReaction:
    """reactant -> products for a volume species (rate in s^-1); the
    per-step probability rate is rate * time_unit (pb = time_unit, MCell3
    src/react_util.c:74-78)."""

    reactant: int
    rate: float
    products: List[int] = field(default_factory=list)


@dataclass
class VolSurfReaction:
    """vol + surf -> products (rate in M^-1 s^-1): the volume reactant
    hitting the tile of the surface reactant, which is consumed.  By default
    one surface product that takes the tile with orientation product_orient;
    ``products`` instead lists (species, orientation (+1, -1, 0 none), is
    surface) in the rule's order -- a surface product of the surface
    reactant's species keeps that reactant, volume products start next to
    the wall.  geom_vol / geom_surf are the orientations written in the rule
    (0 none, +1 up ('), -1 down (,))."""

    vol: int
    surf: int
    rate: float
    product: int = -1
    product_orient: int = 1
    geom_vol: int = 0
    geom_surf: int = 0
    products: Optional[List[tuple]] = None

    def product_list(self) -> List[tuple]:
        if self.products is not None:
            return [tuple(p) for p in self.products]
        return [(self.product, self.product_orient, True)]

    def pb_factor(self, part: PartitionConfig, sp_vol: Species) -> float:
        """Vol-surf pb_factor (MCell3 src/react_util.c:132-157): the volume
        reactant's D and time step, doubled for a rule whose two orientations
        are both set and equal or opposite."""
        D = sp_vol.diffusion_constant_3d
        t_step = 1.0 * part.time_unit
        if D <= 0.0:
            return 0.0
        pb = 1.0e11 * part.grid_density / (2.0 * N_AV) * math.sqrt(MY_PI * t_step / D)
        gv, gs = self.geom_vol, self.geom_surf
        if (gv + gs) * (gv - gs) == 0 and gv * gs != 0:
            pb *= 2.0
        return pb


def wall_num_tiles(vertices: np.ndarray, walls: np.ndarray) -> np.ndarray:
    """Tiles per wall of a surface grid (Grid::initialize, src4/wall.cpp:51-56):
    ceil(sqrt(area))^2 with the area in internal units."""
    v0 = vertices[walls[:, 0]]
    e1 = vertices[walls[:, 1]] - v0
    e2 = vertices[walls[:, 2]] - v0
    area = 0.5 * np.linalg.norm(np.cross(e1, e2), axis=1)
    n_axis = np.maximum(np.ceil(np.sqrt(area)), 1).astype(np.int64)
    return n_axis * n_axis


# -------------------------------------------------------------- placement ---
def place_uniform_in_box(n: int, half_edge_um: float, seed: int, part: PartitionConfig) -> np.ndarray:
    """Uniform positions inside the axis-aligned cube [-h, h]^3 (um), returned
    in internal units.  Placement uses numpy's PCG64 seeded with ``seed``."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = rng.uniform(-half_edge_um, half_edge_um, size=(n, 3))
    return p * part.rcp_length_unit


def vertices_internal(objects: Sequence[GeometryObject], part: PartitionConfig):
    """Vertices of all objects in internal units (mcell4_converter.cpp:921-923:
    Vec3(v) * rcp_length_unit) and the walls re-indexed into that list."""
    verts, walls, classes = [], [], []
    base = 0
    for o in objects:
        v = o.vertices() * part.rcp_length_unit
        verts.append(v)
        walls.append(o.walls().astype(np.int64) + base)
        classes.append(np.full(len(o.wall_list), o.surface_class, dtype=np.uint32))
        base += len(v)
    return (np.concatenate(verts).astype(np.float64), np.concatenate(walls).astype(np.uint32),
            np.concatenate(classes))
