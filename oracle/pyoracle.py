"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (liboracle.so).

Imported by tests/, by __graft_entry__.smoke() as the checker, and by the
CPU-baseline leg of bench.py.  Never used on the product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_RNG_PATH = os.path.join(HERE, "_ref", "libref_rng.so")

DRAW_ISAAC = 0
DRAW_PHILOX = 1

_lib = None


def build(quiet=True):
    """Builds liboracle.so (and oracle/_ref when the reference is present)."""
    kw = dict(stdout=subprocess.DEVNULL, stderr=subprocess.STDOUT) if quiet else {}
    subprocess.check_call(["make", "-C", HERE, "-j4"], **kw)
    if os.path.exists("/root/reference/src/rng.h"):
        subprocess.check_call(["make", "-C", HERE, "ref"], **kw)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    d, u32, u64 = ctypes.c_double, ctypes.c_uint32, ctypes.c_uint64
    sig = {
        "orc_error": (ctypes.c_char_p, []),
        "orc_create": (P, [d, d, d, d, d, ctypes.c_int, P]),
        "orc_destroy": (None, [P]),
        "orc_get_config": (None, [P, P, P]),
        "orc_add_vertices_um": (u32, [P, P, u32]),
        "orc_add_walls": (None, [P, P, u32, u32, u32, u32]),
        "orc_finalize_walls": (None, [P]),
        "orc_num_walls": (u32, [P]),
        "orc_get_walls": (None, [P, P]),
        "orc_get_vertices": (None, [P, P]),
        "orc_get_subpart_walls": (u64, [P, P, P]),
        "orc_add_species": (u32, [P, d, ctypes.c_int, u32, d]),
        "orc_get_birthdays": (None, [P, P]),
        "orc_get_species": (None, [P, u32, P]),
        "orc_add_bimol_rxn": (ctypes.c_int, [P, u32, u32, d, P, u32]),
        "orc_get_rxn": (None, [P, ctypes.c_int, P]),
        "orc_add_unimol_rxn": (None, [P, u32, d, P, u32]),
        "orc_unimol_reactions": (u64, [P]),
        "orc_add_volsurf_rxn": (ctypes.c_int, [P, u32, u32, ctypes.c_int, ctypes.c_int, d, u32, P, P, P]),
        "orc_set_counted_volumes": (None, [P, u32, P, P]),
        "orc_set_molecule_volumes": (None, [P, u32, P, P]),
        "orc_get_molecule_volumes": (None, [P, P]),
        "orc_get_volsurf_rxn": (None, [P, ctypes.c_int, P]),
        "orc_add_surface_molecules": (ctypes.c_int, [P, u32, u32, P, P, P, P]),
        "orc_get_surface_state": (None, [P, P, P, P]),
        "orc_surface_stats": (None, [P, P]),
        "orc_surface_net": (ctypes.c_int, [P, P, P]),
        "orc_edge_point": (None, [P, u32, P, P, P]),
        "orc_uv2grid": (u32, [P, u32, d, d]),
        "orc_grid2uv": (None, [P, u32, u32, P]),
        "orc_grid_info": (None, [P, u32, P]),
        "orc_xyz2grid": (ctypes.c_int64, [P, u32, P]),
        "orc_get_tiles": (None, [P, u32, P]),
        "orc_volsurf_reactions": (u64, [P]),
        "orc_add_molecules": (ctypes.c_int, [P, u32, u32, P]),
        "orc_set_rng": (None, [P, ctypes.c_int, u64, u32]),
        "orc_set_snapshot": (None, [P, ctypes.c_int]),
        "orc_set_trace": (None, [P, ctypes.c_int]),
        "orc_set_barrier_interval": (None, [P, u32]),
        "orc_set_barriers": (None, [P, P, u32]),
        "orc_get_rxn_log": (u64, [P, P, P, P]),
        "orc_clear_rxn_log": (None, [P]),
        "orc_exact_disk": (d, [P, P, P, u32]),
        "orc_set_surface_classes": (None, [P, u32, P]),
        "orc_set_wall_classes": (None, [P, u32, u32, u32]),
        "orc_get_wall_stats": (None, [P, P]),
        "orc_run": (ctypes.c_int, [P, u32, P]),
        "orc_iteration": (u64, [P]),
        "orc_num_molecules": (u32, [P]),
        "orc_get_molecules": (None, [P, P, P, P, P, P]),
        "orc_get_hits": (u64, [P, P, P]),
        "orc_clear_hits": (None, [P]),
        "orc_get_stats": (None, [P, P]),
        "orc_isaac64_u32": (None, [u32, u64, P]),
        "orc_gauss_isaac": (None, [u32, u64, P]),
        "orc_philox_block": (None, [P, P, P]),
        "orc_gauss_from_tape": (ctypes.c_int64, [P, u64, u64, P]),
        "orc_philox_stream": (None, [u64, u32, u32, u32, u32, u64, P]),
        "orc_collide_wall": (ctypes.c_int, [P, P, u32, P, P, P]),
        "orc_crossed_subparts": (u32, [P, P, P, P, u32]),
        "orc_subpart_index": (u32, [P, P]),
        "orc_collide_wall_rays": (None, [P, u32, u64, P, P, P, P, P]),
        "orc_test_bimolecular_u": (None, [d, P, u32, d, u64, P, P]),
        "orc_get_wall_vertices": (None, [P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def ref_rng():
    """The reference's own RNG compiled from /root/reference (or None)."""
    if not os.path.exists(REF_RNG_PATH):
        return None
    lib = ctypes.CDLL(REF_RNG_PATH)
    for name in ("ref_isaac64_u32", "ref_isaac64_dbl", "ref_rng_gauss"):
        getattr(lib, name).restype = None
        getattr(lib, name).argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p]
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class OracleWorld:
    """One oracle World (a partition with geometry, species and molecules)."""

    def __init__(self, cfg, bbox_um, has_bimol):
        self.lib = load()
        bb = np.asarray(list(bbox_um[0]) + list(bbox_um[1]), np.float64)
        ir = -1.0 if cfg.interaction_radius is None else cfg.interaction_radius
        self.w = self.lib.orc_create(cfg.time_step, cfg.surface_grid_density, ir, cfg.partition_dimension,
                                     cfg.subpartition_dimension, 1 if has_bimol else 0, _p(bb))

    def close(self):
        if self.w:
            self.lib.orc_destroy(self.w)
            self.w = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config(self):
        d = np.zeros(14)
        u = np.zeros(4, np.uint32)
        self.lib.orc_get_config(self.w, _p(d), _p(u))
        keys = ["time_unit", "length_unit", "rcp_length_unit", "grid_density", "rxn_radius_3d",
                "partition_edge_length", "ox", "oy", "oz", "qx", "qy", "qz", "subpart_edge_length",
                "subpart_edge_length_rcp"]
        out = dict(zip(keys, d.tolist()))
        out.update(n=int(u[0]), n2=int(u[1]), n3=int(u[2]), use_expanded_list=int(u[3]))
        return out

    def add_object(self, obj, object_id=0):
        v = np.ascontiguousarray(obj.vertices(), np.float64)
        t = np.ascontiguousarray(obj.walls(), np.uint32)
        first = self.lib.orc_add_vertices_um(self.w, _p(v), len(v))
        self.lib.orc_add_walls(self.w, _p(t), len(t), first, object_id, obj.surface_class)

    def set_surface_classes(self, action):
        """action: uint8 array [num_classes, num_species], front | back << 2."""
        a = np.ascontiguousarray(action, np.uint8)
        self.lib.orc_set_surface_classes(self.w, a.shape[0], _p(a))

    def wall_stats(self):
        out = np.zeros(2, np.uint64)
        self.lib.orc_get_wall_stats(self.w, _p(out))
        return {"transparent_crossings": int(out[0]), "absorbed": int(out[1])}

    def finalize_walls(self):
        self.lib.orc_finalize_walls(self.w)

    def walls(self):
        n = self.lib.orc_num_walls(self.w)
        out = np.zeros((n, 16))
        self.lib.orc_get_walls(self.w, _p(out))
        return out

    def vertices(self, n):
        out = np.zeros((n, 3))
        self.lib.orc_get_vertices(self.w, _p(out))
        return out

    def subpart_walls(self, n3):
        cnt = self.lib.orc_get_subpart_walls(self.w, None, None)
        off = np.zeros(n3 + 1, np.uint32)
        idx = np.zeros(max(cnt, 1), np.uint32)
        self.lib.orc_get_subpart_walls(self.w, _p(off), _p(idx))
        return off, idx[:cnt]

    def add_species(self, D, is_vol=True, flags=0, custom_time_step=0.0):
        """custom_time_step: seconds (0: the global time step)."""
        return self.lib.orc_add_species(self.w, D, 1 if is_vol else 0, flags, custom_time_step)

    def species(self, s):
        out = np.zeros(4)
        self.lib.orc_get_species(self.w, s, _p(out))
        return out

    def add_bimol_rxn(self, a, b, rate, products):
        p = np.asarray(products, np.uint32)
        return self.lib.orc_add_bimol_rxn(self.w, a, b, rate, _p(p), len(p))

    def add_unimol_rxn(self, species, rate, products):
        p = np.ascontiguousarray(products, np.uint32)
        self.lib.orc_add_unimol_rxn(self.w, species, rate, _p(p) if len(p) else None, len(p))

    def add_volsurf_rxn(self, vol, surf, geom_vol, geom_surf, rate, products):
        """products: [(species, orientation, is_surface), ...] in the rule's order."""
        sp = np.ascontiguousarray([p[0] for p in products] or [0], np.uint32)
        orient = np.ascontiguousarray([p[1] for p in products] or [0], np.int32)
        surf_ = np.ascontiguousarray([1 if p[2] else 0 for p in products] or [0], np.uint32)
        return self.lib.orc_add_volsurf_rxn(self.w, vol, surf, geom_vol, geom_surf, rate, len(products), _p(sp),
                                            _p(orient), _p(surf_))

    def volsurf_rxn(self, r):
        out = np.zeros(2)
        self.lib.orc_get_volsurf_rxn(self.w, r, _p(out))
        return out

    def add_surface_molecules(self, species, walls, tiles, orients, uv=None):
        """uv: (n, 2) positions on the walls, or None for the tiles' centres."""
        wa = np.ascontiguousarray(walls, np.uint32)
        ta = np.ascontiguousarray(tiles, np.uint32)
        oa = np.ascontiguousarray(orients, np.int32)
        ua = None if uv is None else np.ascontiguousarray(uv, np.float64).reshape(-1, 2)
        if self.lib.orc_add_surface_molecules(self.w, species, len(wa), _p(wa), _p(ta), _p(oa),
                                              None if ua is None else _p(ua)) != 0:
            raise RuntimeError(self.lib.orc_error().decode())

    def surface_state(self):
        """Per molecule id: wall, tile (0xFFFFFFFF for volume molecules) and uv."""
        n = self.lib.orc_num_molecules(self.w)
        wa = np.zeros(n, np.uint32)
        ta = np.zeros(n, np.uint32)
        uv = np.zeros((n, 2))
        self.lib.orc_get_surface_state(self.w, _p(wa), _p(ta), _p(uv))
        return wa, ta, uv

    def surface_stats(self):
        out = np.zeros(2, np.uint64)
        self.lib.orc_surface_stats(self.w, _p(out))
        return {"moves_between_walls": int(out[0]), "retries": int(out[1])}

    def surface_net(self, num_walls):
        """Per wall and edge: the wall across (-1: none) and the transform
        (cos, sin, translate u, translate v, forward flag)."""
        nb = np.zeros((num_walls, 3), np.uint32)
        tr = np.zeros((num_walls, 3, 5))
        if self.lib.orc_surface_net(self.w, _p(nb), _p(tr)) != 0:
            raise RuntimeError(self.lib.orc_error().decode())
        nbi = nb.astype(np.int64)
        nbi[nb == 0xFFFFFFFF] = -1
        return nbi, tr

    def edge_point(self, wall, loc, disp):
        out = np.zeros(6)
        self.lib.orc_edge_point(self.w, wall, _p(np.asarray(loc, np.float64)), _p(np.asarray(disp, np.float64)),
                                _p(out))
        return int(out[0]), out[1:3], int(out[3]), out[4:6]

    def uv2grid(self, wall, u, v):
        return self.lib.orc_uv2grid(self.w, wall, u, v)

    def grid2uv(self, wall, tile):
        out = np.zeros(2)
        self.lib.orc_grid2uv(self.w, wall, tile, _p(out))
        return out

    def grid_info(self, wall):
        out = np.zeros(8)
        self.lib.orc_grid_info(self.w, wall, _p(out))
        return out

    def xyz2grid(self, wall, pos):
        return int(self.lib.orc_xyz2grid(self.w, wall, _p(np.asarray(pos, np.float64))))

    def tiles(self, wall):
        n = int(self.grid_info(wall)[1])
        out = np.zeros(n, np.uint32)
        self.lib.orc_get_tiles(self.w, wall, _p(out))
        return out

    def volsurf_reactions(self):
        return int(self.lib.orc_volsurf_reactions(self.w))

    def unimol_reactions(self):
        return int(self.lib.orc_unimol_reactions(self.w))

    def rxn(self, r):
        out = np.zeros(2)
        self.lib.orc_get_rxn(self.w, r, _p(out))
        return out

    def add_molecules(self, species, pos_lu):
        pos = np.ascontiguousarray(pos_lu, np.float64).reshape(-1, 3)
        if self.lib.orc_add_molecules(self.w, species, len(pos), _p(pos)) != 0:
            raise RuntimeError(self.lib.orc_error().decode())

    def set_rng(self, mode, seed, replica=0):
        self.lib.orc_set_rng(self.w, mode, seed, replica)

    def set_snapshot(self, on):
        self.lib.orc_set_snapshot(self.w, 1 if on else 0)

    def set_trace(self, on):
        self.lib.orc_set_trace(self.w, 1 if on else 0)

    def set_barrier_interval(self, b):
        self.lib.orc_set_barrier_interval(self.w, b)

    def set_barriers(self, barriers):
        """Scheduler barriers at the listed iterations (ascending) instead of
        every barrier_interval iterations."""
        a = np.ascontiguousarray(sorted(barriers), dtype=np.uint64)
        self.lib.orc_set_barriers(self.w, a.ctypes.data_as(ctypes.c_void_p), len(a))

    def run(self, n):
        secs = ctypes.c_double()
        if self.lib.orc_run(self.w, n, ctypes.byref(secs)) != 0:
            raise RuntimeError(self.lib.orc_error().decode())
        return secs.value

    def set_counted_volumes(self, wall_inside, wall_outside):
        wi = np.ascontiguousarray(wall_inside, np.uint32)
        wo = np.ascontiguousarray(wall_outside, np.uint32)
        self.lib.orc_set_counted_volumes(self.w, len(wi), _p(wi), _p(wo))

    def set_molecule_volumes(self, ids, cvols):
        ids = np.ascontiguousarray(ids, np.uint32)
        cv = np.ascontiguousarray(cvols, np.uint32)
        self.lib.orc_set_molecule_volumes(self.w, len(ids), _p(ids), _p(cv))

    def molecule_volumes(self):
        n = self.lib.orc_num_molecules(self.w)
        out = np.zeros(max(n, 1), np.uint32)
        self.lib.orc_get_molecule_volumes(self.w, _p(out))
        return out[:n]

    def molecules(self):
        n = self.lib.orc_num_molecules(self.w)
        sp = np.zeros(n, np.uint32)
        fl = np.zeros(n, np.uint32)
        pos = np.zeros((n, 3))
        sub = np.zeros(n, np.uint32)
        dt = np.zeros(n)
        self.lib.orc_get_molecules(self.w, _p(sp), _p(fl), _p(pos), _p(sub), _p(dt))
        bd = np.zeros(n)
        self.lib.orc_get_birthdays(self.w, _p(bd))
        return {"species": sp, "flags": fl, "pos": pos, "subpart": sub, "diffusion_time": dt, "birthday": bd}

    def hits(self):
        n = self.lib.orc_get_hits(self.w, None, None)
        u = np.zeros((max(n, 1), 5), np.uint32)
        t = np.zeros(max(n, 1))
        self.lib.orc_get_hits(self.w, _p(u), _p(t))
        return u[:n], t[:n]

    def clear_hits(self):
        self.lib.orc_clear_hits(self.w)

    def rxn_log(self):
        n = self.lib.orc_get_rxn_log(self.w, None, None, None)
        u = np.zeros((max(n, 1), 4), np.uint32)
        rx = np.zeros(max(n, 1), np.int32)
        d = np.zeros((max(n, 1), 4))
        self.lib.orc_get_rxn_log(self.w, _p(u), _p(rx), _p(d))
        return u[:n], rx[:n], d[:n]

    def stats(self):
        out = np.zeros(9, np.uint64)
        self.lib.orc_get_stats(self.w, _p(out))
        keys = ["ray_voxel_tests", "ray_polygon_tests", "ray_polygon_colls", "mol_wall_reflections",
                "vol_mol_vol_mol_collisions", "diffuse_3d_calls", "redo_count", "isaac_draws", "partition_clips"]
        return dict(zip(keys, out.tolist()))

    def collide_wall(self, pos, wall, move):
        pos = np.asarray(pos, np.float64)
        move = np.asarray(move, np.float64)
        t = np.zeros(1)
        hit = np.zeros(3)
        ct = self.lib.orc_collide_wall(self.w, _p(pos), wall, _p(move), _p(t), _p(hit))
        return ct, float(t[0]), hit

    def crossed_subparts(self, pos, disp, cap=4096):
        pos = np.asarray(pos, np.float64)
        disp = np.asarray(disp, np.float64)
        out = np.zeros(cap, np.uint32)
        n = self.lib.orc_crossed_subparts(self.w, _p(pos), _p(disp), _p(out), cap)
        return out[:min(n, cap)]

    def collide_wall_rays(self, wall, pos, move):
        """collide_wall with the ray trace's REDO handling: types (0 MISS,
        1 FRONT, 2 BACK, 3 REDO), t and hit point per ray."""
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        move = np.ascontiguousarray(move, np.float64).reshape(-1, 3)
        n = len(pos)
        types = np.zeros(n, np.int32)
        t = np.zeros(n)
        hit = np.zeros((n, 3))
        self.lib.orc_collide_wall_rays(self.w, wall, n, _p(pos), _p(move), _p(types), _p(t), _p(hit))
        return types, t, hit

    def wall_vertex_indices(self):
        out = np.zeros((self.lib.orc_num_walls(self.w), 3), np.uint32)
        self.lib.orc_get_wall_vertices(self.w, _p(out))
        return out

    def subpart_index(self, pos):
        return self.lib.orc_subpart_index(self.w, _p(np.asarray(pos, np.float64)))


def format_count(d, form):
    """C++ iostream's text for d: form 0 a default stream, 1 std::scientific
    at precision 8."""
    lib = load()
    buf = ctypes.create_string_buffer(64)
    lib.orc_format_count.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    n = lib.orc_format_count(d, form, buf, 64)
    assert n >= 0
    return buf.value.decode()


def test_bimolecular_u(max_fixed_p, cum_probs, scaling, u):
    """The oracle's test_bimolecular for given draws u: pathway or -1."""
    lib = load()
    cum = np.ascontiguousarray(cum_probs, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    out = np.zeros(len(u), np.int32)
    lib.orc_test_bimolecular_u(max_fixed_p, _p(cum), len(cum), scaling, len(u), _p(u), _p(out))
    return out


def world_from_workload(wl, draw_mode=DRAW_PHILOX, seed=None, replica=0):
    """Builds an OracleWorld with the same inputs as ``wl`` for one replica:
    geometry (um vertices, converted by the oracle itself), species,
    reactions, and that replica's molecules."""
    from mcell_amd import model as M
    has_bimol = len(wl.rxns) > 0
    w = OracleWorld(wl.config, M.bounding_box(wl.objects), has_bimol)
    for k, obj in enumerate(wl.objects):
        w.add_object(obj, k)
    w.finalize_walls()
    species = wl.species_for(replica) if hasattr(wl, "species_for") else wl.species
    for s in species:
        w.add_species(s.diffusion_constant_3d, is_vol=not getattr(s, "surface", False),
                      custom_time_step=getattr(s, "custom_time_step", 0.0))
    for r in wl.rxns:
        w.add_bimol_rxn(r.reactant_a, r.reactant_b, r.rate, r.products)
    for u in getattr(wl, "unimol_rxns", []) or []:
        w.add_unimol_rxn(u.reactant, u.rate, u.products)
    for v in getattr(wl, "volsurf_rxns", []) or []:
        w.add_volsurf_rxn(v.vol, v.surf, v.geom_vol, v.geom_surf, v.rate, v.product_list())
    if getattr(wl, "surface_classes", None) is not None:
        w.set_surface_classes(wl.surface_classes)
    sel = wl.replica == replica
    ids = wl.ids[sel]
    assert np.array_equal(ids, np.arange(len(ids))), "oracle ids are insertion order"
    sp = wl.mol_species[sel]
    pos = wl.positions[sel]
    # molecules are added in id order; consecutive runs of one species at once
    start = 0
    while start < len(ids):
        end = start
        while end < len(ids) and sp[end] == sp[start]:
            end += 1
        w.add_molecules(int(sp[start]), pos[start:end])
        start = end
    cv = getattr(wl, "counted_volumes", None)
    if cv is not None:
        w.set_counted_volumes(cv["wall_inside"], cv["wall_outside"])
        w.set_molecule_volumes(ids, cv["molecules"][sel])
    sm = getattr(wl, "surface_mols", None)
    if sm is not None:
        ssel = np.nonzero(sm["replica"] == replica)[0]
        assert np.array_equal(sm["ids"][ssel], len(ids) + np.arange(len(ssel))), "surface ids follow the volume ids"
        start = 0
        while start < len(ssel):
            end = start
            while end < len(ssel) and sm["species"][ssel[end]] == sm["species"][ssel[start]]:
                end += 1
            k = ssel[start:end]
            w.add_surface_molecules(int(sm["species"][k[0]]), sm["walls"][k], sm["tiles"][k], sm["orients"][k],
                                    uv=sm["uv"][k] if "uv" in sm else None)
            start = end
    w.set_rng(draw_mode, wl.config.seed if seed is None else seed, replica)
    return w
