"""Python side of the diffuse-react engine: a thin owner of an ``mcg_ctx``.

This plays the role MCell4's DiffuseReactEvent plays for World: it is created
once per partition (src4/world.cpp:280-282), receives the flattened
geometry/species tables, and is stepped once per iteration
(DiffuseReactEvent::step, src4/diffuse_react_event.cpp:53-64).  All compute
runs in libmcg.so on the GPU; nothing here computes positions.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .model import PartitionConfig


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class DiffuseReactEngine:
    def __init__(self, part: PartitionConfig, precision: str = "fp64", seed: int = 1, num_replicas: int = 1,
                 capacity: int = 1 << 20, trace: bool = False, barrier_interval: int = 1,
                 propose_only: bool = False, start_iteration: int = 0, barriers: Optional[Sequence[int]] = None,
                 drain: str = "auto"):
        """barrier_interval / barriers: the scheduler's barriers (count
        events): every barrier_interval iterations, or at the iterations
        listed in barriers.  step() hands the engine the time up to the next
        one, as Scheduler::handle_next_event does (scheduler.cpp:303-307)."""
        self.lib = _lib.load()
        self.part = part
        cfg = _lib.McgConfig()
        cfg.time_unit = part.time_unit
        cfg.length_unit = part.length_unit
        for k in range(3):
            cfg.origin[k] = part.origin[k]
        cfg.partition_edge_length = part.partition_edge_length
        cfg.subpart_edge_length = part.subpart_edge_length
        cfg.subpart_edge_length_rcp = part.subpart_edge_length_rcp
        cfg.num_subparts_per_edge = part.num_subparts_per_edge
        cfg.use_expanded_list = part.use_expanded_list
        cfg.rxn_radius_3d = part.rxn_radius_3d
        cfg.precision = {"fp64": _lib.MCG_PREC_FP64, "fp32": _lib.MCG_PREC_FP32_FALLBACK}[precision]
        cfg.num_replicas = num_replicas
        cfg.seed = seed
        cfg.molecule_capacity = capacity
        cfg.flags = ((_lib.MCG_FLAG_TRACE_HITS if trace else 0) | (_lib.MCG_FLAG_PROPOSE_ONLY if propose_only else 0) |
                     {"auto": 0, "inline": _lib.MCG_FLAG_DRAIN_INLINE, "defer": _lib.MCG_FLAG_DRAIN_DEFER}[drain])
        cfg.start_iteration = start_iteration
        self.barrier_interval = barrier_interval
        self.barriers = None if barriers is None else sorted(int(b) for b in barriers)
        self.num_replicas = num_replicas
        self.ctx = ctypes.c_void_p()
        _lib.check(self.lib.mcg_create(ctypes.byref(cfg), ctypes.byref(self.ctx)))
        self.trace = trace

    @classmethod
    def from_workload(cls, wl, precision: str = "fp64", seed: Optional[int] = None, capacity: Optional[int] = None,
                      propose_only: bool = False, log_capacity: int = 0, trace: bool = False,
                      add_molecules: bool = True, start_iteration: int = 0, barriers=None, drain: str = "auto"):
        """An engine holding the workload's geometry, species, reaction
        classes and (with add_molecules) molecules of every replica."""
        from .workloads import rxn_table, species_steps, unimol_table, volsurf_table
        cap = capacity if capacity is not None else max(2 * wl.num_molecules, 1024)
        eng = cls(wl.part, precision=precision, seed=wl.config.seed if seed is None else seed,
                  num_replicas=wl.num_replicas, capacity=cap, trace=trace, barrier_interval=wl.barrier_interval,
                  propose_only=propose_only, start_iteration=start_iteration, barriers=barriers, drain=drain)
        eng.set_geometry(wl.vertices, wl.walls, wl.wall_class)
        eng.set_species(wl.species_steps)
        if wl.surface_classes is not None:
            eng.set_surface_classes(wl.surface_classes)
        if wl.rxns:
            eng.set_bimol_rxns(rxn_table(wl))
        if wl.unimol_rxns:
            eng.set_unimol_rxns(unimol_table(wl))
        if wl.volsurf_rxns:
            eng.set_volsurf_rxns(volsurf_table(wl))
        if wl.replica_D is not None:
            for r in range(wl.num_replicas):
                eng.set_replica_params(r, species_steps(wl, r), rxn_table(wl, r) if wl.rxns else None)
        if log_capacity:
            eng.set_rxn_log(log_capacity)
        cv = getattr(wl, "counted_volumes", None)
        if cv is not None:
            eng.set_counted_volumes(cv["num"], cv["wall_inside"], cv["wall_outside"])
        if not add_molecules:
            return eng
        for r in range(wl.num_replicas):
            sel = wl.replica == r
            if sel.any():
                eng.add_volume_molecules(r, wl.ids[sel], wl.mol_species[sel], wl.positions[sel],
                                         counted_volumes=None if cv is None else cv["molecules"][sel])
        sm = wl.surface_mols
        if sm is not None:
            for r in range(wl.num_replicas):
                sel = sm["replica"] == r
                if sel.any():
                    eng.add_surface_molecules(r, sm["ids"][sel], sm["species"][sel], sm["walls"][sel],
                                              sm["tiles"][sel], sm["orients"][sel],
                                              uv=sm["uv"][sel] if "uv" in sm else None)
        return eng

    def close(self):
        if self.ctx:
            self.lib.mcg_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------- set-up --
    def set_geometry(self, vertices: np.ndarray, walls: np.ndarray, wall_class: Optional[np.ndarray] = None):
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        w = np.ascontiguousarray(walls, dtype=np.uint32).reshape(-1, 3)
        c = None if wall_class is None else np.ascontiguousarray(wall_class, dtype=np.uint32)
        _lib.check(self.lib.mcg_set_geometry(self.ctx, _ptr(v), len(v), _ptr(w), len(w), _ptr(c)))

    def set_species(self, species: Sequence[tuple]):
        """species: (space_step, time_step, flags) per species id."""
        arr = (_lib.McgSpecies * len(species))()
        for i, (ss, ts, fl) in enumerate(species):
            arr[i].space_step = ss
            arr[i].time_step = ts
            arr[i].flags = fl
        _lib.check(self.lib.mcg_set_species(self.ctx, arr, len(species)))

    def set_bimol_rxns(self, rxns):
        """rxns: (reactant_a, reactant_b, cum_probs, products per pathway)."""
        _lib.check(self.lib.mcg_set_bimol_rxns(self.ctx, self._rxn_array(rxns), len(rxns)))

    def set_surface_classes(self, action: np.ndarray):
        """action: uint8 [num_classes, num_species], front | back << 2."""
        a = np.ascontiguousarray(action, np.uint8)
        _lib.check(self.lib.mcg_set_surface_classes(self.ctx, a.shape[0], _ptr(a)))

    @staticmethod
    def _rxn_array(rxns):
        arr = (_lib.McgBimolRxn * max(len(rxns), 1))()
        for i, (a, b, cum, prods) in enumerate(rxns):
            arr[i].reactant_a = a
            arr[i].reactant_b = b
            arr[i].num_pathways = len(cum)
            for k, p in enumerate(cum):
                arr[i].cum_probs[k] = p
                arr[i].num_products[k] = len(prods[k])
                for q, ps in enumerate(prods[k]):
                    arr[i].products[k][q] = ps
        return arr

    def set_replica_params(self, replica: int, species: Sequence[tuple], rxns=None):
        sp = (_lib.McgSpecies * len(species))()
        for i, (ss, ts, fl) in enumerate(species):
            sp[i].space_step = ss
            sp[i].time_step = ts
            sp[i].flags = fl
        arr = None if rxns is None else self._rxn_array(rxns)
        _lib.check(self.lib.mcg_set_replica_params(self.ctx, replica, sp, arr))

    def set_unimol_rxns(self, rxns):
        """rxns: (species, cumulative rates per time unit, products per pathway)."""
        arr = (_lib.McgUnimolRxn * max(len(rxns), 1))()
        for i, (sp, cum, prods) in enumerate(rxns):
            arr[i].reactant = sp
            arr[i].num_pathways = len(cum)
            for k, r in enumerate(cum):
                arr[i].cum_rates[k] = r
                arr[i].num_products[k] = len(prods[k])
                for q, ps in enumerate(prods[k]):
                    arr[i].products[k][q] = ps
        _lib.check(self.lib.mcg_set_unimol_rxns(self.ctx, arr, len(rxns)))

    def set_volsurf_rxns(self, rxns):
        """rxns: (vol, surf, geom_vol, geom_surf, cum_probs, products per
        pathway as [(species, orientation, is_surface), ...]) per class."""
        arr = (_lib.McgVolSurfRxn * max(len(rxns), 1))()
        for i, (v, sf, gv, gs, cum, prods) in enumerate(rxns):
            arr[i].vol_species = v
            arr[i].surf_species = sf
            arr[i].geom_vol = gv
            arr[i].geom_surf = gs
            arr[i].num_pathways = len(cum)
            for k, p in enumerate(cum):
                arr[i].cum_probs[k] = p
                arr[i].num_products[k] = len(prods[k])
                for q, (ps, o, surf) in enumerate(prods[k]):
                    arr[i].products[k][q] = ps
                    arr[i].product_orient[k][q] = o
                    if surf:
                        arr[i].surf_products[k] |= 1 << q
        _lib.check(self.lib.mcg_set_volsurf_rxns(self.ctx, arr, len(rxns)))

    def add_surface_molecules(self, replica: int, ids, species, walls, tiles, orients, uv=None,
                              diffusion_times=None, flags=None, birthdays=None):
        """uv: (n, 2) positions in the walls' frames (None: the tiles'
        centres); diffusion_times and birthdays: absolute (None: the clock);
        flags: MCG_MOLF_MATURE or 0 (None: 0)."""
        a = [np.ascontiguousarray(x, dtype=np.uint32) for x in (ids, species, walls, tiles)]
        o = np.ascontiguousarray(orients, dtype=np.int32)
        assert all(len(x) == len(o) for x in a)
        u = None if uv is None else np.ascontiguousarray(uv, dtype=np.float64).reshape(-1, 2)
        dt = None if diffusion_times is None else np.ascontiguousarray(diffusion_times, dtype=np.float64)
        fl = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint32)
        bd = None if birthdays is None else np.ascontiguousarray(birthdays, dtype=np.float64)
        assert all(x is None or len(x) == len(o) for x in (u, dt, fl, bd))
        _lib.check(self.lib.mcg_add_surface_molecules(self.ctx, replica, _ptr(a[0]), _ptr(a[1]), _ptr(a[2]),
                                                      _ptr(a[3]), _ptr(o), _ptr(u), _ptr(dt), _ptr(fl), _ptr(bd),
                                                      len(o)))

    def surface_molecules(self):
        """The surface molecules in (replica, wall, tile) order."""
        cnt = ctypes.c_uint64()
        _lib.check(self.lib.mcg_get_surface_molecules(self.ctx, None, None, None, None, None, None, None, None,
                                                      None, None, 0, ctypes.byref(cnt)))
        n = cnt.value
        out = {k: np.zeros(n, np.uint32) for k in ("id", "species", "replica", "wall", "tile", "flags")}
        out["orient"] = np.zeros(n, np.int32)
        out["uv"] = np.zeros((n, 2), np.float64)
        out["diffusion_time"] = np.zeros(n, np.float64)
        out["birthday"] = np.zeros(n, np.float64)
        _lib.check(self.lib.mcg_get_surface_molecules(self.ctx, _ptr(out["id"]), _ptr(out["species"]),
                                                      _ptr(out["replica"]), _ptr(out["wall"]), _ptr(out["tile"]),
                                                      _ptr(out["orient"]), _ptr(out["uv"]),
                                                      _ptr(out["diffusion_time"]), _ptr(out["flags"]),
                                                      _ptr(out["birthday"]), n, ctypes.byref(cnt)))
        return out

    def set_rxn_log(self, capacity: int):
        _lib.check(self.lib.mcg_set_rxn_log(self.ctx, capacity))

    def rxn_log(self) -> np.ndarray:
        """Events of the last step call (structured array, mcg_rxn_event)."""
        cnt = ctypes.c_uint64()
        _lib.check(self.lib.mcg_get_rxn_log(self.ctx, None, 0, ctypes.byref(cnt)))
        out = np.zeros(cnt.value, dtype=_lib.RXN_EVENT_DTYPE)
        if cnt.value:
            _lib.check(self.lib.mcg_get_rxn_log(self.ctx, _ptr(out), cnt.value, ctypes.byref(cnt)))
        return out

    def count_species(self, num_replicas: int, num_species: int) -> np.ndarray:
        out = np.zeros((num_replicas, num_species), np.uint64)
        _lib.check(self.lib.mcg_count_species(self.ctx, _ptr(out)))
        return out

    def count_species_noalloc(self, out: np.ndarray) -> np.ndarray:
        """The same counts into out (uint64 [replicas, species]) without any
        device allocation (mcg_count_species_noalloc)."""
        assert out.dtype == np.uint64 and out.flags.c_contiguous
        _lib.check(self.lib.mcg_count_species_noalloc(self.ctx, _ptr(out)))
        return out

    def add_volume_molecules(self, replica: int, ids: np.ndarray, species: np.ndarray, positions: np.ndarray,
                             diffusion_times=None, unimol_times=None, flags=None, counted_volumes=None,
                             birthdays=None):
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        sp = np.ascontiguousarray(species, dtype=np.uint32)
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        assert len(ids) == len(sp) == len(pos)
        dt = None if diffusion_times is None else np.ascontiguousarray(diffusion_times, dtype=np.float64)
        ut = None if unimol_times is None else np.ascontiguousarray(unimol_times, dtype=np.float64)
        fl = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint32)
        cv = None if counted_volumes is None else np.ascontiguousarray(counted_volumes, dtype=np.uint32)
        bd = None if birthdays is None else np.ascontiguousarray(birthdays, dtype=np.float64)
        for a in (dt, ut, fl, cv, bd):
            assert a is None or len(a) == len(ids)
        _lib.check(self.lib.mcg_add_volume_molecules(self.ctx, replica, _ptr(ids), _ptr(sp), _ptr(pos), _ptr(dt),
                                                     _ptr(ut), _ptr(fl), _ptr(cv), _ptr(bd), len(ids)))

    def set_counted_volumes(self, num_volumes: int, wall_inside, wall_outside):
        """Per wall: the counted volume index inside / outside its object
        (0xFFFFFFFF for both: not a counted object)."""
        wi = np.ascontiguousarray(wall_inside, dtype=np.uint32)
        wo = np.ascontiguousarray(wall_outside, dtype=np.uint32)
        _lib.check(self.lib.mcg_set_counted_volumes(self.ctx, num_volumes, _ptr(wi), _ptr(wo)))
        self.num_volumes = num_volumes

    def rxn_counts(self) -> np.ndarray:
        """Reactions carried out, [replica, counted volume, pathway slot]
        (slots: bimolecular class * 4 + pathway, then unimolecular classes
        by species, then volume-surface classes)."""
        dims = np.zeros(3, np.uint32)
        _lib.check(self.lib.mcg_get_rxn_counts(self.ctx, None, _ptr(dims)))
        out = np.zeros(tuple(int(d) for d in dims), np.uint64)
        _lib.check(self.lib.mcg_get_rxn_counts(self.ctx, _ptr(out), None))
        return out

    def count_species_by_volume(self, num_replicas: int, num_volumes: int, num_species: int) -> np.ndarray:
        out = np.zeros((num_replicas, num_volumes, num_species), np.uint64)
        _lib.check(self.lib.mcg_count_species_by_volume(self.ctx, _ptr(out)))
        return out

    def next_molecule_ids(self) -> np.ndarray:
        out = np.zeros(self.num_replicas, np.uint32)
        _lib.check(self.lib.mcg_get_next_molecule_ids(self.ctx, _ptr(out)))
        return out

    def set_next_molecule_ids(self, next_ids):
        a = np.ascontiguousarray(next_ids, dtype=np.uint32)
        assert len(a) == self.num_replicas
        _lib.check(self.lib.mcg_set_next_molecule_ids(self.ctx, _ptr(a)))

    # --------------------------------------------------------------- run --
    def next_barrier(self, it: int) -> Optional[int]:
        """The first barrier strictly after iteration it (Calendar::
        get_time_up_to_next_barrier, scheduler.cpp:212-242), or None."""
        if self.barriers is not None:
            for b in self.barriers:
                if b > it:
                    return b
            return None
        b = max(int(self.barrier_interval), 1)
        return (it // b + 1) * b

    def step(self, first_iteration: int, num_iterations: int = 1) -> _lib.McgStepStats:
        """Runs iterations first_iteration .. first_iteration + num_iterations
        - 1, one mcg_step call per stretch between barriers, each handed the
        time up to its barrier.  Returns the summed statistics."""
        total = _lib.McgStepStats()
        it, end = first_iteration, first_iteration + num_iterations
        while it < end:
            nb = self.next_barrier(it)
            time_up = (nb - it) if nb is not None else (end - it) + 100
            chunk = min(end - it, time_up)
            st = _lib.McgStepStats()
            _lib.check(self.lib.mcg_step(self.ctx, it, chunk, float(time_up), ctypes.byref(st)))
            for f in ("molecule_steps", "wall_hits", "reactions", "kernel_ms", "wall_ms", "kernel_launches",
                      "transparent_crossings", "settle_rounds"):
                setattr(total, f, getattr(total, f) + getattr(st, f))
            total.slow_path_last = st.slow_path_last
            total.drain_form = st.drain_form
            it += chunk
        return total

    def step_raw(self, first_iteration: int, num_iterations: int, time_up_to_next_barrier: float):
        """One mcg_step call as the C ABI takes it."""
        st = _lib.McgStepStats()
        _lib.check(self.lib.mcg_step(self.ctx, first_iteration, num_iterations, float(time_up_to_next_barrier),
                                     ctypes.byref(st)))
        return st

    def sync(self):
        _lib.check(self.lib.mcg_sync(self.ctx))

    def last_kernel_time(self):
        ms = ctypes.c_double()
        n = ctypes.c_uint32()
        _lib.check(self.lib.mcg_last_kernel_time(self.ctx, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    # ----------------------------------------------------------- readback --
    def num_molecules(self) -> int:
        n = ctypes.c_uint64()
        _lib.check(self.lib.mcg_num_molecules(self.ctx, ctypes.byref(n)))
        return n.value

    def molecules(self):
        n = self.num_molecules()
        ids = np.zeros(n, np.uint32)
        sp = np.zeros(n, np.uint32)
        rep = np.zeros(n, np.uint32)
        pos = np.zeros((n, 3), np.float64)
        sub = np.zeros(n, np.uint32)
        dt = np.zeros(n, np.float64)
        ut = np.zeros(n, np.float64)
        fl = np.zeros(n, np.uint32)
        cv = np.zeros(n, np.uint32)
        bd = np.zeros(n, np.float64)
        cnt = ctypes.c_uint64()
        _lib.check(self.lib.mcg_get_volume_molecules(self.ctx, _ptr(ids), _ptr(sp), _ptr(rep), _ptr(pos), _ptr(sub),
                                                     _ptr(dt), _ptr(ut), _ptr(fl), _ptr(cv), _ptr(bd), n,
                                                     ctypes.byref(cnt)))
        return {"id": ids, "species": sp, "replica": rep, "pos": pos, "subpart": sub, "diffusion_time": dt,
                "unimol_time": ut, "flags": fl, "counted_volume": cv, "birthday": bd}

    def walls(self):
        cnt = ctypes.c_uint64()
        _lib.check(self.lib.mcg_get_walls(self.ctx, None, None, None, ctypes.byref(cnt)))
        nsub = self.part.num_subparts_per_edge ** 3
        off = np.zeros(nsub + 1, np.uint32)
        idx = np.zeros(max(cnt.value, 1), np.uint32)
        return off, idx, cnt.value

    def wall_data(self, num_walls: int):
        consts = np.zeros((num_walls, 16), np.float64)
        cnt = ctypes.c_uint64()
        _lib.check(self.lib.mcg_get_walls(self.ctx, None, None, None, ctypes.byref(cnt)))
        nsub = self.part.num_subparts_per_edge ** 3
        off = np.zeros(nsub + 1, np.uint32)
        idx = np.zeros(max(cnt.value, 1), np.uint32)
        _lib.check(self.lib.mcg_get_walls(self.ctx, _ptr(consts), _ptr(off), _ptr(idx), ctypes.byref(cnt)))
        return consts, off, idx[:cnt.value]

    def hit_trace(self):
        n = self.num_molecules()
        walls = np.zeros(n * _lib.TRACE_HITS, np.uint32)
        counts = np.zeros(n, np.uint32)
        _lib.check(self.lib.mcg_get_hit_trace(self.ctx, _ptr(walls), _ptr(counts), n))
        return walls.reshape(n, _lib.TRACE_HITS), counts
