"""ctypes binding of the engine's C ABI (include/mcg.h).

The library is loaded from mcell_amd/lib/libmcg.so (built in place by
``__graft_entry__.build()`` / ``make -C mcell_amd/csrc``).  There is no
fallback: if the library is missing the import fails loudly.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmcg.so")

MCG_OK = 0
MCG_PREC_FP64 = 0
MCG_PREC_FP32_FALLBACK = 1
MCG_FLAG_TRACE_HITS = 1
MCG_FLAG_PROPOSE_ONLY = 2
MCG_FLAG_DRAIN_INLINE = 4
MCG_FLAG_DRAIN_DEFER = 8

MCG_SPF_CAN_VOLVOL = 1
MCG_SPF_CAN_VOLWALL = 2
MCG_SPF_CAN_VOLSURF = 4
MCG_SPF_TARGET_ONLY = 8
MCG_SPF_SURFACE = 16

MCG_WALL_REFLECT = 0
MCG_WALL_TRANSPARENT = 1
MCG_WALL_ABSORB = 2

TRACE_HITS = 8

# every symbol declared in include/mcg.h
EXPORTED = [
    "mcg_last_error", "mcg_abi_version", "mcg_create", "mcg_destroy", "mcg_set_geometry", "mcg_set_species",
    "mcg_add_volume_molecules", "mcg_step", "mcg_sync", "mcg_num_molecules", "mcg_get_volume_molecules",
    "mcg_get_walls", "mcg_get_hit_trace", "mcg_last_kernel_time", "mcg_set_bimol_rxns", "mcg_set_rxn_log",
    "mcg_get_rxn_log", "mcg_count_species", "mcg_set_replica_params", "mcg_set_surface_classes",
    "mcg_set_unimol_rxns", "mcg_set_volsurf_rxns", "mcg_add_surface_molecules", "mcg_get_surface_molecules",
    "mcg_get_next_molecule_ids", "mcg_set_next_molecule_ids", "mcg_count_species_noalloc",
    "mcg_set_counted_volumes", "mcg_count_species_by_volume", "mcg_get_rxn_counts",
]

MCG_MOLF_SCHEDULE_UNIMOL = 1
MCG_MOLF_MATURE = 2
TIME_INVALID = -256.0


class McgConfig(ctypes.Structure):
    _fields_ = [
        ("time_unit", ctypes.c_double),
        ("length_unit", ctypes.c_double),
        ("origin", ctypes.c_double * 3),
        ("partition_edge_length", ctypes.c_double),
        ("subpart_edge_length", ctypes.c_double),
        ("subpart_edge_length_rcp", ctypes.c_double),
        ("num_subparts_per_edge", ctypes.c_uint32),
        ("use_expanded_list", ctypes.c_int32),
        ("rxn_radius_3d", ctypes.c_double),
        ("precision", ctypes.c_int32),
        ("num_replicas", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("molecule_capacity", ctypes.c_uint64),
        ("flags", ctypes.c_uint32),
        ("start_iteration", ctypes.c_uint32),
    ]


class McgSpecies(ctypes.Structure):
    _fields_ = [
        ("space_step", ctypes.c_double),
        ("time_step", ctypes.c_double),
        ("flags", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class McgBimolRxn(ctypes.Structure):
    _fields_ = [
        ("reactant_a", ctypes.c_uint32),
        ("reactant_b", ctypes.c_uint32),
        ("num_pathways", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("cum_probs", ctypes.c_double * 4),
        ("num_products", ctypes.c_uint32 * 4),
        ("products", (ctypes.c_uint32 * 3) * 4),
    ]


class McgUnimolRxn(ctypes.Structure):
    _fields_ = [
        ("reactant", ctypes.c_uint32),
        ("num_pathways", ctypes.c_uint32),
        ("cum_rates", ctypes.c_double * 4),
        ("num_products", ctypes.c_uint32 * 4),
        ("products", (ctypes.c_uint32 * 3) * 4),
    ]


class McgVolSurfRxn(ctypes.Structure):
    _fields_ = [
        ("vol_species", ctypes.c_uint32),
        ("surf_species", ctypes.c_uint32),
        ("geom_vol", ctypes.c_int32),
        ("geom_surf", ctypes.c_int32),
        ("num_pathways", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("cum_probs", ctypes.c_double * 4),
        ("num_products", ctypes.c_uint32 * 4),
        ("products", (ctypes.c_uint32 * 3) * 4),
        ("product_orient", (ctypes.c_int32 * 3) * 4),
        ("surf_products", ctypes.c_uint32 * 4),
    ]


class McgRxnEvent(ctypes.Structure):
    _fields_ = [
        ("time", ctypes.c_double),
        ("pos", ctypes.c_double * 3),
        ("iteration", ctypes.c_uint32),
        ("replica", ctypes.c_uint32),
        ("mover_id", ctypes.c_uint32),
        ("partner_id", ctypes.c_uint32),
        ("rxn", ctypes.c_uint32),
        ("pathway", ctypes.c_uint32),
        ("first_product_id", ctypes.c_uint32),
        ("num_products", ctypes.c_uint32),
    ]


# numpy view of mcg_rxn_event
RXN_EVENT_DTYPE = [("time", "<f8"), ("pos", "<f8", (3,)), ("iteration", "<u4"), ("replica", "<u4"),
                   ("mover_id", "<u4"), ("partner_id", "<u4"), ("rxn", "<u4"), ("pathway", "<u4"),
                   ("first_product_id", "<u4"), ("num_products", "<u4")]


class McgStepStats(ctypes.Structure):
    _fields_ = [
        ("molecule_steps", ctypes.c_uint64),
        ("wall_hits", ctypes.c_uint64),
        ("reactions", ctypes.c_uint64),
        ("kernel_ms", ctypes.c_double),
        ("wall_ms", ctypes.c_double),
        ("kernel_launches", ctypes.c_uint32),
        ("slow_path_last", ctypes.c_uint32),
        ("transparent_crossings", ctypes.c_uint64),
        ("settle_rounds", ctypes.c_uint32),
        ("drain_form", ctypes.c_uint32),
    ]


class McgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("mcg status %d: %s" % (status, msg))
        self.status = status


_lib = None


def load():
    """Loads libmcg.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libmcg.so not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                          "or `make -C mcell_amd/csrc` (%s missing)" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    lib.mcg_last_error.restype = ctypes.c_char_p
    lib.mcg_abi_version.restype = ctypes.c_int
    lib.mcg_create.argtypes = [ctypes.POINTER(McgConfig), ctypes.POINTER(P)]
    lib.mcg_destroy.argtypes = [P]
    lib.mcg_set_geometry.argtypes = [P, P, ctypes.c_uint32, P, ctypes.c_uint32, P]
    lib.mcg_set_species.argtypes = [P, ctypes.POINTER(McgSpecies), ctypes.c_uint32]
    lib.mcg_add_volume_molecules.argtypes = [P, ctypes.c_uint32, P, P, P, P, P, P, P, P, ctypes.c_uint64]
    lib.mcg_step.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.POINTER(McgStepStats)]
    lib.mcg_sync.argtypes = [P]
    lib.mcg_num_molecules.argtypes = [P, ctypes.POINTER(ctypes.c_uint64)]
    lib.mcg_get_volume_molecules.argtypes = [P, P, P, P, P, P, P, P, P, P, P, ctypes.c_uint64,
                                             ctypes.POINTER(ctypes.c_uint64)]
    lib.mcg_set_counted_volumes.argtypes = [P, ctypes.c_uint32, P, P]
    lib.mcg_count_species_by_volume.argtypes = [P, P]
    lib.mcg_get_rxn_counts.argtypes = [P, P, P]
    lib.mcg_get_next_molecule_ids.argtypes = [P, P]
    lib.mcg_set_next_molecule_ids.argtypes = [P, P]
    lib.mcg_get_walls.argtypes = [P, P, P, P, ctypes.POINTER(ctypes.c_uint64)]
    lib.mcg_get_hit_trace.argtypes = [P, P, P, ctypes.c_uint64]
    lib.mcg_last_kernel_time.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint32)]
    lib.mcg_set_bimol_rxns.argtypes = [P, ctypes.POINTER(McgBimolRxn), ctypes.c_uint32]
    lib.mcg_set_rxn_log.argtypes = [P, ctypes.c_uint64]
    lib.mcg_get_rxn_log.argtypes = [P, P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    lib.mcg_count_species.argtypes = [P, P]
    lib.mcg_count_species_noalloc.argtypes = [P, P]
    lib.mcg_set_surface_classes.argtypes = [P, ctypes.c_uint32, P]
    lib.mcg_set_unimol_rxns.argtypes = [P, ctypes.POINTER(McgUnimolRxn), ctypes.c_uint32]
    lib.mcg_set_replica_params.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(McgSpecies), P]
    lib.mcg_set_volsurf_rxns.argtypes = [P, ctypes.POINTER(McgVolSurfRxn), ctypes.c_uint32]
    lib.mcg_add_surface_molecules.argtypes = [P, ctypes.c_uint32, P, P, P, P, P, P, P, P, P, ctypes.c_uint64]
    lib.mcg_get_surface_molecules.argtypes = [P, P, P, P, P, P, P, P, P, P, P, ctypes.c_uint64,
                                              ctypes.POINTER(ctypes.c_uint64)]
    for name in EXPORTED:
        if not name.startswith("mcg_last_error") and name != "mcg_abi_version":
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def check(status):
    if status != MCG_OK:
        raise McgError(status, load().mcg_last_error().decode())
