"""The C ABI library loads and exports every entry point include/mcg.h
declares (no compute calls: these run without a GPU)."""
import ctypes
import os
import re

from mcell_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "mcg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mcg_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert sorted(_lib.EXPORTED) == declared_functions()


def test_abi_version_and_struct_sizes(tmp_path):
    import shutil
    import subprocess
    lib = _lib.load()
    assert lib.mcg_abi_version() == 9
    sizes = (ctypes.sizeof(_lib.McgConfig), ctypes.sizeof(_lib.McgSpecies), ctypes.sizeof(_lib.McgStepStats),
             ctypes.sizeof(_lib.McgBimolRxn), ctypes.sizeof(_lib.McgRxnEvent), ctypes.sizeof(_lib.McgUnimolRxn),
             ctypes.sizeof(_lib.McgVolSurfRxn))
    import numpy as np
    assert np.dtype(_lib.RXN_EVENT_DTYPE).itemsize == ctypes.sizeof(_lib.McgRxnEvent)
    if shutil.which("gcc") is None:
        assert sizes == (112, 24, 64, 112, 64, 104, 184)
        return
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include "mcg.h"\nint main(){printf("%zu %zu %zu %zu %zu %zu %zu", '
                   'sizeof(mcg_config), sizeof(mcg_species), sizeof(mcg_step_stats), sizeof(mcg_bimol_rxn), '
                   'sizeof(mcg_rxn_event), sizeof(mcg_unimol_rxn), sizeof(mcg_volsurf_rxn));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    c_sizes = tuple(int(v) for v in subprocess.check_output([str(exe)]).split())
    assert sizes == c_sizes


def test_create_rejects_bad_config_without_touching_a_device():
    lib = _lib.load()
    cfg = _lib.McgConfig()
    ctx = ctypes.c_void_p()
    cfg.num_subparts_per_edge = 0  # invalid before any HIP call
    assert lib.mcg_create(ctypes.byref(cfg), ctypes.byref(ctx)) == 1
    assert b"num_subparts_per_edge" in lib.mcg_last_error()
    assert not ctx.value


def test_create_enforces_the_radius_rule_without_touching_a_device():
    """MCell4Converter's check (mcell4_converter.cpp:121-127): the reaction
    radius times sqrt(2) must be below half the subpartition edge."""
    lib = _lib.load()
    cfg = _lib.McgConfig()
    cfg.num_subparts_per_edge = 10
    cfg.subpart_edge_length = 1.5
    cfg.subpart_edge_length_rcp = 1 / 1.5
    cfg.partition_edge_length = 15.0
    cfg.rxn_radius_3d = 0.56419
    cfg.precision = 0
    cfg.num_replicas = 1
    cfg.molecule_capacity = 10
    ctx = ctypes.c_void_p()
    assert lib.mcg_create(ctypes.byref(cfg), ctypes.byref(ctx)) == 1
    assert b"sqrt(2)" in lib.mcg_last_error()
    assert not ctx.value


def test_model_layer_enforces_the_radius_rule():
    import pytest
    from mcell_amd import model as M
    box = M.create_box("box", edge_dimension=1.0)
    with pytest.raises(ValueError):
        M.derive_partition(M.Config(partition_dimension=1.2, subpartition_dimension=0.015), M.bounding_box([box]), True)
    M.derive_partition(M.Config(subpartition_dimension=0.02), M.bounding_box([box]), True)
