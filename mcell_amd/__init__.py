"""mcell_amd: an MI355X-native diffuse-react engine for MCell4.

The compute path is libmcg.so (HIP kernels for gfx950 behind the C ABI in
include/mcg.h); this package holds the host-side model description
(``model``), the engine owner (``engine``) and the workload builders
(``workloads``).
"""
from . import model  # noqa: F401

__all__ = ["model", "engine", "workloads"]
