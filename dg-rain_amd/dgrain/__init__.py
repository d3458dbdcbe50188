"""dgrain: host side of the MI355X-native DG-RAIN 2D shallow-water step.

The compute lives in libdgrain.so (gfx950 HIP kernels behind the C ABI of
include/dgrain.h); this package binds it (native), builds meshes in the
reference's fort.14 variant (mesh) and generates the synthetic workloads
(configs).
"""
from . import native, mesh, configs  # noqa: F401

__all__ = ["native", "mesh", "configs"]
