"""shatterforge_amd: the per-frame physics hot path of the Shatterforge engine as hand-written gfx950 HIP kernels
behind a C ABI (include/sf.h, libsf_hip.so).

Modules:
    world   box world (rigid step) and the reference's box-core exports
    load    load path (scalar-potential solve, joint verdicts)
    scenes  seeded scenes
    structure  structural analysis (spring network, multigrid-preconditioned conjugate gradients) over a brick store
    hulls   the convex-chunk library
"""
from ._native import lib, SfError, LIB_PATH  # noqa: F401
from . import scenes, world, load, structure, hulls  # noqa: F401
from .world import World  # noqa: F401
from .load import LoadBatch  # noqa: F401

__all__ = ["lib", "SfError", "World", "LoadBatch", "scenes", "world", "load", "structure", "hulls"]
