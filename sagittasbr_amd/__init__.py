"""sagittasbr_amd — MI355X-native shooting-and-bouncing-rays RCS engine.

The compute path is libsbr.so (HIP kernels for gfx950 behind the C ABI in
include/sbr.h); this package is its Python binding plus the procedural
meshes and the Mie series used to validate it.
"""
from ._native import (  # noqa: F401
    SBR_BATCHED,
    SBR_FP32,
    SBR_FP64,
    SBR_MIXED,
    SBR_PER_ANGLE,
    BvhOptions,
    Context,
    Job,
    Plan,
    SbrError,
    SweepParams,
    angle_frame,
    direction_jobs,
    header_functions,
    host_bvh,
    lib,
    load_model,
    obj_parse_number,
    sweep_jobs,
)

__all__ = [
    "SBR_BATCHED", "SBR_FP32", "SBR_FP64", "SBR_MIXED", "SBR_PER_ANGLE", "BvhOptions", "Context", "Job", "Plan",
    "SbrError", "SweepParams", "angle_frame", "direction_jobs", "header_functions", "host_bvh", "lib", "load_model", "obj_parse_number", "sweep_jobs",
]
