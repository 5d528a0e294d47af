"""arcadia_amd — MI355X-native 64-tree DDA voxel ray traversal (Arcadia's trace_world_64tree).

The compute path is the HIP library libarcadia_hip.so (C ABI in include/); this package is
its Python plumbing: ctypes bindings, world/ray helpers and device-memory glue for tests and
benchmarks.
"""
from .native import (  # noqa: F401
    ARC_GEN_CAVES,
    ARC_GEN_HEIGHTMAP,
    ARC_GEN_SPARSE,
    ARC_SEM_CHUNK,
    ARC_SEM_WORLD,
    ARC_SHADOW_BLOCKED,
    ARC_SHADOW_BUDGET_DARK,
    ARC_SHADOW_BUDGET_LIT,
    ARC_SHADOW_CLEAR,
    ARC_SHADOW_DI,
    ARC_SHADOW_EMITTER,
    ARC_SHADOW_MISS_WORLD,
    ARC_SHADOW_SUN,
    BRICK_DTYPE,
    HIT_DTYPE,
    NODE_DTYPE,
    RAY_DTYPE,
    RESULT_DTYPE,
    SHADOW_REC_DTYPE,
    SHARC_PROBE_DTYPE,
    SHARC_SAMPLE_DTYPE,
    ArcError,
    Batch,
    Context,
    GbufferParams,
    HostWorld,
    Nrc,
    NrcParams,
    PATH_SAMPLE_DTYPE,
    ShadeParams,
    ShadowParams,
    Sharc,
    TraceParams,
    bounce_rays,
    bounce_rays_dense,
    camera_look,
    estimate,
    gen_params,
    lib,
    lod_payload,
    palette_material,
    primary_rays,
    view_rays,
)

__all__ = [n for n in dir() if not n.startswith("_")]
