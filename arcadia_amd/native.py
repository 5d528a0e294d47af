"""ctypes bindings of libarcadia_hip.so (C ABI: include/arcadia_hip.h, include/arcadia_world.h).

The library is the product: there is no Python or CPU fallback for the traversal. Loading
fails loudly when the shared library has not been built (python -m arcadia_amd.build).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libarcadia_hip.so"

ARC_OK = 0
ARC_ERR_INVALID = -1
ARC_ERR_HIP = -2
ARC_ERR_NO_WORLD = -3
ARC_ERR_CAPACITY = -4
ARC_ERR_IO = -5
ARC_ERR_CORRUPT = -6

ARC_SEM_WORLD = 0
ARC_SEM_CHUNK = 1
ARC_LAUNCH_CLASS_BOUNCE = 3  # arc_trace_bounce_device launch shape slot
ARC_LAUNCH_CLASS_SHADOW = 4  # arc_trace_shadow_device launch shape slot
ARC_GEN_HEIGHTMAP = 0
ARC_GEN_CAVES = 1
ARC_GEN_SPARSE = 2
ARC_FACE_NONE = 255
ARC_HIT_FLAG_HIT = 1
ARC_HIT_FLAG_LOD = 2
ARC_SHADOW_SUN = 0
ARC_SHADOW_DI = 1
ARC_SHADOW_MISS_WORLD = 0
ARC_SHADOW_CLEAR = 1
ARC_SHADOW_BLOCKED = 2
ARC_SHADOW_EMITTER = 3
ARC_SHADOW_BUDGET_DARK = 4
ARC_SHADOW_BUDGET_LIT = 5
ARC_IS_LEAF_FLAG = 0x80000000
ARC_IS_UNIFORM_FLAG = 0x40000000
ARC_PTR_INDEX_MASK = 0x3FFFFFFF
ARC_EMPTY_CHUNK_SENTINEL = 0xFFFFFFFF
EMPTY_CHUNK_SENTINEL = 0xFFFFFFFF

NODE_DTYPE = np.dtype([("mask_low", "<u4"), ("mask_high", "<u4"), ("child_ptr", "<u4"), ("lod_payload", "<u4")])
BRICK_DTYPE = np.dtype([("packed_materials", "<u4", (16,))])
RAY_DTYPE = np.dtype([("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"), ("t_start", "<f4"),
                      ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("t_max", "<f4")])
HIT_DTYPE = np.dtype([("t_voxels", "<f4"), ("vx", "<u2"), ("vy", "<u2"), ("vz", "<u2"), ("mat", "u1"),
                      ("face", "u1"), ("steps", "<u2"), ("flags", "<u2")])
RESULT_DTYPE = np.dtype([("t_voxels", "<f4"), ("hit_pos_m", "<f4", (3,)), ("normal", "<f4", (3,)),
                         ("lod_rgb", "<f4", (3,)), ("contact_ao", "<f4"), ("mat", "<u4"), ("voxel_hash", "<u4"),
                         ("steps", "<u4"), ("hit", "<u4"), ("vx", "<u4"), ("vy", "<u4"), ("vz", "<u4"),
                         ("reserved", "<u4", (2,))])
assert NODE_DTYPE.itemsize == 16 and BRICK_DTYPE.itemsize == 64
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16 and RESULT_DTYPE.itemsize == 80


SHADOW_REC_DTYPE = np.dtype([("t_voxels", "<f4"), ("steps", "<u4"), ("vx", "<u2"), ("vy", "<u2"), ("vz", "<u2"),
                             ("outcome", "u1"), ("reserved", "u1")])


RIGID_BODY_DTYPE = np.dtype([("world_to_local", "<f4", (16,)), ("local_to_world", "<f4", (16,)),
                             ("prev_local_to_world", "<f4", (16,)), ("local_bounds_and_root", "<f4", (4,)),
                             ("world_sphere", "<f4", (4,))])
GI_RESERVOIR_DTYPE = np.dtype([("sample_pos", "<f4", (3,)), ("norm_m_flags", "<u4"), ("lo_rg", "<u4"),
                               ("lo_b_w", "<u4"), ("primary_depth", "<f4"), ("shadow_vis_m", "<u4")])
DI_RESERVOIR_DTYPE = np.dtype([("light_idx_seed", "<u4"), ("w_sum", "<f4"), ("m", "<f4"), ("w", "<f4")])
EMISSIVE_DTYPE = np.dtype([("pos", "<f4", (3,)), ("radius", "<f4"), ("intensity", "<f4", (3,)),
                           ("normal_packed", "<u4")])
ALIAS_DTYPE = np.dtype([("prob", "<f4"), ("alias_idx", "<u4")])
ARC_MAX_RIGID_BODIES = 32  # trace_rigid_bodies reads min(count, 32) instances
RIGID_HIT_DTYPE = np.dtype([("t_voxels", "<f4"), ("hit_pos_m", "<f4", (3,)), ("prev_hit_pos_m", "<f4", (3,)),
                            ("normal", "<f4", (3,)), ("contact_ao", "<f4"), ("mat", "<u4"), ("voxel_hash", "<u4"),
                            ("hit", "<u4"), ("body", "<u4"), ("reserved", "<u4")])


class ShadowParams(C.Structure):
    """arc_shadow_params: mode ARC_SHADOW_SUN (trace_shadow_vis) / ARC_SHADOW_DI (trace_shadow_vis_di),
    max_steps = the quality's shadow_max_steps (72 at high quality)."""
    _fields_ = [("mode", C.c_uint32), ("max_steps", C.c_uint32)]


class TraceParams(C.Structure):
    """arc_trace_params: the traversal fields of CameraUniforms (types.wgsl:19-63)."""

    _fields_ = [("semantics", C.c_uint32), ("lod_bias", C.c_uint32), ("ignore_water", C.c_uint32),
                ("max_steps", C.c_uint32 * 3), ("lod_dist_m", C.c_float * 3), ("lod_step", C.c_uint32 * 3),
                ("ao_max_dist_bits", C.c_float), ("reserved", C.c_uint32 * 3)]

    @classmethod
    def exact(cls, lod_bias: int = 0, ignore_water: bool = False, max_steps: int = 4096) -> "TraceParams":
        """Full traversal: LOD off (distance thresholds +inf, step thresholds 2^32 - 1: no distance or
        step reaches them) and a step cap no ray reaches."""
        p = cls()
        p.semantics = ARC_SEM_WORLD
        p.lod_bias = lod_bias
        p.ignore_water = int(ignore_water)
        for i in range(3):
            p.max_steps[i] = max_steps
            p.lod_dist_m[i] = float("inf")
            p.lod_step[i] = 0xFFFFFFFF
        p.ao_max_dist_bits = 24.0
        return p

    @classmethod
    def reference_default(cls, lod_bias: int = 0, ignore_water: bool = False) -> "TraceParams":
        """RaymarchQualitySettings::high_quality (crates/arcadia-render/src/quality.rs:101-134)."""
        p = cls()
        p.semantics = ARC_SEM_WORLD
        p.lod_bias = lod_bias
        p.ignore_water = int(ignore_water)
        p.max_steps[0], p.max_steps[1], p.max_steps[2] = 200, 72, 48
        p.lod_dist_m[0], p.lod_dist_m[1], p.lod_dist_m[2] = 120.0, 28.0, 24.0
        p.lod_step[0], p.lod_step[1], p.lod_step[2] = 155, 60, 48
        p.ao_max_dist_bits = 24.0
        return p

    @classmethod
    def chunk(cls) -> "TraceParams":
        """raycast_chunk_tree semantics (crates/arcadia-core/src/raycast_cpu.rs:28-218)."""
        p = cls.exact()
        p.semantics = ARC_SEM_CHUNK
        return p


class LaunchConfig(C.Structure):
    _fields_ = [("blocks_per_cu", C.c_uint32), ("block_threads", C.c_uint32), ("refill_lanes", C.c_uint32),
                ("batch_rays", C.c_uint32)]


class GenParams(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("octaves", C.c_uint32), ("seed", C.c_uint64), ("height_min", C.c_float),
                ("height_max", C.c_float), ("wavelength", C.c_float), ("density_threshold", C.c_float),
                ("n_cuboids", C.c_uint32), ("cuboid_min", C.c_uint32), ("cuboid_max", C.c_uint32),
                ("mat_min", C.c_uint32), ("mat_max", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("forward", C.c_float * 3), ("right", C.c_float * 3), ("up", C.c_float * 3),
                ("tan_half_fov", C.c_float), ("aspect", C.c_float), ("jitter", C.c_float * 2)]


class RestirParams(C.Structure):
    """arc_restir_params: the CameraUniforms / EmissiveParams fields the ReSTIR passes read."""
    _fields_ = [("cam", Camera), ("prev_eye", C.c_float * 3), ("frame_index", C.c_uint32),
                ("sun_dir", C.c_float * 3), ("sun_intensity", C.c_float), ("width", C.c_uint32),
                ("height", C.c_uint32), ("gi_max_dist", C.c_float), ("shadow_max_dist", C.c_float),
                ("shadow_jitter", C.c_float), ("shadow_mode", C.c_uint32), ("shadow_max_steps", C.c_uint32),
                ("denoiser_taps", C.c_uint32), ("target_valid", C.c_uint32), ("n_emissive", C.c_uint32),
                ("total_flux", C.c_float), ("reserved", C.c_uint32)]

    @classmethod
    def make(cls, cam: "Camera", width: int, height: int, sun_dir=(0.35, 0.8, 0.45), sun_intensity: float = 3.0,
             frame: int = 0, prev_eye=None, shadow_mode: int = 1, denoiser_taps: int = 16, n_emissive: int = 0,
             total_flux: float = 0.0, gi_max_dist: float = 26.0, shadow_max_dist: float = 88.0,
             shadow_jitter: float = 0.018, shadow_max_steps: int = 72) -> "RestirParams":
        """The high-quality settings of quality.rs:101-134 by default."""
        p = cls()
        p.cam = cam
        p.prev_eye[:] = tuple(cam.eye) if prev_eye is None else tuple(prev_eye)
        p.frame_index = frame
        sd = np.asarray(sun_dir, np.float64)
        p.sun_dir[:] = tuple((sd / np.linalg.norm(sd)).astype(np.float32))
        p.sun_intensity = sun_intensity
        p.width, p.height = width, height
        p.gi_max_dist, p.shadow_max_dist, p.shadow_jitter = gi_max_dist, shadow_max_dist, shadow_jitter
        p.shadow_mode, p.shadow_max_steps, p.denoiser_taps = shadow_mode, shadow_max_steps, denoiser_taps
        p.n_emissive, p.total_flux = n_emissive, total_flux
        return p


class ArcView(C.Structure):
    _fields_ = [("cam", Camera), ("width", C.c_uint32), ("height", C.c_uint32), ("max_dist_m", C.c_float),
                ("reserved", C.c_uint32)]


class GpuPatch(C.Structure):
    _fields_ = [("chunk_flat_idx", C.c_uint32), ("new_root_node_idx", C.c_uint32), ("node_offset", C.c_uint32),
                ("n_nodes", C.c_uint32), ("nodes", C.c_void_p), ("brick_offset", C.c_uint32),
                ("n_bricks", C.c_uint32), ("bricks", C.c_void_p)]


class EditReport(C.Structure):
    """arc_edit_report (ChunkEditReport, set.rs:17-80)."""
    _fields_ = [("min_dirty_node", C.c_uint64), ("max_dirty_node", C.c_uint64), ("min_dirty_brick", C.c_uint64),
                ("max_dirty_brick", C.c_uint64), ("added_voxels", C.c_uint32), ("had_structural_growth", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]

    def has_dirty_nodes(self) -> bool:
        return self.min_dirty_node <= self.max_dirty_node

    def has_dirty_bricks(self) -> bool:
        return self.min_dirty_brick <= self.max_dirty_brick


class ShadeParams(C.Structure):
    _fields_ = [("sun_dir", C.c_float * 3), ("emission_scale", C.c_float), ("sky_scale", C.c_float),
                ("reserved", C.c_uint32 * 3)]

    @classmethod
    def default(cls) -> "ShadeParams":
        s = np.array([0.35, 0.80, 0.45])
        s /= np.linalg.norm(s)
        return cls((C.c_float * 3)(*s.astype(np.float32).tolist()), 1.0, 1.0, (C.c_uint32 * 3)())


class NrcParams(C.Structure):
    _fields_ = [("bbox_min", C.c_float * 3), ("bbox_size", C.c_float), ("log2_table", C.c_uint32),
                ("base_res", C.c_float), ("max_res", C.c_float), ("lr_grid", C.c_float), ("lr_mlp", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint64)]

    @classmethod
    def default(cls, bbox_min=(0.0, 0.0, 0.0), bbox_size=64.0, log2_table=19, seed=1) -> "NrcParams":
        return cls((C.c_float * 3)(*bbox_min), bbox_size, log2_table, 16.0, 2048.0, 1e-2, 1e-3, 0.9, 0.99, 1e-10,
                   seed)


class GbufferParams(C.Structure):
    """arc_gbuffer_params: camera basis, jitter, previous view-projection (column-major)."""
    _fields_ = [("eye", C.c_float * 3), ("forward", C.c_float * 3), ("right", C.c_float * 3),
                ("up", C.c_float * 3), ("tan_half_fov", C.c_float), ("aspect", C.c_float),
                ("jitter", C.c_float * 2), ("prev_view_proj", C.c_float * 16), ("width", C.c_uint32),
                ("height", C.c_uint32), ("max_dist_m", C.c_float), ("reserved", C.c_uint32)]

    @classmethod
    def from_camera(cls, cam: "Camera", width: int, height: int, prev_view_proj=None,
                    max_dist_m: float = 250.0) -> "GbufferParams":
        g = cls()
        for f in ("eye", "forward", "right", "up", "jitter"):
            setattr(g, f, getattr(cam, f))
        g.tan_half_fov, g.aspect = cam.tan_half_fov, cam.aspect
        m = np.eye(4, dtype=np.float32) if prev_view_proj is None else np.asarray(prev_view_proj, np.float32)
        g.prev_view_proj = (C.c_float * 16)(*m.reshape(-1).tolist())
        g.width, g.height, g.max_dist_m = width, height, max_dist_m
        return g


_lib = None

# (name, restype, argtypes)
_SIGS = [
    ("arc_create", C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    ("arc_destroy", None, [C.c_void_p]),
    ("arc_last_error", C.c_char_p, [C.c_void_p]),
    ("arc_device_info", C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("arc_world_upload", C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.c_void_p, C.c_void_p, C.c_uint64,
                                   C.c_uint64, C.c_void_p, C.c_uint64, C.c_uint64]),
    ("arc_world_patch", C.c_int, [C.c_void_p, C.POINTER(GpuPatch), C.c_uint64]),
    ("arc_world_device_pools", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                         C.POINTER(C.c_void_p)]),
    ("arc_world_patch_host", C.c_int, [C.POINTER(GpuPatch), C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p,
                                       C.c_uint64, C.c_void_p, C.c_uint64]),
    ("arc_world_read_pools", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p,
                                       C.c_uint64, C.c_uint64]),
    ("arc_trace", C.c_int, [C.c_void_p, C.POINTER(TraceParams), C.c_void_p, C.c_uint64, C.c_void_p]),
    ("arc_trace_rigid_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64, C.c_void_p,
                                         C.c_void_p]),
    ("arc_query_voxels_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]),
    ("arc_trace_shadow_device", C.c_int, [C.c_void_p, C.POINTER(ShadowParams), C.c_void_p, C.c_uint64, C.c_void_p,
                                          C.c_void_p, C.c_void_p]),
    ("arc_trace_device", C.c_int, [C.c_void_p, C.POINTER(TraceParams), C.c_void_p, C.c_uint64, C.c_void_p,
                                   C.c_void_p]),
    ("arc_trace_full_device", C.c_int, [C.c_void_p, C.POINTER(TraceParams), C.c_void_p, C.c_uint64, C.c_void_p,
                                        C.c_void_p]),
    ("arc_primary_gbuffer", C.c_int, [C.c_void_p, C.POINTER(TraceParams), C.POINTER(GbufferParams), C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p]),
    ("arc_batch_create", C.c_int, [C.c_void_p, C.POINTER(ArcView), C.c_uint32, C.POINTER(TraceParams),
                                   C.POINTER(TraceParams), C.c_uint32, C.c_float, C.c_uint64,
                                   C.POINTER(C.c_void_p)]),
    ("arc_batch_destroy", None, [C.c_void_p]),
    ("arc_batch_slots", C.c_uint64, [C.c_void_p]),
    ("arc_batch_view_first", C.c_uint64, [C.c_void_p, C.c_uint32]),
    ("arc_batch_trace", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("arc_batch_capture", C.c_int, [C.c_void_p, C.c_void_p]),
    ("arc_batch_replay", C.c_int, [C.c_void_p, C.c_void_p]),
    ("arc_batch_set_launch_config", C.c_int, [C.c_void_p, C.POINTER(LaunchConfig)]),
    ("arc_batch_get_launch_config", C.c_int, [C.c_void_p, C.POINTER(LaunchConfig)]),
    ("arc_batch_autotune", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(LaunchConfig),
                                     C.POINTER(C.c_float)]),
    ("arc_sharc_create", C.c_int, [C.c_int, C.c_uint32, C.POINTER(C.c_void_p)]),
    ("arc_sharc_destroy", None, [C.c_void_p]),
    ("arc_sharc_info", C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                 C.POINTER(C.c_void_p)]),
    ("arc_sharc_clear", C.c_int, [C.c_void_p, C.c_void_p]),
    ("arc_sharc_read", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("arc_sharc_deposit", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_float), C.c_void_p]),
    ("arc_sharc_resolve", C.c_int, [C.c_void_p, C.c_void_p]),
    ("arc_sharc_invalidate", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    ("arc_sharc_deposit_paths", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_float), C.c_void_p]),
    ("arc_sharc_query_paths", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_float), C.c_void_p,
                                        C.c_void_p]),
    ("arc_sharc_query", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_float), C.c_void_p,
                                  C.c_void_p]),
    ("arc_batch_set_seed", C.c_int, [C.c_void_p, C.c_uint64]),
    ("arc_batch_shade", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(ShadeParams), C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p]),
    ("arc_estimate", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p]),
    ("arc_nrc_create", C.c_int, [C.c_int, C.POINTER(NrcParams), C.POINTER(C.c_void_p)]),
    ("arc_nrc_destroy", None, [C.c_void_p]),
    ("arc_nrc_reset", C.c_int, [C.c_void_p, C.c_uint64]),
    ("arc_nrc_infer", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]),
    ("arc_nrc_train", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    ("arc_nrc_backward", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    ("arc_nrc_zero_grads", C.c_int, [C.c_void_p]),
    ("arc_nrc_loss", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int]),
    ("arc_nrc_info", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32), C.POINTER(C.c_float),
                               C.POINTER(C.c_uint32)]),
    ("arc_nrc_params_io", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    ("arc_trace_pools_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_uint64, C.c_void_p, C.c_void_p]),
    ("arc_set_launch_config", C.c_int, [C.c_void_p, C.POINTER(LaunchConfig)]),
    ("arc_get_launch_config", C.c_int, [C.c_void_p, C.POINTER(LaunchConfig)]),
    ("arc_set_launch_config_class", C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(LaunchConfig)]),
    ("arc_get_launch_config_class", C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(LaunchConfig)]),
    ("arc_trace_bounce_device", C.c_int, [C.c_void_p, C.POINTER(TraceParams), C.POINTER(TraceParams), C.c_void_p,
                                         C.c_uint64, C.c_uint64, C.c_float, C.c_void_p, C.c_void_p]),
    ("arc_autotune_bounce", C.c_int, [C.c_void_p, C.POINTER(TraceParams), C.POINTER(TraceParams), C.c_void_p,
                                     C.c_uint64, C.c_uint64, C.c_float, C.c_void_p, C.c_void_p, C.c_uint32,
                                     C.POINTER(LaunchConfig), C.POINTER(C.c_float)]),
    ("arc_autotune", C.c_int, [C.c_void_p, C.POINTER(TraceParams), C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                               C.c_uint32, C.POINTER(LaunchConfig), C.POINTER(C.c_float)]),
    ("arc_autotune_shadow", C.c_int, [C.c_void_p, C.POINTER(ShadowParams), C.c_void_p, C.c_uint64, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(LaunchConfig),
                                      C.POINTER(C.c_float)]),
    ("arc_world_host_create", C.c_int, [C.POINTER(C.c_uint32), C.POINTER(C.c_void_p)]),
    ("arc_world_host_destroy", None, [C.c_void_p]),
    ("arc_world_host_set_chunk_dense", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    ("arc_world_host_set_chunk_sparse", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64]),
    ("arc_world_host_chunk_tree", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(C.c_uint64), C.c_void_p,
                                            C.POINTER(C.c_uint64)]),
    ("arc_world_host_pack", C.c_int, [C.c_void_p]),
    ("arc_world_host_pools", C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_void_p),
                                       C.POINTER(C.c_void_p), C.POINTER(C.c_uint64), C.POINTER(C.c_void_p),
                                       C.POINTER(C.c_uint64)]),
    ("arc_palette_material", C.c_int, [C.c_uint32, C.POINTER(C.c_float)]),
    ("arc_palette_lod_payload", C.c_uint32, [C.c_uint32]),
    ("arc_world_generate", C.c_int, [C.c_void_p, C.POINTER(GenParams), C.c_int]),
    ("arc_gen_voxel", C.c_uint32, [C.c_void_p, C.POINTER(GenParams), C.c_uint32, C.c_uint32, C.c_uint32]),
    ("arc_chunk_set_voxels", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64, C.POINTER(EditReport)]),
    ("arc_chunk_clear_voxels", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint32),
                                         C.POINTER(EditReport)]),
    ("arc_chunk_carve_sphere", C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_float), C.c_float, C.c_int,
                                         C.POINTER(C.c_uint32), C.POINTER(EditReport)]),
    ("arc_chunk_compact", C.c_int, [C.c_void_p, C.c_uint32]),
    ("arc_chunk_info", C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    ("arc_world_host_voxel", C.c_uint32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    ("arc_world_host_sync_chunk", C.c_int, [C.c_void_p, C.c_uint32]),
    ("arc_world_host_sync_chunk_partial", C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(EditReport)]),
    ("arc_world_host_set_voxels", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint32)]),
    ("arc_world_host_clear_voxels", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint32)]),
    ("arc_world_host_carve_crater", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_float, C.c_int,
                                              C.POINTER(C.c_uint32)]),
    ("arc_world_host_pending_patches", C.c_int, [C.c_void_p, C.POINTER(C.POINTER(GpuPatch)),
                                                 C.POINTER(C.c_uint64)]),
    ("arc_world_host_clear_patches", C.c_int, [C.c_void_p]),
    ("arc_camera_look", C.c_int, [C.POINTER(C.c_float), C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_uint32,
                                  C.POINTER(Camera)]),
    ("arc_gen_primary_rays", C.c_int, [C.POINTER(Camera), C.c_uint32, C.c_uint32, C.c_float, C.c_void_p,
                                       C.c_void_p]),
    ("arc_gen_view_rays", C.c_int64, [C.POINTER(Camera), C.c_uint32, C.c_uint32, C.c_float, C.c_void_p]),
    ("arc_gen_bounce_rays_dense", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_float,
                                            C.c_void_p]),
    ("arc_snapshot_write", C.c_int, [C.c_void_p, C.c_char_p]),
    ("arc_snapshot_open", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("arc_snapshot_close", None, [C.c_void_p]),
    ("arc_snapshot_info", C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("arc_snapshot_verify", C.c_int, [C.c_void_p, C.c_int]),
    ("arc_snapshot_load_host", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    ("arc_snapshot_last_error", C.c_char_p, []),
    ("arc_snapshot_upload", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_int]),
    ("arc_crc32", C.c_uint32, [C.c_void_p, C.c_uint64, C.c_uint32]),
    ("arc_restir_temporal", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    ("arc_restir_spatial", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("arc_gen_bounce_rays", C.c_int64, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_float, C.c_void_p,
                                        C.c_void_p]),
]

EXPORTED = [name for name, _, _ in _SIGS]


def lib() -> C.CDLL:
    """Load the native library (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m arcadia_amd.build`")
        # One HIP runtime per process: torch ships its own libamdhip64 (same soname). Loading
        # torch first lets this library bind to that copy, so device pointers and streams from
        # torch and from the C ABI live in the same runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        h = C.CDLL(str(LIB_PATH))
        for name, res, args in _SIGS:
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


class ArcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"arcadia error {code}: {msg}")
        self.code = code


def _ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


# ---------------------------------------------------------------------------------------------
# host world
# ---------------------------------------------------------------------------------------------
def gen_params(kind: int, seed: int = 1, **kw) -> GenParams:
    """Generator parameters with the workload defaults of each world kind."""
    p = GenParams()
    p.kind = kind
    p.seed = seed
    p.octaves = 5
    p.mat_min, p.mat_max = 1, 15
    if kind == ARC_GEN_HEIGHTMAP:
        p.height_min, p.height_max, p.wavelength = 32.0, 200.0, 128.0
    elif kind == ARC_GEN_CAVES:
        p.height_min, p.height_max, p.wavelength = 192.0, 704.0, 384.0
        p.density_threshold = 0.45
        p.n_cuboids, p.cuboid_min, p.cuboid_max = 2000, 2, 64
    else:
        p.height_min, p.height_max, p.wavelength = 32.0, 224.0, 512.0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class HostWorld:
    """CPU-side world: per-chunk 64-trees and the packed pools (WorldMap, world_map.rs:30-205)."""

    def __init__(self, grid=(1, 1, 1)):
        self.grid = tuple(int(g) for g in grid)
        h = C.c_void_p()
        g = (C.c_uint32 * 3)(*self.grid)
        rc = lib().arc_world_host_create(g, C.byref(h))
        if rc != ARC_OK:
            raise ArcError(rc, "arc_world_host_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().arc_world_host_destroy(self._h)
            self._h = None

    __del__ = close

    def chunk_index(self, cx: int, cy: int, cz: int) -> int:
        return cx + cy * self.grid[0] + cz * self.grid[0] * self.grid[1]

    def set_chunk_dense(self, idx: int, dense: np.ndarray) -> None:
        """dense: uint8 [256,256,256] indexed [z, y, x]."""
        d = np.ascontiguousarray(dense, dtype=np.uint8)
        assert d.shape == (256, 256, 256)
        rc = lib().arc_world_host_set_chunk_dense(self._h, idx, _ptr(d))
        if rc != ARC_OK:
            raise ArcError(rc, "set_chunk_dense")

    def set_chunk_sparse(self, idx: int, voxels) -> None:
        """voxels: iterable of (lx, ly, lz, mat) — ChunkTree64::from_sparse_voxels."""
        v = np.ascontiguousarray(np.asarray(voxels, dtype=np.uint32).reshape(-1, 4))
        rc = lib().arc_world_host_set_chunk_sparse(self._h, idx, _ptr(v) if v.size else None, v.shape[0])
        if rc != ARC_OK:
            raise ArcError(rc, "set_chunk_sparse")

    def chunk_tree(self, idx: int):
        nn, nb = C.c_uint64(0), C.c_uint64(0)
        rc = lib().arc_world_host_chunk_tree(self._h, idx, None, C.byref(nn), None, C.byref(nb))
        if rc != ARC_OK:
            raise ArcError(rc, "chunk_tree")
        nodes = np.zeros(nn.value, NODE_DTYPE)
        bricks = np.zeros(max(nb.value, 1), BRICK_DTYPE)
        rc = lib().arc_world_host_chunk_tree(self._h, idx, _ptr(nodes), C.byref(nn), _ptr(bricks), C.byref(nb))
        if rc != ARC_OK:
            raise ArcError(rc, "chunk_tree")
        return nodes, bricks[: nb.value]

    def generate(self, params: GenParams, nthreads: int | None = None) -> None:
        nt = nthreads or min(16, os.cpu_count() or 1)
        rc = lib().arc_world_generate(self._h, C.byref(params), nt)
        if rc != ARC_OK:
            raise ArcError(rc, "arc_world_generate")

    def gen_voxel(self, params: GenParams, x: int, y: int, z: int) -> int:
        return int(lib().arc_gen_voxel(self._h, C.byref(params), x, y, z))

    def pack(self) -> None:
        rc = lib().arc_world_host_pack(self._h)
        if rc != ARC_OK:
            raise ArcError(rc, "pack")

    @classmethod
    def _adopt(cls, handle, grid) -> "HostWorld":
        w = cls.__new__(cls)
        w.grid = tuple(int(g) for g in grid)
        w._h = handle
        return w

    def save(self, path) -> None:
        """Write the chunk trees as a world snapshot (arc_snapshot_write)."""
        rc = lib().arc_snapshot_write(self._h, str(path).encode())
        if rc != ARC_OK:
            raise ArcError(rc, f"arc_snapshot_write: {lib().arc_snapshot_last_error().decode()}")

    # ---- in-place edits (world_edit.cpp) ----
    def _rc(self, rc, what):
        if rc != ARC_OK:
            raise ArcError(rc, what)

    def set_voxels_local(self, idx: int, voxels) -> EditReport:
        """ChunkTree64::set_voxels_local_tracked on chunk `idx`; voxels = (x, y, z, mat) rows."""
        v = np.ascontiguousarray(np.asarray(voxels, dtype=np.uint32).reshape(-1, 4))
        rep = EditReport()
        self._rc(lib().arc_chunk_set_voxels(self._h, idx, _ptr(v) if v.size else None, v.shape[0], C.byref(rep)),
                 "arc_chunk_set_voxels")
        return rep

    def clear_voxels_local(self, idx: int, voxels):
        v = np.ascontiguousarray(np.asarray(voxels, dtype=np.uint32).reshape(-1, 3))
        rep, k = EditReport(), C.c_uint32()
        self._rc(lib().arc_chunk_clear_voxels(self._h, idx, _ptr(v) if v.size else None, v.shape[0], C.byref(k),
                                              C.byref(rep)), "arc_chunk_clear_voxels")
        return k.value, rep

    def carve_sphere_local(self, idx: int, center, radius: float, scorch_rim: bool = False):
        c = (C.c_float * 3)(*center)
        d, rep = C.c_uint32(), EditReport()
        self._rc(lib().arc_chunk_carve_sphere(self._h, idx, c, radius, int(scorch_rim), C.byref(d), C.byref(rep)),
                 "arc_chunk_carve_sphere")
        return d.value, rep

    def compact_chunk(self, idx: int) -> None:
        self._rc(lib().arc_chunk_compact(self._h, idx), "arc_chunk_compact")

    def chunk_info(self, idx: int) -> dict:
        nn, nb, sv = C.c_uint64(), C.c_uint64(), C.c_uint64()
        al = (C.c_uint32 * 4)()
        self._rc(lib().arc_chunk_info(self._h, idx, C.byref(nn), C.byref(nb), C.byref(sv), al), "arc_chunk_info")
        return {"n_nodes": nn.value, "n_bricks": nb.value, "solid": sv.value, "alloc": tuple(al)}

    def voxel(self, x: int, y: int, z: int) -> int:
        return int(lib().arc_world_host_voxel(self._h, x, y, z))

    def sync_chunk(self, idx: int) -> None:
        self._rc(lib().arc_world_host_sync_chunk(self._h, idx), "sync_chunk")

    def sync_chunk_partial(self, idx: int, rep: EditReport) -> None:
        self._rc(lib().arc_world_host_sync_chunk_partial(self._h, idx, C.byref(rep)), "sync_chunk_partial")

    def set_voxels(self, voxels) -> int:
        """WorldMap::set_voxels_global; voxels = (x, y, z, mat) rows in global voxel coordinates."""
        v = np.ascontiguousarray(np.asarray(voxels, dtype=np.int32).reshape(-1, 4))
        k = C.c_uint32()
        self._rc(lib().arc_world_host_set_voxels(self._h, _ptr(v) if v.size else None, v.shape[0], C.byref(k)),
                 "set_voxels")
        return k.value

    def clear_voxels(self, voxels) -> int:
        v = np.ascontiguousarray(np.asarray(voxels, dtype=np.int32).reshape(-1, 3))
        k = C.c_uint32()
        self._rc(lib().arc_world_host_clear_voxels(self._h, _ptr(v) if v.size else None, v.shape[0], C.byref(k)),
                 "clear_voxels")
        return k.value

    def carve_crater(self, center_m, radius_m: float, scorch_rim: bool = False) -> int:
        c = (C.c_float * 3)(*center_m)
        k = C.c_uint32()
        self._rc(lib().arc_world_host_carve_crater(self._h, c, radius_m, int(scorch_rim), C.byref(k)), "carve_crater")
        return k.value

    def pending_patches(self):
        """Queued GpuPatch views (ctypes array, valid until the next edit / clear_patches)."""
        p, n = C.POINTER(GpuPatch)(), C.c_uint64()
        self._rc(lib().arc_world_host_pending_patches(self._h, C.byref(p), C.byref(n)), "pending_patches")
        return p, n.value

    def patch_list(self):
        """Copies of the queued patches as dicts (nodes / bricks as numpy arrays)."""
        p, n = self.pending_patches()
        out = []
        for i in range(n):
            q = p[i]
            nodes = (np.frombuffer((C.c_char * (q.n_nodes * 16)).from_address(q.nodes), NODE_DTYPE).copy()
                     if q.n_nodes else np.zeros(0, NODE_DTYPE))
            bricks = (np.frombuffer((C.c_char * (q.n_bricks * 64)).from_address(q.bricks), BRICK_DTYPE).copy()
                      if q.n_bricks else np.zeros(0, BRICK_DTYPE))
            out.append({"chunk_flat_idx": q.chunk_flat_idx, "new_root_node_idx": q.new_root_node_idx,
                        "node_offset": q.node_offset, "nodes": nodes, "brick_offset": q.brick_offset,
                        "bricks": bricks})
        return out

    def apply_patches_to(self, roots: np.ndarray, nodes: np.ndarray, bricks: np.ndarray) -> int:
        """Apply the queued patches to host copies of the pools (arc_world_patch_host, in place)."""
        p, n = self.pending_patches()
        if n:
            self._rc(lib().arc_world_patch_host(p, n, len(roots), _ptr(roots), _ptr(nodes), len(nodes), _ptr(bricks),
                                                len(bricks)), "arc_world_patch_host")
        return n

    def clear_patches(self) -> None:
        self._rc(lib().arc_world_host_clear_patches(self._h), "clear_patches")

    def pools(self, copy: bool = True):
        """(grid, chunk_roots, node_pool, brick_pool) after pack(): copies, or (copy=False) views of
        the host pools valid until the next edit."""
        g = (C.c_uint32 * 3)()
        r, n, b = C.c_void_p(), C.c_void_p(), C.c_void_p()
        nn, nb = C.c_uint64(), C.c_uint64()
        rc = lib().arc_world_host_pools(self._h, g, C.byref(r), C.byref(n), C.byref(nn), C.byref(b), C.byref(nb))
        if rc != ARC_OK:
            raise ArcError(rc, "pools (call pack() first)")
        nch = g[0] * g[1] * g[2]
        roots = np.ctypeslib.as_array(C.cast(r, C.POINTER(C.c_uint32)), (nch,))
        nodes = np.frombuffer((C.c_char * (nn.value * 16)).from_address(n.value), NODE_DTYPE)
        bricks = np.frombuffer((C.c_char * (nb.value * 64)).from_address(b.value), BRICK_DTYPE)
        if copy:
            roots, nodes, bricks = roots.copy(), nodes.copy(), bricks.copy()
        return (g[0], g[1], g[2]), roots, nodes, bricks


def palette_material(mat: int) -> np.ndarray:
    out = (C.c_float * 8)()
    rc = lib().arc_palette_material(mat, out)
    if rc != ARC_OK:
        raise ArcError(rc, "palette")
    return np.array(out[:], np.float32)


def lod_payload(mat: int) -> int:
    return int(lib().arc_palette_lod_payload(mat))


# ---------------------------------------------------------------------------------------------
# rays
# ---------------------------------------------------------------------------------------------
def camera_look(eye, yaw: float, pitch: float, fov_y_deg: float, width: int, height: int) -> Camera:
    cam = Camera()
    e = (C.c_float * 3)(*eye)
    rc = lib().arc_camera_look(e, yaw, pitch, fov_y_deg, width, height, C.byref(cam))
    if rc != ARC_OK:
        raise ArcError(rc, "camera_look")
    return cam


def primary_rays(cam: Camera, width: int, height: int, max_dist_m: float = 250.0):
    """(rays, pixel_of_ray) in 8x8-tile order."""
    rays = np.zeros(width * height, RAY_DTYPE)
    pix = np.zeros(width * height, np.uint32)
    rc = lib().arc_gen_primary_rays(C.byref(cam), width, height, max_dist_m, _ptr(rays), _ptr(pix))
    if rc != ARC_OK:
        raise ArcError(rc, "primary_rays")
    return rays, pix


def view_rays(cam: Camera, width: int, height: int, max_dist_m: float = 250.0) -> np.ndarray:
    """Padded 8x8-tile-order primary rays of a whole view (empty rays in the padding slots)."""
    n = ((width + 7) // 8) * ((height + 7) // 8) * 64
    rays = np.zeros(n, RAY_DTYPE)
    k = lib().arc_gen_view_rays(C.byref(cam), width, height, max_dist_m, _ptr(rays))
    if k != n:
        raise ArcError(int(k), "view_rays")
    return rays


def bounce_rays_dense(prays: np.ndarray, hits: np.ndarray, index_base: int = 0, seed: int = 1,
                      max_dist_m: float = 26.0) -> np.ndarray:
    """One bounce per parent ray (empty ray where the parent missed)."""
    prays = np.ascontiguousarray(prays, RAY_DTYPE)
    hits = np.ascontiguousarray(hits, HIT_DTYPE)
    out = np.zeros(len(prays), RAY_DTYPE)
    if len(prays):
        rc = lib().arc_gen_bounce_rays_dense(_ptr(prays), _ptr(hits), len(prays), index_base, seed, max_dist_m,
                                             _ptr(out))
        if rc != ARC_OK:
            raise ArcError(rc, "bounce_rays_dense")
    return out


def bounce_rays(prays: np.ndarray, hits: np.ndarray, seed: int = 1, max_dist_m: float = 26.0):
    """(rays, src_index): one cosine bounce per hit primary ray."""
    n = len(prays)
    out = np.zeros(max(n, 1), RAY_DTYPE)
    src = np.zeros(max(n, 1), np.uint32)
    k = lib().arc_gen_bounce_rays(_ptr(prays), _ptr(hits), n, seed, max_dist_m, _ptr(out), _ptr(src))
    if k < 0:
        raise ArcError(int(k), "bounce_rays")
    return out[:k].copy(), src[:k].copy()


def trace_pools_noalloc(grid, d_roots: int, d_nodes: int, d_bricks: int, params: TraceParams, d_rays: int, n: int,
                        d_hits: int, stream: int = 0) -> int:
    """arc_trace_pools_noalloc: compact hit records of `n` rays on caller-owned device pools, with no
    context and no allocation (device pointers, asynchronous on `stream`); returns the status code."""
    g = (C.c_uint32 * 3)(*(int(v) for v in grid))
    return int(lib().arc_trace_pools_noalloc(g, d_roots, d_nodes, d_bricks, C.byref(params), d_rays, n, d_hits,
                                             stream or None))


# ---------------------------------------------------------------------------------------------
# device context
# ---------------------------------------------------------------------------------------------
class Snapshot:
    """An open world snapshot file (arc_snapshot_open): header, index and checksums checked."""

    def __init__(self, path):
        self.path = str(path)
        h = C.c_void_p()
        rc = lib().arc_snapshot_open(self.path.encode(), C.byref(h))
        if rc != ARC_OK:
            raise ArcError(rc, f"arc_snapshot_open: {lib().arc_snapshot_last_error().decode()}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().arc_snapshot_close(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def info(self) -> dict:
        g = (C.c_uint32 * 3)()
        n, tn, tb = C.c_uint32(), C.c_uint64(), C.c_uint64()
        lib().arc_snapshot_info(self._h, g, C.byref(n), C.byref(tn), C.byref(tb))
        return {"grid": tuple(g), "chunks": n.value, "nodes": tn.value, "bricks": tb.value}

    def verify(self, nthreads: int = 8) -> None:
        rc = lib().arc_snapshot_verify(self._h, nthreads)
        if rc != ARC_OK:
            raise ArcError(rc, f"arc_snapshot_verify: {lib().arc_snapshot_last_error().decode()}")

    def load_host(self, nthreads: int = 8) -> HostWorld:
        h = C.c_void_p()
        rc = lib().arc_snapshot_load_host(self._h, nthreads, C.byref(h))
        if rc != ARC_OK:
            raise ArcError(rc, f"arc_snapshot_load_host: {lib().arc_snapshot_last_error().decode()}")
        return HostWorld._adopt(h, self.info()["grid"])


def crc32(data, crc: int = 0) -> int:
    """The snapshot checksum (zlib.crc32) computed by the library."""
    b = np.frombuffer(bytes(data), np.uint8) if not isinstance(data, np.ndarray) else np.ascontiguousarray(data)
    return int(lib().arc_crc32(b.ctypes.data if b.nbytes else None, b.nbytes, crc))


class Context:
    """A device context holding the world pools (Renderer's Group0Resources equivalent)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib().arc_create(device, C.byref(h))
        if rc != ARC_OK:
            raise ArcError(rc, "arc_create (no usable HIP device?)")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().arc_destroy(self._h)
            self._h = None

    __del__ = close

    def _check(self, rc: int, what: str):
        if rc != ARC_OK:
            raise ArcError(rc, f"{what}: {lib().arc_last_error(self._h).decode()}")

    def device_info(self):
        cu, lds = C.c_uint32(), C.c_uint32()
        self._check(lib().arc_device_info(self._h, C.byref(cu), C.byref(lds)), "device_info")
        return cu.value, lds.value

    def upload(self, grid, roots, nodes, bricks, node_capacity: int = 0, brick_capacity: int = 0):
        g = (C.c_uint32 * 3)(*grid)
        roots = np.ascontiguousarray(roots, np.uint32)
        nodes = np.ascontiguousarray(nodes)
        bricks = np.ascontiguousarray(bricks)
        self._check(lib().arc_world_upload(self._h, g, _ptr(roots), _ptr(nodes), len(nodes), node_capacity,
                                           _ptr(bricks), len(bricks), brick_capacity), "arc_world_upload")

    def upload_world(self, world: HostWorld, **kw):
        self.upload(*world.pools(), **kw)

    def restir_temporal(self, p: RestirParams, sun_trace: TraceParams, gi_trace: TraceParams, d_depth_mat: int,
                        d_norm_motion: int, d_gi_in: int, d_gi_out: int, d_di_in: int, d_di_out: int,
                        d_sources: int = 0, d_alias: int = 0, sharc=None, bodies=None, stream: int = 0) -> None:
        """Pass 2A (arc_restir_temporal) on device buffers; `sharc` a Sharc or None, `bodies` host
        RIGID_BODY_DTYPE instances or None."""
        b = None if bodies is None or len(bodies) == 0 else np.ascontiguousarray(bodies, RIGID_BODY_DTYPE)
        self._check(lib().arc_restir_temporal(self._h, C.byref(p), C.byref(sun_trace), C.byref(gi_trace), d_depth_mat,
                                              d_norm_motion, d_gi_in, d_gi_out, d_di_in, d_di_out, d_sources or None,
                                              d_alias or None, sharc._h if sharc is not None else None,
                                              b.ctypes.data if b is not None else None, 0 if b is None else len(b),
                                              stream or None), "arc_restir_temporal")

    def restir_spatial(self, p: RestirParams, d_depth_mat: int, d_norm_motion: int, d_gi_in: int, d_gi_out: int,
                       d_di_in: int, d_di_out: int, d_sources: int = 0, stream: int = 0) -> None:
        """Pass 2B (arc_restir_spatial) on device buffers."""
        self._check(lib().arc_restir_spatial(self._h, C.byref(p), d_depth_mat, d_norm_motion, d_gi_in, d_gi_out,
                                             d_di_in, d_di_out, d_sources or None, stream or None),
                    "arc_restir_spatial")

    def upload_snapshot(self, snap: Snapshot, node_capacity: int = 0, brick_capacity: int = 0,
                        nthreads: int = 8) -> None:
        """Stream a snapshot's chunks into the device pools (arc_snapshot_upload)."""
        self._check(lib().arc_snapshot_upload(self._h, snap._h, node_capacity, brick_capacity, nthreads),
                    "arc_snapshot_upload")

    def batch(self, views, primary: TraceParams, bounce: TraceParams, n_bounces: int = 2,
              bounce_max_dist_m: float = 26.0, seed: int = 1) -> "Batch":
        return Batch(self, views, primary, bounce, n_bounces, bounce_max_dist_m, seed)

    def read_pools(self, n_roots: int, n_nodes: int, n_bricks: int):
        """Device pools read back: (chunk_roots, nodes[:n_nodes], bricks[:n_bricks])."""
        roots = np.zeros(n_roots, np.uint32)
        nodes = np.zeros(n_nodes, NODE_DTYPE)
        bricks = np.zeros(n_bricks, BRICK_DTYPE)
        self._check(lib().arc_world_read_pools(self._h, _ptr(roots), _ptr(nodes) if n_nodes else None, 0, n_nodes,
                                               _ptr(bricks) if n_bricks else None, 0, n_bricks), "read_pools")
        return roots, nodes, bricks

    def patch_world(self, world: "HostWorld", clear: bool = True) -> int:
        """Uploads the world's queued patches (Renderer::apply_world_patches) and drains the queue."""
        p, n = world.pending_patches()
        if n:
            self._check(lib().arc_world_patch(self._h, p, n), "arc_world_patch")
        if clear:
            world.clear_patches()
        return n

    def patch(self, patches) -> None:
        arr = (GpuPatch * max(len(patches), 1))()
        keep = []
        for i, p in enumerate(patches):
            nodes = np.ascontiguousarray(p["nodes"])
            bricks = np.ascontiguousarray(p["bricks"])
            keep += [nodes, bricks]
            arr[i] = GpuPatch(p["chunk_flat_idx"], p["new_root_node_idx"], p["node_offset"], len(nodes),
                              _ptr(nodes) if len(nodes) else None, p["brick_offset"], len(bricks),
                              _ptr(bricks) if len(bricks) else None)
        self._check(lib().arc_world_patch(self._h, arr, len(patches)), "arc_world_patch")

    def set_launch_config(self, blocks_per_cu=0, block_threads=0, refill_lanes=0, batch_rays=0):
        cfg = LaunchConfig(blocks_per_cu, block_threads, refill_lanes, batch_rays)
        self._check(lib().arc_set_launch_config(self._h, C.byref(cfg)), "set_launch_config")

    def launch_config(self, lod_class: int = 0) -> dict:
        cfg = LaunchConfig()
        self._check(lib().arc_get_launch_config_class(self._h, lod_class, C.byref(cfg)), "get_launch_config")
        return {k: getattr(cfg, k) for k, _ in LaunchConfig._fields_}

    def set_launch_config_class(self, lod_class: int, blocks_per_cu=0, block_threads=0, refill_lanes=0,
                                batch_rays=0):
        cfg = LaunchConfig(blocks_per_cu, block_threads, refill_lanes, batch_rays)
        self._check(lib().arc_set_launch_config_class(self._h, lod_class, C.byref(cfg)), "set_launch_config_class")

    def autotune(self, params: TraceParams, d_rays: int, n: int, d_hits: int, stream: int = 0, reps: int = 3):
        """Native autotuner: picks the fastest launch shape for the ray class of `params`."""
        cfg, ms = LaunchConfig(), C.c_float()
        self._check(lib().arc_autotune(self._h, C.byref(params), d_rays, n, d_hits, stream or None, reps,
                                       C.byref(cfg), C.byref(ms)), "arc_autotune")
        return {k: getattr(cfg, k) for k, _ in LaunchConfig._fields_}, ms.value

    def trace_bounce_device(self, primary: TraceParams, bounce: TraceParams, d_rays: int, n: int, seed: int,
                            bounce_max_dist_m: float, d_hits: int, stream: int = 0) -> None:
        """arc_trace_bounce_device: the n rays and one diffuse bounce off each hit in one launch; d_hits
        receives 2n compact records (the rays', then the bounce of ray i at n + i)."""
        self._check(lib().arc_trace_bounce_device(self._h, C.byref(primary), C.byref(bounce), d_rays, n, seed,
                                                  bounce_max_dist_m, d_hits, stream or None),
                    "arc_trace_bounce_device")

    def autotune_bounce(self, primary: TraceParams, bounce: TraceParams, d_rays: int, n: int, seed: int,
                        bounce_max_dist_m: float, d_hits: int, stream: int = 0, reps: int = 3):
        """Native autotuner for arc_trace_bounce_device's launch shape on this ray set."""
        cfg, ms = LaunchConfig(), C.c_float()
        self._check(lib().arc_autotune_bounce(self._h, C.byref(primary), C.byref(bounce), d_rays, n, seed,
                                              bounce_max_dist_m, d_hits, stream or None, reps, C.byref(cfg),
                                              C.byref(ms)), "arc_autotune_bounce")
        return {k: getattr(cfg, k) for k, _ in LaunchConfig._fields_}, ms.value

    def trace(self, params: TraceParams, rays: np.ndarray) -> np.ndarray:
        """Host arrays in, compact hits out (synchronous)."""
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        if len(rays):
            self._check(lib().arc_trace(self._h, C.byref(params), _ptr(rays), len(rays), _ptr(hits)), "arc_trace")
        return hits

    def trace_device(self, params: TraceParams, d_rays: int, n: int, d_hits: int, stream: int = 0) -> None:
        """Device pointers (e.g. torch tensor data_ptr()), asynchronous on `stream`."""
        self._check(lib().arc_trace_device(self._h, C.byref(params), d_rays, n, d_hits, stream or None),
                    "arc_trace_device")

    def trace_shadow_device(self, mode: int, max_steps: int, d_rays: int, n: int, d_vis: int, d_rec: int = 0,
                            stream: int = 0) -> None:
        """Shadow any-hit walker over device rays: d_vis f32 per ray, d_rec (optional) SHADOW_REC_DTYPE."""
        sp = ShadowParams(mode, max_steps)
        self._check(lib().arc_trace_shadow_device(self._h, C.byref(sp), d_rays, n, d_vis, d_rec or None,
                                                  stream or None), "arc_trace_shadow_device")

    def autotune_shadow(self, mode: int, max_steps: int, d_rays: int, n: int, d_vis: int, d_rec: int = 0,
                        stream: int = 0, reps: int = 3):
        """Native autotuner for the shadow walkers' launch shape (class ARC_LAUNCH_CLASS_SHADOW) on this ray set."""
        sp = ShadowParams(mode, max_steps)
        cfg, ms = LaunchConfig(), C.c_float()
        self._check(lib().arc_autotune_shadow(self._h, C.byref(sp), d_rays, n, d_vis, d_rec or None, stream or None,
                                              reps, C.byref(cfg), C.byref(ms)), "arc_autotune_shadow")
        return {k: getattr(cfg, k) for k, _ in LaunchConfig._fields_}, ms.value

    def trace_rigid_device(self, bodies: np.ndarray, d_rays: int, n: int, d_out: int, stream: int = 0) -> None:
        """trace_rigid_bodies over device rays; `bodies` is a host RIGID_BODY_DTYPE array (<= 32 used),
        d_out receives RIGID_HIT_DTYPE records."""
        b = np.ascontiguousarray(bodies, RIGID_BODY_DTYPE)
        self._check(lib().arc_trace_rigid_device(self._h, b.ctypes.data if len(b) else None, len(b), d_rays, n, d_out,
                                                 stream or None), "arc_trace_rigid_device")

    def query_voxels_device(self, d_points: int, n: int, d_mat: int, stream: int = 0) -> None:
        """query_voxel_world_meters at n device points (3 f32 metres each) -> u32 materials."""
        self._check(lib().arc_query_voxels_device(self._h, d_points, n, d_mat, stream or None),
                    "arc_query_voxels_device")

    def trace_full_device(self, params: TraceParams, d_rays: int, n: int, d_out: int, stream: int = 0) -> None:
        self._check(lib().arc_trace_full_device(self._h, C.byref(params), d_rays, n, d_out, stream or None),
                    "arc_trace_full_device")

    def primary_gbuffer(self, params: TraceParams, g: GbufferParams, d_depth_mat: int, d_norm_motion: int,
                        d_albedo_ao: int, stream: int = 0) -> None:
        """Primary G-buffer pass into device buffers of 2, 2 and 1 u32 per pixel (asynchronous)."""
        self._check(lib().arc_primary_gbuffer(self._h, C.byref(params), C.byref(g), d_depth_mat, d_norm_motion,
                                              d_albedo_ao, stream or None), "arc_primary_gbuffer")


class Batch:
    """Batched views traced with their bounces in one persistent launch (arc_batch_*)."""

    def __init__(self, ctx: Context, views, primary: TraceParams, bounce: TraceParams, n_bounces: int = 2,
                 bounce_max_dist_m: float = 26.0, seed: int = 1):
        """views: iterable of (Camera, width, height, max_dist_m)."""
        self.ctx = ctx
        views = list(views)
        arr = (ArcView * len(views))()
        for i, (cam, w, h, md) in enumerate(views):
            arr[i] = ArcView(cam, w, h, md, 0)
        self.views = views
        self.n_bounces = n_bounces
        h = C.c_void_p()
        ctx._check(lib().arc_batch_create(ctx._h, arr, len(views), C.byref(primary), C.byref(bounce), n_bounces,
                                          bounce_max_dist_m, seed, C.byref(h)), "arc_batch_create")
        self._h = h
        self.slots = int(lib().arc_batch_slots(h))

    def close(self):
        if getattr(self, "_h", None):
            lib().arc_batch_destroy(self._h)
            self._h = None

    __del__ = close

    def view_first(self, v: int) -> int:
        return int(lib().arc_batch_view_first(self._h, v))

    def records(self) -> int:
        return (1 + self.n_bounces) * self.slots

    def trace(self, d_hits: int, stream: int = 0) -> None:
        self.ctx._check(lib().arc_batch_trace(self._h, d_hits, stream or None), "arc_batch_trace")

    def capture(self, d_hits: int) -> None:
        self.ctx._check(lib().arc_batch_capture(self._h, d_hits), "arc_batch_capture")

    def replay(self, stream: int = 0) -> None:
        self.ctx._check(lib().arc_batch_replay(self._h, stream or None), "arc_batch_replay")

    def set_launch_config(self, blocks_per_cu=0, block_threads=0, refill_lanes=0, batch_rays=0):
        cfg = LaunchConfig(blocks_per_cu, block_threads, refill_lanes, batch_rays)
        self.ctx._check(lib().arc_batch_set_launch_config(self._h, C.byref(cfg)), "arc_batch_set_launch_config")

    def launch_config(self) -> dict:
        cfg = LaunchConfig()
        self.ctx._check(lib().arc_batch_get_launch_config(self._h, C.byref(cfg)), "arc_batch_get_launch_config")
        return {k: getattr(cfg, k) for k, _ in LaunchConfig._fields_}

    def set_seed(self, seed: int) -> None:
        self.ctx._check(lib().arc_batch_set_seed(self._h, seed), "arc_batch_set_seed")

    def shade(self, d_hits: int, params: "ShadeParams", d_samples: int = 0, d_radiance: int = 0, d_accum: int = 0,
              stream: int = 0) -> None:
        self.ctx._check(lib().arc_batch_shade(self._h, d_hits, C.byref(params), d_samples or None, d_radiance or None,
                                              d_accum or None, stream or None), "arc_batch_shade")

    def autotune(self, d_hits: int, stream: int = 0, reps: int = 3):
        cfg, ms = LaunchConfig(), C.c_float()
        self.ctx._check(lib().arc_batch_autotune(self._h, d_hits, stream or None, reps, C.byref(cfg), C.byref(ms)),
                        "arc_batch_autotune")
        return {k: getattr(cfg, k) for k, _ in LaunchConfig._fields_}, ms.value


SHARC_SAMPLE_DTYPE = np.dtype([("pos", "<f4", (3,)), ("normal", "<f4", (3,)), ("radiance", "<f4", (3,)),
                               ("reserved", "<u4")])
SHARC_PROBE_DTYPE = np.dtype([("pos", "<f4", (3,)), ("normal", "<f4", (3,)), ("jitter", "<f4", (2,))])


class Sharc:
    """Device SHARC cache (arc_sharc_*); sample / probe buffers are device pointers."""

    def __init__(self, log2_buckets: int = 15, device: int = 0):
        h = C.c_void_p()
        rc = lib().arc_sharc_create(device, log2_buckets, C.byref(h))
        if rc != ARC_OK:
            raise ArcError(rc, "arc_sharc_create")
        self._h = h
        self.log2 = log2_buckets
        n, k, a, r = C.c_uint32(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        lib().arc_sharc_info(h, C.byref(n), C.byref(k), C.byref(a), C.byref(r))
        self.slots, self.d_keys, self.d_accum, self.d_resolved = n.value, k.value, a.value, r.value

    def close(self):
        if getattr(self, "_h", None):
            lib().arc_sharc_destroy(self._h)
            self._h = None

    __del__ = close

    def _rc(self, rc, what):
        if rc != ARC_OK:
            raise ArcError(rc, what)

    def clear(self, stream: int = 0):
        self._rc(lib().arc_sharc_clear(self._h, stream or None), "arc_sharc_clear")

    def read(self):
        """(keys, accum, resolved) copied to host uint32 arrays."""
        k = np.zeros(self.slots, np.uint32)
        a = np.zeros(self.slots * 4, np.uint32)
        r = np.zeros(self.slots * 4, np.uint32)
        self._rc(lib().arc_sharc_read(self._h, _ptr(k), _ptr(a), _ptr(r)), "arc_sharc_read")
        return k, a, r

    def deposit(self, d_samples: int, n: int, cam, stream: int = 0):
        self._rc(lib().arc_sharc_deposit(self._h, d_samples, n, (C.c_float * 3)(*cam), stream or None),
                 "arc_sharc_deposit")

    def resolve(self, stream: int = 0):
        self._rc(lib().arc_sharc_resolve(self._h, stream or None), "arc_sharc_resolve")

    def invalidate(self, spheres, stream: int = 0):
        """spheres: (n, 4) x, y, z, radius in meters."""
        sp = np.ascontiguousarray(np.asarray(spheres, np.float32).reshape(-1, 4))
        self._rc(lib().arc_sharc_invalidate(self._h, _ptr(sp) if len(sp) else None, len(sp), stream or None),
                 "arc_sharc_invalidate")

    def deposit_paths(self, d_samples: int, n: int, cam, stream: int = 0):
        self._rc(lib().arc_sharc_deposit_paths(self._h, d_samples, n, (C.c_float * 3)(*cam), stream or None),
                 "arc_sharc_deposit_paths")

    def query_paths(self, d_samples: int, n: int, cam, d_out: int, stream: int = 0):
        self._rc(lib().arc_sharc_query_paths(self._h, d_samples, n, (C.c_float * 3)(*cam), d_out, stream or None),
                 "arc_sharc_query_paths")

    def query(self, d_probes: int, n: int, cam, d_out: int, stream: int = 0):
        self._rc(lib().arc_sharc_query(self._h, d_probes, n, (C.c_float * 3)(*cam), d_out, stream or None),
                 "arc_sharc_query")


PATH_SAMPLE_DTYPE = np.dtype([("pos", "<f4", (3,)), ("flags", "<u4"), ("dir", "<f4", (3,)), ("pad0", "<u4"),
                              ("le0", "<f4", (3,)), ("pad1", "<u4"), ("a0", "<f4", (3,)), ("pad2", "<u4"),
                              ("base", "<f4", (3,)), ("pad3", "<u4"), ("target", "<f4", (3,)), ("pad4", "<u4")])
PS_RAY, PS_X0_HIT, PS_X1_HIT = 1, 2, 4


def estimate(d_samples: int, n: int, d_cache: int, mode: int, d_out: int, stream: int = 0) -> None:
    rc = lib().arc_estimate(d_samples, d_cache or None, n, mode, d_out, stream or None)
    if rc != ARC_OK:
        raise ArcError(rc, "arc_estimate")


class Nrc:
    """Neural radiance cache (arc_nrc_*)."""

    def __init__(self, params: NrcParams, device: int = 0):
        h = C.c_void_p()
        rc = lib().arc_nrc_create(device, C.byref(params), C.byref(h))
        if rc != ARC_OK:
            raise ArcError(rc, "arc_nrc_create")
        self._h = h
        self.params = params
        gp, mp = C.c_uint64(), C.c_uint32()
        sc, dn = (C.c_float * 16)(), (C.c_uint32 * 16)()
        lib().arc_nrc_info(h, C.byref(gp), C.byref(mp), sc, dn)
        self.grid_params, self.mlp_params = gp.value, mp.value
        self.scale = np.array(sc[:], np.float32)
        self.dense = np.array(dn[:], np.uint32)

    def close(self):
        if getattr(self, "_h", None):
            lib().arc_nrc_destroy(self._h)
            self._h = None

    __del__ = close

    def _rc(self, rc, what):
        if rc != ARC_OK:
            raise ArcError(rc, what)

    def reset(self, seed: int):
        self._rc(lib().arc_nrc_reset(self._h, seed), "arc_nrc_reset")

    def infer(self, d_samples: int, n: int, d_out: int, stream: int = 0):
        self._rc(lib().arc_nrc_infer(self._h, d_samples, n, d_out, stream or None), "arc_nrc_infer")

    def train(self, d_samples: int, n: int, stream: int = 0):
        self._rc(lib().arc_nrc_train(self._h, d_samples, n, stream or None), "arc_nrc_train")

    def backward(self, d_samples: int, n: int, stream: int = 0):
        self._rc(lib().arc_nrc_backward(self._h, d_samples, n, stream or None), "arc_nrc_backward")

    def zero_grads(self):
        self._rc(lib().arc_nrc_zero_grads(self._h), "arc_nrc_zero_grads")

    def loss(self, reset: bool = True) -> float:
        v = C.c_float()
        self._rc(lib().arc_nrc_loss(self._h, C.byref(v), int(reset)), "arc_nrc_loss")
        return v.value

    def get(self):
        """(grid, mlp, grid_grad, mlp_grad) FP32 host copies."""
        g = np.zeros(self.grid_params, np.float32)
        m = np.zeros(self.mlp_params, np.float32)
        gg = np.zeros(self.grid_params, np.float32)
        mg = np.zeros(self.mlp_params, np.float32)
        self._rc(lib().arc_nrc_params_io(self._h, _ptr(g), _ptr(m), _ptr(gg), _ptr(mg), 0), "arc_nrc_params_io")
        return g, m, gg, mg

    def set(self, grid: np.ndarray = None, mlp: np.ndarray = None):
        g = None if grid is None else np.ascontiguousarray(grid, np.float32)
        m = None if mlp is None else np.ascontiguousarray(mlp, np.float32)
        self._rc(lib().arc_nrc_params_io(self._h, _ptr(g) if g is not None else None,
                                         _ptr(m) if m is not None else None, None, None, 1), "arc_nrc_params_io")
