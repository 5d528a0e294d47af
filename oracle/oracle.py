"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) — TEST INFRASTRUCTURE ONLY.

Used by tests/, the smoke check and bench.py's CPU-baseline leg as the checker of the HIP
traversal. Nothing in the product path imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

NODE_DTYPE = np.dtype([("mask_low", "<u4"), ("mask_high", "<u4"), ("child_ptr", "<u4"), ("lod_payload", "<u4")])
RAY_DTYPE = np.dtype([("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"), ("t_start", "<f4"),
                      ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("t_max", "<f4")])
HIT_DTYPE = np.dtype([("t_voxels", "<f4"), ("vx", "<u2"), ("vy", "<u2"), ("vz", "<u2"), ("mat", "u1"),
                      ("face", "u1"), ("steps", "<u2"), ("flags", "<u2")])
RESULT_DTYPE = np.dtype([("t_voxels", "<f4"), ("hit_pos_m", "<f4", (3,)), ("normal", "<f4", (3,)),
                         ("lod_rgb", "<f4", (3,)), ("contact_ao", "<f4"), ("mat", "<u4"), ("voxel_hash", "<u4"),
                         ("steps", "<u4"), ("hit", "<u4"), ("vx", "<u4"), ("vy", "<u4"), ("vz", "<u4"),
                         ("reserved", "<u4", (2,))])
COUNTER_DTYPE = np.dtype([("node_loads", "<u4"), ("brick_loads", "<u4"), ("root_loads", "<u4"), ("steps", "<u4"),
                          ("level_loads", "<u4", (4,)), ("restarts", "<u4", (4,)), ("empty_chunk_steps", "<u4")])
SHADOW_REC_DTYPE = np.dtype([("t_voxels", "<f4"), ("steps", "<u4"), ("vx", "<u2"), ("vy", "<u2"), ("vz", "<u2"),
                             ("outcome", "u1"), ("reserved", "u1")])
BRUTE_DTYPE = np.dtype([("t", "<f8"), ("vx", "<i4"), ("vy", "<i4"), ("vz", "<i4"), ("hit", "<i4"), ("face", "<i4"),
                        ("tie", "<i4"), ("mat", "<i4"), ("pad", "<i4")])


class OrcWorld(C.Structure):
    _fields_ = [("grid", C.c_uint32 * 3), ("roots", C.c_void_p), ("nodes", C.c_void_p), ("n_nodes", C.c_uint64),
                ("bricks", C.c_void_p), ("n_bricks", C.c_uint64)]


class OrcParams(C.Structure):  # layout of arc_trace_params
    _fields_ = [("semantics", C.c_uint32), ("lod_bias", C.c_uint32), ("ignore_water", C.c_uint32),
                ("max_steps", C.c_uint32 * 3), ("lod_dist_m", C.c_float * 3), ("lod_step", C.c_uint32 * 3),
                ("ao_max_dist_bits", C.c_float), ("reserved", C.c_uint32 * 3)]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        h.orc_trace_world.argtypes = [P, P, P, C.c_uint64, P, P, C.c_int]
        h.orc_trace_chunk.argtypes = [P, P, C.c_uint64, P, P, C.c_int]
        h.orc_trace_meters.argtypes = [P, P, C.c_uint64, P, C.c_int]
        h.orc_compact.argtypes = [P, C.c_uint64, C.c_int, P]
        h.orc_query_voxel.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32]
        h.orc_query_voxel.restype = C.c_uint32
        h.orc_decode_dense.argtypes = [P, P, C.c_int]
        h.orc_brute_dense.argtypes = [P, P, P, C.c_uint64, C.c_int, C.c_double, P, C.c_int]
        h.orc_primary_gbuffer.argtypes = [P, P, P, P, P, P, P, C.c_int]
        h.orc_primary_gbuffer.restype = None
        h.orc_f16_to_f32.argtypes = [C.c_uint16]
        h.orc_f16_to_f32.restype = C.c_float
        h.orc_sharc_make_key.argtypes = [P, P, P, C.c_float, C.c_float, C.c_uint32, P, P]
        h.orc_sharc_deposit.argtypes = [C.c_uint32, P, P, P, C.c_uint64, P]
        h.orc_sharc_resolve.argtypes = [C.c_uint32, P, P, P]
        h.orc_sharc_query.argtypes = [C.c_uint32, P, P, C.c_uint64, P, P]
        for f in (h.orc_sharc_make_key, h.orc_sharc_deposit, h.orc_sharc_resolve, h.orc_sharc_query):
            f.restype = None
        h.orc_trace_rigid.argtypes = [P, P, C.c_uint32, P, C.c_uint64, P]
        h.orc_trace_rigid.restype = None
        h.orc_query_point.argtypes = [P, P]
        h.orc_query_point.restype = C.c_uint32
        h.orc_trace_shadow.argtypes = [P, P, P, C.c_uint64, P, P, C.c_int]
        h.orc_trace_shadow.restype = None
        h.orc_f32_to_f16.argtypes = [C.c_float]
        h.orc_f32_to_f16.restype = C.c_uint16
        for f in (h.orc_trace_world, h.orc_trace_chunk, h.orc_trace_meters, h.orc_compact, h.orc_decode_dense,
                  h.orc_brute_dense):
            f.restype = None
        _lib = h
    return _lib


def _threads(n):
    return n or min(16, os.cpu_count() or 1)


class World:
    """Oracle view of packed pools (keeps the arrays alive)."""

    def __init__(self, grid, roots, nodes, bricks):
        self.grid = tuple(int(g) for g in grid)
        self.roots = np.ascontiguousarray(roots, np.uint32)
        self.nodes = np.ascontiguousarray(nodes).view(NODE_DTYPE) if nodes.dtype != NODE_DTYPE else \
            np.ascontiguousarray(nodes)
        self.bricks = np.ascontiguousarray(bricks)
        self._s = OrcWorld((C.c_uint32 * 3)(*self.grid), self.roots.ctypes.data, self.nodes.ctypes.data,
                           len(self.nodes), self.bricks.ctypes.data, len(self.bricks))

    @property
    def ref(self):
        return C.byref(self._s)


def _params(p) -> OrcParams:
    q = OrcParams()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(OrcParams))
    return q


def trace_world(world: World, params, rays: np.ndarray, nthreads: int = 0):
    """trace_world_64tree_from over `rays` -> (full results, per-ray counters)."""
    rays = np.ascontiguousarray(rays).view(RAY_DTYPE)
    out = np.zeros(len(rays), RESULT_DTYPE)
    cnt = np.zeros(len(rays), COUNTER_DTYPE)
    q = _params(params)
    if len(rays):
        lib().orc_trace_world(world.ref, C.byref(q), rays.ctypes.data, len(rays), out.ctypes.data, cnt.ctypes.data,
                              _threads(nthreads))
    return out, cnt


def trace_chunk(world: World, rays: np.ndarray, nthreads: int = 0):
    """raycast_chunk_tree on chunk 0 -> (full results, counters)."""
    rays = np.ascontiguousarray(rays).view(RAY_DTYPE)
    out = np.zeros(len(rays), RESULT_DTYPE)
    cnt = np.zeros(len(rays), COUNTER_DTYPE)
    if len(rays):
        lib().orc_trace_chunk(world.ref, rays.ctypes.data, len(rays), out.ctypes.data, cnt.ctypes.data,
                              _threads(nthreads))
    return out, cnt


class ShadowParams(C.Structure):  # layout of arc_shadow_params
    _fields_ = [("mode", C.c_uint32), ("max_steps", C.c_uint32)]


def trace_shadow(world: World, mode: int, max_steps: int, rays: np.ndarray, nthreads: int = 0):
    """trace_shadow_vis (mode 0) / trace_shadow_vis_di (mode 1) -> (visibility f32, records)."""
    rays = np.ascontiguousarray(rays).view(RAY_DTYPE)
    vis = np.zeros(len(rays), np.float32)
    rec = np.zeros(len(rays), SHADOW_REC_DTYPE)
    sp = ShadowParams(mode, max_steps)
    if len(rays):
        lib().orc_trace_shadow(world.ref, C.byref(sp), rays.ctypes.data, len(rays), vis.ctypes.data, rec.ctypes.data,
                               _threads(nthreads))
    return vis, rec


RIGID_BODY_DTYPE = np.dtype([("world_to_local", "<f4", (16,)), ("local_to_world", "<f4", (16,)),
                             ("prev_local_to_world", "<f4", (16,)), ("local_bounds_and_root", "<f4", (4,)),
                             ("world_sphere", "<f4", (4,))])
RIGID_HIT_DTYPE = np.dtype([("t_voxels", "<f4"), ("hit_pos_m", "<f4", (3,)), ("prev_hit_pos_m", "<f4", (3,)),
                            ("normal", "<f4", (3,)), ("contact_ao", "<f4"), ("mat", "<u4"), ("voxel_hash", "<u4"),
                            ("hit", "<u4"), ("body", "<u4"), ("reserved", "<u4")])


def trace_rigid(world: World, bodies: np.ndarray, rays: np.ndarray) -> np.ndarray:
    """trace_rigid_bodies over `rays` with the first min(len(bodies), 32) instances."""
    rays = np.ascontiguousarray(rays).view(RAY_DTYPE)
    b = np.ascontiguousarray(bodies, RIGID_BODY_DTYPE)
    out = np.zeros(len(rays), RIGID_HIT_DTYPE)
    if len(rays):
        lib().orc_trace_rigid(world.ref, b.ctypes.data if len(b) else None, len(b), rays.ctypes.data, len(rays),
                              out.ctypes.data)
    return out


def query_points(world: World, points: np.ndarray) -> np.ndarray:
    """query_voxel_world_meters at each (x, y, z) metres."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    return np.array([lib().orc_query_point(world.ref, p.ctypes.data) for p in pts], np.uint32)


def trace_meters(world: World, rays: np.ndarray, nthreads: int = 0):
    """WorldMap::raycast_meters -> full results (t in voxel units)."""
    rays = np.ascontiguousarray(rays).view(RAY_DTYPE)
    out = np.zeros(len(rays), RESULT_DTYPE)
    if len(rays):
        lib().orc_trace_meters(world.ref, rays.ctypes.data, len(rays), out.ctypes.data, _threads(nthreads))
    return out


def compact(full: np.ndarray) -> np.ndarray:
    out = np.zeros(len(full), HIT_DTYPE)
    if len(full):
        lib().orc_compact(np.ascontiguousarray(full).ctypes.data, len(full), 1, out.ctypes.data)
    return out


def query_voxel(world: World, x, y, z) -> int:
    return int(lib().orc_query_voxel(world.ref, x, y, z))


def decode_dense(world: World, nthreads: int = 0) -> np.ndarray:
    X, Y, Z = (g * 256 for g in world.grid)
    d = np.zeros((Z, Y, X), np.uint8)
    lib().orc_decode_dense(world.ref, d.ctypes.data, _threads(nthreads))
    return d


def brute(dense: np.ndarray, rays: np.ndarray, world_units: bool = True, max_t_scale: float = 1.0,
          nthreads: int = 0) -> np.ndarray:
    """Independent dense unit-voxel walk (double precision)."""
    dense = np.ascontiguousarray(dense, np.uint8)
    Z, Y, X = dense.shape
    dims = (C.c_uint32 * 3)(X, Y, Z)
    rays = np.ascontiguousarray(rays).view(RAY_DTYPE)
    out = np.zeros(len(rays), BRUTE_DTYPE)
    if len(rays):
        lib().orc_brute_dense(dense.ctypes.data, dims, rays.ctypes.data, len(rays), int(world_units), max_t_scale,
                              out.ctypes.data, _threads(nthreads))
    return out


def primary_gbuffer(world: World, params, g, palette: np.ndarray, nthreads: int = 0):
    """primary_raymarch opaque path -> (depth_mat (H,W,2) u32, norm_motion (H,W,2) u32, albedo_ao (H,W) u32).

    `g` is an arc_gbuffer_params ctypes structure, `palette` the 256x8 float32 material table."""
    W, H = int(g.width), int(g.height)
    dm = np.zeros((H, W, 2), np.uint32)
    nm = np.zeros((H, W, 2), np.uint32)
    aa = np.zeros((H, W), np.uint32)
    pal = np.ascontiguousarray(palette, np.float32)
    assert pal.shape == (256, 8)
    q = _params(params)
    lib().orc_primary_gbuffer(world.ref, C.byref(q), C.byref(g), pal.ctypes.data, dm.ctypes.data, nm.ctypes.data,
                              aa.ctypes.data, _threads(nthreads))
    return dm, nm, aa


def f32_to_f16_bits(x: float) -> int:
    return int(lib().orc_f32_to_f16(float(x)))


SHARC_SAMPLE_DTYPE = np.dtype([("pos", "<f4", (3,)), ("normal", "<f4", (3,)), ("radiance", "<f4", (3,)),
                               ("reserved", "<u4")])
SHARC_PROBE_DTYPE = np.dtype([("pos", "<f4", (3,)), ("normal", "<f4", (3,)), ("jitter", "<f4", (2,))])


class Sharc:
    """Sequential SHARC cache (keys, accum, resolved as numpy arrays)."""

    def __init__(self, log2_buckets: int = 15):
        self.log2 = log2_buckets
        self.slots = 4 << log2_buckets
        self.keys = np.zeros(self.slots, np.uint32)
        self.accum = np.zeros(self.slots * 4, np.uint32)
        self.resolved = np.zeros(self.slots * 4, np.uint32)

    def make_key(self, pos, normal, cam, jitter=(0.0, 0.0)):
        b, k = C.c_uint32(), C.c_uint32()
        lib().orc_sharc_make_key((C.c_float * 3)(*pos), (C.c_float * 3)(*normal), (C.c_float * 3)(*cam), jitter[0],
                                 jitter[1], (1 << self.log2) - 1, C.byref(b), C.byref(k))
        return b.value, k.value

    def deposit(self, samples: np.ndarray, cam) -> None:
        s = np.ascontiguousarray(samples, SHARC_SAMPLE_DTYPE)
        if len(s):
            lib().orc_sharc_deposit(self.log2, self.keys.ctypes.data, self.accum.ctypes.data, s.ctypes.data, len(s),
                                    (C.c_float * 3)(*cam))

    def resolve(self) -> None:
        lib().orc_sharc_resolve(self.slots, self.keys.ctypes.data, self.accum.ctypes.data, self.resolved.ctypes.data)

    def query(self, probes: np.ndarray, cam) -> np.ndarray:
        q = np.ascontiguousarray(probes, SHARC_PROBE_DTYPE)
        out = np.zeros((len(q), 4), np.float32)
        if len(q):
            lib().orc_sharc_query(self.log2, self.resolved.ctypes.data, q.ctypes.data, len(q), (C.c_float * 3)(*cam),
                                  out.ctypes.data)
        return out


# ---- ReSTIR GI / DI passes (restir_oracle.c) ----------------------------------------------------
GI_RESERVOIR_DTYPE = np.dtype([("sample_pos", "<f4", (3,)), ("norm_m_flags", "<u4"), ("lo_rg", "<u4"),
                               ("lo_b_w", "<u4"), ("primary_depth", "<f4"), ("shadow_vis_m", "<u4")])
DI_RESERVOIR_DTYPE = np.dtype([("light_idx_seed", "<u4"), ("w_sum", "<f4"), ("m", "<f4"), ("w", "<f4")])


class _RestirIn(C.Structure):
    _fields_ = [("p", C.c_void_p), ("sun_trace", C.c_void_p), ("gi_trace", C.c_void_p), ("depth_mat", C.c_void_p),
                ("norm_motion", C.c_void_p), ("gi_in", C.c_void_p), ("di_in", C.c_void_p), ("sources", C.c_void_p),
                ("alias", C.c_void_p), ("palette", C.c_void_p), ("sharc_log2_buckets", C.c_uint32),
                ("sharc_resolved", C.c_void_p), ("bodies", C.c_void_p), ("n_bodies", C.c_uint32)]


def _restir(fn, world: World, params, depth_mat, norm_motion, gi_in, di_in, palette, sun_trace=None, gi_trace=None,
            sources=None, alias=None, sharc=None, bodies=None, nthreads: int = 0):
    n = params.width * params.height
    keep = []

    def arr(a, dtype=None):
        if a is None:
            return None
        a = np.ascontiguousarray(a if dtype is None else np.asarray(a).view(dtype) if np.asarray(a).dtype != dtype
                                 else a)
        keep.append(a)
        return a.ctypes.data if a.size else None

    inp = _RestirIn()
    inp.p = C.cast(C.pointer(params), C.c_void_p)
    keep.append(params)
    if sun_trace is not None:
        inp.sun_trace = C.cast(C.pointer(sun_trace), C.c_void_p)
    if gi_trace is not None:
        inp.gi_trace = C.cast(C.pointer(gi_trace), C.c_void_p)
    inp.depth_mat = arr(np.asarray(depth_mat, np.uint32))
    inp.norm_motion = arr(np.asarray(norm_motion, np.uint32))
    inp.gi_in = arr(np.asarray(gi_in).view(GI_RESERVOIR_DTYPE))
    inp.di_in = arr(np.asarray(di_in).view(DI_RESERVOIR_DTYPE))
    inp.sources = arr(sources)
    inp.alias = arr(alias)
    inp.palette = arr(np.asarray(palette, np.float32))
    if sharc is not None:
        inp.sharc_log2_buckets = sharc[0]
        inp.sharc_resolved = arr(np.asarray(sharc[1], np.uint32))
    if bodies is not None and len(bodies):
        inp.bodies = arr(np.ascontiguousarray(bodies, RIGID_BODY_DTYPE))
        inp.n_bodies = len(bodies)
    gi_out = np.zeros(n, GI_RESERVOIR_DTYPE)
    di_out = np.zeros(n, DI_RESERVOIR_DTYPE)
    f = getattr(lib(), fn)
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    f.restype = None
    f(world.ref, C.byref(inp), gi_out.ctypes.data, di_out.ctypes.data, _threads(nthreads))
    return gi_out, di_out


def restir_temporal(world: World, params, sun_trace, gi_trace, depth_mat, norm_motion, gi_in, di_in, palette,
                    sources=None, alias=None, sharc=None, bodies=None, nthreads: int = 0):
    """temporal_main over every pixel -> (GI reservoirs, DI reservoirs). sharc = (log2_buckets,
    resolved u32 table) or None; palette = 256 x 8 f32."""
    return _restir("orc_restir_temporal", world, params, depth_mat, norm_motion, gi_in, di_in, palette, sun_trace,
                   gi_trace, sources, alias, sharc, bodies, nthreads)


def restir_spatial(world: World, params, depth_mat, norm_motion, gi_in, di_in, palette, sources=None,
                   nthreads: int = 0):
    """spatial_main over every pixel -> (GI reservoirs, DI reservoirs)."""
    return _restir("orc_restir_spatial", world, params, depth_mat, norm_motion, gi_in, di_in, palette,
                   sources=sources, nthreads=nthreads)
