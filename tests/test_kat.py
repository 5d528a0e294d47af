"""Known-answer tests of the reference CPU raycaster (crates/arcadia-core/src/raycast_cpu.rs:334-446),
run on the oracle (both reference semantics) — and, marked gpu, on the HIP kernels."""
from __future__ import annotations

import numpy as np
import pytest

import arcadia_amd as A
from oracle import oracle as O
from tests import worlds as W

MAT_STONE_DEEP = 2


def _floor_world():
    """from_generator with y < 32 stone, air above (raycast_cpu.rs:337-349)."""
    dense = np.zeros((256, 256, 256), np.uint8)
    dense[:, :32, :] = MAT_STONE_DEEP
    w = A.HostWorld((1, 1, 1))
    w.set_chunk_dense(0, dense)
    w.pack()
    return w.pools()


def _ray(o, d, max_dist):
    r = np.zeros(1, A.RAY_DTYPE)
    r["ox"], r["oy"], r["oz"] = o
    r["dx"], r["dy"], r["dz"] = d
    r["t_max"] = max_dist
    return r


def _glam_normalize(v):
    v = np.asarray(v, np.float32)
    dot = np.float32(v[0] * v[0]) + np.float32(v[1] * v[1])
    dot = np.float32(dot + np.float32(v[2] * v[2]))
    rcp = np.float32(1.0) / np.sqrt(dot, dtype=np.float32)
    return tuple(np.float32(x * rcp) for x in v)


def _both(pools, o, d, max_dist):
    """Trace one voxel-unit ray with raycast_chunk_tree semantics and the same ray in meters with
    trace_world_64tree_from semantics; returns (chunk record, world record)."""
    ow = O.World(*pools)
    rc = _ray(o, d, max_dist)
    full_c, _ = O.trace_chunk(ow, rc)
    rw = _ray(tuple(np.float32(x) / np.float32(16.0) for x in o), d, np.float32(max_dist) / np.float32(16.0))
    full_w, _ = O.trace_world(ow, A.TraceParams.exact(), rw)
    return O.compact(full_c)[0], O.compact(full_w)[0]


FACE = {(1, 0): 0, (-1, 0): 1, (1, 1): 2, (-1, 1): 3, (1, 2): 4, (-1, 2): 5}  # (sign, axis) -> face code


@pytest.mark.parametrize("sem", ["chunk", "world"])
def test_cpu_raycast_hit(sem):  # raycast_cpu.rs:334-359
    hc, hw = _both(_floor_world(), (128.5, 50.0, 128.5), (0.0, -1.0, 0.0), 100.0)
    h = hc if sem == "chunk" else hw
    assert h["flags"] & 1
    assert h["mat"] == MAT_STONE_DEEP
    assert h["vy"] == 31
    assert h["face"] == FACE[(1, 1)]  # normal +Y
    assert abs(float(h["t_voxels"]) - 18.0) < 0.05


@pytest.mark.parametrize("sem", ["chunk", "world"])
def test_grazing_angle_face_and_edge(sem):  # raycast_cpu.rs:361-392
    pools = _floor_world()
    hc, hw = _both(pools, (10.0, 32.0001, 10.0), (1.0, 0.0, 0.0), 100.0)
    assert not ((hc if sem == "chunk" else hw)["flags"] & 1)
    d = _glam_normalize((1.0, -0.005, 0.0))
    hc, hw = _both(pools, (10.0, 32.5, 10.0), d, 200.0)
    h = hc if sem == "chunk" else hw
    assert h["flags"] & 1 and h["mat"] == MAT_STONE_DEEP and h["vy"] == 31
    assert h["face"] == FACE[(1, 1)]


@pytest.mark.parametrize("sem", ["chunk", "world"])
def test_nibble_row_skipping_sparse_and_semidense(sem):  # raycast_cpu.rs:394-428
    pools = W.sparse_world([(3, 0, 0, MAT_STONE_DEEP)])
    pick = (lambda hc, hw: hc) if sem == "chunk" else (lambda hc, hw: hw)
    h = pick(*_both(pools, (0.1, 0.5, 0.5), (1.0, 0.0, 0.0), 10.0))
    assert (h["vx"], h["vy"], h["vz"]) == (3, 0, 0)
    assert h["face"] == FACE[(-1, 0)]
    assert abs(float(h["t_voxels"]) - 2.9) < 1e-3
    h = pick(*_both(pools, (3.9, 0.5, 0.5), (-1.0, 0.0, 0.0), 10.0))
    assert (h["vx"], h["vy"], h["vz"]) == (3, 0, 0)
    assert h["face"] == FACE[(1, 0)]
    pools = W.sparse_world([(3, 0, 0, MAT_STONE_DEEP), (0, 0, 0, MAT_STONE_DEEP)])
    h = pick(*_both(pools, (1.1, 0.5, 0.5), (1.0, 0.0, 0.0), 10.0))
    assert (h["vx"], h["vy"], h["vz"]) == (3, 0, 0)
    assert h["face"] == FACE[(-1, 0)]
    assert abs(float(h["t_voxels"]) - 1.9) < 1e-3


@pytest.mark.parametrize("sem", ["chunk", "world"])
def test_multilevel_ancestor_caching_across_bricks(sem):  # raycast_cpu.rs:430-446
    pools = W.sparse_world([(5, 0, 0, MAT_STONE_DEEP)])
    hc, hw = _both(pools, (0.5, 0.5, 0.5), (1.0, 0.0, 0.0), 20.0)
    h = hc if sem == "chunk" else hw
    assert (h["vx"], h["vy"], h["vz"]) == (5, 0, 0)
    assert h["face"] == FACE[(-1, 0)]
    assert abs(float(h["t_voxels"]) - 4.5) < 1e-3


KAT_CASES = [
    ("floor", (128.5, 50.0, 128.5), (0.0, -1.0, 0.0), 100.0),
    ("skim", (10.0, 32.0001, 10.0), (1.0, 0.0, 0.0), 100.0),
    ("grazing", (10.0, 32.5, 10.0), _glam_normalize((1.0, -0.005, 0.0)), 200.0),
]


@pytest.mark.gpu
def test_kats_on_gpu(ctx):
    """The same KAT rays through the HIP kernels (both semantics), bit-exact against the oracle."""
    pools = _floor_world()
    ow = O.World(*pools)
    ctx.upload(*pools)
    for name, o, d, m in KAT_CASES:
        rc = _ray(o, d, m)
        got = ctx.trace(A.TraceParams.chunk(), rc)
        ref = O.compact(O.trace_chunk(ow, rc)[0])
        assert got.tobytes() == ref.tobytes(), name
        rw = _ray(tuple(np.float32(x) / np.float32(16.0) for x in o), d, np.float32(m) / np.float32(16.0))
        got = ctx.trace(A.TraceParams.exact(), rw)
        ref = O.compact(O.trace_world(ow, A.TraceParams.exact(), rw)[0])
        assert got.tobytes() == ref.tobytes(), name
    for vox, o, d, m in [([(3, 0, 0, 2)], (0.1, 0.5, 0.5), (1.0, 0.0, 0.0), 10.0),
                         ([(3, 0, 0, 2)], (3.9, 0.5, 0.5), (-1.0, 0.0, 0.0), 10.0),
                         ([(3, 0, 0, 2), (0, 0, 0, 2)], (1.1, 0.5, 0.5), (1.0, 0.0, 0.0), 10.0),
                         ([(5, 0, 0, 2)], (0.5, 0.5, 0.5), (1.0, 0.0, 0.0), 20.0)]:
        pools = W.sparse_world(vox)
        ctx.upload(*pools)
        rc = _ray(o, d, m)
        got = ctx.trace(A.TraceParams.chunk(), rc)
        ref = O.compact(O.trace_chunk(O.World(*pools), rc)[0])
        assert got.tobytes() == ref.tobytes()
        assert got["flags"][0] & 1


def test_meters_direction_follows_glam_normalize_or_zero():
    """raycast_meters normalizes its direction with glam 0.29.3's Vec3::normalize_or_zero
    (world_map.rs:470, Cargo.lock pins glam 0.29.3): rcp = 1 / length(); self * rcp when rcp is
    finite and > 0, else ZERO, and a zero direction returns None (:471-473). A zero, NaN or
    infinite direction and one whose squared length overflows therefore miss; a scaled direction
    traces exactly as the unit direction the rule produces."""
    pools = _floor_world()
    ow = O.World(*pools)
    o = (8.0, 3.0, 8.0)  # metres: 48 voxels up, the floor's top face is at y = 32 voxels
    unit = O.trace_meters(ow, _ray(o, (0.0, -1.0, 0.0), 10.0))[0]
    assert unit["hit"] and unit["vy"] == 31 and unit["mat"] == MAT_STONE_DEEP
    assert abs(float(unit["t_voxels"]) - 16.0) < 0.05  # 16 voxels (1 m) to the top face
    for d in ((0.0, 0.0, 0.0), (np.nan, -1.0, 0.0), (0.0, -np.inf, 0.0), (1e20, -1e20, 0.0)):
        r = O.trace_meters(ow, _ray(o, d, 10.0))[0]
        assert not r["hit"], d
    # (1e19, -1e19, 0): the squares stay finite (2e38 < FLT_MAX), so it normalizes and hits
    assert O.trace_meters(ow, _ray(o, (1e19, -1e19, 0.0), 10.0))[0]["hit"]
    # a scaled direction gives the record of the normalized one, bit for bit: by a power of two
    # always (d * 2^k scales the dot product by 4^k, its root by 2^k and rcp by 2^-k, all exact),
    # by other factors whenever the rule maps both to the same unit vector
    rng = np.random.default_rng(7)
    same = 0
    for _ in range(50):
        d = rng.normal(size=3).astype(np.float32)
        d[1] = -abs(d[1]) - np.float32(0.2)
        raw = O.trace_meters(ow, _ray(o, tuple(d), 50.0))
        assert raw[0]["hit"]
        for scale in (np.float32(0.25), np.float32(8.0)):
            a = O.trace_meters(ow, _ray(o, tuple(np.float32(x * scale) for x in d), 50.0))
            assert a.tobytes() == raw.tobytes()
        for scale in (np.float32(0.37), np.float32(5.0)):
            if _glam_normalize(tuple(np.float32(x * scale) for x in d)) == _glam_normalize(d):
                same += 1
                a = O.trace_meters(ow, _ray(o, tuple(np.float32(x * scale) for x in d), 50.0))
                assert a.tobytes() == raw.tobytes()
    assert same >= 5
