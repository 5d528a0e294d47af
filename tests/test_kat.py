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
