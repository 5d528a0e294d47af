"""Pins of the CPU oracle beyond the reference KATs.

- an independent dense brute-force walker (double precision, shares no code) agrees on hit,
  voxel, face and material for every ray that is not a floating-point tie, and on t to 1e-5;
- closed forms for axis-aligned rays onto a heightmap;
- the two reference semantics coincide where the reference says they must (one chunk, no LOD);
- a second, pure-Python restatement agrees bit for bit, step count included — and planted
  mistakes in it are caught.
"""
from __future__ import annotations

import numpy as np
import pytest

import arcadia_amd as A
from oracle import oracle as O
from oracle import py_trace
from tests import worlds as W


def test_brute_force_agreement_heightmap():
    pools = W.heightmap_world(1)
    ow = O.World(*pools)
    rays = W.poses_rays(W.c1_poses(1), 160, 90)
    hits = O.compact(O.trace_world(ow, A.TraceParams.exact(), rays)[0])
    b = O.brute(O.decode_dense(ow), rays, world_units=True)
    nt = b["tie"] == 0
    assert nt.mean() > 0.6
    assert np.array_equal(b["hit"][nt], (hits["flags"][nt] & 1).astype(np.int32))
    both = nt & (b["hit"] == 1)
    for k in ("vx", "vy", "vz", "mat", "face"):
        assert np.array_equal(b[k][both], hits[k][both].astype(np.int32)), k
    rel = np.abs(b["t"][both] - hits["t_voxels"][both]) / np.maximum(1.0, b["t"][both])
    assert rel.max() < 1e-5


def test_brute_force_agreement_caves_random_rays():
    pools = W.caves_world(1)
    ow = O.World(*pools)
    rays = W.random_rays(20000, (2, 2, 2), seed=9, inside=False)
    hits = O.compact(O.trace_world(ow, A.TraceParams.exact(), rays)[0])
    b = O.brute(O.decode_dense(ow), rays, world_units=True)
    nt = b["tie"] == 0
    assert np.array_equal(b["hit"][nt], (hits["flags"][nt] & 1).astype(np.int32))
    both = nt & (b["hit"] == 1)
    assert both.sum() > 1000
    for k in ("vx", "vy", "vz", "mat", "face"):
        assert np.array_equal(b[k][both], hits[k][both].astype(np.int32)), k


def test_axis_aligned_closed_form():
    """A ray straight down from above the terrain at a voxel centre hits the top solid voxel of
    its column at t = o_y - h (voxel units) with normal +Y."""
    pools = W.heightmap_world(1)
    ow = O.World(*pools)
    dense = O.decode_dense(ow)
    rng = np.random.default_rng(3)
    xs, zs = rng.integers(0, 256, 300), rng.integers(0, 256, 300)
    r = np.zeros(300, A.RAY_DTYPE)
    r["ox"], r["oy"], r["oz"] = (xs + 0.5) / 16.0, 15.5, (zs + 0.5) / 16.0
    r["dy"] = -1.0
    r["t_max"] = 100.0
    h = O.compact(O.trace_world(ow, A.TraceParams.exact(), r)[0])
    for i in range(300):
        col = dense[zs[i], :, xs[i]]
        top = int(np.nonzero(col)[0].max())
        assert h["flags"][i] & 1 and h["vy"][i] == top and h["face"][i] == 2
        assert h["t_voxels"][i] == np.float32(15.5 * 16.0) - np.float32(top + 1)


def test_world_and_chunk_semantics_coincide_on_one_chunk():
    pools = W.heightmap_world(2)
    ow = O.World(*pools)
    rays = W.poses_rays(W.c1_poses(2), 160, 90)
    hw = O.compact(O.trace_world(ow, A.TraceParams.exact(max_steps=512), rays)[0])
    rv = rays.copy()
    for k in ("ox", "oy", "oz"):
        rv[k] = rv[k] * np.float32(16.0)
    rv["t_max"] = rv["t_max"] * np.float32(16.0)
    hc = O.compact(O.trace_chunk(ow, rv)[0])
    hit = (hw["flags"] & 1) == 1
    assert np.array_equal(hit, (hc["flags"] & 1) == 1)
    for k in ("vx", "vy", "vz", "mat", "face", "steps", "t_voxels"):
        assert np.array_equal(hw[k][hit], hc[k][hit]), k


def _py_vs_c(pools, params, rays, mutate=None):
    full, cnt = O.trace_world(O.World(*pools), params, rays)
    c = O.compact(full)
    g, roots, nodes, bricks = pools
    mism = []
    for i in range(len(rays)):
        p = py_trace.trace(g, roots, nodes, bricks, params, rays[i], mutate=mutate)
        ref = (int(c["flags"][i] & 1), int(c["steps"][i]))
        got = (p["hit"], p["steps"])
        if p["hit"]:
            ref += (int(c["vx"][i]), int(c["vy"][i]), int(c["vz"][i]), int(c["face"][i]), int(c["mat"][i]),
                    float(c["t_voxels"][i]))
            got += (p["vx"], p["vy"], p["vz"], p["face"], p["mat"], float(p["t"]))
        if ref != got or (mutate is None and p["node_loads"] != int(cnt["node_loads"][i])):
            mism.append(i)
    return mism


def test_python_restatement_matches_c_oracle():
    pools = W.caves_world(1)
    rays = np.concatenate([W.random_rays(150, (2, 2, 2), seed=4, inside=False),
                           W.poses_rays([((13.7, 30.2, 11.9), 0.6, -0.55)], 16, 9, fov=65.0)])
    assert _py_vs_c(pools, A.TraceParams.exact(), rays) == []
    p = A.TraceParams.reference_default(2, ignore_water=True)
    p.lod_dist_m[2], p.lod_step[2] = 5.0, 12
    assert _py_vs_c(pools, p, rays) == []


def edge_rays(n=40):
    """Shallow rays that land on the y = 32 floor 1e-5 voxel past an x-plane of the cell the
    descent stops in (the solid floor collapses to 16^3 uniform cells, so a multiple of 16): the
    cell's x and y entry planes differ by far less than the face tie-break epsilon, so the normal
    comes from the tie-break toward the larger |dir| component (traversal64.wgsl:477-497)."""
    r = np.zeros(n, A.RAY_DTYPE)
    d = np.array([1.0, -0.5, 0.0])
    d /= np.linalg.norm(d)
    for i in range(n):
        x_hit = 16 * (2 + i % 13) + 1e-5
        r[i]["ox"], r[i]["oy"], r[i]["oz"] = (x_hit - 16.0) / 16.0, 40.0 / 16.0, (5 + i + 0.5) / 16.0
        r[i]["dx"], r[i]["dy"], r[i]["dz"] = d
        r[i]["t_max"] = 50.0
    return r


def floor_pools():
    dense = np.zeros((256, 256, 256), np.uint8)
    dense[:, :32, :] = 2
    w = A.HostWorld((1, 1, 1))
    w.set_chunk_dense(0, dense)
    w.pack()
    return w.pools()


def test_edge_ties_resolve_toward_larger_direction_component():
    h = O.compact(O.trace_world(O.World(*floor_pools()), A.TraceParams.exact(), edge_rays())[0])
    assert (h["flags"] & 1).all()
    assert (h["face"] == 1).mean() > 0.9  # -X (|dx| > |dy|), not the +Y face the ray enters through
    assert _py_vs_c(floor_pools(), A.TraceParams.exact(), edge_rays()) == []


@pytest.mark.parametrize("mutate", ["eps0", "steps+1"])
def test_python_cross_check_catches_planted_mistakes(mutate):
    pools = floor_pools()
    rays = np.concatenate([edge_rays(), W.random_rays(50, (1, 1, 1), seed=1)])
    assert len(_py_vs_c(pools, A.TraceParams.exact(), rays, mutate=mutate)) > 0


def test_ancestor_restart_does_not_change_results():
    """The register-cached ancestor restart is a pure optimisation: a restatement that always
    restarts at the root gives identical outputs (node-load counts differ)."""
    pools = W.caves_world(1)
    rays = W.random_rays(120, (2, 2, 2), seed=8, inside=False)
    assert _py_vs_c(pools, A.TraceParams.exact(), rays, mutate="no_restart") == []


def test_oracle_counters_are_consistent():
    pools = W.caves_world(1)
    rays = W.random_rays(5000, (2, 2, 2), seed=2, inside=False)
    full, cnt = O.trace_world(O.World(*pools), A.TraceParams.exact(), rays)
    assert np.array_equal(cnt["level_loads"].sum(1), cnt["node_loads"])
    assert np.array_equal(cnt["steps"], full["steps"])
    assert np.array_equal(cnt["restarts"].sum(1) + cnt["empty_chunk_steps"], cnt["steps"])
