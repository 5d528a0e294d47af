"""Shadow any-hit walkers: trace_shadow_vis / trace_shadow_vis_di (traversal64.wgsl:571-773 /
781-989).

CPU: the oracle restatement is pinned by closed-form known answers on constructed worlds (the
LOD cell rules at 48 / 192 voxels, the water skip with the sun walker reading the node's LOD
material and the DI walker the brick, the budget heuristics, the DI emitter window) and by
equivalence with the exact world traversal (itself pinned against the reference's KATs and a
brute-force walker) wherever the walkers' min_shift stays 0.
GPU: every record (t, steps, voxel, outcome) and visibility of the kernel equals the oracle's.
"""
from __future__ import annotations

import numpy as np
import pytest

import arcadia_amd as A
from arcadia_amd import shadow_rays as S
from oracle import oracle as O
from tests import worlds as W

SUN, DI = A.ARC_SHADOW_SUN, A.ARC_SHADOW_DI


def box(x0, y0, z0, x1, y1, z1, m):
    z, y, x = np.mgrid[z0:z1, y0:y1, x0:x1]
    return np.stack([x.ravel(), y.ravel(), z.ravel(), np.full(x.size, m)], 1).astype(np.int32)


def vox(*pts):
    return np.array(pts, np.int32).reshape(-1, 4)


def world_of(grid, voxels):
    w = A.HostWorld(grid)
    w.set_voxels(np.ascontiguousarray(voxels, np.int32))
    w.pack()
    return w.pools()


def rays_x(starts, t_max_m, d=(1.0, 0.0, 0.0)):
    """Rays from voxel-unit origins along d (world metres = voxels / 16)."""
    s = np.asarray(starts, np.float64)
    r = np.zeros(len(s), A.RAY_DTYPE)
    r["ox"], r["oy"], r["oz"] = s[:, 0] / 16.0, s[:, 1] / 16.0, s[:, 2] / 16.0
    r["dx"], r["dy"], r["dz"] = d
    r["t_max"] = t_max_m
    return r


def shadow(pools, mode, steps, rays):
    return O.trace_shadow(O.World(*pools), mode, steps, rays)


# ---- closed-form known answers ---------------------------------------------------------------

def lod_world():
    # S: stone voxel at (90, 40, 40); T: stone voxel at (216, 40, 40), in a 256^3 chunk.
    return world_of((1, 1, 1), vox((90, 40, 40, 5), (216, 40, 40, 5)))


def test_sun_walker_tests_4_cubed_cells_beyond_48_voxels_and_16_cubed_beyond_192():
    pools = lod_world()
    r = rays_x([(8.5, 40.5, 40.5), (8.5, 42.5, 40.5), (8.5, 44.5, 40.5), (8.5, 48.5, 40.5)], 14.0)
    vis, rec = shadow(pools, SUN, 72, r)
    # through S: blocked where its 4^3 cell starts (t = 88 - 8.5, min_shift 2 beyond t = 48)
    assert rec["outcome"][0] == A.ARC_SHADOW_BLOCKED and (rec["vx"][0], rec["vy"][0]) == (88, 40)
    assert rec["t_voxels"][0] == np.float32(79.5)
    # misses S but crosses its 4^3 cell: still blocked there (the exact traversal misses)
    assert rec["outcome"][1] == A.ARC_SHADOW_BLOCKED and (rec["vx"][1], rec["vy"][1]) == (88, 42)
    # outside S's 4^3 cell; crosses T's 16^3 cell beyond t = 192: blocked at x = 208 (t = 199.5)
    assert rec["outcome"][2] == A.ARC_SHADOW_BLOCKED and (rec["vx"][2], rec["vy"][2]) == (208, 44)
    assert rec["t_voxels"][2] == np.float32(199.5)
    # clear of both 16^3 cells: reaches max_dist (14 m = 224 voxels)
    assert rec["outcome"][3] == A.ARC_SHADOW_CLEAR and rec["t_voxels"][3] >= np.float32(224.0)
    assert vis.tolist() == [0.0, 0.0, 0.0, 1.0]
    full, _ = O.trace_world(O.World(*pools), A.TraceParams.exact(0), r)
    assert full["hit"][0] & 1 and full["vx"][0] == 90 and not (full["hit"][1] & 1) and not (full["hit"][2] & 1)


def test_di_walker_keeps_voxel_precision_near_the_target():
    pools = lod_world()
    # The ray beside S through S's 4^3 cell, target beyond T. Steps: x = 8 (chunk root, empty
    # 64^3 block), 64 (empty 16^3 block), 80, 84 (empty 4^3 cells), 88 (S's cell, step index 4:
    # voxel level, the DI rule needs step > 6), 89, 90, then 91 at step index 7 with > 32 voxels
    # left: min_shift 2, the restart lands on S's cell, blocked at voxel 91 (the sun walker,
    # whose rule is t > 48 or step > 6, is blocked at 88). With the target 20 voxels past the
    # cell's start the walker stays at voxel level there and passes.
    r = rays_x([(8.5, 42.5, 40.5), (8.5, 42.5, 40.5)], 14.0)
    r["t_max"][1] = (88.0 - 8.5 + 20.0) / 16.0
    vis, rec = shadow(pools, DI, 72, r)
    assert rec["outcome"][0] == A.ARC_SHADOW_BLOCKED and rec["vx"][0] == 91 and rec["steps"][0] == 8
    assert rec["outcome"][1] == A.ARC_SHADOW_CLEAR and vis[1] == 1.0


def test_water_is_stepped_over_sun_reads_lod_material_di_reads_the_brick():
    # uniform water leaf [64,68)^3; mixed leaf [100,104)x[64,68)^2 of water with one stone voxel
    pools = world_of((1, 1, 1), np.concatenate([box(64, 64, 64, 68, 68, 68, 18), box(100, 64, 64, 104, 68, 68, 18),
                                               vox((101, 65, 65, 7))]))
    g, roots, nodes, bricks = pools
    nd = nodes.view(O.NODE_DTYPE)
    mixed = [i for i in range(len(nd)) if (int(nd["child_ptr"][i]) >> 30) == 2]  # leaf, not uniform
    assert len(mixed) == 1 and int(nd["lod_payload"][mixed[0]]) >> 24 == 18  # the leaf's LOD material: water
    r = rays_x([(60.5, 65.5, 65.5)], 3.0)
    vis_s, rec_s = shadow(pools, SUN, 72, r)
    vis_d, rec_d = shadow(pools, DI, 72, r)
    assert rec_s["outcome"][0] == A.ARC_SHADOW_CLEAR and vis_s[0] == 1.0  # stone voxel taken as its leaf's water
    assert rec_d["outcome"][0] == A.ARC_SHADOW_BLOCKED and (rec_d["vx"][0], rec_d["vy"][0]) == (101, 65)
    assert rec_d["t_voxels"][0] == np.float32(101.0 - 60.5)
    full, _ = O.trace_world(O.World(*pools), A.TraceParams.exact(0), r)
    assert full["hit"][0] & 1 and full["vx"][0] == 64 and full["mat"][0] == 18  # the exact trace stops at water


def test_sun_budget_heuristic_is_dark_below_32_m_and_lit_above():
    # 48 m tall world; one voxel per chunk far from the rays keeps every chunk root allocated
    pools = world_of((1, 3, 1), vox((250, 10, 250, 3), (250, 300, 250, 3), (250, 600, 250, 3)))
    r = rays_x([(8.5, 160.5, 8.5), (8.5, 640.5, 8.5)], 15.0)  # 10 m and 40 m high
    vis, rec = shadow(pools, SUN, 2, r)
    assert rec["outcome"].tolist() == [A.ARC_SHADOW_BUDGET_DARK, A.ARC_SHADOW_BUDGET_LIT]
    assert rec["steps"].tolist() == [2, 2] and vis.tolist() == [0.0, 1.0]
    assert rec["t_voxels"][0] == np.float32(128.0 - 8.5)  # two 64-voxel empty blocks crossed
    vis, rec = shadow(pools, SUN, 72, r)
    assert rec["outcome"].tolist() == [A.ARC_SHADOW_CLEAR, A.ARC_SHADOW_CLEAR]


def test_di_emitter_window_and_step_budget():
    # emitter E (mat 19) at (120, 40, 40); stone markers at (x, 43, 40) for x in [8, 48) keep the
    # walker at voxel level near the origin (their 4^3 cells are the ray's cells)
    marks = vox(*[(x, 43, 40, 4) for x in range(8, 48, 4)])
    pools = world_of((1, 1, 1), np.concatenate([vox((120, 40, 40, 19)), marks]))
    r = rays_x([(52.5, 40.5, 40.5)] * 3, 1.0)
    t_e = 120.0 - 52.5
    r["t_max"] = np.array([t_e + 10.0, t_e + 20.0, t_e + 10.0]) / 16.0
    r["oy"][2] = 41.5 / 16.0  # passes beside E
    vis, rec = shadow(pools, DI, 72, r)
    assert rec["outcome"].tolist() == [A.ARC_SHADOW_EMITTER, A.ARC_SHADOW_BLOCKED, A.ARC_SHADOW_CLEAR]
    assert rec["t_voxels"][0] == np.float32(t_e) and vis.tolist() == [1.0, 0.0, 1.0]
    vis_s, rec_s = shadow(pools, SUN, 72, r)  # the sun walker has no emitter rule
    assert rec_s["outcome"][0] == A.ARC_SHADOW_BLOCKED
    # through the markers' cells from x = 8.5: voxel steps (min_shift 0 below t = 48), so the DI
    # walker's 32-step budget (ray_steps.w = 5 clamps to 32) ends far short of the target
    r2 = rays_x([(8.5, 41.5, 40.5)], 15.0)
    vis2, rec2 = shadow(pools, DI, 5, r2)
    assert rec2["outcome"][0] == A.ARC_SHADOW_BUDGET_DARK and rec2["steps"][0] == 32 and vis2[0] == 0.0
    # the sun walker moves to 4^3 cells after 6 steps and stops at a marker's cell
    vis3, rec3 = shadow(pools, SUN, 72, r2)
    assert rec3["outcome"][0] == A.ARC_SHADOW_BLOCKED and rec3["steps"][0] == 8


def test_rays_missing_the_world_are_visible():
    pools = lod_world()
    r = rays_x([(-40.0, 40.5, 40.5), (300.0, 40.5, 40.5)], 14.0)
    r["dx"][1] = 1.0  # starts beyond the box, points away
    r["t_max"][0] = 1.0  # ends before the box
    for mode in (SUN, DI):
        vis, rec = shadow(pools, mode, 72, r)
        assert rec["outcome"].tolist() == [A.ARC_SHADOW_MISS_WORLD] * 2 and vis.tolist() == [1.0, 1.0]


# ---- equivalence with the exact traversal where min_shift stays 0 ----------------------------

@pytest.mark.parametrize("mode,steps,t_max", [(SUN, 7, 3.0), (DI, 32, 2.0)])
def test_walkers_equal_the_exact_traversal_at_voxel_precision(mode, steps, t_max):
    """Caves world (no water, no emissive materials): SUN stays at min_shift 0 while t <= 48 and
    steps <= 6, DI while at most 32 voxels remain; there a walker is blocked exactly where the exact
    traversal (lod 0, same step budget) hits, after the same number of steps."""
    pools = W.caves_world(5, (2, 2, 2), 80)
    rays = W.random_rays(20000, (2, 2, 2), seed=11, inside=True, max_dist=t_max)
    vis, rec = shadow(pools, mode, steps, rays)
    p = A.TraceParams.exact(0)
    p.max_steps[0] = steps
    full, _ = O.trace_world(O.World(*pools), p, rays)
    hit = (full["hit"] & 1) != 0
    blocked = rec["outcome"] == A.ARC_SHADOW_BLOCKED
    assert np.array_equal(hit, blocked)
    assert hit.sum() > 2000 and (~hit).sum() > 2000
    assert np.array_equal(full["steps"], rec["steps"])
    for k in ("vx", "vy", "vz"):
        assert np.array_equal(full[k][hit].astype(np.uint16), rec[k][hit])


# ---- GPU parity ------------------------------------------------------------------------------

def gpu_shadow(ctx, mode, steps, rays):
    import torch

    d_r = torch.from_numpy(np.ascontiguousarray(rays).view(np.uint8)).cuda()
    d_v = torch.full((max(len(rays), 1),), -1.0, dtype=torch.float32, device="cuda")
    d_rec = torch.zeros(max(len(rays), 1) * 16, dtype=torch.uint8, device="cuda")
    ctx.trace_shadow_device(mode, steps, d_r.data_ptr(), len(rays), d_v.data_ptr(), d_rec.data_ptr())
    torch.cuda.synchronize()
    return d_v.cpu().numpy()[: len(rays)], d_rec.cpu().numpy().view(A.SHADOW_REC_DTYPE)[: len(rays)]


def same_records(a, b):
    return np.array_equal(a.view(np.uint32).reshape(-1, 4), b.view(np.uint32).reshape(-1, 4))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [SUN, DI])
def test_gpu_walkers_match_oracle_on_kat_worlds(ctx, mode):
    cases = [(lod_world(), rays_x([(8.5, y + 0.5, 40.5) for y in range(36, 52)], 14.0)),
             (world_of((1, 1, 1), np.concatenate([box(64, 64, 64, 68, 68, 68, 18), box(100, 64, 64, 104, 68, 68, 18),
                                                 vox((101, 65, 65, 7))])), rays_x([(60.5, 65.5, 65.5)], 3.0))]
    for pools, rays in cases:
        ctx.upload(*pools)
        for steps in (2, 7, 72):
            vis, rec = gpu_shadow(ctx, mode, steps, rays)
            ovis, orec = shadow(pools, mode, steps, rays)
            assert np.array_equal(vis, ovis) and same_records(rec, orec), (mode, steps)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [SUN, DI])
def test_gpu_walkers_match_oracle_on_caves_world(ctx, mode):
    pools = W.caves_world(1, (2, 2, 2), 120)
    ctx.upload(*pools)
    rng_rays = [W.random_rays(40000, (2, 2, 2), seed=3, inside=True, max_dist=40.0),
                W.random_rays(20000, (2, 2, 2), seed=4, inside=False, max_dist=60.0)]
    # sun rays from primary hits of a camera looking over the cave world
    cam = A.camera_look((4.0, 22.0, 5.0), 0.7, -0.5, 70.0, 160, 90)
    pr, _ = A.primary_rays(cam, 160, 90, 250.0)
    full, _ = O.trace_world(O.World(*pools), A.TraceParams.exact(0), pr)
    sun = np.array([0.35, 0.8, 0.45])
    sun /= np.linalg.norm(sun)
    hits = O.compact(full)
    rays = np.concatenate(rng_rays + [S.sun_rays(pr, hits, sun, seed=7), S.toward_rays(pr, hits, seed=8)])
    for steps in (72, 36):
        vis, rec = gpu_shadow(ctx, mode, steps, rays)
        ovis, orec = shadow(pools, mode, steps, rays)
        assert np.array_equal(vis, ovis), (mode, steps)
        assert same_records(rec, orec), (mode, steps)
        outcomes = np.bincount(orec["outcome"], minlength=6)
        assert outcomes[A.ARC_SHADOW_BLOCKED] > 1000 and outcomes[A.ARC_SHADOW_CLEAR] > 1000, outcomes


@pytest.mark.gpu
def test_gpu_walker_autotune_and_shapes_keep_records(ctx):
    """The walkers' launch class (ARC_LAUNCH_CLASS_SHADOW): arc_autotune_shadow installs a shape, and
    records under it and under extreme shapes equal the oracle's."""
    import torch

    pools = W.caves_world(1, (2, 2, 2), 120)
    ctx.upload(*pools)
    rays = W.random_rays(30000, (2, 2, 2), seed=5, inside=True, max_dist=40.0)
    saved = ctx.launch_config(A.ARC_LAUNCH_CLASS_SHADOW)
    try:
        for mode in (A.ARC_SHADOW_SUN, A.ARC_SHADOW_DI):
            ovis, orec = shadow(pools, mode, 72, rays)
            n = len(rays)
            d_r = torch.from_numpy(np.ascontiguousarray(rays).view(np.uint8)).cuda()
            d_v = torch.empty(n, dtype=torch.float32, device="cuda")
            d_rec = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
            cfg, ms = ctx.autotune_shadow(mode, 72, d_r.data_ptr(), n, d_v.data_ptr(), d_rec.data_ptr())
            assert ms > 0.0 and ctx.launch_config(A.ARC_LAUNCH_CLASS_SHADOW) == cfg
            for shape in (None, (1, 64, 1, 1), (0, 128, 48, 256)):
                if shape is not None:
                    ctx.set_launch_config_class(A.ARC_LAUNCH_CLASS_SHADOW, *shape)
                vis, rec = gpu_shadow(ctx, mode, 72, rays)
                assert np.array_equal(vis, ovis), (mode, shape)
                assert same_records(rec, orec), (mode, shape)
    finally:
        ctx.set_launch_config_class(A.ARC_LAUNCH_CLASS_SHADOW, *[saved[k] for k in
                                    ("blocks_per_cu", "block_threads", "refill_lanes", "batch_rays")])
