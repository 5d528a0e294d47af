"""Rigid voxel bodies and point queries: trace_rigid_bodies / trace_single_rigid_body
(traversal64.wgsl:1005-1244) and query_voxel_world_meters (:182-220).

CPU: the oracle restatement is pinned by equivalence with the exact world traversal (a body
with the identity transform whose tree is a chunk's tree walks exactly like
trace_world_64tree_from with no LOD and a 64-step cap), by closed-form answers for bodies turned
by quarter turns (exact in binary32), by nearest-body selection and the bounding-sphere reject,
and (point query) by the integer voxel query.
GPU: every field of every record equals the oracle's, for randomly rotated bodies too.
"""
from __future__ import annotations

import numpy as np
import pytest

import arcadia_amd as A
from arcadia_amd import rigid as RB
from oracle import oracle as O
from tests import worlds as W


def body_pools(voxels, dims=(16, 16, 16)):
    w = A.HostWorld((1, 1, 1))
    w.set_voxels(np.ascontiguousarray(voxels, np.int32))
    w.pack()
    return w.pools()


def ray(o, d, t_max):
    r = np.zeros(1, A.RAY_DTYPE)
    r["ox"], r["oy"], r["oz"] = o
    r["dx"], r["dy"], r["dz"] = d
    r["t_max"] = t_max
    return r


def kat_scene():
    """World: one stone voxel far away; body: voxels (2,3,4) mat 7 and (5,3,4) mat 9 in a 16^3 box,
    turned a quarter turn about y (R (x, y, z) = (z, y, -x)) around its centre, centre at (5, 3, 5) m."""
    world = body_pools(np.array([[200, 200, 200, 3]], np.int32))
    pools, root = RB.append_body_tree(world, body_pools(np.array([[2, 3, 4, 7], [5, 3, 4, 9]], np.int32)))
    inst = RB.body_instance(root, (16, 16, 16), RB.rot_y(90.0), (5.0, 3.0, 5.0))
    return pools, inst


def test_identity_body_walks_like_the_exact_world_traversal():
    pools = W.caves_world(2, (1, 1, 1), 60)
    inst = RB.body_instance(int(pools[1][0]), (256, 256, 256), np.eye(3), (8.0, 8.0, 8.0))
    rays = np.concatenate([W.random_rays(15000, (1, 1, 1), seed=5, inside=True, max_dist=40.0),
                           W.random_rays(5000, (1, 1, 1), seed=6, inside=False, max_dist=40.0)])
    ow = O.World(*pools)
    got = O.trace_rigid(ow, np.array([inst]), rays)
    p = A.TraceParams.exact(0)
    p.max_steps[0] = 64
    full, _ = O.trace_world(ow, p, rays)
    hit = (full["hit"] & 1) != 0
    assert np.array_equal(hit, got["hit"] == 1)
    assert hit.sum() > 3000 and (~hit).sum() > 3000
    for k in ("t_voxels", "mat", "voxel_hash"):
        assert np.array_equal(full[k][hit], got[k][hit]), k
    assert np.array_equal(full["hit_pos_m"][hit].view(np.uint32), got["hit_pos_m"][hit].view(np.uint32))
    # the normal rules differ only on edge ties (the world walk prefers the larger direction
    # component within an epsilon, the body walk the larger entry plane): compare away from ties
    same = np.all(full["normal"][hit] == got["normal"][hit], axis=1)
    assert same.mean() > 0.999, same.mean()
    assert np.all(got["body"][hit] == 0) and np.all(got["body"][~hit] == 0xFFFFFFFF)
    # misses: t = max_dist * 16 and the end point of the ray
    assert np.array_equal(got["t_voxels"][~hit], (rays["t_max"][~hit] * np.float32(16.0)).astype(np.float32))


def test_quarter_turned_body_closed_form():
    pools, inst = kat_scene()
    ow = O.World(*pools)
    # straight down onto local voxel (2, 3, 4): world centre (4.78125, 2.71875, 5.34375)
    down = O.trace_rigid(ow, np.array([inst]), ray((4.78125, 4.71875, 5.34375), (0.0, -1.0, 0.0), 10.0))[0]
    assert down["hit"] == 1 and down["mat"] == 7 and down["t_voxels"] == np.float32(31.5)
    assert down["hit_pos_m"].tolist() == [4.78125, 2.75, 5.34375] and down["normal"].tolist() == [0.0, 1.0, 0.0]
    # along +z at the voxels' height: local x decreases, so (5, 3, 4) (mat 9) comes first; its
    # local +x face is the world -z face
    side = O.trace_rigid(ow, np.array([inst]), ray((4.78125, 2.71875, 3.0), (0.0, 0.0, 1.0), 10.0))[0]
    assert side["hit"] == 1 and side["mat"] == 9
    assert side["normal"].tolist() == [0.0, 0.0, -1.0] and side["hit_pos_m"][2] == np.float32(5.125)
    assert side["t_voxels"] == np.float32((5.125 - 3.0) * 16.0)
    # the bounding sphere rejects a ray pointing away; a ray passing beside the voxels misses
    away = O.trace_rigid(ow, np.array([inst]), ray((4.78125, 2.71875, 3.0), (0.0, 0.0, -1.0), 10.0))[0]
    beside = O.trace_rigid(ow, np.array([inst]), ray((4.6, 4.71875, 5.34375), (0.0, -1.0, 0.0), 10.0))[0]
    assert away["hit"] == 0 and beside["hit"] == 0 and beside["body"] == 0xFFFFFFFF


def test_nearest_body_wins_and_the_walk_is_cut_at_it():
    pools, inst = kat_scene()
    far = inst.copy()
    far_inst = RB.body_instance(int(inst["local_bounds_and_root"].view(np.uint32)[3]), (16, 16, 16), RB.rot_y(90.0),
                                (5.0, 3.0, 8.0))
    r = ray((4.78125, 2.71875, 3.0), (0.0, 0.0, 1.0), 20.0)
    ow = O.World(*pools)
    for bodies, want in (([inst, far_inst], 0), ([far_inst, inst], 1)):
        h = O.trace_rigid(ow, np.array(bodies), r)[0]
        assert h["hit"] == 1 and h["body"] == want and h["t_voxels"] == np.float32((5.125 - 3.0) * 16.0)
    # only the far body: hit 3 m further
    h = O.trace_rigid(ow, np.array([far_inst]), r)[0]
    assert h["t_voxels"] == np.float32((8.125 - 3.0) * 16.0)
    # more than 32 instances: the first 32 only
    many = np.array([far.copy() for _ in range(32)] + [inst])
    many["world_sphere"][:32] = [100.0, 100.0, 100.0, 0.5]  # all far away
    assert O.trace_rigid(ow, many, r)[0]["hit"] == 0


def test_point_query_matches_integer_voxel_query():
    pools = W.caves_world(1, (2, 2, 2), 60)
    ow = O.World(*pools)
    rng = np.random.default_rng(3)
    pts = rng.uniform(-1.0, 33.0, (4000, 3)).astype(np.float32)
    got = O.query_points(ow, pts)
    vp = pts.astype(np.float32) * np.float32(16.0)
    inside = np.all((vp >= 0.0) & (vp < 512.0), axis=1)
    want = np.zeros(len(pts), np.uint32)
    for i in np.nonzero(inside)[0]:
        x, y, z = (int(v) for v in vp[i])
        want[i] = O.lib().orc_query_voxel(ow.ref, x, y, z)
    assert np.array_equal(got, want)
    assert (got[inside] != 0).sum() > 500 and (got[~inside] == 0).all()


# ---- GPU parity ------------------------------------------------------------------------------

def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def rotated_scene(seed=1, n_bodies=12):
    rng = np.random.default_rng(seed)
    world = W.caves_world(3, (2, 2, 2), 60)
    shapes = []
    for k in range(3):
        vox = rng.integers(0, 24, (300, 3))
        vox = np.concatenate([vox, rng.integers(1, 30, (300, 1))], 1)
        shapes.append(body_pools(vox.astype(np.int32)))
    pools, roots = world, []
    for sp in shapes:
        pools, r = RB.append_body_tree(pools, sp)
        roots.append(r)
    bodies = []
    for b in range(n_bodies):
        pos = rng.uniform(4.0, 28.0, 3)
        bodies.append(RB.body_instance(roots[b % 3], (24, 24, 24), random_rotation(rng), pos,
                                       prev_rot=random_rotation(rng), prev_pos_m=pos + rng.normal(size=3) * 0.1))
    return pools, np.array(bodies)


def gpu_rigid(ctx, bodies, rays):
    import torch

    d_r = torch.from_numpy(np.ascontiguousarray(rays).view(np.uint8)).cuda()
    d_o = torch.zeros(len(rays) * 64, dtype=torch.uint8, device="cuda")
    ctx.trace_rigid_device(bodies, d_r.data_ptr(), len(rays), d_o.data_ptr())
    torch.cuda.synchronize()
    return d_o.cpu().numpy().view(A.RIGID_HIT_DTYPE)


@pytest.mark.gpu
def test_gpu_rigid_matches_oracle(ctx):
    pools, inst = kat_scene()
    ctx.upload(*pools)
    kat_rays = np.concatenate([ray((4.78125, 4.71875, 5.34375), (0.0, -1.0, 0.0), 10.0),
                               ray((4.78125, 2.71875, 3.0), (0.0, 0.0, 1.0), 10.0),
                               ray((4.78125, 2.71875, 3.0), (0.0, 0.0, -1.0), 10.0)])
    got = gpu_rigid(ctx, np.array([inst]), kat_rays)
    want = O.trace_rigid(O.World(*pools), np.array([inst]), kat_rays)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    pools, bodies = rotated_scene()
    ctx.upload(*pools)
    rng = np.random.default_rng(9)
    rays = W.random_rays(30000, (2, 2, 2), seed=10, inside=True, max_dist=30.0)
    # rays aimed at the bodies' centres as well
    c = bodies["world_sphere"][:, :3]
    o = rng.uniform(0.0, 32.0, (20000, 3))
    d = c[rng.integers(0, len(c), 20000)] + rng.normal(size=(20000, 3)) * 0.4 - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    aimed = np.zeros(20000, A.RAY_DTYPE)
    aimed["ox"], aimed["oy"], aimed["oz"] = o[:, 0], o[:, 1], o[:, 2]
    aimed["dx"], aimed["dy"], aimed["dz"] = d[:, 0], d[:, 1], d[:, 2]
    aimed["t_max"] = 40.0
    rays = np.concatenate([rays, aimed])
    for nb in (len(bodies), 1, 0):
        got = gpu_rigid(ctx, bodies[:nb], rays)
        want = O.trace_rigid(O.World(*pools), bodies[:nb], rays)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), nb
    assert (want["hit"] == 1).sum() == 0 and (got["hit"] == 1).sum() == 0
    want = O.trace_rigid(O.World(*pools), bodies, rays)
    assert (want["hit"] == 1).sum() > 3000


@pytest.mark.gpu
def test_gpu_point_query_matches_oracle(ctx):
    import torch

    pools = W.caves_world(1, (2, 2, 2), 60)
    ctx.upload(*pools)
    rng = np.random.default_rng(4)
    pts = rng.uniform(-1.0, 33.0, (20000, 3)).astype(np.float32)
    pts[:5] = [[np.nan, 1.0, 1.0], [0.0, 0.0, 0.0], [32.0, 1.0, 1.0], [31.99999, 1.0, 1.0], [-0.0, 5.0, 5.0]]
    d_p = torch.from_numpy(pts).cuda()
    d_m = torch.zeros(len(pts), dtype=torch.int32, device="cuda")
    ctx.query_voxels_device(d_p.data_ptr(), len(pts), d_m.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(d_m.cpu().numpy().view(np.uint32), O.query_points(O.World(*pools), pts))
