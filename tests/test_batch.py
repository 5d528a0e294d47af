"""Batched uneven views traced with their diffuse bounces in one persistent launch.

CPU tests pin the shared ray generators: padded view rays against the compact primary-ray
generator, dense bounces against the compact bounce generator, and the bounce directions against a
float64 recomputation with the true sin / cos (the generators use their own polynomial so host and
device agree bit for bit). GPU tests compare every record of a batch launch (and of hipGraph
replays of it) with the oracle tracing the host-generated rays depth by depth.
"""
from __future__ import annotations

import numpy as np
import pytest

import arcadia_amd as A
from oracle import oracle as O
from tests import worlds as W

MASK64 = (1 << 64) - 1


def splitmix64(z: int) -> int:
    z = (z + 0x9E3779B97F4A7C15) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def test_view_rays_are_padded_primary_rays():
    cam = A.camera_look((3.0, 14.0, 2.0), 0.4, -0.4, 70.0, 101, 59)
    vr = A.view_rays(cam, 101, 59)
    pr, _ = A.primary_rays(cam, 101, 59)
    live = vr["t_max"] > 0
    assert len(vr) == 13 * 8 * 64 and live.sum() == 101 * 59
    assert np.array_equal(vr[live].view(np.uint32), pr.view(np.uint32))
    assert np.all(vr["t_max"][~live] == -1.0)


def _fake_hits(n, rng):
    hits = np.zeros(n, A.HIT_DTYPE)
    hits["flags"] = rng.integers(0, 2, n)
    hits["face"] = rng.integers(0, 6, n)
    hits["t_voxels"] = rng.uniform(1.0, 500.0, n).astype(np.float32)
    return hits


def test_dense_bounces_match_compact_bounces():
    rng = np.random.default_rng(3)
    rays = W.random_rays(5000, (2, 2, 2), seed=4)
    hits = _fake_hits(5000, rng)
    dense = A.bounce_rays_dense(rays, hits, index_base=0, seed=9)
    comp, src = A.bounce_rays(rays, hits, seed=9)
    hit = (hits["flags"] & 1) == 1
    assert np.array_equal(src, np.nonzero(hit)[0])
    assert np.array_equal(dense[hit].view(np.uint32), comp.view(np.uint32))
    assert np.all(dense["t_max"][~hit] == -1.0)
    shifted = A.bounce_rays_dense(rays, hits, index_base=5000, seed=9)
    assert not np.array_equal(shifted[hit]["dx"], dense[hit]["dx"])  # the index feeds the hash


def test_bounce_directions_match_float64_trig():
    """Directions from the polynomial sin/cos agree with a float64 recomputation (true trig)."""
    rng = np.random.default_rng(5)
    n = 3000
    rays = W.random_rays(n, (2, 2, 2), seed=6)
    hits = _fake_hits(n, rng)
    hits["flags"] = 1
    seed = 77
    got = A.bounce_rays_dense(rays, hits, index_base=123, seed=seed)
    for i in range(0, n, 7):
        face = int(hits["face"][i])
        nrm = np.zeros(3)
        nrm[face >> 1] = -1.0 if face & 1 else 1.0
        sn = nrm + np.array([0.0, 0.22, 0.0])
        sn /= np.linalg.norm(sn)
        up = np.array([0.0, 0.0, 1.0]) if abs(sn[2]) < 0.999 else np.array([1.0, 0.0, 0.0])
        t = np.cross(up, sn)
        t /= np.linalg.norm(t)
        b = np.cross(sn, t)
        h = splitmix64(seed ^ splitmix64(123 + i + 0x51ED270B))
        u1 = (h & 0xFFFFFF) / 16777216.0
        u2 = ((h >> 24) & 0xFFFFFF) / 16777216.0
        phi = float(np.float32(6.2831853) * np.float32(u2))
        d = t * (np.sqrt(u1) * np.cos(phi)) + b * (np.sqrt(u1) * np.sin(phi)) + sn * np.sqrt(1.0 - u1)
        g = np.array([got["dx"][i], got["dy"][i], got["dz"][i]], np.float64)
        assert np.max(np.abs(g - d)) < 2e-6, (i, g, d)
        assert abs(np.linalg.norm(g) - 1.0) < 1e-5


def host_batch(pools, views, p0, pb, n_bounces, bmax, seed):
    """The batch's records from host-generated rays traced by the oracle, depth by depth."""
    rays = np.concatenate([A.view_rays(cam, w, h, md) for cam, w, h, md in views])
    n = len(rays)
    world = O.World(*pools)
    out = []
    cur = rays
    for d in range(n_bounces + 1):
        full, _ = O.trace_world(world, p0 if d == 0 else pb, cur)
        hits = O.compact(full)
        out.append(hits)
        if d < n_bounces:
            cur = A.bounce_rays_dense(cur, hits, index_base=d * n, seed=seed, max_dist_m=bmax)
    return np.concatenate(out), n


def small_views(seed=0, k=5, base=(13.7, 30.2, 11.9)):
    rng = np.random.default_rng(seed)
    views = []
    for i in range(k):
        w = int(rng.integers(24, 330))
        h = max(1, w * 9 // 16 + int(rng.integers(-3, 4)))
        eye = tuple(float(v) for v in np.array(base) + rng.uniform(-6, 6, 3))
        cam = A.camera_look(eye, float(rng.uniform(0, 6.28)), float(rng.uniform(-0.9, -0.2)), 70.0, w, h)
        views.append((cam, w, h, 250.0))
    return views


@pytest.mark.gpu
def test_gpu_batch_bit_exact_and_graph_replays(ctx):
    import torch

    pools = W.caves_world(1)
    ctx.upload(*pools)
    views = small_views(0, 6)
    p0, pb = A.TraceParams.exact(0), A.TraceParams.exact(2, ignore_water=True)
    ref, n = host_batch(pools, views, p0, pb, 2, 26.0, 21)
    b = ctx.batch(views, p0, pb, n_bounces=2, bounce_max_dist_m=26.0, seed=21)
    assert b.slots == n and b.view_first(len(views)) == n
    print(f"[batch] {len(views)} views, {n} slots", flush=True)
    d = torch.zeros(b.records() * 16, dtype=torch.uint8, device="cuda")
    b.trace(d.data_ptr())
    torch.cuda.synchronize()
    got = d.cpu().numpy().view(A.HIT_DTYPE)
    bad = W.diff_records(got, ref)
    assert len(bad) == 0, f"{len(bad)} of {len(ref)} records differ, first {bad[:5]}"
    assert (ref["flags"][n:2 * n] & 1).sum() > 1000  # bounces were traced
    # graph: replays interleaved with eager launches into another buffer, each checked
    g = torch.full((b.records() * 16,), 7, dtype=torch.uint8, device="cuda")
    b.capture(g.data_ptr())
    e = torch.zeros_like(d)
    for _ in range(3):
        g.fill_(7)
        b.replay()
        b.trace(e.data_ptr())
        torch.cuda.synchronize()
        assert len(W.diff_records(g.cpu().numpy().view(A.HIT_DTYPE), ref)) == 0
        assert len(W.diff_records(e.cpu().numpy().view(A.HIT_DTYPE), ref)) == 0
    for cfg in ({"block_threads": 128, "refill_lanes": 8, "batch_rays": 32}, {"blocks_per_cu": 2, "refill_lanes": 48},
                {"block_threads": 1024, "refill_lanes": 24}):  # 1024: one workgroup per CU, every chunk root in LDS
        b.set_launch_config(**cfg)
        e.zero_()
        b.trace(e.data_ptr())
        torch.cuda.synchronize()
        assert len(W.diff_records(e.cpu().numpy().view(A.HIT_DTYPE), ref)) == 0, cfg
    ctx.upload(*pools)  # pools moved: the captured graph must be refused
    with pytest.raises(A.ArcError):
        b.replay()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("classes", [("ref", 0, "ref", 1), ("exact", 0, "ref", 2)])
def test_gpu_batch_lod_params(ctx, classes):
    """The path kernel with the reference-default quality settings (LOD floors and step caps on),
    and with LOD on for the bounces only: the LOD tests then run per lane with each depth's own
    thresholds."""
    import torch

    def params(kind, cls, water):
        return (A.TraceParams.reference_default if kind == "ref" else A.TraceParams.exact)(cls, ignore_water=water)

    pools = W.caves_world(1)
    ctx.upload(*pools)
    views = small_views(4, 5)
    p0, pb = params(classes[0], classes[1], False), params(classes[2], classes[3], True)
    ref, n = host_batch(pools, views, p0, pb, 2, 26.0, 33)
    b = ctx.batch(views, p0, pb, n_bounces=2, bounce_max_dist_m=26.0, seed=33)
    d = torch.zeros(b.records() * 16, dtype=torch.uint8, device="cuda")
    b.trace(d.data_ptr())
    torch.cuda.synchronize()
    got = d.cpu().numpy().view(A.HIT_DTYPE)
    bad = W.diff_records(got, ref)
    print(f"[batch] {classes}: {n} slots, {len(bad)} records differ", flush=True)
    assert len(bad) == 0, f"{len(bad)} of {len(ref)} records differ, first {bad[:5]}"
    lod = (ref["flags"] & 2) != 0  # records that stopped at an LOD floor
    assert lod.sum() > 100, lod.sum()
    b.close()


@pytest.mark.gpu
def test_gpu_batch_sparse_world_sampled_parity(ctx):
    """A larger batch on a sparse multi-chunk world: every record of a random sample of slots is
    re-derived on the host (the slot's primary ray and its bounce chain) and traced by the oracle."""
    import torch

    w = A.HostWorld((4, 2, 4))
    w.generate(A.gen_params(A.ARC_GEN_SPARSE, seed=2, height_min=32.0, height_max=300.0, wavelength=256.0))
    w.pack()
    pools = w.pools()
    ctx.upload(*pools)
    views = small_views(3, 16, base=(32.0, 25.0, 32.0))
    p0, pb = A.TraceParams.exact(0), A.TraceParams.exact(2, ignore_water=True)
    b = ctx.batch(views, p0, pb, n_bounces=2, bounce_max_dist_m=26.0, seed=5)
    n = b.slots
    d = torch.zeros(b.records() * 16, dtype=torch.uint8, device="cuda")
    b.trace(d.data_ptr())
    torch.cuda.synchronize()
    got = d.cpu().numpy().view(A.HIT_DTYPE)
    rays = np.concatenate([A.view_rays(cam, vw, vh, md) for cam, vw, vh, md in views])
    sample = np.sort(np.random.default_rng(1).choice(n, min(n, 20000), replace=False))
    world = O.World(*pools)
    cur = rays[sample]
    for depth in range(3):
        full, _ = O.trace_world(world, p0 if depth == 0 else pb, cur)
        hits = O.compact(full)
        bad = W.diff_records(got[depth * n + sample], hits)
        assert len(bad) == 0, f"depth {depth}: {len(bad)} differ"
        if depth < 2:  # each slot's bounce is hashed with its own slot index
            cur = np.concatenate([A.bounce_rays_dense(cur[k:k + 1], hits[k:k + 1], index_base=depth * n + int(s),
                                                      seed=5) for k, s in enumerate(sample)])
