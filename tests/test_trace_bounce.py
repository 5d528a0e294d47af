"""GPU parity of arc_trace_bounce_device: a ray set and one diffuse bounce off each hit in one
persistent launch. Every record is compared with the oracle: the primary records with the oracle's
trace of the rays, and record n + i with the oracle's trace of arc_gen_bounce_rays_dense's bounce
of ray i off the oracle's own hit (the empty-ray record where ray i missed)."""
from __future__ import annotations

import numpy as np
import pytest

import arcadia_amd as A
from tests import worlds as W

pytestmark = pytest.mark.gpu

SEED = 5
BOUNCE_M = 26.0


def expected(pools, p0, pb, rays, seed=SEED, bounce_m=BOUNCE_M):
    ref_p, _, _ = W.oracle_hits(pools, p0, rays)
    brays = A.bounce_rays_dense(rays, ref_p, index_base=0, seed=seed, max_dist_m=bounce_m)
    ref_b, _, _ = W.oracle_hits(pools, pb, brays)
    return np.concatenate([ref_p, ref_b])


def run(ctx, p0, pb, rays, seed=SEED, bounce_m=BOUNCE_M):
    import torch

    n = len(rays)
    d_r = torch.from_numpy(np.ascontiguousarray(rays).view(np.uint8)).cuda()
    d_h = torch.full((2 * n * 16,), 0xAB, dtype=torch.uint8, device="cuda")  # every record must be written
    ctx.trace_bounce_device(p0, pb, d_r.data_ptr(), n, seed, bounce_m, d_h.data_ptr())
    torch.cuda.synchronize()
    return d_h.cpu().numpy().view(A.HIT_DTYPE)


def check(label, got, ref):
    bad = W.diff_records(got, ref)
    print(f"[bounce] {label}: {len(got)} records, {len(bad)} differ", flush=True)
    assert len(bad) == 0, f"{label}: first differing records {bad[:5]} got {got[bad[:2]]} ref {ref[bad[:2]]}"


def caves_rays(w=480, h=270):
    cam = A.camera_look((13.7, 30.2, 11.9), 0.6, -0.55, 65.0, w, h)
    rays, _ = A.primary_rays(cam, w, h, 250.0)
    return rays


def test_bounce_launch_caves_exact(ctx):
    pools = W.caves_world(1)
    ctx.upload(*pools)
    rays = caves_rays()
    p0, pb = A.TraceParams.exact(0), A.TraceParams.exact(2, ignore_water=True)
    ref = expected(pools, p0, pb, rays)
    n = len(rays)
    hits = (ref["flags"][:n] & 1).sum()
    assert hits > 0.5 * n
    for k in range(2):  # the message words return to zero: a second launch gives the same records
        check(f"caves 2x2x2 480x270, launch {k}", run(ctx, p0, pb, rays), ref)
    m = 1000  # a shorter set after a longer one: records of the first m rays
    got = run(ctx, p0, pb, rays[:m])
    check("first 1000 rays", got, np.concatenate([ref[:m], ref[n:n + m]]))


def test_bounce_launch_reference_quality(ctx):
    """LOD cut-offs and step caps of the reference's high-quality settings: the two classes then
    trace with different parameters per lane."""
    pools = W.caves_world(1)
    ctx.upload(*pools)
    rays = caves_rays(320, 180)
    p0, pb = A.TraceParams.reference_default(0), A.TraceParams.reference_default(2, ignore_water=True)
    check("reference-default quality", run(ctx, p0, pb, rays), expected(pools, p0, pb, rays))


@pytest.mark.parametrize("shape", [(1, 64, 1, 1), (1, 256, 64, 7), (0, 128, 8, 32), (6, 256, 48, 128)])
def test_bounce_launch_shapes(ctx, shape):
    """Launch shapes from one 64-lane workgroup per CU with windows of 1 ray (bounce lanes then
    wait for primaries still in flight in their own wave) to large windows."""
    pools = W.caves_world(1)
    ctx.upload(*pools)
    rays = caves_rays(320, 180)
    p0, pb = A.TraceParams.exact(0), A.TraceParams.exact(2, ignore_water=True)
    saved = ctx.launch_config(A.ARC_LAUNCH_CLASS_BOUNCE)
    ctx.set_launch_config_class(A.ARC_LAUNCH_CLASS_BOUNCE, *shape)
    try:
        check(f"shape {shape}", run(ctx, p0, pb, rays), expected(pools, p0, pb, rays))
    finally:
        ctx.set_launch_config_class(A.ARC_LAUNCH_CLASS_BOUNCE, *[saved[k] for k in
                                    ("blocks_per_cu", "block_threads", "refill_lanes", "batch_rays")])


def test_bounce_launch_small_and_missing_sets(ctx):
    pools = W.caves_world(1)
    ctx.upload(*pools)
    p0, pb = A.TraceParams.exact(0), A.TraceParams.exact(2, ignore_water=True)
    rays = caves_rays(160, 90)
    for n in (1, 37, 64, 65):
        check(f"{n} rays", run(ctx, p0, pb, rays[:n]), expected(pools, p0, pb, rays[:n]))
    # every ray misses (pointing up and out of the world, or empty): every bounce slot holds the
    # empty-ray record
    up = rays[:500].copy()
    up["dx"], up["dy"], up["dz"] = 0.0, 1.0, 0.0
    up["oy"] = 40.0
    up["t_max"][::3] = -1.0
    got = run(ctx, p0, pb, up)
    check("rays that all miss", got, expected(pools, p0, pb, up))
    assert (got["flags"] & 1).sum() == 0 and np.all(got["t_voxels"][500:] == -16.0)


def test_bounce_launch_random_rays_and_seeds(ctx):
    """Rays from inside and around the world in random directions, two seeds of the bounce hash."""
    pools = W.caves_world(2)
    ctx.upload(*pools)
    p0, pb = A.TraceParams.exact(0), A.TraceParams.exact(2, ignore_water=True)
    rays = np.concatenate([W.random_rays(20000, (2, 2, 2), seed=3), W.random_rays(5000, (2, 2, 2), 4, inside=False)])
    for seed in (1, 2**40 + 3):
        check(f"random rays, seed {seed}", run(ctx, p0, pb, rays, seed=seed),
              expected(pools, p0, pb, rays, seed=seed))


def test_bounce_launch_water_world(ctx):
    """A world that holds water (the kernel's water-test instance): primary rays that stop at water
    and bounce rays that step through it (ignore_water), from inside and around the world."""
    pools = W.water_world(True).pools()
    assert (np.asarray(pools[3]).view(np.uint8) == 18).any()
    ctx.upload(*pools)
    rays = W.random_rays(30000, (2, 2, 2), seed=43)
    for p0, pb in ((A.TraceParams.exact(0), A.TraceParams.exact(2, ignore_water=True)),
                   (A.TraceParams.exact(0, ignore_water=True), A.TraceParams.exact(2, ignore_water=False))):
        check(f"water world, primary ignore_water {p0.ignore_water}, bounce {pb.ignore_water}",
              run(ctx, p0, pb, rays), expected(pools, p0, pb, rays))


def test_bounce_autotune(ctx):
    """arc_autotune_bounce picks a shape for this ray set and installs it: the next launch uses it
    and its records are exact."""
    import torch

    pools = W.caves_world(1)
    ctx.upload(*pools)
    rays = caves_rays(320, 180)
    p0, pb = A.TraceParams.exact(0), A.TraceParams.exact(2, ignore_water=True)
    n = len(rays)
    d_r = torch.from_numpy(np.ascontiguousarray(rays).view(np.uint8)).cuda()
    d_h = torch.zeros(2 * n * 16, dtype=torch.uint8, device="cuda")
    saved = ctx.launch_config(A.ARC_LAUNCH_CLASS_BOUNCE)
    try:
        cfg, ms = ctx.autotune_bounce(p0, pb, d_r.data_ptr(), n, SEED, BOUNCE_M, d_h.data_ptr())
        print(f"[bounce] autotuned {cfg} {ms:.4f} ms", flush=True)
        assert ms > 0.0 and cfg["block_threads"] in (128, 256) and cfg["refill_lanes"] > 0 and cfg["batch_rays"] > 0
        assert ctx.launch_config(A.ARC_LAUNCH_CLASS_BOUNCE) == cfg
        check("after autotune", run(ctx, p0, pb, rays), expected(pools, p0, pb, rays))
    finally:
        ctx.set_launch_config_class(A.ARC_LAUNCH_CLASS_BOUNCE, *[saved[k] for k in
                                    ("blocks_per_cu", "block_threads", "refill_lanes", "batch_rays")])


def test_bounce_launch_repeats_are_identical(ctx):
    """Launches repeated across launch shapes and seeds: every launch of a seed gives the records of
    its first launch (the message words and queue counters return to zero between launches, and no
    result depends on the order in which waves publish or take messages)."""
    pools = W.caves_world(1)
    ctx.upload(*pools)
    rays = caves_rays(480, 270)
    p0, pb = A.TraceParams.exact(0), A.TraceParams.exact(2, ignore_water=True)
    saved = ctx.launch_config(A.ARC_LAUNCH_CLASS_BOUNCE)
    first = {}
    try:
        for k, shape in enumerate([(0, 256, 24, 64), (1, 64, 1, 1), (4, 256, 8, 32), (0, 128, 48, 128), (6, 256, 64, 256)] * 3):
            ctx.set_launch_config_class(A.ARC_LAUNCH_CLASS_BOUNCE, *shape)
            for seed in (SEED, 2**40 + 3):
                got = run(ctx, p0, pb, rays, seed=seed)
                if seed not in first:
                    first[seed] = got
                    check(f"seed {seed}, first launch", got, expected(pools, p0, pb, rays, seed=seed))
                else:
                    assert np.array_equal(got.view(np.uint32), first[seed].view(np.uint32)), (k, shape, seed)
    finally:
        ctx.set_launch_config_class(A.ARC_LAUNCH_CLASS_BOUNCE, *[saved[k] for k in
                                    ("blocks_per_cu", "block_threads", "refill_lanes", "batch_rays")])
