"""GPU parity: the HIP traversal against the CPU oracle, bit for bit (hit voxel, face, t,
material, step count, flags), through the C ABI."""
from __future__ import annotations

import numpy as np
import pytest

import arcadia_amd as A
from oracle import oracle as O
from tests import worlds as W

pytestmark = pytest.mark.gpu


def _check(ctx, pools, params, rays, label):
    print(f"[parity] {label}: {len(rays)} rays", flush=True)
    ctx.upload(*pools)
    got = ctx.trace(params, rays)
    ref, _, _ = W.oracle_hits(pools, params, rays)
    bad = W.diff_records(got, ref)
    assert len(bad) == 0, (
        f"{label}: {len(bad)}/{len(rays)} records differ; first idx {bad[:5]} got {got[bad[:3]]} ref {ref[bad[:3]]}")
    return got


def test_heightmap_primary_exact(ctx):
    pools = W.heightmap_world(1)
    rays = W.poses_rays(W.c1_poses(1), 640, 360)
    hits = _check(ctx, pools, A.TraceParams.exact(), rays, "256^3 heightmap, 4 poses 640x360")
    assert (hits["flags"] & 1).mean() > 0.3


def test_heightmap_chunk_semantics(ctx):
    pools = W.heightmap_world(2)
    rays = W.poses_rays(W.c1_poses(2), 320, 180)
    for k in ("ox", "oy", "oz"):
        rays[k] *= 16.0
    rays["t_max"] = 250.0 * 16.0
    print(f"[parity] chunk semantics: {len(rays)} rays", flush=True)
    ctx.upload(*pools)
    got = ctx.trace(A.TraceParams.chunk(), rays)
    full, _ = O.trace_chunk(O.World(*pools), rays)
    ref = O.compact(full)
    bad = W.diff_records(got, ref)
    assert len(bad) == 0, f"{len(bad)} differ, first {bad[:5]}"


def test_caves_primary_and_bounce(ctx):
    pools = W.caves_world(1)
    cam = A.camera_look((13.7, 30.2, 11.9), 0.6, -0.55, 65.0, 480, 270)
    prays, _ = A.primary_rays(cam, 480, 270, 250.0)
    ph = _check(ctx, pools, A.TraceParams.exact(0), prays, "caves 2x2x2 primary")
    brays, _ = A.bounce_rays(prays, ph, seed=5, max_dist_m=26.0)
    assert len(brays) > 1000
    _check(ctx, pools, A.TraceParams.exact(2, ignore_water=True), brays, "caves 2x2x2 diffuse bounce")


def test_reference_default_quality(ctx):
    """Step caps and LOD cut-offs of RaymarchQualitySettings::high_quality, all three ray classes."""
    pools = W.caves_world(1)
    rays = W.random_rays(40000, (2, 2, 2), seed=11, inside=False, max_dist=120.0)
    for bias in (0, 1, 2):
        p = A.TraceParams.reference_default(bias, ignore_water=bias > 0)
        # LOD distances scaled into the small world so LOD stops actually happen
        p.lod_dist_m[0], p.lod_dist_m[1], p.lod_dist_m[2] = 6.0, 3.0, 5.0
        p.lod_step[0], p.lod_step[1], p.lod_step[2] = 20, 15, 12
        _check(ctx, pools, p, rays, f"reference-default lod_bias={bias}")


def test_random_rays_inside_and_outside(ctx):
    pools = W.caves_world(2, grid=(3, 2, 2), n_cuboids=60)
    for inside in (True, False):
        rays = W.random_rays(50000, (3, 2, 2), seed=3 + inside, inside=inside)
        _check(ctx, pools, A.TraceParams.exact(), rays, f"random rays inside={inside}")


def test_axis_aligned_and_tiny_components(ctx):
    pools = W.heightmap_world(1)
    rng = np.random.default_rng(7)
    n = 6000
    r = np.zeros(n, A.RAY_DTYPE)
    r["ox"], r["oy"], r["oz"] = rng.uniform(0.3, 15.7, n), rng.uniform(13.0, 15.9, n), rng.uniform(0.3, 15.7, n)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    axis = rng.integers(0, 3, n)
    kind = rng.integers(0, 4, n)
    for i in range(n):
        if kind[i] == 0:      # exact zero component
            d[i, axis[i]] = 0.0
        elif kind[i] == 1:    # below the 1e-7 threshold
            d[i, axis[i]] = np.float32(3e-8) * (1 if i % 2 else -1)
        elif kind[i] == 2:    # axis aligned
            d[i] = 0.0
            d[i, 1] = -1.0
    d /= np.linalg.norm(d, axis=1, keepdims=True).astype(np.float32)
    r["dx"], r["dy"], r["dz"] = d[:, 0], d[:, 1], d[:, 2]
    r["t_max"] = 250.0
    _check(ctx, pools, A.TraceParams.exact(), r, "zero/tiny/axis-aligned directions")


def test_empty_and_sparse_worlds(ctx):
    empty = W.sparse_world([], grid=(2, 1, 2))
    rays = W.random_rays(5000, (2, 1, 2), seed=1, inside=False)
    hits = _check(ctx, empty, A.TraceParams.exact(), rays, "empty world")
    assert not (hits["flags"] & 1).any()
    sp = W.sparse_world([(3, 0, 0, 2), (0, 0, 0, 2), (200, 17, 33, 5), (255, 255, 255, 9)])
    rays = W.random_rays(20000, (1, 1, 1), seed=2, inside=True)
    _check(ctx, sp, A.TraceParams.exact(), rays, "sparse voxels")


def test_full_records(ctx):
    torch = pytest.importorskip("torch")
    pools = W.heightmap_world(1)
    rays = W.poses_rays(W.c1_poses(3), 160, 90)
    params = A.TraceParams.exact(0)
    ctx.upload(*pools)
    d_rays = torch.from_numpy(rays.view(np.uint8)).cuda()
    d_out = torch.zeros(len(rays) * A.RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    ctx.trace_full_device(params, d_rays.data_ptr(), len(rays), d_out.data_ptr(),
                          torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = d_out.cpu().numpy().view(A.RESULT_DTYPE)
    ref, _ = O.trace_world(O.World(*pools), params, rays)
    bad = W.diff_records(got, ref)
    assert len(bad) == 0, f"{len(bad)} full records differ, first {bad[:5]}: {got[bad[:2]]} vs {ref[bad[:2]]}"


def test_device_entry_matches_host_entry(ctx):
    torch = pytest.importorskip("torch")
    pools = W.heightmap_world(1)
    rays = W.poses_rays(W.c1_poses(4), 320, 180)
    params = A.TraceParams.exact()
    ctx.upload(*pools)
    host = ctx.trace(params, rays)
    d_rays = torch.from_numpy(rays.view(np.uint8)).cuda()
    d_hits = torch.zeros(len(rays) * 16, dtype=torch.uint8, device="cuda")
    for _ in range(3):  # repeated launches reuse the self-resetting queue
        d_hits.zero_()
        ctx.trace_device(params, d_rays.data_ptr(), len(rays), d_hits.data_ptr(),
                         torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert d_hits.cpu().numpy().tobytes() == host.tobytes()


@pytest.mark.parametrize("cfg", [dict(blocks_per_cu=1, block_threads=64, refill_lanes=64),
                                 dict(blocks_per_cu=4, block_threads=128, refill_lanes=1),
                                 dict(blocks_per_cu=16, block_threads=256, refill_lanes=32)])
def test_launch_shapes_do_not_change_results(ctx, cfg):
    pools = W.caves_world(1)
    rays = W.random_rays(30000, (2, 2, 2), seed=21, inside=False)
    ctx.set_launch_config(**cfg)
    try:
        _check(ctx, pools, A.TraceParams.exact(), rays, f"launch {cfg}")
    finally:
        ctx.set_launch_config()


@pytest.mark.parametrize("n,batch", [(1, 64), (7, 1), (63, 3), (8 * 64 - 1, 64), (8 * 64 + 1, 64),
                                     (5000, 1 << 20), (5000, 7)])
def test_queue_hands_out_every_ray_once(ctx, n, batch):
    """Every ray of a launch gets its record whatever the window size against the eight claim
    counters: fewer rays than one window, windows of one ray, fewer windows than counters (counters
    empty from the start), sizes just off a multiple of windows x counters, one window for all."""
    pools = W.caves_world(1)
    rays = W.random_rays(n, (2, 2, 2), seed=100 + n, inside=False)
    ctx.set_launch_config(blocks_per_cu=2, block_threads=256, refill_lanes=16, batch_rays=batch)
    try:
        _check(ctx, pools, A.TraceParams.exact(), rays, f"queue n={n} batch={batch}")
    finally:
        ctx.set_launch_config()


def test_autotuner_keeps_results_bit_exact(ctx):
    torch = pytest.importorskip("torch")
    pools = W.caves_world(1)
    rays = W.random_rays(60000, (2, 2, 2), seed=31, inside=False)
    params = A.TraceParams.exact(2, ignore_water=True)
    ctx.upload(*pools)
    d_rays = torch.from_numpy(rays.view(np.uint8)).cuda()
    d_hits = torch.zeros(len(rays) * 16, dtype=torch.uint8, device="cuda")
    try:
        cfg, ms = ctx.autotune(params, d_rays.data_ptr(), len(rays), d_hits.data_ptr(),
                               torch.cuda.current_stream().cuda_stream, reps=1)
        assert ms > 0 and cfg["block_threads"] in (128, 256)
        assert ctx.launch_config(2) == cfg
        got = ctx.trace(params, rays)
    finally:
        ctx.set_launch_config()
    ref, _, _ = W.oracle_hits(pools, params, rays)
    assert len(W.diff_records(got, ref)) == 0


@pytest.mark.parametrize("ignore_water", [True, False])
def test_water_world(ctx, ignore_water):
    """The ignore_water rules (uniform water nodes stepped over whole, water voxels stepped through,
    traversal64.wgsl:398-417, 451-465) on a world that holds water, all three ray classes."""
    pools = W.water_world(True).pools()
    assert (np.asarray(pools[3]).view(np.uint8) == 18).any()
    rays = W.random_rays(60000, (2, 2, 2), seed=41)
    for bias in (0, 1, 2):
        _check(ctx, pools, A.TraceParams.exact(bias, ignore_water=ignore_water), rays,
               f"water world lod_bias {bias} ignore_water {ignore_water}")


def test_water_patched_into_dry_world(ctx):
    """A world without water traces with ignore_water set (the pools cannot hold water, so the
    kernels skip its tests); water patched in afterwards must be honoured at once."""
    w = W.water_world(False)
    ctx.upload(*w.pools(copy=False))
    w.clear_patches()
    params = A.TraceParams.exact(2, ignore_water=True)
    rays = W.random_rays(60000, (2, 2, 2), seed=42)
    dry = ctx.trace(params, rays)
    ref, _, _ = W.oracle_hits(w.pools(), params, rays)
    assert len(W.diff_records(dry, ref)) == 0
    z, y, x = np.mgrid[0:64, 0:24, 0:64]
    w.set_voxels(np.stack([x.ravel() + 128, y.ravel() + 200, z.ravel() + 128, np.full(x.size, 18)], 1).astype(np.int32))
    ctx.patch_world(w)
    wet = ctx.trace(params, rays)
    ref2, _, _ = W.oracle_hits(w.pools(), params, rays)
    assert len(W.diff_records(wet, ref2)) == 0
    assert len(W.diff_records(wet, dry)) > 0  # the lake changed some records
