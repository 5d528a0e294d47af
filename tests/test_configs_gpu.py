"""The five benchmark configurations at their stated sizes on the GPU, each checked against the oracle.

  * full frame: 1024^3 cave world (4x4x4 chunks, 3D fBm at 0.45 + 2000 cuboids), 1920x1080
    primary rays and one cosine bounce per hit, traced as two launches and as one
    (arc_trace_bounce_device); strided samples compared with the oracle record for record (every
    8th primary ray, every 8th bounce ray);
  * batched views: 64 views of uneven sizes (width log-uniform in 320..2560) over the 4096^3 sparse
    world, primary + 2 bounces in one persistent launch, replayed from its hipGraph; 20,000 random
    slots re-derived on the host and traced by the oracle at every depth;
  * learned cache: 1920x1080 on the full-frame world with 8 emissive clusters, SHARC with 2^22
    entries, 2^16 four-bounce training paths per frame, judged against the FP64 1024-spp path
    tracer: the neural cache must beat no cache and match or beat SHARC's relative MSE;
  * destruction: 30 frames on the full-frame world, each carving >= 1% of the solid voxels in
    seeded spheres, patching the device pools and tracing the 1080p primary + diffuse frame; 4096
    random pixels per frame compared with the oracle on the edited host pools, both depths.
The single-chunk configuration (256^3 heightmap, 4 poses at 640x360) runs at its size in
test_gpu_parity.py.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import arcadia_amd as A
from oracle import oracle as O
from tests import worlds as W

pytestmark = pytest.mark.gpu

THREADS = min(16, len(os.sched_getaffinity(0)))


def full_frame_world(seed=1):
    w = A.HostWorld((4, 4, 4))
    w.generate(A.gen_params(A.ARC_GEN_CAVES, seed=seed), nthreads=THREADS)
    return w


FULL_CAM = dict(eye=(31.3, 58.0, 27.1), yaw=0.7, pitch=-0.5, fov=65.0)


def full_cam():
    return A.camera_look(FULL_CAM["eye"], FULL_CAM["yaw"], FULL_CAM["pitch"], FULL_CAM["fov"], 1920, 1080)


def test_full_frame_primary_and_bounce(ctx):
    import torch

    w = full_frame_world()
    w.pack()
    pools = w.pools()
    assert len(pools[2]) > 5_000_000  # the stated world, not a stand-in
    ctx.upload(*pools)
    prays, _ = A.primary_rays(full_cam(), 1920, 1080, 250.0)
    print(f"[config] full frame: {len(pools[2])} nodes, {len(prays)} primary rays", flush=True)
    p0, pb = A.TraceParams.exact(0), A.TraceParams.exact(2, ignore_water=True)
    d_r = torch.from_numpy(prays.view(np.uint8)).cuda()
    d_h = torch.zeros(len(prays) * 16, dtype=torch.uint8, device="cuda")
    ctx.trace_device(p0, d_r.data_ptr(), len(prays), d_h.data_ptr())
    torch.cuda.synchronize()
    ph = d_h.cpu().numpy().view(A.HIT_DTYPE)
    brays, bsrc = A.bounce_rays(prays, ph, seed=1, max_dist_m=26.0)
    assert len(brays) > 1_000_000
    d_b = torch.from_numpy(brays.view(np.uint8)).cuda()
    d_bh = torch.zeros(len(brays) * 16, dtype=torch.uint8, device="cuda")
    ctx.trace_device(pb, d_b.data_ptr(), len(brays), d_bh.data_ptr())
    torch.cuda.synchronize()
    bh = d_bh.cpu().numpy().view(A.HIT_DTYPE)
    # the same frame as one launch: primaries, then each hit's bounce generated on the device
    n = len(prays)
    d_f = torch.zeros(2 * n * 16, dtype=torch.uint8, device="cuda")
    ctx.trace_bounce_device(p0, pb, d_r.data_ptr(), n, 1, 26.0, d_f.data_ptr())
    torch.cuda.synchronize()
    fh = d_f.cpu().numpy().view(A.HIT_DTYPE)
    miss = np.ones(n, bool)
    miss[bsrc] = False
    assert np.all(fh[n:][miss]["t_voxels"] == -16.0) and np.all(fh[n:][miss]["steps"] == 0)
    ow = O.World(*pools)
    for label, params, rays, got, fused in (("primary", p0, prays, ph, fh[:n]), ("bounce", pb, brays, bh, fh[n:][bsrc])):
        idx = np.arange(0, len(rays), 8)
        full, _ = O.trace_world(ow, params, rays[idx], THREADS)
        ref = O.compact(full)
        bad = W.diff_records(got[idx], ref)
        bad_f = W.diff_records(fused[idx], ref)
        print(f"[config] full frame {label}: {len(idx)} sampled records, {len(bad)} differ (two launches), "
              f"{len(bad_f)} differ (one launch)", flush=True)
        assert len(bad) == 0, f"{label}: first differing sample indices {idx[bad[:5]]}"
        assert len(bad_f) == 0, f"{label}, one launch: first differing sample indices {idx[bad_f[:5]]}"


def batched_views(seed=1):
    """The 64 views of the batched configuration (width log-uniform in 320..2560, 16:9)."""
    rng = np.random.default_rng(seed)
    views = []
    for _ in range(64):
        wd = int(round(float(np.exp(rng.uniform(np.log(320.0), np.log(2560.0))))))
        ht = max(1, wd * 9 // 16)
        eye = (float(rng.uniform(16.0, 240.0)), float(rng.uniform(16.0, 26.0)), float(rng.uniform(16.0, 240.0)))
        cam = A.camera_look(eye, float(rng.uniform(0.0, 6.2831853)), float(rng.uniform(-0.6, -0.15)), 70.0, wd, ht)
        views.append((cam, wd, ht, 250.0))
    return views


def test_batched_views_full_size(ctx):
    import torch

    w = A.HostWorld((16, 16, 16))
    w.generate(A.gen_params(A.ARC_GEN_SPARSE, seed=1), nthreads=THREADS)
    w.pack()
    pools = w.pools()
    del w
    ctx.upload(*pools)
    views = batched_views(1)
    p0, pb = A.TraceParams.exact(0), A.TraceParams.exact(2, ignore_water=True)
    b = ctx.batch(views, p0, pb, n_bounces=2, bounce_max_dist_m=26.0, seed=1)
    n = b.slots
    print(f"[config] batched: {len(pools[2])} nodes, {len(views)} views, {n} slots", flush=True)
    hits = torch.full((b.records() * 16,), 7, dtype=torch.uint8, device="cuda")
    b.capture(hits.data_ptr())
    for _ in range(2):
        b.replay()
    torch.cuda.synchronize()
    rng = np.random.default_rng(3)
    sample = np.sort(rng.choice(n, 20000, replace=False))
    rays = np.zeros(len(sample), A.RAY_DTYPE)
    first = 0
    for cam, wd, ht, md in views:
        nv = ((wd + 7) // 8) * ((ht + 7) // 8) * 64
        sel = (sample >= first) & (sample < first + nv)
        if sel.any():
            rays[sel] = A.view_rays(cam, wd, ht, md)[sample[sel] - first]
        first += nv
    assert first == n
    idx = torch.from_numpy(np.concatenate([d * n + sample for d in range(3)]).astype(np.int64)).cuda()
    got = hits.view(torch.int32).reshape(-1, 4)[idx].cpu().numpy().view(A.HIT_DTYPE).reshape(3, -1)
    ow = O.World(*pools)
    cur, traced = rays, 0
    for depth in range(3):
        full, _ = O.trace_world(ow, p0 if depth == 0 else pb, cur, THREADS)
        ref = O.compact(full)
        bad = W.diff_records(got[depth], ref)
        traced += int((ref["t_voxels"] != -16.0).sum())
        print(f"[config] batched depth {depth}: {len(bad)} of {len(sample)} sampled slots differ", flush=True)
        assert len(bad) == 0, f"depth {depth}: first slots {sample[bad[:5]]}"
        if depth < 2:
            cur = np.concatenate([A.bounce_rays_dense(cur[k:k + 1], ref[k:k + 1], index_base=depth * n + int(s),
                                                      seed=1, max_dist_m=26.0) for k, s in enumerate(sample)])
    assert traced > 30000  # bounces were traced, not only primaries
    b.close()


def test_learned_cache_full_size(ctx):
    from arcadia_amd import lightcache as LC

    w = full_frame_world()
    LC.add_emissive_clusters(w, (4, 4, 4), seed=1, count=8)
    w.pack()
    ctx.upload(*w.pools())
    del w
    run = LC.CacheRun(ctx, full_cam(), 1920, 1080, train_size=(344, 192), sharc_log2=20,
                      nrc=A.NrcParams.default(bbox_size=64.0, log2_table=14), seed=1)
    assert run.n_train >= 1 << 16 and run.n_eval >= 1920 * 1080
    print(f"[config] learned cache: {run.n_train} training paths per frame, judge 1024 spp", flush=True)
    run.reference(1024)
    for f in range(60):
        run.train_paths(f)
        run.train_sharc()
        run.train_nrc(1)
    err = run.errors(seed=(1 << 40) + 1)
    print(f"[config] learned cache rel-MSE: {err}", flush=True)
    assert err["nrc"] < err["no_cache"]
    assert err["nrc"] <= err["sharc"]
    run.close()


def test_destruction_sequence_full_size(ctx):
    import torch

    w = full_frame_world()
    w.pack()
    ctx.upload(*w.pools(copy=False))
    w.clear_patches()
    cam = full_cam()
    p0, pb = A.TraceParams.exact(0), A.TraceParams.exact(2, ignore_water=True)
    frame = ctx.batch([(cam, 1920, 1080, 250.0)], p0, pb, n_bounces=1, bounce_max_dist_m=26.0, seed=1)
    hits = torch.empty(frame.records() * 16, dtype=torch.uint8, device="cuda")
    sharc = A.Sharc(20)
    rays0 = A.view_rays(cam, 1920, 1080, 250.0)
    n = frame.slots
    rng = np.random.default_rng(6)
    for f in range(30):
        solid = sum(w.chunk_info(i)["solid"] for i in range(64))
        done, spheres = 0, []
        while done < solid // 100:
            c = rng.uniform(0.0, 64.0, 3)
            r = rng.uniform(4.0, 32.0) / 16.0
            done += w.carve_crater(tuple(c), r)
            spheres.append((c[0], c[1], c[2], r))
        ctx.patch_world(w)
        frame.set_seed(7919 + f)
        frame.trace(hits.data_ptr())
        sharc.invalidate(np.array(spheres, np.float32))
        torch.cuda.synchronize()
        got = hits.cpu().numpy().view(A.HIT_DTYPE).reshape(2, n)
        sample = np.sort(rng.choice(n, 4096, replace=False))
        ow = O.World(*w.pools(copy=False))
        h0 = O.compact(O.trace_world(ow, p0, rays0[sample], THREADS)[0])
        br = np.concatenate([A.bounce_rays_dense(rays0[s:s + 1], h0[k:k + 1], index_base=int(s), seed=7919 + f,
                                                 max_dist_m=26.0) for k, s in enumerate(sample)])
        h1 = O.compact(O.trace_world(ow, pb, br, THREADS)[0])
        bad0, bad1 = W.diff_records(got[0][sample], h0), W.diff_records(got[1][sample], h1)
        print(f"[config] destruction frame {f}: carved {done} voxels in {len(spheres)} spheres, "
              f"{len(bad0)} + {len(bad1)} of 4096 + 4096 records differ", flush=True)
        assert len(bad0) == 0 and len(bad1) == 0, f"frame {f}"
    sharc.close()
    frame.close()
