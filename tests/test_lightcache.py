"""Path shading and the radiance-cache workload.

GPU: the device shading of a path launch against the binary32 restatement (host-generated rays,
oracle-traced records), bit for bit; the FP64 judge equals the float64 sum of the per-launch
radiances; and on a small scene both caches, trained online, lower the error of the cached
estimate below the estimate without a cache (what the learned cache gains).
"""
from __future__ import annotations

import numpy as np
import pytest

import arcadia_amd as A
from arcadia_amd import lightcache as LC
from oracle import oracle as O
from oracle import pt_ref as P

pytestmark = pytest.mark.gpu


def small_scene():
    w = A.HostWorld((2, 2, 2))
    w.generate(A.gen_params(A.ARC_GEN_CAVES, seed=3, n_cuboids=60, height_min=96.0, height_max=352.0,
                            wavelength=192.0))
    clusters = LC.add_emissive_clusters(w, (2, 2, 2), seed=3, count=4, r_min=5, r_max=9)
    w.pack()
    return w.pools(), clusters


def test_gpu_shading_matches_restatement(ctx):
    import torch

    pools, _ = small_scene()
    ctx.upload(*pools)
    cam = A.camera_look((16.0, 26.0, 16.0), 0.8, -0.6, 70.0, 97, 55)
    p0, pb = A.TraceParams.exact(0), A.TraceParams.exact(2, ignore_water=True)
    b = ctx.batch([(cam, 97, 55, 250.0)], p0, pb, n_bounces=4, bounce_max_dist_m=250.0, seed=5)
    n = b.slots
    hits = torch.zeros(b.records() * 16, dtype=torch.uint8, device="cuda")
    samples = torch.zeros(n * 96, dtype=torch.uint8, device="cuda")
    rad = torch.zeros(n * 3, dtype=torch.float32, device="cuda")
    sp = A.ShadeParams.default()
    b.trace(hits.data_ptr())
    b.shade(hits.data_ptr(), sp, d_samples=samples.data_ptr(), d_radiance=rad.data_ptr())
    torch.cuda.synchronize()
    got_s = samples.cpu().numpy().view(A.PATH_SAMPLE_DTYPE)
    got_r = rad.cpu().numpy().reshape(n, 3)
    # host rays, oracle hits, depth by depth
    world = O.World(*pools)
    rays = [A.view_rays(cam, 97, 55, 250.0)]
    hs = []
    for d in range(5):
        full, _ = O.trace_world(world, p0 if d == 0 else pb, rays[d])
        hs.append(O.compact(full))
        if d < 4:
            rays.append(A.bounce_rays_dense(rays[d], hs[d], index_base=d * n, seed=5, max_dist_m=250.0))
    dev = hits.cpu().numpy().view(A.HIT_DTYPE).reshape(5, n)
    for d in range(5):
        assert np.array_equal(dev[d].view(np.uint32), hs[d].view(np.uint32)), d
    pal = np.stack([A.palette_material(m) for m in range(256)])
    L, out, flags = P.shade(rays, hs, pal, list(sp.sun_dir), sp.emission_scale, sp.sky_scale)
    assert np.array_equal(got_r.view(np.uint32), L.view(np.uint32))
    assert np.array_equal(got_s["flags"], flags)
    for k in ("pos", "dir", "le0", "a0", "base", "target"):
        assert np.array_equal(got_s[k].view(np.uint32), out[k].view(np.uint32)), k
    assert (flags & 4).sum() > 1000 and (got_s["target"] > 0).any()
    # FP64 judge = float64 sum of the per-launch radiances
    acc = torch.zeros(n * 3, dtype=torch.float64, device="cuda")
    want = np.zeros(n * 3)
    for s in range(6):
        b.set_seed(100 + s)
        b.trace(hits.data_ptr())
        b.shade(hits.data_ptr(), sp, d_radiance=rad.data_ptr(), d_accum=acc.data_ptr())
        torch.cuda.synchronize()
        want += rad.cpu().numpy().astype(np.float64)
    assert np.array_equal(acc.cpu().numpy(), want)
    b.close()


def test_gpu_caches_lower_the_error_of_the_estimate(ctx):
    pools, clusters = small_scene()
    ctx.upload(*pools)
    c = np.array(clusters[0][0], np.float64) / 16.0
    eye = np.clip(c + np.array([6.0, 5.0, 6.0]), 1.0, 31.0)
    d = c - eye
    cam = A.camera_look(tuple(eye), float(np.arctan2(d[0], d[2])), float(np.arctan2(d[1], np.hypot(d[0], d[2]))),
                        70.0, 256, 144)
    run = LC.CacheRun(ctx, cam, 256, 144, train_size=(128, 72), sharc_log2=16,
                      nrc=A.NrcParams.default(bbox_size=32.0, log2_table=16))
    run.reference(256)
    for f in range(80):
        run.train_paths(f)
        run.train_sharc()
        run.train_nrc(1)
    err = run.errors(seed=99)
    print(f"[cache] {err}", flush=True)
    assert err["sharc"] < err["no_cache"]
    assert err["nrc"] < err["no_cache"]
    ms = run.cache_ms(reps=3)
    print(f"[cache] {ms}", flush=True)
    run.close()
