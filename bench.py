"""Benchmark of the 64-tree DDA traversal on one MI355X.

Workload (one "step"): the full-frame traversal of a seeded 1024^3 world (4x4x4 chunks of
256^3, 3D-fBm terrain with caves thresholded at density 0.45 plus 2000 random axis-aligned
cuboids of edge 2..64 voxels, random materials 1..15): 1920x1080 primary rays from a pinhole
camera (lod_bias 0, 250 m), then one cosine-weighted diffuse bounce ray per hit pixel
(lod_bias 2, ignore_water, 26 m). Exact traversal (no LOD cut-off, no step truncation).
The primary rays are resident in HBM before timing; a step is one launch of
arc_trace_bounce_device, which traces them and generates and traces each hit's bounce on the
device. The same rays traced as two launches (primary, then the host-generated bounce rays) are
timed beside it under config.split.

Prints one JSON line (metric = Grays/s over primary+diffuse rays). A second workload is reported
beside it under "batched": 64 camera views of uneven sizes (width log-uniform in 320..2560,
16:9) over a 4096^3 sparse world (16^3 chunks, fBm heightmap terrain in the bottom layer), each
pixel's primary ray plus 2 diffuse bounces, all in one persistent launch replayed as a hipGraph.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Grays/s 64-tree DDA (primary+diffuse); cache rel-MSE vs FP64 oracle at fixed ms"
HBM_PEAK_GBS = 8000.0


def measured_traffic(kind: str):
    """HBM bytes per launch of the traversal kernel of one ray class ("primary" / "bounce") on
    this workload, from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary in profiles/ (scaled by the
    copy-kernel factors recorded there); None when the summary is absent."""
    p = ROOT / "profiles" / "k_trace_fullframe_traffic.txt"
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith(f"{kind}_traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def measured_counters() -> dict:
    """L2 hit rate, VALU busy and lanes active of the traversal kernels on this workload, from the
    rocprofv3 counter summary in profiles/ (key: value lines); empty when the summary is absent."""
    p = ROOT / "profiles" / "k_trace_fullframe_counters.txt"
    out = {}
    if not p.exists():
        return out
    for line in p.read_text().splitlines():
        if line.startswith("#") or ":" not in line:
            continue
        k, v = (x.strip() for x in line.split(":", 1))
        try:
            out[k] = float(v)
        except ValueError:
            out[k] = v
    out["source"] = "profiles/k_trace_fullframe_counters.txt"
    return out


def build_workload(seed: int, nthreads: int):
    import arcadia_amd as A

    t0 = time.time()
    world = A.HostWorld((4, 4, 4))
    world.generate(A.gen_params(A.ARC_GEN_CAVES, seed=seed), nthreads=nthreads)
    world.pack()
    pools = world.pools()
    cam = A.camera_look((31.3, 58.0, 27.1), 0.7, -0.5, 65.0, 1920, 1080)
    prays, _ = A.primary_rays(cam, 1920, 1080, 250.0)
    return world, pools, prays, time.time() - t0


def host_cpu_info():
    """Host cores the CPU baseline runs on: the affinity mask, capped by the box's declared CPU share
    (OMP_NUM_THREADS, set on GPU boxes to the cores one GPU's jobs may use), plus nproc and the model."""
    aff = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(aff, int(share)) if share.isdigit() and int(share) > 0 else aff
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, {"nproc": os.cpu_count(), "affinity_cores": aff, "omp_num_threads": share or None,
                     "model": model}


def time_cpu(fn, rays, budget_s: float):
    """CPU baseline protocol: one full warm-up pass over the sample, then the median of 5 timed
    passes. The sample is the whole ray set when 6 passes fit `budget_s`, else an even stride of it.
    Returns (rays/s, stride, sample size)."""
    probe = rays[:: max(1, len(rays) // 50000)]
    t = time.perf_counter()
    fn(probe)
    rate = len(probe) / max(time.perf_counter() - t, 1e-6)
    n = int(min(len(rays), max(len(probe), rate * budget_s / 6.0)))
    stride = max(1, -(-len(rays) // n))
    sample = np.ascontiguousarray(rays[::stride])
    fn(sample)  # warm-up pass
    times = []
    for _ in range(5):
        t = time.perf_counter()
        fn(sample)
        times.append(time.perf_counter() - t)
    return len(sample) / float(np.median(times)), stride, len(sample)


def cpu_baseline(pools, prays, brays, p_primary, p_bounce, threads: int, cpu_info: dict, budget_s: float):
    """Reference CPU traversal (WorldMap::raycast_meters restated in C, the oracle) on the same rays;
    also the per-ray load counters that give the algorithmic bytes of the traversal
    (trace_world_64tree_from restated in C)."""
    from oracle import oracle as O

    ow = O.World(*pools)
    allr = np.concatenate([prays, brays])
    cpu_rate, stride, n_sample = time_cpu(lambda r: O.trace_meters(ow, r, threads), allr, budget_s)
    # algorithmic bytes per ray from the WGSL-semantics counters (strided samples per class); the
    # same oracle records are kept to check the timed launches' records bit for bit
    def bytes_per_ray(rays, params):
        idx = np.arange(0, len(rays), max(1, len(rays) // 200000))
        full, cnt = O.trace_world(ow, params, rays[idx], threads)
        b = 16.0 * cnt["node_loads"] + 4.0 * cnt["brick_loads"] + 4.0 * cnt["root_loads"] + 32.0 + 16.0
        return float(b.mean()), float(cnt["steps"].mean()), (idx, O.compact(full))
    bp, sp, ref_p = bytes_per_ray(prays, p_primary)
    bb, sb, ref_b = bytes_per_ray(brays, p_bounce)
    return {
        "value": cpu_rate / 1e9,
        "unit": "Grays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n_sample} rays (every {stride}th of the {len(allr)} primary+bounce rays), "
                  f"WorldMap::raycast_meters restated in C (-O3 -ffp-contract=off), one warm-up pass, median of 5",
        "host": cpu_info,
    }, {"primary": bp, "bounce": bb, "steps_primary": sp, "steps_bounce": sb, "ref_primary": ref_p, "ref_bounce": ref_b}


def c1_poses(seed: int):
    """4 seeded poses above the 256^3 terrain: eye y in [13, 15.5] m, pitch -20..-60 deg."""
    rng = np.random.default_rng(seed)
    poses = []
    for _ in range(4):
        eye = (float(rng.uniform(2.0, 14.0)), float(rng.uniform(13.0, 15.5)), float(rng.uniform(2.0, 14.0)))
        poses.append((eye, float(rng.uniform(0, 2 * np.pi)), float(np.deg2rad(rng.uniform(-60.0, -20.0)))))
    return poses


def c1_leg(args, stream, threads, cpu_info):
    """Plain traversal, small: one 256^3 chunk (4-level tree) from a seeded fBm heightmap, 4 poses of
    640x360 primary rays, raycast_chunk_tree semantics (512-step cap, voxel units) on the GPU, and the
    same rays through the C restatement of raycast_chunk_tree on the host cores."""
    import torch

    import arcadia_amd as A
    from oracle import oracle as O

    w = A.HostWorld((1, 1, 1))
    w.generate(A.gen_params(A.ARC_GEN_HEIGHTMAP, seed=args.seed), nthreads=threads)
    w.pack()
    pools = w.pools()
    rays = []
    for eye, yaw, pitch in c1_poses(args.seed):
        cam = A.camera_look(eye, yaw, pitch, 90.0, 640, 360)
        r, _ = A.primary_rays(cam, 640, 360, 250.0)
        rays.append(r)
    rays = np.concatenate(rays)
    for k in ("ox", "oy", "oz"):
        rays[k] *= 16.0
    rays["t_max"] = 250.0 * 16.0
    n = len(rays)
    params = A.TraceParams.chunk()
    ctx = A.Context(0)
    ctx.upload(*pools)
    sptr = stream.cuda_stream
    d_rays = torch.from_numpy(rays.view(np.uint8)).cuda()
    d_hits = torch.zeros(n * 16, dtype=torch.uint8, device="cuda")
    if not args.no_autotune:
        ctx.autotune(params, d_rays.data_ptr(), n, d_hits.data_ptr(), sptr)
    for _ in range(3):
        ctx.trace_device(params, d_rays.data_ptr(), n, d_hits.data_ptr(), sptr)
    torch.cuda.synchronize()
    k = max(5, args.steps)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(k):
        ctx.trace_device(params, d_rays.data_ptr(), n, d_hits.data_ptr(), sptr)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / k
    got = d_hits.cpu().numpy().view(A.HIT_DTYPE)
    ow = O.World(*pools)
    full, cnt = O.trace_chunk(ow, rays, threads)
    exact = bool(np.array_equal(got.view(np.uint32), O.compact(full).view(np.uint32)))
    out = {
        "metric": "Grays/s raycast_chunk_tree (256^3 heightmap chunk, 4 poses 640x360, 512-step cap)",
        "value": round(n / (ms * 1e-3) / 1e9, 4),
        "unit": "Grays/s",
        "ms": round(ms, 4),
        "rays": n,
        "hit_fraction": round(float((got["flags"] & 1).mean()), 4),
        "steps_per_ray": round(float(cnt["steps"].mean()), 2),
        "oracle_exact": exact,
        "launch": ctx.launch_config(0),
    }
    if not args.no_cpu_baseline:
        rate, stride, n_sample = time_cpu(lambda r: O.trace_chunk(ow, r, threads), rays, args.cpu_budget)
        out["cpu_baseline"] = {
            "value": rate / 1e9, "unit": "Grays/s", "cores": threads, "kind": "port",
            "sample": f"{n_sample} rays (every {stride}th of {n}), raycast_chunk_tree restated in C "
                      f"(-O3 -ffp-contract=off), one warm-up pass, median of 5", "host": cpu_info}
    ctx.close()
    return out


def batched_leg(args, threads, stream):
    """64 uneven views + 2 bounces per hit over a 4096^3 sparse world, one graph replay per frame."""
    import torch

    import arcadia_amd as A

    t0 = time.time()
    world = A.HostWorld((16, 16, 16))
    world.generate(A.gen_params(A.ARC_GEN_SPARSE, seed=args.seed), nthreads=threads)
    world.pack()
    pools = world.pools()
    solid = sum(world.chunk_info(i)["solid"] for i in range(16 ** 3))
    del world
    t_build = time.time() - t0
    rng = np.random.default_rng(args.seed)
    views = []
    for _ in range(64):
        w = int(round(float(np.exp(rng.uniform(np.log(320.0), np.log(2560.0))))))
        h = max(1, w * 9 // 16)
        eye = (float(rng.uniform(16.0, 240.0)), float(rng.uniform(16.0, 26.0)), float(rng.uniform(16.0, 240.0)))
        cam = A.camera_look(eye, float(rng.uniform(0.0, 6.2831853)), float(rng.uniform(-0.6, -0.15)), 70.0, w, h)
        views.append((cam, w, h, 250.0))
    ctx = A.Context(0)
    ctx.upload(*pools)
    p0, pb = A.TraceParams.exact(0), A.TraceParams.exact(2, ignore_water=True)
    b = ctx.batch(views, p0, pb, n_bounces=2, bounce_max_dist_m=26.0, seed=args.seed)
    hits = torch.empty(b.records() * 16, dtype=torch.uint8, device="cuda")
    sptr = stream.cuda_stream
    print(f"[bench] batched: world {t_build:.1f}s, {len(pools[2])} nodes, {len(pools[3])} bricks, "
          f"{b.slots} slots", file=sys.stderr, flush=True)
    tuned = None
    if not args.no_autotune:
        tuned = b.autotune(hits.data_ptr(), sptr, reps=1)
        print(f"[bench] batched autotuned: {tuned[0]} ({tuned[1]:.2f} ms)", file=sys.stderr, flush=True)
    b.capture(hits.data_ptr())
    for _ in range(2):
        b.replay(sptr)
    torch.cuda.synchronize()
    k = max(3, min(args.steps, 10))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(k):
        b.replay(sptr)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / k
    t = hits.view(torch.float32)[0::4].reshape(3, b.slots)
    traced = (t != -16.0).sum(dim=1).cpu().numpy().astype(np.int64)
    n_rays = int(traced.sum())
    exact = batched_sample_check(pools, views, hits, b.slots, p0, pb, args.seed, 2, 26.0, 20000, threads)
    print(f"[bench] batched: sampled slots vs oracle {exact}", file=sys.stderr, flush=True)
    out = {
        "metric": "Grays/s batched views (primary + 2 diffuse bounces, one persistent launch)",
        "value": round(n_rays / (ms * 1e-3) / 1e9, 4),
        "unit": "Grays/s",
        "ms_per_frame": round(ms, 4),
        "frames": k,
        "rays_per_frame": n_rays,
        "primary_rays": int(traced[0]),
        "bounce1_rays": int(traced[1]),
        "bounce2_rays": int(traced[2]),
        "views": len(views),
        "slots": int(b.slots),
        "world": "4096^3 (16^3 chunks), fBm heightmap terrain 32..224 voxels, random mats 1..15",
        "occupancy": round(solid / float(4096 ** 3), 4),
        "world_nodes": int(len(pools[2])),
        "world_bricks": int(len(pools[3])),
        "launch": b.launch_config(),
        "autotuned": tuned is not None,
        "graph": "hipGraph of the single path launch, replayed per frame",
        "oracle_exact": exact,
    }
    b.close()
    ctx.close()
    return out


def batched_sample_check(pools, views, hits, n, p0, pb, seed, n_bounces, bounce_max_dist_m, count, threads):
    """Every record of `count` random pixel slots of the last replay against the oracle: each slot's
    primary ray is regenerated on the host (its view's padded 8x8-tile order), traced by the oracle,
    and its bounce chain re-derived from the oracle's hit with the slot's own bounce index."""
    import torch

    import arcadia_amd as A
    from oracle import oracle as O

    rng = np.random.default_rng(seed + 7)
    sample = np.sort(rng.choice(n, min(n, count), replace=False))
    rays = np.zeros(len(sample), A.RAY_DTYPE)
    first = 0
    for cam, w, h, md in views:
        nv = ((w + 7) // 8) * ((h + 7) // 8) * 64
        sel = (sample >= first) & (sample < first + nv)
        if sel.any():
            rays[sel] = A.view_rays(cam, w, h, md)[sample[sel] - first]
        first += nv
    idx = torch.from_numpy(np.concatenate([d * n + sample for d in range(n_bounces + 1)]).astype(np.int64)).cuda()
    got = hits.view(torch.int32).reshape(-1, 4)[idx].cpu().numpy().view(A.HIT_DTYPE).reshape(-1)
    ow = O.World(*pools)
    cur, exact = rays, {}
    for depth in range(n_bounces + 1):
        full, _ = O.trace_world(ow, p0 if depth == 0 else pb, cur, threads)
        ref = O.compact(full)
        exact[f"depth{depth}"] = bool(np.array_equal(got[depth * len(sample):(depth + 1) * len(sample)].view(np.uint32),
                                                     ref.view(np.uint32)))
        if depth < n_bounces:
            cur = np.concatenate([A.bounce_rays_dense(cur[k:k + 1], ref[k:k + 1], index_base=depth * n + int(s),
                                                      seed=seed, max_dist_m=bounce_max_dist_m)
                                  for k, s in enumerate(sample)])
    exact["slots"] = int(len(sample))
    return exact


def shadow_leg(args, ctx, pools, prays, phits, stream, threads):
    """Shadow any-hit walkers on the full-frame world: one sun visibility ray per primary hit
    (trace_shadow_vis, 72 steps, 88 m) and one visibility ray per hit towards another pixel's hit
    point (trace_shadow_vis_di); device records checked against the oracle on a sample."""
    import torch

    import arcadia_amd as A
    from arcadia_amd import shadow_rays as S

    sun = (0.35, 0.8, 0.45)
    out = {"metric": "Grays/s shadow any-hit walkers (full-frame world, rays from the 1080p primary hits)"}
    sptr = stream.cuda_stream
    for name, mode, rays in (("sun", A.ARC_SHADOW_SUN, S.sun_rays(prays, phits, sun, seed=args.seed)),
                             ("di", A.ARC_SHADOW_DI, S.toward_rays(prays, phits, seed=args.seed))):
        n = len(rays)
        d_r = torch.from_numpy(rays.view(np.uint8)).cuda()
        d_v = torch.empty(n, dtype=torch.float32, device="cuda")
        d_rec = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
        if not args.no_autotune:  # the walkers' launch shape, tuned on this mode's rays (untimed setup)
            cfg, tms = ctx.autotune_shadow(mode, 72, d_r.data_ptr(), n, d_v.data_ptr(), d_rec.data_ptr(), sptr)
            out.setdefault("launch", {})[name] = cfg
            print(f"[bench] shadow {name} autotuned: {cfg} ({tms:.3f} ms)", file=sys.stderr, flush=True)
        for _ in range(2):
            ctx.trace_shadow_device(mode, 72, d_r.data_ptr(), n, d_v.data_ptr(), d_rec.data_ptr(), sptr)
        k = max(3, min(args.steps, 10))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(k):
            ctx.trace_shadow_device(mode, 72, d_r.data_ptr(), n, d_v.data_ptr(), d_rec.data_ptr(), sptr)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / k
        vis = d_v.cpu().numpy()
        rec = d_rec.cpu().numpy().view(A.SHADOW_REC_DTYPE)
        sample = np.arange(0, n, max(1, n // 8192))
        exact = None
        if not args.no_cpu_baseline:
            from oracle import oracle as O

            ovis, orec = O.trace_shadow(O.World(*pools), mode, 72, rays[sample], threads)
            exact = bool(np.array_equal(ovis, vis[sample]) and
                         np.array_equal(orec.view(np.uint32), rec[sample].view(np.uint32)))
        out[name] = {
            "rays": int(n),
            "ms": round(ms, 4),
            "grays": round(n / (ms * 1e-3) / 1e9, 4),
            "visible_fraction": round(float(vis.mean()), 4),
            "outcomes": np.bincount(rec["outcome"], minlength=6).tolist(),
            "mean_steps": round(float(rec["steps"].mean()), 2),
            "oracle_sample_exact": exact,
            "sample_rays": int(len(sample)),
        }
        print(f"[bench] shadow {name}: {n} rays {ms:.3f} ms -> {out[name]['grays']} Grays/s, exact sample {exact}",
              file=sys.stderr, flush=True)
    return out


def rigid_leg(args, ctx, pools, prays, phits, stream, threads):
    """Rigid voxel bodies over the full-frame world, as the primary G-buffer pass traces them
    (primary_gbuffer.wgsl:123-125): every 1080p primary ray walks the 32 bodies (the instance
    limit) up to its world hit distance (250 m without one). The bodies are 32^3-voxel trees of
    three shapes (a solid rock, a hollow crate, scattered debris) at random orientations, each
    floating 1.5 m above a random visible surface point 8..48 m from the camera. The device
    records are checked against the oracle on a sample. The world pools are re-uploaded with the
    body trees appended."""
    import torch

    import arcadia_amd as A
    from arcadia_amd import rigid as RB
    from arcadia_amd import shadow_rays as S

    rng = np.random.default_rng(args.seed + 7)
    g = np.indices((32, 32, 32)).reshape(3, -1).T
    r2 = ((g - 15.5) ** 2).sum(1)
    shell = np.any((g == 1) | (g == 30), axis=1) & np.all((g >= 1) & (g <= 30), axis=1)
    shapes = [g[r2 < 14.0 ** 2], g[shell], rng.integers(0, 32, (3000, 3))]
    bpools, roots = pools, []
    for k, vox in enumerate(shapes):
        w = A.HostWorld((1, 1, 1))
        mats = rng.integers(1, 16, (len(vox), 1))
        w.set_voxels(np.ascontiguousarray(np.concatenate([vox, mats], 1), np.int32))
        w.pack()
        bpools, root = RB.append_body_tree(bpools, w.pools())
        roots.append(root)
    hidx, hp, hn = S.hit_points(prays, phits)
    dist = phits["t_voxels"][hidx] / 16.0
    pick = rng.choice(np.nonzero((dist > 8.0) & (dist < 48.0))[0], A.ARC_MAX_RIGID_BODIES, replace=False)
    bodies = []
    for i, j in enumerate(pick):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        a, b, c, d = q
        rot = np.array([[1 - 2 * (c * c + d * d), 2 * (b * c - d * a), 2 * (b * d + c * a)],
                        [2 * (b * c + d * a), 1 - 2 * (b * b + d * d), 2 * (c * d - b * a)],
                        [2 * (b * d - c * a), 2 * (c * d + b * a), 1 - 2 * (b * b + c * c)]])
        bodies.append(RB.body_instance(roots[i % 3], (32, 32, 32), rot, hp[j] + hn[j] * 1.5))
    bodies = np.array(bodies)
    rays = prays.copy()
    hit = (phits["flags"] & 1) != 0
    rays["t_max"] = np.where(hit, phits["t_voxels"] * np.float32(1.0 / 16.0), rays["t_max"]).astype(np.float32)
    ctx.upload(*bpools)
    n = len(rays)
    sptr = stream.cuda_stream
    d_r = torch.from_numpy(rays.view(np.uint8)).cuda()
    d_o = torch.empty(n * 64, dtype=torch.uint8, device="cuda")
    for _ in range(2):
        ctx.trace_rigid_device(bodies, d_r.data_ptr(), n, d_o.data_ptr(), sptr)
    k = max(3, min(args.steps, 10))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(k):
        ctx.trace_rigid_device(bodies, d_r.data_ptr(), n, d_o.data_ptr(), sptr)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / k
    out_h = d_o.cpu().numpy().view(A.RIGID_HIT_DTYPE)
    sample = np.arange(0, n, max(1, n // 20000))
    exact = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        want = O.trace_rigid(O.World(*bpools), bodies, rays[sample])
        exact = bool(np.array_equal(want.view(np.uint32), out_h[sample].view(np.uint32)))
    res = {
        "metric": "Grays/s rigid-body walker (32 bodies of 32^3 voxels, 1080p primary rays cut at the world hit)",
        "rays": int(n),
        "bodies": int(len(bodies)),
        "ms": round(ms, 4),
        "grays": round(n / (ms * 1e-3) / 1e9, 4),
        "hit_fraction": round(float((out_h["hit"] == 1).mean()), 4),
        "oracle_sample_exact": exact,
        "sample_rays": int(len(sample)),
    }
    print(f"[bench] rigid: {n} rays x {len(bodies)} bodies {ms:.3f} ms -> {res['grays']} Grays/s, "
          f"hits {res['hit_fraction']}, exact sample {exact}", file=sys.stderr, flush=True)
    return res


def snapshot_leg(args, ctx, world, pools, threads):
    """World snapshot of the full-frame world: written once (untimed), then loaded three ways,
    each timed to resident device pools: the streaming pool builder (arc_snapshot_upload, file
    pages already cached), the same after a host load + pack, and arc_world_upload of the
    already-packed host pools. The streamed pools are read back and compared with the host pools."""
    import tempfile

    import arcadia_amd as A

    grid, roots, nodes, bricks = pools
    pool_bytes = len(nodes) * 16 + len(bricks) * 64
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "world.arcw")
        t = time.perf_counter()
        world.save(path)
        t_write = time.perf_counter() - t
        file_bytes = os.path.getsize(path)
        res = {"metric": "world snapshot load to resident device pools (full-frame world)",
               "pool_bytes": pool_bytes, "file_bytes": file_bytes, "write_s": round(t_write, 3)}
        with A.Snapshot(path) as s:
            s.verify(threads)  # also pulls the file into the page cache
            times = []
            for _ in range(3):
                t = time.perf_counter()
                ctx.upload_snapshot(s, nthreads=threads)
                times.append(time.perf_counter() - t)
            res["stream_s"] = round(float(np.median(times)), 4)
            r2, n2, b2 = ctx.read_pools(len(roots), len(nodes), len(bricks))
            res["stream_pools_equal"] = bool(np.array_equal(r2, roots) and
                                             np.array_equal(n2.view(np.uint32), nodes.view(np.uint32)) and
                                             np.array_equal(b2.view(np.uint32), bricks.view(np.uint32)))
            t = time.perf_counter()
            w2 = s.load_host(threads)
            w2.pack()
            ctx.upload_world(w2)
            res["host_load_pack_upload_s"] = round(time.perf_counter() - t, 4)
            w2.close()
    times = []
    for _ in range(3):
        t = time.perf_counter()
        ctx.upload(grid, roots, nodes, bricks)
        times.append(time.perf_counter() - t)
    res["upload_packed_pools_s"] = round(float(np.median(times)), 4)
    res["stream_gbs"] = round(pool_bytes / res["stream_s"] / 1e9, 2)
    print(f"[bench] snapshot: {pool_bytes / 1e6:.0f} MB pools; stream {res['stream_s']} s "
          f"({res['stream_gbs']} GB/s, equal {res['stream_pools_equal']}), host load+pack+upload "
          f"{res['host_load_pack_upload_s']} s, upload of packed pools {res['upload_packed_pools_s']} s",
          file=sys.stderr, flush=True)
    return res


def restir_leg(args, ctx, pools, prays, phits, stream, threads):
    """ReSTIR GI / DI passes at 1080p on the full-frame world: G-buffer of the still camera, then
    temporal + spatial per frame with the history carried over, 64 emissive sources 0.6 m above
    visible surfaces (DI), no SHARC (every GI hit takes the bounce-sun shadow ray), 72-step any-hit
    shadows, the high-quality settings. Frames 2.. are timed with HIP events; frame 0 is checked
    against the oracle over every pixel."""
    import torch

    import arcadia_amd as A
    from arcadia_amd import restir as R
    from arcadia_amd import shadow_rays as S

    W, H = 1920, 1080
    cam = A.camera_look((31.3, 58.0, 27.1), 0.7, -0.5, 65.0, W, H)
    rng = np.random.default_rng(args.seed + 11)
    _, hp, hn = S.hit_points(prays, phits)
    pick = rng.choice(len(hp), 64, replace=False)
    src = R.emissive_sources((hp[pick] + hn[pick] * 0.6).astype(np.float32), 0.25, rng.uniform(1.0, 6.0, (64, 3)),
                             normals=hn[pick], eta=rng.uniform(0.0, 1.0, 64))
    lum = src["intensity"].astype(np.float64) @ np.array([0.2126, 0.7152, 0.0722])
    alias = R.alias_table(lum)
    n = W * H
    g = A.GbufferParams.from_camera(cam, W, H, prev_view_proj=R.view_proj(cam))
    d_dm = torch.zeros(n * 2, dtype=torch.int32, device="cuda")
    d_nm = torch.zeros(n * 2, dtype=torch.int32, device="cuda")
    d_aa = torch.zeros(n, dtype=torch.int32, device="cuda")
    sptr = stream.cuda_stream
    ctx.primary_gbuffer(A.TraceParams.exact(0), g, d_dm.data_ptr(), d_nm.data_ptr(), d_aa.data_ptr(), sptr)
    # the G-buffer pass itself (ray generation, full-record trace, packing), timed over 5 passes
    eg = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
    eg[0].record(stream)
    for _ in range(5):
        ctx.primary_gbuffer(A.TraceParams.exact(0), g, d_dm.data_ptr(), d_nm.data_ptr(), d_aa.data_ptr(), sptr)
    eg[1].record(stream)
    torch.cuda.synchronize()
    gbuffer_ms = eg[0].elapsed_time(eg[1]) / 5
    gi = [torch.zeros(n * 32, dtype=torch.uint8, device="cuda") for _ in range(3)]
    di = [torch.zeros(n * 16, dtype=torch.uint8, device="cuda") for _ in range(3)]
    gi[0].view(torch.float32).view(-1, 8)[:, 6] = 1e5  # empty history: primary_depth 1e5
    di[0].view(torch.int32).view(-1, 4)[:, 0] = -1      # light_idx_seed 0xFFFFFFFF
    d_src = torch.from_numpy(src.view(np.uint8).copy()).cuda()
    d_al = torch.from_numpy(alias.view(np.uint8).copy()).cuda()
    sun_t, gi_t = A.TraceParams.exact(1, True), A.TraceParams.exact(2, True)
    k = max(4, min(args.steps, 10))
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(k)]
    hist = 0  # gi[hist] / di[hist] = the history the next temporal pass reads
    frame0 = None
    for f in range(k):
        p = A.RestirParams.make(cam, W, H, frame=f, n_emissive=64, total_flux=float(lum.sum()), sun_intensity=3.0)
        tmp, out = (hist + 1) % 3, (hist + 2) % 3
        ev[f][0].record(stream)
        ctx.restir_temporal(p, sun_t, gi_t, d_dm.data_ptr(), d_nm.data_ptr(), gi[hist].data_ptr(), gi[tmp].data_ptr(),
                            di[hist].data_ptr(), di[tmp].data_ptr(), d_src.data_ptr(), d_al.data_ptr(), stream=sptr)
        ev[f][1].record(stream)
        ctx.restir_spatial(p, d_dm.data_ptr(), d_nm.data_ptr(), gi[tmp].data_ptr(), gi[out].data_ptr(),
                           di[tmp].data_ptr(), di[out].data_ptr(), d_src.data_ptr(), stream=sptr)
        ev[f][2].record(stream)
        if f == 0:
            torch.cuda.synchronize()
            frame0 = (p, [x.cpu().numpy() for x in (d_dm, d_nm, gi[hist], di[hist], gi[tmp], di[tmp], gi[out], di[out])])
        hist = out
    torch.cuda.synchronize()
    t_ms = [ev[f][0].elapsed_time(ev[f][1]) for f in range(2, k)]
    s_ms = [ev[f][1].elapsed_time(ev[f][2]) for f in range(2, k)]
    go = gi[hist].cpu().numpy().view(A.GI_RESERVOIR_DTYPE)
    exact = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        p, (dm, nm, gi0, di0, gt, dt, g1, d1) = frame0
        dm, nm = dm.view(np.uint32).reshape(-1, 2), nm.view(np.uint32).reshape(-1, 2)
        ow = O.World(*pools)
        ot, odt = O.restir_temporal(ow, p, sun_t, gi_t, dm, nm, gi0.view(A.GI_RESERVOIR_DTYPE),
                                    di0.view(A.DI_RESERVOIR_DTYPE), R.palette(), sources=src, alias=alias,
                                    nthreads=threads)
        oo, odo = O.restir_spatial(ow, p, dm, nm, ot, odt, R.palette(), sources=src, nthreads=threads)
        exact = bool(all(np.array_equal(a.view(np.uint32), np.ascontiguousarray(b).view(np.uint32))
                         for a, b in ((gt, ot), (dt, odt), (g1, oo), (d1, odo))))
    res = {
        "metric": "ms per 1080p frame, ReSTIR GI/DI temporal + spatial (full-frame world, 64 emitters, no SHARC); "
                  "gbuffer_ms: the primary G-buffer pass that feeds them",
        "temporal_ms": round(float(np.median(t_ms)), 4),
        "spatial_ms": round(float(np.median(s_ms)), 4),
        "gbuffer_ms": round(float(gbuffer_ms), 4),
        "frames_timed": len(t_ms),
        "pixels": n,
        "mean_M": round(float(((go["norm_m_flags"] >> 16) & 0xFF).mean()), 3),
        "frame0_oracle_exact": exact,
    }
    print(f"[bench] restir: temporal {res['temporal_ms']} ms, spatial {res['spatial_ms']} ms per 1080p frame, "
          f"mean M {res['mean_M']}, frame 0 exact {exact}; G-buffer pass {res['gbuffer_ms']} ms", file=sys.stderr,
          flush=True)
    return res


def cache_leg(args, threads):
    """Learned radiance cache vs SHARC (2^22 entries) vs no cache, rel-MSE against the FP64
    1024-spp path-traced judge at 1920x1080 on the full-frame world with 8 emissive clusters, both
    caches trained online on 2^16 four-bounce paths per frame; then the same comparison at a fixed
    per-frame GPU budget (the neural cache's frame, with SHARC given as many paths as fit)."""
    import arcadia_amd as A
    from arcadia_amd import lightcache as LC

    t0 = time.time()
    world = A.HostWorld((4, 4, 4))
    world.generate(A.gen_params(A.ARC_GEN_CAVES, seed=args.seed), nthreads=threads)
    clusters = LC.add_emissive_clusters(world, (4, 4, 4), seed=args.seed, count=8)
    world.pack()
    pools = world.pools()
    del world
    ctx = A.Context(0)
    ctx.upload(*pools)
    cam = A.camera_look((31.3, 58.0, 27.1), 0.7, -0.5, 65.0, 1920, 1080)
    run = LC.CacheRun(ctx, cam, 1920, 1080, train_size=(344, 192), sharc_log2=20,
                      nrc=A.NrcParams.default(bbox_size=64.0, log2_table=14), seed=args.seed)
    print(f"[bench] cache: scene {time.time() - t0:.1f}s, {run.n_eval} eval slots, {run.n_train} training paths "
          f"per frame", file=sys.stderr, flush=True)
    ref_s = run.reference(args.cache_ref_spp, progress=lambda k: print(f"[bench] cache: reference {k} spp",
                                                                        file=sys.stderr, flush=True))
    for f in range(args.cache_frames):
        run.train_paths(f)
        run.train_sharc()
        run.train_nrc(1)
    eval_seed = (1 << 40) + args.seed
    err = run.errors(seed=eval_seed)
    ms = run.cache_ms(reps=5)
    # Rel-MSE at a fixed per-frame GPU budget: each cache's frame = tracing and shading its training
    # paths + its own work. The budget is the neural cache's frame (2^16 paths); SHARC spends it on
    # as many training paths as fit beside its deposit, resolve and query.
    t_train = run.train_trace_ms()
    nrc_frame, sharc_frame = t_train + ms["nrc_ms"], t_train + ms["sharc_ms"]
    at = run.sharc_at_budget(nrc_frame, args.cache_frames, 20, eval_seed, t_train / run.n_train, ms["sharc_ms"])
    print(f"[bench] cache at {nrc_frame:.3f} ms/frame: nrc {err['nrc']:.5f}, sharc {at['rel_mse']:.5f} "
          f"({at['paths_per_frame']} paths/frame, measured {at['ms_per_frame']:.3f} ms)", file=sys.stderr, flush=True)
    out = {
        "metric": "rel-MSE vs FP64 1024-spp path-traced judge at 1920x1080 (lower is better), and at a fixed ms budget",
        "nrc": round(err["nrc"], 5),
        "sharc": round(err["sharc"], 5),
        "no_cache": round(err["no_cache"], 5),
        "pt_1spp": round(err["pt_1spp"], 5),
        "nrc_ms_per_frame": round(ms["nrc_ms"], 4),
        "sharc_ms_per_frame": round(ms["sharc_ms"], 4),
        "training_trace_ms_per_frame": round(t_train, 4),
        "at_fixed_ms": {
            "budget_ms": round(nrc_frame, 4),
            "nrc": round(err["nrc"], 5),
            "sharc": round(at["rel_mse"], 5),
            "sharc_paths_per_frame": at["paths_per_frame"],
            "sharc_measured_ms_per_frame": round(at["ms_per_frame"], 4),
            "note": "per-frame GPU ms of tracing + shading the training paths plus the cache's own work; "
                    "the budget is the neural cache's frame at 2^16 paths, SHARC's paths scaled to fill it",
        },
        "frame_ms": {"nrc": round(nrc_frame, 4), "sharc": round(sharc_frame, 4)},
        "frames": args.cache_frames,
        "training_paths_per_frame": int(run.n_train),
        "ref_spp": run.ref_spp,
        "ref_seconds": round(ref_s, 2),
        "scene": "full-frame world + 8 emissive clusters (mats 19-22, r 6..14 voxels), analytic sky, 4-bounce paths",
        "sharc_entries": 1 << 22,
        "nrc_model": {"hash_levels": 16, "table": 1 << 14, "mlp": "38->64->64->64->3 FP16 MFMA, FP32 accumulate"},
    }
    run.close()
    ctx.close()
    return out


def destruction_leg(args, threads):
    """30 frames over the full-frame world; each frame carves >= 1% of the solid voxels with
    seeded spheres (radius 4..32 voxels), syncs the touched nodes into GPU patches, traces the
    1080p primary + diffuse frame in one launch (arc_trace_bounce_device), checks a sample of it bit-exact against the
    oracle on the edited host pools, invalidates SHARC surfels in the carved spheres and takes a
    neural-cache training step on fresh paths."""
    import torch

    import arcadia_amd as A
    from oracle import oracle as O

    world = A.HostWorld((4, 4, 4))
    world.generate(A.gen_params(A.ARC_GEN_CAVES, seed=args.seed), nthreads=threads)
    world.pack()
    ctx = A.Context(0)
    ctx.upload(*world.pools(copy=False))
    world.clear_patches()
    cam = A.camera_look((31.3, 58.0, 27.1), 0.7, -0.5, 65.0, 1920, 1080)
    p0, pb = A.TraceParams.exact(0), A.TraceParams.exact(2, ignore_water=True)
    tcam = A.camera_look((0.0, 0.0, 0.0), 0.0, 0.0, 60.0, 341, 192)
    for f in ("eye", "forward", "right", "up"):
        setattr(tcam, f, getattr(cam, f))
    tcam.tan_half_fov, tcam.aspect = cam.tan_half_fov, 341 / 192
    train = ctx.batch([(tcam, 341, 192, 250.0)], p0, pb, n_bounces=4, bounce_max_dist_m=250.0, seed=args.seed)
    rays0 = A.view_rays(cam, 1920, 1080, 250.0)
    n = len(rays0)
    d_rays0 = torch.from_numpy(rays0.view(np.uint8)).cuda()
    hits = torch.empty(2 * n * 16, dtype=torch.uint8, device="cuda")
    thits = torch.empty(train.records() * 16, dtype=torch.uint8, device="cuda")
    tsamp = torch.empty(train.slots * 96, dtype=torch.uint8, device="cuda")
    sharc = A.Sharc(20)
    nrc = A.Nrc(A.NrcParams.default(bbox_size=64.0, log2_table=14))
    shade = A.ShadeParams.default()
    rng = np.random.default_rng(args.seed + 5)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    t_edit, t_patch, t_trace, t_cache, carved, exact, patched_mb = [], [], [], [], [], 0, []
    for f in range(args.destruction_frames):
        solid = sum(world.chunk_info(i)["solid"] for i in range(64))
        quota = solid // 100
        t0 = time.perf_counter()
        done, spheres = 0, []
        while done < quota:
            c = rng.uniform(0.0, 64.0, 3)
            r = rng.uniform(4.0, 32.0) / 16.0
            done += world.carve_crater(tuple(c), r)
            spheres.append((c[0], c[1], c[2], r))
        t1 = time.perf_counter()
        pl = world.patch_list() if f == 0 else None
        if pl is not None:
            patched_mb.append(sum(p["nodes"].nbytes + p["bricks"].nbytes for p in pl) / 2**20)
        ctx.patch_world(world)
        t2 = time.perf_counter()
        ev[0].record()
        # the frame's primaries (8x8-tile order) and their bounces in one launch
        ctx.trace_bounce_device(p0, pb, d_rays0.data_ptr(), n, args.seed * 7919 + f, 26.0, hits.data_ptr())
        ev[1].record()
        sharc.invalidate(np.array(spheres, np.float32))
        train.set_seed(args.seed * 104729 + f)
        train.trace(thits.data_ptr())
        train.shade(thits.data_ptr(), shade, d_samples=tsamp.data_ptr())
        sharc.deposit_paths(tsamp.data_ptr(), train.slots, tuple(cam.eye))
        sharc.resolve()
        nrc.train(tsamp.data_ptr(), train.slots)
        ev[2].record()
        torch.cuda.synchronize()
        t_edit.append(t1 - t0)
        t_patch.append(t2 - t1)
        t_trace.append(ev[0].elapsed_time(ev[1]))
        t_cache.append(ev[1].elapsed_time(ev[2]))
        carved.append(done)
        # bit-exact check of a sample of this frame against the oracle on the edited host pools
        got = hits.cpu().numpy().view(A.HIT_DTYPE).reshape(2, n)
        sample = np.sort(rng.choice(n, 4096, replace=False))
        pools = world.pools(copy=False)
        ow = O.World(*pools)
        full, _ = O.trace_world(ow, p0, rays0[sample], threads)
        h0 = O.compact(full)
        ok = len(np.nonzero((got[0][sample].view(np.uint32).reshape(-1, 4) != h0.view(np.uint32).reshape(-1, 4)).any(1))[0]) == 0
        br = np.zeros(len(sample), A.RAY_DTYPE)
        for k, s_ in enumerate(sample):
            br[k] = A.bounce_rays_dense(rays0[s_:s_ + 1], h0[k:k + 1], index_base=int(s_),
                                        seed=args.seed * 7919 + f, max_dist_m=26.0)[0]
        full1, _ = O.trace_world(ow, pb, br, threads)
        h1 = O.compact(full1)
        ok = ok and len(np.nonzero((got[1][sample].view(np.uint32).reshape(-1, 4) !=
                                    h1.view(np.uint32).reshape(-1, 4)).any(1))[0]) == 0
        exact += int(ok)
        print(f"[bench] destruction frame {f}: carved {done} ({len(spheres)} spheres) edit {t1 - t0:.3f}s "
              f"patch {t2 - t1:.3f}s trace {t_trace[-1]:.3f}ms cache {t_cache[-1]:.3f}ms exact {ok}",
              file=sys.stderr, flush=True)
    t = hits.view(torch.float32)[0::4].reshape(2, n)
    rays = int(((t != -16.0).sum()).item())
    out = {
        "metric": "destruction sequence: per-frame edit + patch + 1080p primary/diffuse traversal + cache update",
        "frames": args.destruction_frames,
        "frames_bit_exact": exact,
        "carved_voxels_per_frame": int(np.mean(carved)),
        "host_edit_ms": round(1e3 * float(np.mean(t_edit)), 2),
        "patch_upload_ms": round(1e3 * float(np.mean(t_patch)), 2),
        "first_frame_patch_mib": round(patched_mb[0], 1) if patched_mb else None,
        "trace_ms": round(float(np.mean(t_trace)), 4),
        "trace_grays": round(rays / (float(np.mean(t_trace)) * 1e-3) / 1e9, 4),
        "cache_ms": round(float(np.mean(t_cache)), 4),
        "check": "4096 random pixels per frame: primary and bounce records vs the oracle on the edited host pools",
    }
    sharc.close()
    nrc.close()
    train.close()
    ctx.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=10.0, help="seconds of CPU work for the baseline sample")
    ap.add_argument("--profile", action="store_true",
                    help="5 timed steps after 1 warm-up, no CPU leg and no other workloads (for rocprofv3); the "
                         "autotuner still runs, so the timed launches use the same shapes as a full run")
    ap.add_argument("--no-autotune", action="store_true", help="use the default launch shapes")
    ap.add_argument("--launch-primary", default="", help="fixed launch shape of the primary-ray launches, "
                    "blocks_per_cu,block_threads,refill_lanes,batch_rays (no autotuning of that class)")
    ap.add_argument("--launch-bounce", default="", help="the same for the bounce-ray launches")
    ap.add_argument("--launch-fused", default="", help="the same for the one-launch primary + bounce step")
    ap.add_argument("--no-c1", action="store_true", help="skip the single-chunk raycast_chunk_tree workload")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-views workload")
    ap.add_argument("--no-shadow", action="store_true", help="skip the shadow-walker workload")
    ap.add_argument("--no-rigid", action="store_true", help="skip the rigid-body workload")
    ap.add_argument("--no-snapshot", action="store_true", help="skip the world-snapshot workload")
    ap.add_argument("--no-restir", action="store_true", help="skip the ReSTIR pass workload")
    ap.add_argument("--no-cache", action="store_true", help="skip the radiance-cache workload")
    ap.add_argument("--cache-ref-spp", type=int, default=1024)
    ap.add_argument("--cache-frames", type=int, default=60)
    ap.add_argument("--no-destruction", action="store_true", help="skip the destruction sequence")
    ap.add_argument("--destruction-frames", type=int, default=30)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run one device copy of this many MiB (scales rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this traversal benchmark runs on exactly one GPU; --gpus must be 1 "
                                   "(rays shard by screen tile across GPUs with a full world replica each)"}))
        return 2
    if args.profile:
        args.steps, args.warmup, args.no_cpu_baseline = min(args.steps, 5), 1, True
        args.no_batched = True
        args.no_c1 = True
        args.no_shadow = True
        args.no_rigid = True
        args.no_snapshot = True
        args.no_restir = True
        args.no_cache = True
        args.no_destruction = True

    import torch

    import arcadia_amd as A

    threads, cpu_info = host_cpu_info()
    world, pools, prays, t_build = build_workload(args.seed, threads)
    print(f"[bench] world built in {t_build:.1f}s: {len(pools[2])} nodes, {len(pools[3])} bricks", file=sys.stderr,
          flush=True)
    ctx = A.Context(0)
    ctx.upload(*pools)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream
    p_primary = A.TraceParams.exact(0, ignore_water=False)
    p_bounce = A.TraceParams.exact(2, ignore_water=True)

    d_prays = torch.from_numpy(prays.view(np.uint8)).cuda()
    d_phits = torch.zeros(len(prays) * 16, dtype=torch.uint8, device="cuda")
    ctx.trace_device(p_primary, d_prays.data_ptr(), len(prays), d_phits.data_ptr(), sptr)
    torch.cuda.synchronize()
    phits = d_phits.cpu().numpy().view(A.HIT_DTYPE)
    brays, bsrc = A.bounce_rays(prays, phits, seed=args.seed, max_dist_m=26.0)
    d_brays = torch.from_numpy(brays.view(np.uint8)).cuda()
    d_bhits = torch.zeros(max(len(brays), 1) * 16, dtype=torch.uint8, device="cuda")
    n_p, n_b = len(prays), len(brays)
    print(f"[bench] {n_p} primary rays, {n_b} bounce rays", file=sys.stderr, flush=True)

    d_fhits = torch.zeros(2 * n_p * 16, dtype=torch.uint8, device="cuda")
    bounce_seed, bounce_m = args.seed, 26.0

    tuned = {}
    fixed = {}
    for cls, spec in ((0, args.launch_primary), (2, args.launch_bounce), (A.ARC_LAUNCH_CLASS_BOUNCE, args.launch_fused)):
        if spec:
            bpc, bt, rf, bat = (int(x) for x in spec.split(","))
            ctx.set_launch_config_class(cls, blocks_per_cu=bpc, block_threads=bt, refill_lanes=rf, batch_rays=bat)
            fixed[cls] = True
    if not args.no_autotune:  # launch shapes from the native autotuner (untimed setup)
        for cls, label, params, d_r, n_r, d_h in ((0, "primary", p_primary, d_prays, n_p, d_phits),
                                                  (2, "bounce", p_bounce, d_brays, n_b, d_bhits)):
            if cls in fixed:
                continue
            cfg, ms = ctx.autotune(params, d_r.data_ptr(), n_r, d_h.data_ptr(), sptr)
            tuned[label] = cfg
            print(f"[bench] autotuned {label}: {cfg} ({ms:.3f} ms)", file=sys.stderr, flush=True)
        if A.ARC_LAUNCH_CLASS_BOUNCE not in fixed:
            cfg, ms = ctx.autotune_bounce(p_primary, p_bounce, d_prays.data_ptr(), n_p, bounce_seed, bounce_m,
                                          d_fhits.data_ptr(), sptr)
            tuned["fused"] = cfg
            print(f"[bench] autotuned fused: {cfg} ({ms:.3f} ms)", file=sys.stderr, flush=True)

    def step(ev0=None, ev1=None):
        if ev0 is not None:
            ev0.record(stream)
        ctx.trace_bounce_device(p_primary, p_bounce, d_prays.data_ptr(), n_p, bounce_seed, bounce_m,
                                d_fhits.data_ptr(), sptr)
        if ev1 is not None:
            ev1.record(stream)

    def split_step(evs=None):
        if evs is not None:
            evs[0].record(stream)
        ctx.trace_device(p_primary, d_prays.data_ptr(), n_p, d_phits.data_ptr(), sptr)
        if evs is not None:
            evs[1].record(stream)
        ctx.trace_device(p_bounce, d_brays.data_ptr(), n_b, d_bhits.data_ptr(), sptr)
        if evs is not None:
            evs[2].record(stream)

    if args.copy_probe_mb:
        a = torch.ones(args.copy_probe_mb << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for k in range(args.steps):
        step(*evs[k])
    end.record(stream)
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(end)
    ms_launch = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    ms_step = total_ms / args.steps
    rays_per_step = n_p + n_b
    value = rays_per_step / (ms_step * 1e-3) / 1e9
    # the same rays as two launches (bounce rays generated on the host beforehand), for comparison
    for _ in range(args.warmup):
        split_step()
    torch.cuda.synchronize()
    sevs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    for k in range(args.steps):
        split_step(sevs[k])
    torch.cuda.synchronize()
    ms_p = float(np.mean([e[0].elapsed_time(e[1]) for e in sevs]))
    ms_b = float(np.mean([e[1].elapsed_time(e[2]) for e in sevs]))
    print(f"[bench] {args.steps} steps: {ms_step:.3f} ms/step (one launch {ms_launch:.3f} ms) -> {value:.3f} Grays/s; "
          f"two launches: primary {ms_p:.3f} ms + bounce {ms_b:.3f} ms -> {rays_per_step / ((ms_p + ms_b) * 1e-3) / 1e9:.3f} "
          f"Grays/s", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "full-frame traversal: 1024^3 world (4x4x4 chunks, 3D fBm caves @0.45 + 2000 cuboids), "
                        "1920x1080 primary + 1 cosine diffuse bounce per hit, exact (no LOD, no step cap); "
                        "one launch per step (arc_trace_bounce_device: bounces generated on the device)",
            "world_nodes": int(len(pools[2])),
            "world_bricks": int(len(pools[3])),
            "primary_rays": int(n_p),
            "bounce_rays": int(n_b),
            "ms_launch": round(ms_launch, 4),
            "launch": ctx.launch_config(A.ARC_LAUNCH_CLASS_BOUNCE),
            "primary_grays": round(n_p / (ms_p * 1e-3) / 1e9, 4),
            "bounce_grays": round(n_b / (ms_b * 1e-3) / 1e9, 4),
            "split": {
                "note": "the same rays as two launches of k_trace (primary, then the bounce rays generated on "
                        "the host before timing)",
                "grays": round(rays_per_step / ((ms_p + ms_b) * 1e-3) / 1e9, 4),
                "ms_primary": round(ms_p, 4),
                "ms_bounce": round(ms_b, 4),
                "launch": {"primary": ctx.launch_config(0), "bounce": ctx.launch_config(2)},
            },
            "autotuned": sorted(tuned),
            "parallelism": "single GPU",
        },
    }
    if not args.no_cpu_baseline:
        cpu, bytes_per = cpu_baseline(pools, prays, brays, p_primary, p_bounce, threads, cpu_info,
                                      args.cpu_budget)
        # One kernel runs per step: k_trace_bounce on the primary rays and their bounces. Its
        # algorithmic bytes are the primary rays' and the bounce rays' (the oracle's per-ray load
        # counts); the two k_trace launches of the same rays sit beside it under "split".
        def roof(kind, n, ms):
            achieved = bytes_per[kind] * n / (ms * 1e-3) / 1e9
            return {"achieved": round(achieved, 2), "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic(kind), "bytes_per_ray": round(bytes_per[kind], 2),
                    "avg_launch_ms": round(ms, 4)}
        rp, rb = roof("primary", n_p, ms_p), roof("bounce", n_b, ms_b)
        ev = measured_counters()
        bytes_launch = bytes_per["primary"] * n_p + bytes_per["bounce"] * n_b
        achieved = bytes_launch / (ms_launch * 1e-3) / 1e9
        traffic = measured_traffic("fused")
        out["roofline"] = {
            # what bounds the kernel, from its rocprofv3 counters: VALU busy near saturation with HBM
            # traffic a few percent of peak is issue-bound; `achieved` / `frac` stay the algorithmic
            # (logical) bytes per launch over HBM peak
            "bound": ev.get("fused_bound", "unmeasured"),
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "hbm_gbs": None if traffic is None else round(traffic / (ms_launch * 1e-3) / 1e9, 1),
            "l2_hit": ev.get("fused_l2_hit"),
            "valu_busy": ev.get("fused_valu_busy"),
            "lanes_active": ev.get("fused_lanes_active"),
            "kernel": "k_trace_bounce<water off, off32, uniform params>: primary rays and their on-device bounces",
            "bytes_per_launch": int(bytes_launch),
            "bytes_per_ray": {"primary": rp["bytes_per_ray"], "bounce": rb["bytes_per_ray"]},
            "steps_per_ray": {"primary": round(bytes_per["steps_primary"], 2),
                              "bounce": round(bytes_per["steps_bounce"], 2)},
            "avg_launch_ms": round(ms_launch, 4),
            "split": {
                "primary": dict(rp, kernel="k_trace<WORLD, compact, ignore_water off> on the primary rays",
                                hbm_gbs=None if rp["traffic"] is None else round(rp["traffic"] / (ms_p * 1e-3) / 1e9, 1),
                                l2_hit=ev.get("primary_l2_hit"), valu_busy=ev.get("primary_valu_busy"),
                                lanes_active=ev.get("primary_lanes_active")),
                "bounce": dict(rb, kernel="k_trace<WORLD, compact, ignore_water on> on the bounce rays",
                               hbm_gbs=None if rb["traffic"] is None else round(rb["traffic"] / (ms_b * 1e-3) / 1e9, 1),
                               l2_hit=ev.get("bounce_l2_hit"), valu_busy=ev.get("bounce_valu_busy"),
                               lanes_active=ev.get("bounce_lanes_active")),
            },
            "evidence": ev.get("source"),
        }
        out["cpu_baseline"] = cpu
        # the records of the last timed launches against the oracle on the strided samples: the
        # one-launch step's (primary i at i, its bounce at n + i) and the two launches'
        got_f = d_fhits.cpu().numpy().view(A.HIT_DTYPE)
        got_p = d_phits.cpu().numpy().view(A.HIT_DTYPE)
        got_b = d_bhits.cpu().numpy().view(A.HIT_DTYPE)[:n_b]
        exact = {}
        idx_p, ref_p = bytes_per["ref_primary"]
        idx_b, ref_b = bytes_per["ref_bounce"]
        for kind, got, idx, ref in (("primary", got_f[:n_p], idx_p, ref_p), ("bounce", got_f[n_p:][bsrc], idx_b, ref_b),
                                    ("split_primary", got_p, idx_p, ref_p), ("split_bounce", got_b, idx_b, ref_b)):
            exact[kind] = bool(np.array_equal(got[idx].view(np.uint32), ref.view(np.uint32)))
            exact[kind + "_rays"] = int(len(idx))
        miss = np.ones(n_p, bool)
        miss[bsrc] = False
        empty = got_f[n_p:][miss]
        exact["bounce_of_misses_empty"] = bool(np.all(empty["t_voxels"] == -16.0) and np.all(empty["steps"] == 0)
                                               and np.all((empty["flags"] & 1) == 0))
        out["oracle_sample_exact"] = exact
        print(f"[bench] timed records vs oracle: {exact}", file=sys.stderr, flush=True)
    if not args.no_c1:
        out["c1"] = c1_leg(args, stream, threads, cpu_info)
        print(f"[bench] c1: {out['c1']['value']} Grays/s ({out['c1']['ms']} ms), exact {out['c1']['oracle_exact']}",
              file=sys.stderr, flush=True)
    if not args.no_shadow:
        out["shadow"] = shadow_leg(args, ctx, pools, prays, phits, stream, threads)
    if not args.no_rigid:
        out["rigid"] = rigid_leg(args, ctx, pools, prays, phits, stream, threads)
    if not args.no_snapshot:
        out["snapshot"] = snapshot_leg(args, ctx, world, pools, threads)
    if not args.no_restir:
        out["restir"] = restir_leg(args, ctx, pools, prays, phits, stream, threads)
    ctx.close()
    del d_prays, d_phits, d_brays, d_bhits, d_fhits
    if not args.no_batched:
        out["batched"] = batched_leg(args, threads, stream)
        print(f"[bench] batched: {out['batched']['value']} Grays/s, {out['batched']['ms_per_frame']} ms/frame",
              file=sys.stderr, flush=True)
    if not args.no_cache:
        out["cache"] = cache_leg(args, threads)
        c = out["cache"]
        print(f"[bench] cache: rel-MSE nrc {c['nrc']} sharc {c['sharc']} none {c['no_cache']}; "
              f"ms nrc {c['nrc_ms_per_frame']} sharc {c['sharc_ms_per_frame']}", file=sys.stderr, flush=True)
    if not args.no_destruction:
        out["destruction"] = destruction_leg(args, threads)
    failures = parity_failures(out)
    if failures:
        out["parity_failures"] = failures
    print(json.dumps(out), flush=True)
    if failures:
        print(f"[bench] records differ from the oracle: {failures}", file=sys.stderr, flush=True)
        return 1
    return 0


def parity_failures(out: dict) -> list:
    """Every oracle comparison the run made that did not match (the run then exits non-zero);
    comparisons a run skips (None) are not failures."""
    bad = []
    for k, v in out.get("oracle_sample_exact", {}).items():
        if v is False:
            bad.append(f"full-frame {k}")
    if out.get("c1", {}).get("oracle_exact") is False:
        bad.append("c1")
    for walker in ("sun", "di"):
        if out.get("shadow", {}).get(walker, {}).get("oracle_sample_exact") is False:
            bad.append(f"shadow {walker}")
    if out.get("rigid", {}).get("oracle_sample_exact") is False:
        bad.append("rigid")
    if out.get("snapshot", {}).get("stream_pools_equal") is False:
        bad.append("snapshot pools")
    if out.get("restir", {}).get("frame0_oracle_exact") is False:
        bad.append("restir frame 0")
    if "destruction" in out and out["destruction"]["frames_bit_exact"] != out["destruction"]["frames"]:
        bad.append("destruction")
    for k, v in out.get("batched", {}).get("oracle_exact", {}).items():
        if v is False:
            bad.append(f"batched {k}")
    return bad


if __name__ == "__main__":
    sys.exit(main())
