"""Radiance-cache workload: the learned cache and SHARC, trained online from path samples the
traversal generates, judged against an FP64-accumulated path-traced reference.

A frame:
  1. training paths: one 4-bounce path per pixel of a ~2^14-pixel view of the camera
     (arc_batch_trace, one persistent launch), shaded into path samples (arc_batch_shade);
  2. SHARC: deposit each sample's radiance beyond x1 at x1's surfel, resolve;
     neural cache: one Adam step on the samples;
  3. evaluation (1920x1080 by default): one path per pixel; the image estimate at the primary hit
     x0 is le0 + a0 (base + cache(x1)) with the cache answering for the rest of the path.
The judge is the mean of `ref_spp` full 4-bounce paths per pixel, summed in FP64. Error is the
relative MSE mean((est - ref)^2 / (ref^2 + 0.01)) over pixels and RGB.

No reference counterpart: Arcadia's SHARC update traces its own paths in sharc_update.wgsl;
here both caches learn from the same traced samples.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import native as N


def add_emissive_clusters(world: N.HostWorld, grid, seed: int = 1, count: int = 8, r_min: int = 6,
                          r_max: int = 14) -> list:
    """Seeded emissive blobs (mats 19..22, the palette's glowing crystals and lava) centred on the
    terrain surface of random columns, written into the host world before packing; returns their
    (center, radius, material) in voxels."""
    rng = np.random.default_rng(seed + 0xC1)
    ext = np.array(grid) * 256
    out = []
    for k in range(count):
        r = int(rng.integers(r_min, r_max + 1))
        cx, cz = int(rng.integers(r, ext[0] - r)), int(rng.integers(r, ext[2] - r))
        top = next((y for y in range(ext[1] - 1, -1, -1) if world.voxel(cx, y, cz) != 0), ext[1] // 2)
        c = np.array([cx, min(top, ext[1] - r - 1), cz])
        m = 19 + (k % 4)
        z, y, x = np.mgrid[-r:r + 1, -r:r + 1, -r:r + 1]
        inside = x * x + y * y + z * z <= r * r
        pts = np.stack([x[inside] + c[0], y[inside] + c[1], z[inside] + c[2], np.full(inside.sum(), m)], 1)
        world.set_voxels(pts.astype(np.int32))
        out.append((tuple(int(v) for v in c), r, m))
    return out


class CacheRun:
    """Buffers and launches of the workload on one camera."""

    def __init__(self, ctx: N.Context, cam: N.Camera, width: int, height: int, train_size=(171, 96),
                 shade: N.ShadeParams | None = None, sharc_log2: int = 20, nrc: N.NrcParams | None = None,
                 max_dist_m: float = 250.0, seed: int = 1):
        import torch

        self.torch = torch
        self.ctx = ctx
        self.cam = cam
        self.shade = shade or N.ShadeParams.default()
        self.cam_pos = tuple(float(v) for v in cam.eye)
        p0, pb = N.TraceParams.exact(0), N.TraceParams.exact(2, ignore_water=True)
        self.eval = ctx.batch([(cam, width, height, max_dist_m)], p0, pb, n_bounces=4, bounce_max_dist_m=max_dist_m,
                              seed=seed)
        self.max_dist_m = max_dist_m
        self.seed = seed
        self.train = self.train_batch(train_size)
        dev = "cuda"
        ne, nt = self.eval.slots, self.train.slots
        self.n_eval, self.n_train = ne, nt
        self.eval_hits = torch.empty(self.eval.records() * 16, dtype=torch.uint8, device=dev)
        self.train_hits = torch.empty(self.train.records() * 16, dtype=torch.uint8, device=dev)
        self.eval_samples = torch.empty(ne * 96, dtype=torch.uint8, device=dev)
        self.train_samples = torch.empty(nt * 96, dtype=torch.uint8, device=dev)
        self.ref = torch.zeros(ne * 3, dtype=torch.float64, device=dev)
        self.ref_spp = 0
        self.pt = torch.empty(ne * 3, dtype=torch.float32, device=dev)
        self.cache_out = torch.empty(ne * 4, dtype=torch.float32, device=dev)
        self.est = torch.empty(ne * 3, dtype=torch.float32, device=dev)
        self.sharc = N.Sharc(sharc_log2)
        self.nrc = N.Nrc(nrc or N.NrcParams.default())
        self.seed = seed

    def train_batch(self, size):
        """One 4-bounce path per pixel of a `size` view of the camera (same basis and vertical FOV)."""
        tw, th = size
        tcam = N.camera_look((0.0, 0.0, 0.0), 0.0, 0.0, 60.0, tw, th)
        for f in ("eye", "forward", "right", "up"):
            setattr(tcam, f, getattr(self.cam, f))
        tcam.tan_half_fov = self.cam.tan_half_fov
        tcam.aspect = tw / th
        p0, pb = N.TraceParams.exact(0), N.TraceParams.exact(2, ignore_water=True)
        return self.ctx.batch([(tcam, tw, th, self.max_dist_m)], p0, pb, n_bounces=4,
                              bounce_max_dist_m=self.max_dist_m, seed=self.seed)

    def close(self):
        self.sharc.close()
        self.nrc.close()
        self.eval.close()
        self.train.close()

    # ---- judge -------------------------------------------------------------------------------
    def reference(self, spp: int, seed_base: int = 1 << 32, progress=None) -> float:
        """Adds `spp` full 4-bounce paths per pixel to the FP64 sum; returns GPU seconds."""
        t0 = time.perf_counter()
        for s in range(spp):
            self.eval.set_seed(seed_base + self.ref_spp)
            self.eval.trace(self.eval_hits.data_ptr())
            self.eval.shade(self.eval_hits.data_ptr(), self.shade, d_accum=self.ref.data_ptr())
            self.ref_spp += 1
            if progress and (s + 1) % 128 == 0:
                self.torch.cuda.synchronize()
                progress(s + 1)
        self.torch.cuda.synchronize()
        return time.perf_counter() - t0

    # ---- one training frame ----------------------------------------------------------------
    def train_paths(self, frame: int) -> None:
        self.train.set_seed(self.seed * 1000003 + frame)
        self.train.trace(self.train_hits.data_ptr())
        self.train.shade(self.train_hits.data_ptr(), self.shade, d_samples=self.train_samples.data_ptr())

    def train_sharc(self) -> None:
        self.sharc.deposit_paths(self.train_samples.data_ptr(), self.n_train, self.cam_pos)
        self.sharc.resolve()

    def train_nrc(self, steps: int = 1) -> None:
        per = (self.n_train + steps - 1) // steps
        for k in range(steps):
            n = min(per, self.n_train - k * per)
            self.nrc.train(self.train_samples.data_ptr() + k * per * 96, n)

    # ---- evaluation ------------------------------------------------------------------------
    def eval_paths(self, seed: int) -> None:
        self.eval.set_seed(seed)
        self.eval.trace(self.eval_hits.data_ptr())
        self.eval.shade(self.eval_hits.data_ptr(), self.shade, d_samples=self.eval_samples.data_ptr(),
                        d_radiance=self.pt.data_ptr())

    def query_sharc(self) -> None:
        self.sharc.query_paths(self.eval_samples.data_ptr(), self.n_eval, self.cam_pos, self.cache_out.data_ptr())

    def query_nrc(self) -> None:
        self.nrc.infer(self.eval_samples.data_ptr(), self.n_eval, self.cache_out.data_ptr())

    def rel_mse(self, img) -> float:
        torch = self.torch
        ref = (self.ref / max(self.ref_spp, 1)).view(-1, 3)
        s = self.eval_samples.view(torch.int32).view(-1, 24)
        valid = (s[:, 3] & N.PS_RAY) != 0
        e = img.view(-1, 3).double() - ref
        r = (e * e / (ref * ref + 0.01))[valid]
        return float(r.mean().item())

    def errors(self, seed: int) -> dict:
        """rel-MSE of: 1-spp path tracing, no cache, SHARC, neural cache (same paths)."""
        self.eval_paths(seed)
        out = {"pt_1spp": self.rel_mse(self.pt)}
        N.estimate(self.eval_samples.data_ptr(), self.n_eval, 0, 0, self.est.data_ptr())
        out["no_cache"] = self.rel_mse(self.est)
        self.query_sharc()
        N.estimate(self.eval_samples.data_ptr(), self.n_eval, self.cache_out.data_ptr(), 1, self.est.data_ptr())
        out["sharc"] = self.rel_mse(self.est)
        self.query_nrc()
        N.estimate(self.eval_samples.data_ptr(), self.n_eval, self.cache_out.data_ptr(), 1, self.est.data_ptr())
        out["nrc"] = self.rel_mse(self.est)
        return out

    def train_trace_ms(self, batch=None, reps: int = 5) -> float:
        """GPU ms to trace and shade one frame's training paths (median of `reps`)."""
        torch = self.torch
        b = batch or self.train
        hits = torch.empty(b.records() * 16, dtype=torch.uint8, device="cuda")
        samples = torch.empty(b.slots * 96, dtype=torch.uint8, device="cuda")
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        t = []
        for r in range(reps + 1):
            ev[0].record()
            b.trace(hits.data_ptr())
            b.shade(hits.data_ptr(), self.shade, d_samples=samples.data_ptr())
            ev[1].record()
            torch.cuda.synchronize()
            if r:
                t.append(ev[0].elapsed_time(ev[1]))
        return float(np.median(t))

    def sharc_at_budget(self, budget_ms: float, frames: int, sharc_log2: int, eval_seed: int,
                        trace_ms_per_path: float, sharc_ms_base: float) -> dict:
        """SHARC given a per-frame GPU budget: as many training paths as it can trace and shade in
        `budget_ms` beside its own work (deposit + resolve + 1080p query, `sharc_ms_base`), then
        `frames` online frames with that many paths, judged on the same evaluation paths. The
        measured per-frame ms (tracing, shading, deposit, resolve, query) is returned with it."""
        torch = self.torch

        def view_for(n):
            w = max(8, int(round(np.sqrt(n * 16.0 / 9.0) / 8.0)) * 8)
            return w, max(8, int(round(w * 9.0 / 16.0 / 8.0)) * 8)

        # tracing cost per path falls as the batch grows (more waves in flight): size the batch from
        # the small batch's cost, measure it, and rescale once with the measured cost
        n_paths = max(self.n_train, int((budget_ms - sharc_ms_base) / trace_ms_per_path))
        probe = self.train_batch(view_for(n_paths))
        t_probe = self.train_trace_ms(probe, reps=3)
        n_paths = max(self.n_train, int(probe.slots * (budget_ms - sharc_ms_base) / max(t_probe, 1e-6)))
        probe.close()
        tw, th = view_for(n_paths)
        b = self.train_batch((tw, th))
        hits = torch.empty(b.records() * 16, dtype=torch.uint8, device="cuda")
        samples = torch.empty(b.slots * 96, dtype=torch.uint8, device="cuda")
        sharc = N.Sharc(sharc_log2)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        t = []
        for f in range(frames):
            b.set_seed(self.seed * 1000003 + f)
            ev[0].record()
            b.trace(hits.data_ptr())
            b.shade(hits.data_ptr(), self.shade, d_samples=samples.data_ptr())
            sharc.deposit_paths(samples.data_ptr(), b.slots, self.cam_pos)
            sharc.resolve()
            sharc.query_paths(self.eval_samples.data_ptr(), self.n_eval, self.cam_pos, self.cache_out.data_ptr())
            ev[1].record()
            torch.cuda.synchronize()
            t.append(ev[0].elapsed_time(ev[1]))
        self.eval_paths(eval_seed)
        sharc.query_paths(self.eval_samples.data_ptr(), self.n_eval, self.cam_pos, self.cache_out.data_ptr())
        N.estimate(self.eval_samples.data_ptr(), self.n_eval, self.cache_out.data_ptr(), 1, self.est.data_ptr())
        err = self.rel_mse(self.est)
        sharc.close()
        b.close()
        return {"paths_per_frame": int(b.slots), "ms_per_frame": float(np.median(t[1:] if len(t) > 1 else t)),
                "rel_mse": err}

    def cache_ms(self, reps: int = 5, nrc_steps: int = 1) -> dict:
        """GPU ms per frame of each cache's own work (training-path tracing and evaluation tracing
        excluded, they are shared): SHARC deposit + resolve + 1080p query; neural cache training
        step(s) + 1080p inference."""
        torch = self.torch
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        sh, nn = [], []
        for _ in range(reps):
            ev[0].record()
            self.train_sharc()
            self.query_sharc()
            ev[1].record()
            self.train_nrc(nrc_steps)
            self.query_nrc()
            ev[2].record()
            torch.cuda.synchronize()
            sh.append(ev[0].elapsed_time(ev[1]))
            nn.append(ev[1].elapsed_time(ev[2]))
        return {"sharc_ms": float(np.median(sh)), "nrc_ms": float(np.median(nn))}
