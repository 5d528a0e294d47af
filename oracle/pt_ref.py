"""Binary32 restatement of the path shading (radiance, path samples) — TEST INFRASTRUCTURE ONLY.

Given the records of every depth of a path launch (traced by the oracle on the host-generated
rays), walks each slot's path in numpy float32 with the device's operation order: Lambertian
palette albedo, palette emission x emission_scale, escaped rays see eval_sky_radiance
(shaders/common/types.wgsl:500-523) x sky_scale with pow(., 3) and pow(., 14) as products and
mix(a, b, t) = a (1 - t) + b t.
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def mix(a, b, t):
    return (a * (f32(1.0) - t)).astype(f32) + (b * t).astype(f32)


def sky(d, sun):
    """eval_sky_radiance for directions d [n, 3] (float32)."""
    d = d.astype(f32)
    se = f32(sun[1])
    day_t = f32(min(max((se + f32(0.05)) / f32(0.30), f32(0.0)), f32(1.0)))
    twi_t = f32(min(max(f32(1.0) - abs(se - f32(0.02)) / f32(0.22), f32(0.0)), f32(1.0)))
    zen = np.clip(d[:, 1], f32(0.0), f32(1.0)).astype(f32)
    om = (f32(1.0) - zen).astype(f32)
    hor = ((om * om).astype(f32) * om).astype(f32)
    dz, dh = np.array([0.08, 0.22, 0.55], f32), np.array([0.38, 0.50, 0.68], f32)
    nz, nh = np.array([0.0015, 0.0025, 0.0055], f32), np.array([0.0035, 0.0065, 0.0120], f32)
    tw, sc = np.array([0.65, 0.25, 0.10], f32), np.array([0.85, 0.52, 0.24], f32)
    s3 = np.asarray(sun, f32)
    dp = ((d[:, 0] * s3[0]).astype(f32) + (d[:, 1] * s3[1]).astype(f32)).astype(f32) + (d[:, 2] * s3[2]).astype(f32)
    sc0 = np.maximum(dp.astype(f32), f32(0.0))
    x2 = (sc0 * sc0).astype(f32)
    x4 = (x2 * x2).astype(f32)
    x8 = (x4 * x4).astype(f32)
    p14 = ((x8 * x4).astype(f32) * x2).astype(f32)
    tw_w = f32(twi_t * f32(0.75))
    sun_w = f32(max(day_t, twi_t))
    out = np.zeros((len(d), 3), f32)
    for c in range(3):
        skz = mix(nz[c], dz[c], day_t)
        skh = mix(nh[c], dh[c], day_t)
        skh = mix(skh, tw[c], tw_w)
        s = mix(np.full(len(d), skz, f32), np.full(len(d), skh, f32), hor)
        if se > f32(-0.12):
            add = (((sc[c] * p14).astype(f32) * f32(0.35)).astype(f32) * sun_w).astype(f32)
            s = (s + add).astype(f32)
        out[:, c] = s
    return out


def shade(rays_by_depth, hits_by_depth, palette, sun, emission_scale=1.0, sky_scale=1.0):
    """(radiance [n, 3], samples dict) for the path of every slot; rays_by_depth[d] is the ray
    traced at depth d (numpy RAY records), hits_by_depth[d] its compact hit."""
    n = len(rays_by_depth[0])
    D = len(rays_by_depth) - 1
    pal = np.asarray(palette, f32)
    L = np.zeros((n, 3), f32)
    T = np.ones((n, 3), f32)
    T1 = np.ones((n, 3), f32)
    out = {k: np.zeros((n, 3), f32) for k in ("pos", "dir", "le0", "a0", "base", "target")}
    flags = np.zeros(n, np.uint32)
    alive = rays_by_depth[0]["t_max"] > 0
    flags[alive] |= 1
    es, ss = f32(emission_scale), f32(sky_scale)
    for d in range(D + 1):
        r, h = rays_by_depth[d], hits_by_depth[d]
        dirs = np.stack([r["dx"], r["dy"], r["dz"]], 1).astype(f32)
        hit = alive & ((h["flags"] & 1) != 0)
        miss = alive & ~hit
        if miss.any():
            s = (sky(dirs[miss], sun) * ss).astype(f32)
            L[miss] = (L[miss] + (T[miss] * s).astype(f32)).astype(f32)
            idx = np.nonzero(miss)[0]
            if d == 0:
                out["le0"][idx] = s
            elif d == 1:
                out["base"][idx] = s
            else:
                out["target"][idx] = (out["target"][idx] + (T1[idx] * s).astype(f32)).astype(f32)
        mat = h["mat"].astype(np.int64)
        le = (pal[mat, 4:7] * es).astype(f32)
        alb = pal[mat, 0:3]
        idx = np.nonzero(hit)[0]
        L[idx] = (L[idx] + (T[idx] * le[idx]).astype(f32)).astype(f32)
        if d == 0:
            flags[idx] |= 2
            out["le0"][idx] = le[idx]
            out["a0"][idx] = alb[idx]
        elif d == 1:
            flags[idx] |= 4 | (h["face"][idx].astype(np.uint32) << 8)
            ovox = (np.stack([r["ox"], r["oy"], r["oz"]], 1).astype(f32) * f32(16.0)).astype(f32)
            t = h["t_voxels"].astype(f32)[:, None]
            pv = (ovox + (dirs * t).astype(f32)).astype(f32)
            out["pos"][idx] = (pv * f32(0.0625)).astype(f32)[idx]
            out["base"][idx] = le[idx]
        else:
            out["target"][idx] = (out["target"][idx] + (T1[idx] * le[idx]).astype(f32)).astype(f32)
        if d == D:
            break
        if d >= 1:
            T1[idx] = (T1[idx] * alb[idx]).astype(f32)
        T[idx] = (T[idx] * alb[idx]).astype(f32)
        if d == 0:
            nr = rays_by_depth[1]
            out["dir"][idx] = np.stack([nr["dx"], nr["dy"], nr["dz"]], 1).astype(f32)[idx]
        alive = hit
    return L, out, flags
