"""Shadow-ray sets built from traced primary hits, in the shape the reference's lighting passes
build them (host numpy; inputs of the shadow any-hit walkers).

  sun_rays:     the GI pass's sun visibility ray (shaders/lighting_restir_gi.wgsl:134-159):
                origin = hit + 0.08 m along the face normal, direction = the sun jittered in its
                tangent plane by up to `jitter` (0.018 at high quality), max 88 m.
  toward_rays:  the DI visibility ray (:737-744): origin = hit + normal * min(0.08, 0.4 d),
                direction towards a target point, max_dist = max(0.02, d - 0.08); here the
                targets stand in for emitter samples (the reference samples its emissive pool).
"""
from __future__ import annotations

import numpy as np

from .native import RAY_DTYPE


def hit_points(prays: np.ndarray, hits: np.ndarray):
    """(index of each hit ray, hit point in metres, unit face normal) of the compact hits."""
    idx = np.nonzero((hits["flags"] & 1) != 0)[0]
    h, pr = hits[idx], prays[idx]
    o = np.stack([pr["ox"], pr["oy"], pr["oz"]], 1).astype(np.float64)
    d = np.stack([pr["dx"], pr["dy"], pr["dz"]], 1).astype(np.float64)
    p = o + d * (h["t_voxels"].astype(np.float64)[:, None] / 16.0)
    n = np.zeros_like(p)
    face = h["face"].astype(np.int64)
    ok = face < 6
    n[np.nonzero(ok)[0], face[ok] // 2] = np.where(face[ok] % 2 == 0, 1.0, -1.0)
    return idx, p, n


def _pack(org, dirs, tmax):
    r = np.zeros(len(org), RAY_DTYPE)
    r["ox"], r["oy"], r["oz"] = org[:, 0], org[:, 1], org[:, 2]
    r["dx"], r["dy"], r["dz"] = dirs[:, 0], dirs[:, 1], dirs[:, 2]
    r["t_max"] = tmax
    return r


def sun_rays(prays: np.ndarray, hits: np.ndarray, sun, seed: int = 0, max_dist_m: float = 88.0,
             jitter: float = 0.018) -> np.ndarray:
    _, p, n = hit_points(prays, hits)
    s = np.asarray(sun, np.float64)
    s = s / np.linalg.norm(s)
    up = np.array([0.0, 1.0, 0.0]) if abs(s[1]) < 0.9 else np.array([1.0, 0.0, 0.0])
    tan = np.cross(up, s)
    tan /= np.linalg.norm(tan)
    bit = np.cross(s, tan)
    j = np.random.default_rng(seed).uniform(-1.0, 1.0, (len(p), 2)) * jitter
    dirs = s + tan * j[:, :1] + bit * j[:, 1:]
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    return _pack(p + n * 0.08, dirs, max_dist_m)


def toward_rays(prays: np.ndarray, hits: np.ndarray, seed: int = 0) -> np.ndarray:
    """One ray per hit towards the hit point of another, randomly chosen, hit pixel."""
    _, p, n = hit_points(prays, hits)
    tgt = p[np.random.default_rng(seed).integers(0, len(p), len(p))]
    d_init = np.linalg.norm(tgt - p, axis=1)
    org = p + n * np.minimum(0.08, d_init * 0.4)[:, None]
    to = tgt - org
    dist = np.linalg.norm(to, axis=1)
    dirs = np.where((dist > 1e-4)[:, None], to / np.maximum(dist, 1e-4)[:, None], np.array([0.0, 1.0, 0.0]))
    return _pack(org, dirs, np.maximum(0.02, dist - 0.08))
