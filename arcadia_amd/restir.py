"""Host helpers for the ReSTIR GI / DI passes: the inputs a caller's emissive registry and material
table provide (arc_restir_temporal / arc_restir_spatial take them as device buffers).

  palette():            the 256-entry material table (albedo rgb, roughness, emission rgb,
                        variance) as one (256, 8) float32 array (material.rs:99-204).
  pack_emitter_normal:  EmissiveSource.normal_packed as di_unpack_normal_and_eta (di.wgsl:42-53)
                        reads it: the normal scaled by the emitter's directionality eta, each
                        component (v + 1) / 2 quantised to 10 bits.
  alias_table:          a Walker / Vose alias table over per-source weights (flux), the sampling
                        table emissive_alias the DI candidates read (lighting_restir_gi.wgsl:384-389).
  emissive_sources:     EMISSIVE_DTYPE records from per-source arrays.
"""
from __future__ import annotations

import numpy as np

from .native import ALIAS_DTYPE, EMISSIVE_DTYPE, palette_material


def palette() -> np.ndarray:
    return np.stack([palette_material(m) for m in range(256)]).astype(np.float32)


def pack_emitter_normal(n, eta: float = 1.0) -> int:
    v = np.asarray(n, np.float64)
    v = v / np.linalg.norm(v) * float(eta)
    q = np.clip(np.rint((v + 1.0) * 0.5 * 1023.0), 0, 1023).astype(np.uint32)
    return int(q[0] | (q[1] << 10) | (q[2] << 20))


def alias_table(weights) -> np.ndarray:
    """Vose's alias method: column i is kept with probability prob[i], else alias_idx[i] is taken."""
    w = np.asarray(weights, np.float64)
    n = len(w)
    out = np.zeros(n, ALIAS_DTYPE)
    if n == 0:
        return out
    scaled = w * (n / w.sum())
    prob = np.ones(n)
    alias = np.arange(n, dtype=np.uint32)
    small = [i for i in range(n) if scaled[i] < 1.0]
    large = [i for i in range(n) if scaled[i] >= 1.0]
    while small and large:
        s, g = small.pop(), large.pop()
        prob[s] = scaled[s]
        alias[s] = g
        scaled[g] = (scaled[g] + scaled[s]) - 1.0
        (small if scaled[g] < 1.0 else large).append(g)
    for i in small + large:
        prob[i] = 1.0
    out["prob"] = prob.astype(np.float32)
    out["alias_idx"] = alias
    return out


def emissive_sources(pos, radius, intensity, normals=None, eta=None) -> np.ndarray:
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    n = len(pos)
    s = np.zeros(n, EMISSIVE_DTYPE)
    s["pos"] = pos
    s["radius"] = np.broadcast_to(np.asarray(radius, np.float32), (n,))
    s["intensity"] = np.broadcast_to(np.asarray(intensity, np.float32), (n, 3))
    nrm = np.broadcast_to(np.asarray((0.0, 1.0, 0.0) if normals is None else normals, np.float64), (n, 3))
    et = np.broadcast_to(np.asarray(1.0 if eta is None else eta, np.float64), (n,))
    s["normal_packed"] = [pack_emitter_normal(nrm[i], et[i]) for i in range(n)]
    return s


def view_proj(cam) -> np.ndarray:
    """Column-major 4x4 (glam to_cols_array order) whose clip.xy / clip.w is the NDC of the camera's
    pinhole model: the prev_view_proj that gives zero motion for a camera that did not move."""
    r, u, f, e = (np.array(list(getattr(cam, k)), np.float64) for k in ("right", "up", "forward", "eye"))
    m = np.zeros((4, 4))
    m[0, :3] = r / (cam.aspect * cam.tan_half_fov)
    m[0, 3] = -r @ e / (cam.aspect * cam.tan_half_fov)
    m[1, :3] = u / cam.tan_half_fov
    m[1, 3] = -u @ e / cam.tan_half_fov
    m[3, :3] = f
    m[3, 3] = -f @ e
    return m.T.reshape(-1).astype(np.float32)
