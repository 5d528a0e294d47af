"""Float64 restatement of the neural radiance cache's forward and backward pass — TEST
INFRASTRUCTURE ONLY (the checker of the device kernels; no product path imports it).

It follows the device model exactly: 16-level hash grid (2 features per level, trilinear, direct
indexing while (res+1)^3 fits the table, else the spatial hash x ^ 2654435761 y ^ 805459861 z),
input = 32 encoding features + normal + view direction + 1, padded to 48; 3 x 64 ReLU;
3 linear outputs; relative L2 loss mean_c (y_c - t_c)^2 / (y_c^2 + 0.01) with the denominator held
constant. The values the kernels store in FP16 (table entries, weights, inputs, activations,
deltas) are rounded to FP16 at the same points (`f16=True`), so the comparison isolates the FP32
accumulation order; `f16=False` gives the exact float64 function for finite-difference checks.
"""
from __future__ import annotations

import numpy as np

LEVELS, IN, HID, OUT = 16, 48, 64, 16
OFF_W1, OFF_W2 = 0, HID * IN
OFF_W3 = OFF_W2 + HID * HID
OFF_W4 = OFF_W3 + HID * HID
MLP_PARAMS = OFF_W4 + OUT * HID


def r16(x, on=True):
    return x.astype(np.float16).astype(np.float64) if on else x


def unpack_mlp(mlp):
    m = np.asarray(mlp, np.float64)
    return (m[OFF_W1:OFF_W2].reshape(HID, IN), m[OFF_W2:OFF_W3].reshape(HID, HID), m[OFF_W3:OFF_W4].reshape(HID, HID),
            m[OFF_W4:MLP_PARAMS].reshape(OUT, HID))


def corner_index(level, x, y, z, scale, dense, table_mask):
    if dense[level]:
        r = np.uint64(int(scale[level]) + 1)
        return ((x + y * r + z * r * r) & np.uint64(table_mask)).astype(np.int64)
    x32, y32, z32 = (v.astype(np.uint32) for v in (x, y, z))
    h = x32 ^ (y32 * np.uint32(2654435761)) ^ (z32 * np.uint32(805459861))
    return (h & np.uint32(table_mask)).astype(np.int64)


def features(samples, bbox_min, bbox_size):
    """Normalized positions, normals, directions and validity of arc_path_sample records."""
    pos = (samples["pos"].astype(np.float32) - np.asarray(bbox_min, np.float32)) * np.float32(1.0 / bbox_size)
    p = np.clip(pos, 0.0, np.float32(0.99999994)).astype(np.float32)
    face = (samples["flags"] >> 8) & 0xFF
    n = np.zeros((len(samples), 3))
    ok = face < 6
    n[np.nonzero(ok)[0], face[ok] >> 1] = np.where(face[ok] & 1, -1.0, 1.0)
    valid = (samples["flags"] & 4) != 0
    return p, n, samples["dir"].astype(np.float64), valid


def encode(p, table, scale, dense, f16=True):
    """Per sample 32 features and, for the backward pass, the corner indices / weights."""
    n = len(p)
    T = table.shape[1]
    out = np.zeros((n, 2 * LEVELS))
    corners = []
    for level in range(LEVELS):
        x = p.astype(np.float32) * np.float32(scale[level])
        fl = np.floor(x)
        c = fl.astype(np.uint64)
        f = (x - fl).astype(np.float32)
        lev = []
        for k in range(8):
            d = np.array([k & 1, (k >> 1) & 1, k >> 2])
            w = np.ones(n)
            for a in range(3):
                w = w * np.where(d[a], f[:, a], 1.0 - f[:, a])
            idx = corner_index(level, c[:, 0] + np.uint64(d[0]), c[:, 1] + np.uint64(d[1]), c[:, 2] + np.uint64(d[2]),
                               scale, dense, T - 1)
            ent = r16(table[level, idx], f16)
            out[:, 2 * level] += w * ent[:, 0]
            out[:, 2 * level + 1] += w * ent[:, 1]
            lev.append((idx, w))
        corners.append(lev)
    return out, corners


def forward_backward(samples, grid, mlp, scale, dense, bbox_min, bbox_size, f16=True):
    """(outputs y [n,3], summed loss, mlp gradient [MLP_PARAMS], table gradient [16, T, 2])."""
    table = np.asarray(grid, np.float64).reshape(LEVELS, -1, 2)
    W1, W2, W3, W4 = (r16(w, f16) for w in unpack_mlp(mlp))
    p, nrm, d, valid = features(samples, bbox_min, bbox_size)
    enc, corners = encode(p, table, scale, dense, f16)
    n = len(samples)
    X = np.zeros((n, IN))
    X[:, :32] = enc
    X[:, 32:35] = nrm
    X[:, 35:38] = d
    X[:, 38] = 1.0
    X = r16(X, f16)
    a1 = r16(np.maximum(X @ W1.T, 0.0), f16)
    a2 = r16(np.maximum(a1 @ W2.T, 0.0), f16)
    a3 = r16(np.maximum(a2 @ W3.T, 0.0), f16)
    y = a3 @ W4.T
    t = samples["target"].astype(np.float64)
    e = y[:, :3] - t
    den = y[:, :3] ** 2 + 0.01
    loss = float(np.sum((e * e / den)[valid])) / 3.0
    dy = np.zeros((n, OUT))
    dy[:, :3] = np.where(valid[:, None], 2.0 * e / den / 3.0, 0.0)
    dy = r16(dy, f16)
    d3 = r16((dy @ W4) * (a3 > 0), f16)
    d2 = r16((d3 @ W3) * (a2 > 0), f16)
    d1 = r16((d2 @ W2) * (a1 > 0), f16)
    dX = d1 @ W1
    g = np.zeros(MLP_PARAMS)
    g[OFF_W1:OFF_W2] = (d1.T @ X).ravel()
    g[OFF_W2:OFF_W3] = (d2.T @ a1).ravel()
    g[OFF_W3:OFF_W4] = (d3.T @ a2).ravel()
    g[OFF_W4:MLP_PARAMS] = (dy.T @ a3).ravel()
    tg = np.zeros_like(table)
    for level in range(LEVELS):
        for idx, w in corners[level]:
            for f in range(2):
                np.add.at(tg[level, :, f], idx[valid], (w * dX[:, 2 * level + f])[valid])
    return y[:, :3], loss, g, tg


def outputs(samples, grid, mlp, scale, dense, bbox_min, bbox_size, f16=True):
    """Forward only: y [n, 3]."""
    return forward_backward(samples, grid, mlp, scale, dense, bbox_min, bbox_size, f16)[0]


def loss_fixed_den(samples, grid, mlp, scale, dense, bbox_min, bbox_size, den):
    """sum_valid mean_c (y_c - t_c)^2 / den_c in float64 (den given): the function whose exact
    gradient forward_backward(..., f16=False) returns when den = y^2 + 0.01 at the same point."""
    y = outputs(samples, grid, mlp, scale, dense, bbox_min, bbox_size, f16=False)
    p, n, d, valid = features(samples, bbox_min, bbox_size)
    e = y - samples["target"].astype(np.float64)
    return float(np.sum((e * e / den)[valid])) / 3.0
