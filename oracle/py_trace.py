"""Second, independent restatement of trace_world_64tree_from in pure Python — TEST
INFRASTRUCTURE ONLY.

Follows shaders/common/traversal64.wgsl:226-556 line by line with numpy.float32 scalars (every
operation rounded to binary32, no fused multiply-add), written without looking at the C
oracle's code. Used on small worlds to cross-check the C oracle bit for bit, including the
DDA step count, which no reference test pins. Slow: a few hundred rays at a time.

`mutate` plants a deliberate mistake (for the tests that show the cross-check is sensitive):
  "eps0"       face tie-break epsilon forced to zero (traversal64.wgsl:480)
  "steps+1"    step counter off by one
  "no_restart" every iteration restarts at the chunk root (results identical: the ancestor
               cache is a pure optimisation; only node-load counts change)
"""
from __future__ import annotations

import numpy as np

f32 = np.float32
EMPTY = 0xFFFFFFFF
LEAF = 0x80000000
UNIFORM = 0x40000000
PTR = 0x3FFFFFFF
WATER = 18
M32 = 0xFFFFFFFF


def _fmax(a, b):
    return a if a > b else b


def _fmin(a, b):
    return a if a < b else b


def _u32(x: np.float32) -> int:
    if not (x > 0):
        return 0
    if x >= f32(4294967296.0):
        return M32
    return int(x)


def _sign(x):
    return f32(1.0) if x > 0 else (f32(-1.0) if x < 0 else f32(0.0))


def _node(nodes, i):
    n = nodes[i]
    return int(n["mask_low"]), int(n["mask_high"]), int(n["child_ptr"]), int(n["lod_payload"])


def _has(node, b):
    w = node[0] if b < 32 else node[1]
    return (w >> (b & 31)) & 1 == 1


def _child(node, b):
    mask = node[0] | (node[1] << 32)
    return (node[2] & PTR) + bin(mask & ((1 << b) - 1)).count("1")


def trace(grid, roots, nodes, bricks, params, ray, mutate: str | None = None):
    """Returns dict(hit, t, vx, vy, vz, face, mat, steps, node_loads)."""
    o = [f32(ray["ox"]), f32(ray["oy"]), f32(ray["oz"])]
    d = [f32(ray["dx"]), f32(ray["dy"]), f32(ray["dz"])]
    start_t_m, max_dist_m = f32(ray["t_start"]), f32(ray["t_max"])
    lod_bias = int(params.lod_bias)
    ignore_water = bool(params.ignore_water)
    out = dict(hit=0, t=f32(max_dist_m * f32(16.0)), vx=0, vy=0, vz=0, face=255, mat=0, steps=0, node_loads=0)
    vo = [o[a] * f32(16.0) for a in range(3)]
    max_t_v = max_dist_m * f32(16.0)
    start_t_v = _fmax(start_t_m * f32(16.0), f32(0.0))
    safe = [f32(1e-7) if abs(d[a]) < f32(1e-7) else d[a] for a in range(3)]
    inv = [f32(1.0) / safe[a] for a in range(3)]
    wmax = [f32(grid[a]) * f32(256.0) for a in range(3)]
    tmin_v = [((f32(0.0) if inv[a] > 0 else wmax[a]) - vo[a]) * inv[a] for a in range(3)]
    tmax_v = [((wmax[a] if inv[a] > 0 else f32(0.0)) - vo[a]) * inv[a] for a in range(3)]
    wx = _fmax(_fmax(tmin_v[0], tmin_v[1]), tmin_v[2])
    wy = _fmin(_fmin(tmax_v[0], tmax_v[1]), tmax_v[2])
    t_enter = _fmax(wx, start_t_v)
    if wy < t_enter or t_enter >= max_t_v:
        return out
    t = t_enter
    t_limit = _fmin(wy, max_t_v)
    cls = 0 if lod_bias == 0 else (1 if lod_bias == 1 else 2)
    max_steps = int(params.max_steps[cls])
    lod_t = [f32(params.lod_dist_m[k]) * f32(16.0) for k in range(3)]
    lod_s = [int(params.lod_step[k]) for k in range(3)]
    pos = [inv[a] >= 0 for a in range(3)]
    noi = [(-vo[a]) * inv[a] for a in range(3)]
    wmaxi = [grid[a] << 8 for a in range(3)]
    vi = [_u32(_fmin(_fmax(vo[a] + d[a] * (t + f32(1e-4)), f32(0.0)), wmax[a] - f32(1e-3))) for a in range(3)]
    mtf = [(None, EMPTY), (None, EMPTY)]
    prev = [M32, M32, M32]
    cached_shift = 6
    cache = {4: None, 2: None, 0: None}
    loads = 0

    def step_box(bmin, bmax):
        exitc = [bmax[a] if pos[a] else bmin[a] for a in range(3)]
        tm = [f32(exitc[a]) * inv[a] + noi[a] for a in range(3)]
        t_exit = _fmin(_fmin(tm[0], tm[1]), tm[2])
        nt = _fmax(t, t_exit)
        is_x = tm[0] <= tm[1] and tm[0] <= tm[2]
        is_y = (not is_x) and tm[1] <= tm[2]
        ex = [is_x, is_y, (not is_x) and (not is_y)]
        nv = []
        for a in range(3):
            outside = (exitc[a] - (0 if pos[a] else 1)) & M32
            ip = _fmax(vo[a] + d[a] * nt, f32(0.0))
            inside = min(max(_u32(ip), bmin[a]), bmax[a] - 1)
            nv.append(outside if ex[a] else inside)
        return nt, nv

    step = 0
    while step < max_steps:
        if t >= t_limit:
            break
        if any(vi[a] >= wmaxi[a] for a in range(3)):
            break
        steps_done = step + 1
        diff = (prev[0] ^ vi[0]) | (prev[1] ^ vi[1]) | (prev[2] ^ vi[2])
        prev = list(vi)
        root = mtf[0][1]
        if diff >= 256:
            c = (vi[0] >> 8, vi[1] >> 8, vi[2] >> 8)
            cached_shift = 6
            if mtf[1][0] == c:
                root = mtf[1][1]
                mtf = [(c, root), mtf[0]]
            else:
                root = int(roots[c[0] + c[1] * grid[0] + c[2] * grid[0] * grid[1]])
                mtf = [(c, root), mtf[0]]
        if root == EMPTY:
            base = [v & ~255 & M32 for v in vi]
            t, vi = step_box(base, [b + 256 for b in base])
            step += 1
            continue
        min_shift = 0
        if lod_bias == 0:
            if t > lod_t[0] or step > lod_s[0]:
                min_shift = 2
        elif lod_bias == 1:
            if t > lod_t[1] * f32(2.2857) or step > lod_s[1]:
                min_shift = 4
            elif t > lod_t[1]:
                min_shift = 2
        else:
            if t > lod_t[2] or step > lod_s[2]:
                min_shift = 2
        if mutate != "no_restart" and diff < 4 and min_shift == 0 and cached_shift == 0:
            node, shift = cache[0], 0
        elif mutate != "no_restart" and diff < 16 and min_shift <= 2 and cached_shift <= 2:
            node, shift = cache[2], 2
        elif mutate != "no_restart" and diff < 64 and min_shift <= 4 and cached_shift <= 4:
            node, shift = cache[4], 4
        else:
            node, shift = _node(nodes, root), 6
            loads += 1
        cached_shift = shift
        loc = [v & 255 for v in vi]
        while True:
            b = (((loc[2] >> shift) & 3) << 4) | (((loc[1] >> shift) & 3) << 2) | ((loc[0] >> shift) & 3)
            if ignore_water and (node[2] & UNIFORM) and ((node[3] >> 24) & 0xFF) == WATER:
                span = 4 << shift
                base = [v & ~(span - 1) & M32 for v in vi]
                t, vi = step_box(base, [x + span for x in base])
                break
            span = 1 << shift
            cmin = [v & ~(span - 1) & M32 for v in vi]
            cmax = [x + span for x in cmin]
            if not _has(node, b):
                t, vi = step_box(cmin, cmax)
                break
            is_leaf = (node[2] & LEAF) != 0
            if is_leaf or shift <= min_shift:
                mat = (node[3] >> 24) & 0xFF
                if is_leaf and not (node[2] & UNIFORM):
                    word = int(bricks[node[2] & PTR]["packed_materials"][b >> 2])
                    mat = (word >> ((b & 3) * 8)) & 0xFF
                if ignore_water and mat == WATER:
                    t, vi = step_box(cmin, cmax)
                    break
                if mat != 0:
                    tm = [f32(cmin[a] if pos[a] else cmax[a]) * inv[a] + noi[a] for a in range(3)]
                    ce = _fmax(_fmax(tm[0], tm[1]), tm[2])
                    exact = _fmax(ce, f32(0.0))
                    eps = f32(1e-4) * _fmax(ce * f32(0.01), f32(1.0))
                    if mutate == "eps0":
                        eps = f32(0.0)
                    cand = [abs(d[a]) if abs(tm[a] - ce) <= eps else f32(0.0) for a in range(3)]
                    if cand[0] >= cand[1] and cand[0] >= cand[2] and cand[0] > 0:
                        axis = 0
                    elif cand[1] >= cand[2] and cand[1] > 0:
                        axis = 1
                    elif cand[2] > 0:
                        axis = 2
                    elif tm[0] >= tm[1] and tm[0] >= tm[2]:
                        axis = 0
                    elif tm[1] >= tm[2]:
                        axis = 1
                    else:
                        axis = 2
                    nval = -_sign(d[axis])
                    face = 2 * axis if nval > 0 else (2 * axis + 1 if nval < 0 else 6 + axis)
                    out.update(hit=1, t=exact, vx=vi[0], vy=vi[1], vz=vi[2], face=face, mat=mat,
                               steps=steps_done + (1 if mutate == "steps+1" else 0), node_loads=loads)
                    return out
                t, vi = step_box(cmin, cmax)
                break
            node = _node(nodes, _child(node, b))
            loads += 1
            shift -= 2
            cached_shift = shift
            cache[shift] = node
        step += 1
    out.update(steps=step + (1 if mutate == "steps+1" else 0), node_loads=loads)
    return out
