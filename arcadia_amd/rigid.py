"""Host helpers for rigid voxel bodies: place a body's 64-tree in the world pools and build its
GPU instance record.

  append_body_tree: appends the packed tree of a one-chunk HostWorld to the world's node and
                    brick pools, relocating child and brick pointers (the relocation of
                    world_map.rs:705-735, applied to a body), and returns the body's root index.
  body_instance:    the RigidBodyInstance of a body (transform.rs:80-112): local_to_world =
                    [R | pos - R com], world_to_local its inverse, local bounds = dims / 16 m,
                    bounding sphere at pos + R (half_extents - com) with radius
                    |half_extents| + 1/16. Computed in float64 and stored as float32: these
                    are inputs of the walker, not part of its arithmetic.
"""
from __future__ import annotations

import numpy as np

from .native import NODE_DTYPE, RIGID_BODY_DTYPE

LEAF, UNIFORM, PTR_MASK = 0x80000000, 0x40000000, 0x3FFFFFFF


def append_body_tree(pools, body_pools):
    """(world pools with the body's tree appended, root node index of the body)."""
    grid, roots, nodes, bricks = pools
    _, broots, bnodes, bbricks = body_pools
    n0, b0 = len(nodes), len(bricks)
    # the body's pools start with the slot-0 dummies: body node k lands at n0 + k - 1
    bn = np.array(bnodes[1:], NODE_DTYPE, copy=True)
    cp = bn["child_ptr"].astype(np.uint64)
    leaf, uni = (cp & LEAF) != 0, (cp & UNIFORM) != 0
    ptr = cp & PTR_MASK
    inner = ~leaf & ((bn["mask_low"] != 0) | (bn["mask_high"] != 0))
    ptr = np.where(inner, ptr + n0 - 1, ptr)
    ptr = np.where(leaf & ~uni, ptr + b0 - 1, ptr)
    bn["child_ptr"] = ((cp & ~np.uint64(PTR_MASK)) | ptr).astype(np.uint32)
    root = int(broots[0])
    assert root != 0xFFFFFFFF, "empty body"
    new = (grid, roots, np.concatenate([nodes, bn]), np.concatenate([bricks, bbricks[1:]]))
    return new, n0 + root - 1


def rot_y(deg: float) -> np.ndarray:
    a = np.deg2rad(deg)
    c, s = np.cos(a), np.sin(a)
    if abs(deg % 90.0) < 1e-12:  # exact for quarter turns
        c, s = float(round(c)), float(round(s))
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def _l2w(r, pos, com):
    m = np.eye(4)
    m[:3, :3] = r
    m[:3, 3] = np.asarray(pos, np.float64) - r @ np.asarray(com, np.float64)
    return m


def body_instance(root: int, dims_v, rot, pos_m, com_m=None, prev_rot=None, prev_pos_m=None) -> np.ndarray:
    dims = np.asarray(dims_v, np.float64)
    half = dims * (0.5 / 16.0)
    com = half if com_m is None else np.asarray(com_m, np.float64)
    r = np.asarray(rot, np.float64)
    l2w = _l2w(r, pos_m, com)
    w2l = np.eye(4)
    w2l[:3, :3] = r.T
    w2l[:3, 3] = -r.T @ l2w[:3, 3]
    prev = _l2w(np.asarray(prev_rot if prev_rot is not None else r, np.float64),
                prev_pos_m if prev_pos_m is not None else pos_m, com)
    inst = np.zeros(1, RIGID_BODY_DTYPE)
    inst["world_to_local"] = w2l.T.reshape(-1)  # column-major
    inst["local_to_world"] = l2w.T.reshape(-1)
    inst["prev_local_to_world"] = prev.T.reshape(-1)
    lb = np.zeros(4, np.float32)
    lb[:3] = dims / 16.0
    lb[3] = np.array([root], np.uint32).view(np.float32)[0]
    inst["local_bounds_and_root"] = lb
    c = np.asarray(pos_m, np.float64) + r @ (half - com)
    inst["world_sphere"] = [c[0], c[1], c[2], np.linalg.norm(half) + 1.0 / 16.0]
    return inst[0]
