"""Seeded worlds and ray sets shared by the tests (small enough for the oracle)."""
from __future__ import annotations

import functools

import numpy as np

import arcadia_amd as A
from oracle import oracle as O


@functools.lru_cache(maxsize=None)
def heightmap_world(seed: int = 1):
    """256^3 heightmap world: heights 32..200, mats 1..15 per voxel."""
    w = A.HostWorld((1, 1, 1))
    w.generate(A.gen_params(A.ARC_GEN_HEIGHTMAP, seed=seed))
    w.pack()
    return w.pools()


@functools.lru_cache(maxsize=None)
def caves_world(seed: int = 1, grid=(2, 2, 2), n_cuboids: int = 120):
    """Multi-chunk 3D-fBm cave world with random cuboids (the full-frame world at reduced size)."""
    w = A.HostWorld(grid)
    p = A.gen_params(A.ARC_GEN_CAVES, seed=seed, n_cuboids=n_cuboids,
                     height_min=96.0, height_max=352.0, wavelength=192.0)
    w.generate(p)
    w.pack()
    return w.pools()


def sparse_world(voxels, grid=(1, 1, 1), chunk=0):
    w = A.HostWorld(grid)
    w.set_chunk_sparse(chunk, voxels)
    w.pack()
    return w.pools()


def poses_rays(eyes_yaw_pitch, width, height, fov=90.0, max_dist=250.0):
    out = []
    for eye, yaw, pitch in eyes_yaw_pitch:
        cam = A.camera_look(eye, yaw, pitch, fov, width, height)
        r, _ = A.primary_rays(cam, width, height, max_dist)
        out.append(r)
    return np.concatenate(out)


def c1_poses(seed: int = 1):
    """4 seeded poses above the 256^3 terrain: eye y in [13, 15.5] m, pitch -20..-60 deg."""
    rng = np.random.default_rng(seed)
    poses = []
    for _ in range(4):
        eye = (float(rng.uniform(2.0, 14.0)), float(rng.uniform(13.0, 15.5)), float(rng.uniform(2.0, 14.0)))
        yaw = float(rng.uniform(0, 2 * np.pi))
        pitch = float(np.deg2rad(rng.uniform(-60.0, -20.0)))
        poses.append((eye, yaw, pitch))
    return poses


def random_rays(n, grid, seed=0, inside=True, max_dist=250.0):
    """Rays with random origins (inside the world box or around it) and random unit directions."""
    rng = np.random.default_rng(seed)
    ext = np.array(grid, np.float64) * 16.0
    if inside:
        o = rng.uniform(0.0, 1.0, (n, 3)) * ext
    else:
        o = rng.uniform(-0.5, 1.5, (n, 3)) * ext
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros(n, A.RAY_DTYPE)
    r["ox"], r["oy"], r["oz"] = o[:, 0], o[:, 1], o[:, 2]
    r["dx"], r["dy"], r["dz"] = d[:, 0], d[:, 1], d[:, 2]
    r["t_max"] = max_dist
    return r


def oracle_hits(pools, params, rays):
    full, cnt = O.trace_world(O.World(*pools), params, rays)
    return O.compact(full), full, cnt


def diff_records(a: np.ndarray, b: np.ndarray):
    wa = a.view(np.uint32).reshape(len(a), -1)
    wb = b.view(np.uint32).reshape(len(b), -1)
    return np.nonzero((wa != wb).any(1))[0]


def water_world(with_water: bool):
    """The 2x2x2 cave world with a lake: a full 32x16x32 water block (its 16^3 cells collapse to
    uniform water nodes), a thin water layer that leaves mixed water bricks, and water voxels
    scattered through the caves."""
    w = A.HostWorld((2, 2, 2))
    p = A.gen_params(A.ARC_GEN_CAVES, seed=1, n_cuboids=120, height_min=96.0, height_max=352.0, wavelength=192.0)
    w.generate(p)
    if with_water:
        z, y, x = np.mgrid[0:32, 0:16, 0:32]
        lake = np.stack([x.ravel() + 160, y.ravel() + 240, z.ravel() + 160, np.full(x.size, 18)], 1)
        z, x = np.mgrid[0:48, 0:48]
        layer = np.stack([x.ravel() + 100, np.full(x.size, 301), z.ravel() + 300, np.full(x.size, 18)], 1)
        rng = np.random.default_rng(4)
        drops = np.concatenate([rng.integers(0, 512, (4000, 3)), np.full((4000, 1), 18)], 1)
        w.set_voxels(np.concatenate([lake, layer, drops]).astype(np.int32))
    w.pack()
    return w
