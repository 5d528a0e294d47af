"""Primary G-buffer pass (primary_gbuffer.wgsl:93-312, opaque path).

CPU tests pin the oracle's packing against what the mathematics fixes: numpy's float16
conversion, the octahedral encoding of axis normals, a reprojection through the camera's own
view-projection (motion = 0), far-hit albedo = palette albedo, depth = traversal t / 16 on the
same rays as the host ray generator. GPU tests compare the HIP pass bit for bit with the oracle.
"""
from __future__ import annotations

import functools

import numpy as np
import pytest

import arcadia_amd as A
from oracle import oracle as O
from tests import worlds as W


@functools.lru_cache(maxsize=None)
def wide_heightmap():
    """1024 x 256 x 1024 heightmap world (64 m across) so hits reach past the 48 m fade."""
    w = A.HostWorld((4, 1, 4))
    w.generate(A.gen_params(A.ARC_GEN_HEIGHTMAP, seed=4))
    w.pack()
    return w.pools()


def palette():
    return np.stack([A.palette_material(m) for m in range(256)]).astype(np.float32)


def view_proj(cam) -> np.ndarray:
    """Column-major 4x4 whose clip.xy/clip.w is the NDC of the camera's pinhole model."""
    r, u, f, e = (np.array(list(getattr(cam, k)), np.float64) for k in ("right", "up", "forward", "eye"))
    m = np.zeros((4, 4))  # m[row, col]
    m[0, :3] = r / (cam.aspect * cam.tan_half_fov)
    m[0, 3] = -r @ e / (cam.aspect * cam.tan_half_fov)
    m[1, :3] = u / cam.tan_half_fov
    m[1, 3] = -u @ e / cam.tan_half_fov
    m[3, :3] = f
    m[3, 3] = -f @ e
    return m.T.reshape(-1).astype(np.float32)  # column-major storage


def snorm_decode(v):
    s = (v & 0xFFFF).astype(np.uint16).view(np.int16).astype(np.float64)
    t = (v >> 16).astype(np.uint16).view(np.int16).astype(np.float64)
    return np.maximum(s / 32767.0, -1.0), np.maximum(t / 32767.0, -1.0)


def oct_decode(x, y):
    z = 1.0 - np.abs(x) - np.abs(y)
    neg = z < 0
    wx = (1.0 - np.abs(y)) * np.where(x >= 0, 1.0, -1.0)
    wy = (1.0 - np.abs(x)) * np.where(y >= 0, 1.0, -1.0)
    x, y = np.where(neg, wx, x), np.where(neg, wy, y)
    n = np.stack([x, y, z], -1)
    return n / np.linalg.norm(n, axis=-1, keepdims=True)


def half_decode(v):
    return (v & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float64), \
        (v >> 16).astype(np.uint16).view(np.float16).astype(np.float64)


def test_f16_conversion_matches_numpy():
    rng = np.random.default_rng(0)
    bits = np.concatenate([
        rng.integers(0, 2**32, 20000, dtype=np.uint64).astype(np.uint32),
        np.arange(0x33000000, 0x33800000, 997, dtype=np.uint32),   # half subnormal range
        np.arange(0x387FF000, 0x38801000, 3, dtype=np.uint32),     # subnormal / normal boundary
        np.arange(0x477FE000, 0x47802000, 5, dtype=np.uint32),     # overflow boundary
        np.array([0, 0x80000000, 0x7F800000, 0xFF800000, 0x3F800000, 0x33000000, 0x33000001], np.uint32),
    ])
    f = bits.view(np.float32)
    with np.errstate(over="ignore"):
        ref = f.astype(np.float16).view(np.uint16)
    got = np.array([O.f32_to_f16_bits(x) for x in f], np.uint16)
    finite = ~np.isnan(f)
    assert np.array_equal(got[finite], ref[finite])
    assert np.all((got[~finite] & 0x7C00) == 0x7C00) and np.all((got[~finite] & 0x3FF) != 0)


def _oracle_gbuffer(pools, cam, w, h, vp=None):
    g = A.GbufferParams.from_camera(cam, w, h, prev_view_proj=vp)
    return g, O.primary_gbuffer(O.World(*pools), A.TraceParams.exact(), g, palette())


def _pixel_trace(pools, cam, w, h):
    rays, pix = A.primary_rays(cam, w, h, 250.0)
    full, _ = O.trace_world(O.World(*pools), A.TraceParams.exact(), rays)
    img = np.zeros(w * h, O.RESULT_DTYPE)
    img[pix] = full
    dirs = np.zeros((w * h, 3), np.float64)
    dirs[pix] = np.stack([rays["dx"], rays["dy"], rays["dz"]], -1)
    return img.reshape(h, w), dirs.reshape(h, w, 3)


def test_oracle_gbuffer_invariants():
    pools = wide_heightmap()
    w, h = 203, 117  # not multiples of 8
    cam = A.camera_look((3.1, 17.5, 2.7), 0.78, -0.22, 70.0, w, h)
    g, (dm, nm, aa) = _oracle_gbuffer(pools, cam, w, h, view_proj(cam))
    tr, dirs = _pixel_trace(pools, cam, w, h)
    hit = (tr["hit"] & 1) == 1
    assert 0.2 < hit.mean() < 0.99
    depth = dm[..., 0].view(np.float32)
    # depth / material: traversal t in voxels / 16 on the host generator's rays, flags 0
    assert np.array_equal(depth[hit], (tr["t_voxels"][hit] * np.float32(0.0625)).astype(np.float32))
    assert np.array_equal(dm[..., 1][hit], tr["mat"][hit] & 0xFFFF)
    assert np.all(depth[~hit] == np.float32(1e5)) and np.all(dm[..., 1][~hit] == 0)
    # normals: axis normals round-trip exactly through oct + snorm16; misses carry the ray direction
    n = oct_decode(*snorm_decode(nm[..., 0]))
    assert np.allclose(n[hit], tr["normal"][hit], atol=1e-6)
    assert np.allclose(n[~hit], dirs[~hit], atol=2e-4)
    # motion: reprojection through the camera's own view-projection lands on the pixel centre
    mx, my = half_decode(nm[..., 1])
    assert np.max(np.abs(mx[hit])) < 2e-4 and np.max(np.abs(my[hit])) < 2e-4
    assert np.all(mx[~hit] == 0) and np.all(my[~hit] == 0)
    # albedo: far hits (>= 48 m) carry the palette albedo exactly; nearer ones stay within variance
    pal = palette()
    mat = (tr["mat"] & 0xFF).astype(np.int64)
    alb = np.stack([(aa >> (8 * k)) & 0xFF for k in range(3)], -1).astype(np.float64) / 255.0
    base = np.clip(pal[mat][..., :3].astype(np.float64), 0.01, 1.0)
    far = hit & (depth >= 48.0)
    near = hit & (depth < 18.0)
    assert far.sum() > 100 and near.sum() > 100
    assert np.max(np.abs(alb[far] - base[far])) <= 0.5 / 255 + 1e-9
    var = np.abs(pal[mat][..., 7]).astype(np.float64)
    assert np.all(np.abs(alb[hit] - base[hit]) <= (base[hit] * var[hit][:, None]) + 0.5 / 255 + 1e-6)
    ao = ((aa >> 24) & 0xFF).astype(np.float64) / 255.0
    assert np.max(np.abs(ao[hit] - np.clip(tr["contact_ao"][hit], 0, 1))) <= 0.5 / 255 + 1e-9
    assert np.all(ao[~hit] == 1.0) and np.all((aa[~hit] & 0xFFFFFF) == 0)


def test_oracle_gbuffer_motion_identity_matrix():
    """Identity prev_view_proj: motion = (x*0.5+0.5 - u, 0.5 - y*0.5 - v) of the hit position."""
    pools = W.heightmap_world(1)
    w, h = 96, 64
    cam = A.camera_look((8.0, 14.0, 3.0), 0.3, -0.5, 80.0, w, h)
    _, (dm, nm, aa) = _oracle_gbuffer(pools, cam, w, h, None)
    tr, _ = _pixel_trace(pools, cam, w, h)
    hit = (tr["hit"] & 1) == 1
    py, px = np.mgrid[0:h, 0:w]
    pu, pv = (px + 0.5) / w, (py + 0.5) / h
    hp = tr["hit_pos_m"].astype(np.float64)
    ex, ey = hp[..., 0] * 0.5 + 0.5 - pu, 0.5 - hp[..., 1] * 0.5 - pv
    mx, my = half_decode(nm[..., 1])
    tol = lambda e: np.abs(e) * 2.0 ** -10 + 1e-6
    assert np.all(np.abs(mx[hit] - ex[hit]) <= tol(ex[hit]))
    assert np.all(np.abs(my[hit] - ey[hit]) <= tol(ey[hit]))


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["heightmap_203x117", "wide_640x360_jitter", "caves_321x200"])
def test_gpu_gbuffer_bit_exact(ctx, case):
    import torch

    if case == "heightmap_203x117":
        pools, w, h = W.heightmap_world(1), 203, 117
        cam = A.camera_look((8.0, 14.0, 3.0), 0.3, -0.5, 80.0, w, h)
        vp = None
    elif case == "wide_640x360_jitter":
        pools, w, h = wide_heightmap(), 640, 360
        cam = A.camera_look((3.1, 17.5, 2.7), 0.78, -0.22, 70.0, w, h)
        cam.jitter[0], cam.jitter[1] = 0.31, -0.17
        prev = A.camera_look((3.3, 17.4, 2.6), 0.76, -0.21, 70.0, w, h)
        vp = view_proj(prev)
    else:
        pools, w, h = W.caves_world(1), 321, 200
        cam = A.camera_look((13.7, 30.2, 11.9), 0.6, -0.55, 65.0, w, h)
        vp = view_proj(cam)
    print(f"[gbuffer] {case}", flush=True)
    g, (dm, nm, aa) = _oracle_gbuffer(pools, cam, w, h, vp)
    ctx.upload(*pools)
    guard = 64
    sentinel = 0x5A5A5A5A
    bufs = [torch.full((n + guard,), sentinel, dtype=torch.int32, device="cuda")
            for n in (2 * w * h, 2 * w * h, w * h)]
    ctx.primary_gbuffer(A.TraceParams.exact(), g, *(b.data_ptr() for b in bufs))
    torch.cuda.synchronize()
    got = [b.cpu().numpy().view(np.uint32) for b in bufs]
    for name, gb, ref in zip(("depth_mat", "norm_motion", "albedo_ao"), got, (dm, nm, aa)):
        body = gb[:-guard].reshape(ref.shape)
        bad = np.argwhere(body != ref)
        assert len(bad) == 0, f"{case} {name}: {len(bad)} words differ, first {bad[:4].tolist()}"
        assert np.all(gb[-guard:] == sentinel), f"{case} {name}: wrote past the image"
