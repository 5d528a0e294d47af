"""World snapshots: the on-disk format, the host loader and the streaming device pool builder.

The reference has no on-disk world format (ARCHITECTURE.md:243), so nothing of the reference pins
the file; what is pinned:
  * the checksum: CRC-32 of the library against its check value and against zlib.crc32;
  * the layout: an independent Python reader (struct + zlib) parses header, index and payloads
    and finds the chunk trees the world holds, in rebuild_all_gpu_pools' Morton order;
  * the loaders: pools built from a loaded snapshot equal the original world's pools byte for
    byte (host: load + pack; device: streaming upload read back), and traces over them agree;
  * the checks: flipped bytes, truncation, bad magic, a missing file and a tree pointer forged
    with valid checksums are all refused.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np
import pytest

import arcadia_amd as A

HDR = struct.Struct("<8sII3IIQQIIII")
ENTRY = struct.Struct("<IIIIQQ")


def part1by2(n):
    n &= 0x3FF
    n = (n ^ (n << 16)) & 0xFF0000FF
    n = (n ^ (n << 8)) & 0x0300F00F
    n = (n ^ (n << 4)) & 0x030C30C3
    n = (n ^ (n << 2)) & 0x09249249
    return n


def morton(grid, idx):
    cx, cy, cz = idx % grid[0], (idx // grid[0]) % grid[1], idx // (grid[0] * grid[1])
    return part1by2(cx) | (part1by2(cy) << 1) | (part1by2(cz) << 2)


def same_pools(p, q):
    return tuple(p[0]) == tuple(q[0]) and all(
        np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8)) for a, b in zip(p[1:], q[1:]))


def caves(grid=(2, 2, 2), seed=3):
    w = A.HostWorld(grid)
    w.generate(A.gen_params(A.ARC_GEN_CAVES, seed=seed), nthreads=8)
    w.pack()
    return w


def heightmap_with_empty_chunks():
    w = A.HostWorld((3, 2, 2))  # the heightmap fills the bottom layer only
    w.generate(A.gen_params(A.ARC_GEN_HEIGHTMAP, seed=5, height_min=20.0, height_max=120.0), nthreads=8)
    w.pack()
    return w


@pytest.fixture(scope="module")
def world():
    return caves()


def test_crc32_check_value_and_zlib():
    assert A.crc32(b"123456789") == 0xCBF43926
    assert A.crc32(b"") == 0
    rng = np.random.default_rng(0)
    for n in (1, 7, 8, 9, 63, 64, 1000, 4097):
        b = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert A.crc32(b) == zlib.crc32(b)
        assert A.crc32(b[n // 2:], A.crc32(b[:n // 2])) == zlib.crc32(b)  # chained


@pytest.mark.parametrize("make", [caves, heightmap_with_empty_chunks])
def test_round_trip_gives_identical_pools(tmp_path, make):
    w = make()
    path = tmp_path / "w.arcw"
    w.save(path)
    with A.Snapshot(path) as s:
        s.verify()
        w2 = s.load_host()
    w2.pack()
    assert same_pools(w.pools(), w2.pools())


def test_edited_world_round_trip(tmp_path):
    w = caves(seed=4)
    rng = np.random.default_rng(1)
    for _ in range(12):
        w.carve_crater(tuple(rng.uniform(2.0, 30.0, 3)), float(rng.uniform(0.5, 2.0)))
    w.pack()
    w.save(tmp_path / "e.arcw")
    with A.Snapshot(tmp_path / "e.arcw") as s:
        w2 = s.load_host()
    w2.pack()
    assert same_pools(w.pools(), w2.pools())


def test_empty_world(tmp_path):
    w = A.HostWorld((2, 1, 1))
    w.pack()
    w.save(tmp_path / "z.arcw")
    with A.Snapshot(tmp_path / "z.arcw") as s:
        assert s.info() == {"grid": (2, 1, 1), "chunks": 0, "nodes": 0, "bricks": 0}
        w2 = s.load_host()
    w2.pack()
    assert same_pools(w.pools(), w2.pools())


def test_layout_read_independently(tmp_path):
    w = heightmap_with_empty_chunks()
    path = tmp_path / "w.arcw"
    w.save(path)
    data = path.read_bytes()
    magic, version, hbytes, gx, gy, gz, n, tn, tb, align, icrc, reserved, hcrc = HDR.unpack_from(data, 0)
    assert magic == b"ARCW64T\0" and version == 1 and hbytes == 64 and (gx, gy, gz) == w.grid
    assert align == 4096 and reserved == 0 and hcrc == zlib.crc32(data[:60])
    index = data[64:64 + 32 * n]
    assert icrc == zlib.crc32(index) and len(data) % 4096 == 0
    grid = (gx, gy, gz)
    stored = []
    for k in range(n):
        chunk, nn, nb, crc, off, solid = ENTRY.unpack_from(index, 32 * k)
        assert off % 4096 == 0
        payload = data[off:off + 16 * nn + 64 * nb]
        assert crc == zlib.crc32(payload)
        nodes, bricks = w.chunk_tree(chunk)
        assert payload == nodes.tobytes() + bricks.tobytes()
        assert solid == w.chunk_info(chunk)["solid"]
        stored.append(chunk)
    # every non-empty chunk, in ascending Morton order of its coordinates
    nonempty = [i for i in range(gx * gy * gz) if w.chunk_info(i)["solid"] > 0]
    assert stored == sorted(nonempty, key=lambda i: morton(grid, i)) and len(stored) < gx * gy * gz
    assert tn == sum(len(w.chunk_tree(i)[0]) for i in stored)


def rewrite(path, data: bytearray, fix_index=False, fix_header=False):
    if fix_index:
        n = HDR.unpack_from(data, 0)[6]
        struct.pack_into("<I", data, 52, zlib.crc32(bytes(data[64:64 + 32 * n])))
        fix_header = True
    if fix_header:
        struct.pack_into("<I", data, 60, zlib.crc32(bytes(data[:60])))
    path.write_bytes(bytes(data))


def test_corruption_is_refused(tmp_path, world):
    good = tmp_path / "g.arcw"
    world.save(good)
    data = bytearray(good.read_bytes())
    n = HDR.unpack_from(data, 0)[6]
    chunk, nn, nb, crc, off, solid = ENTRY.unpack_from(data, 64 + 32 * (n - 1))
    # a flipped payload byte: the file opens, the chunk check refuses it
    bad = bytearray(data)
    bad[off + 16 * nn + 5] ^= 0x40
    rewrite(tmp_path / "p.arcw", bad)
    with A.Snapshot(tmp_path / "p.arcw") as s:
        for f in (s.verify, s.load_host):
            with pytest.raises(A.ArcError) as e:
                f()
            assert e.value.code == A.ARC_ERR_CORRUPT and "checksum" in str(e.value)
    # header, index, magic, version, truncation
    cases = []
    b = bytearray(data); b[20] ^= 1; cases.append((b, {}))                     # grid, header CRC stale
    b = bytearray(data); b[64 + 8] ^= 1; cases.append((b, {}))                 # index CRC stale
    b = bytearray(data); b[0:8] = b"NOTAWRLD"; cases.append((b, {"fix_header": True}))
    b = bytearray(data); b[8] = 2; cases.append((b, {"fix_header": True}))
    b = bytearray(data[:off + 100]); cases.append((b, {}))                      # payload cut off
    b = bytearray(data); struct.pack_into("<I", b, 64, 7); cases.append((b, {"fix_index": True}))  # Morton order
    for i, (b, kw) in enumerate(cases):
        p = tmp_path / f"c{i}.arcw"
        rewrite(p, b, **kw)
        with pytest.raises(A.ArcError) as e:
            A.Snapshot(p)
        assert e.value.code == A.ARC_ERR_CORRUPT, (i, str(e.value))
    with pytest.raises(A.ArcError) as e:
        A.Snapshot(tmp_path / "missing.arcw")
    assert e.value.code == A.ARC_ERR_IO


def test_wrapping_payload_offset_is_refused(tmp_path, world):
    """The last chunk's offset forged to an aligned value near 2^64, with both CRCs recomputed:
    offset + payload size wraps below the file size, so only a test of each term apart refuses it."""
    good = tmp_path / "g.arcw"
    world.save(good)
    data = bytearray(good.read_bytes())
    n = HDR.unpack_from(data, 0)[6]
    chunk, nn, nb, crc, off, solid = ENTRY.unpack_from(data, 64 + 32 * (n - 1))
    payload = 16 * nn + 64 * nb
    forged = 2 ** 64 - 4096  # aligned and beyond every earlier payload; offset + size wraps
    assert payload > 4096 and forged > off and (forged + payload) % 2 ** 64 < len(data)
    struct.pack_into("<Q", data, 64 + 32 * (n - 1) + 16, forged)
    p = tmp_path / "wrap.arcw"
    rewrite(p, data, fix_index=True)
    with pytest.raises(A.ArcError) as e:
        A.Snapshot(p)
    assert e.value.code == A.ARC_ERR_CORRUPT and "offset" in str(e.value)


def test_forged_tree_pointer_is_refused(tmp_path):
    """A child pointer outside the chunk's node array, with every checksum made valid again."""
    w = heightmap_with_empty_chunks()
    path = tmp_path / "w.arcw"
    w.save(path)
    data = bytearray(path.read_bytes())
    chunk, nn, nb, crc, off, solid = ENTRY.unpack_from(data, 64)
    root = np.frombuffer(bytes(data[off:off + 16]), A.NODE_DTYPE)[0]
    assert root["child_ptr"] & 0x80000000 == 0  # an inner root
    struct.pack_into("<I", data, off + 8, int(root["child_ptr"] & 0xC0000000) | (nn + 5))
    struct.pack_into("<I", data, 64 + 12, zlib.crc32(bytes(data[off:off + 16 * nn + 64 * nb])))
    rewrite(path, data, fix_index=True)
    with A.Snapshot(path) as s:
        with pytest.raises(A.ArcError) as e:
            s.verify()
        assert e.value.code == A.ARC_ERR_CORRUPT and "outside the node pool" in str(e.value)


# ---- GPU: the streaming pool builder ---------------------------------------------------------

@pytest.mark.gpu
def test_gpu_streaming_upload_equals_host_pools(ctx, tmp_path):
    import torch

    world = caves((4, 2, 2), seed=6)
    path = tmp_path / "w.arcw"
    world.save(path)
    grid, roots, nodes, bricks = world.pools()
    assert len(nodes) * 16 + len(bricks) * 64 > 128 << 20  # more than one staging batch
    with A.Snapshot(path) as s:
        ctx.upload_snapshot(s, node_capacity=len(nodes) + 1000, brick_capacity=len(bricks) + 10)
    r2, n2, b2 = ctx.read_pools(len(roots), len(nodes) + 1000, len(bricks) + 10)
    assert np.array_equal(r2, roots)
    assert np.array_equal(n2[:len(nodes)].view(np.uint32), nodes.view(np.uint32))
    assert np.array_equal(b2[:len(bricks)].view(np.uint32), bricks.view(np.uint32))
    assert not n2[len(nodes):].view(np.uint32).any() and not b2[len(bricks):].view(np.uint32).any()
    # traces over the streamed pools equal traces over the regular upload
    cam = A.camera_look((30.0, 30.0, 12.0), 0.6, -0.5, 70.0, 320, 180)
    rays, _ = A.primary_rays(cam, 320, 180, 100.0)
    d_r = torch.from_numpy(rays.view(np.uint8)).cuda()
    out = []
    for up in ("snapshot", "pools"):
        if up == "pools":
            ctx.upload(grid, roots, nodes, bricks)
        d_h = torch.zeros(len(rays) * 16, dtype=torch.uint8, device="cuda")
        ctx.trace_device(A.TraceParams.exact(0), d_r.data_ptr(), len(rays), d_h.data_ptr())
        torch.cuda.synchronize()
        out.append(d_h.cpu().numpy())
    assert np.array_equal(out[0], out[1]) and (out[0].view(A.HIT_DTYPE)["flags"] & 1).sum() > 1000
    # a second upload reuses the context's staging buffers
    with A.Snapshot(path) as s:
        ctx.upload_snapshot(s)
    r3, n3, b3 = ctx.read_pools(len(roots), len(nodes), len(bricks))
    assert np.array_equal(n3.view(np.uint32), nodes.view(np.uint32)) and np.array_equal(r3, roots)


@pytest.mark.gpu
def test_gpu_streaming_upload_refuses_corrupt_chunk(ctx, tmp_path):
    w = heightmap_with_empty_chunks()
    path = tmp_path / "w.arcw"
    w.save(path)
    data = bytearray(path.read_bytes())
    chunk, nn, nb, crc, off, solid = ENTRY.unpack_from(data, 64)
    data[off + 40] ^= 0x10
    path.write_bytes(bytes(data))
    with A.Snapshot(path) as s:
        with pytest.raises(A.ArcError) as e:
            ctx.upload_snapshot(s)
    assert e.value.code == A.ARC_ERR_CORRUPT
    rays = np.zeros(4, A.RAY_DTYPE)
    rays["dy"], rays["t_max"] = -1.0, 10.0
    with pytest.raises(A.ArcError) as e:
        ctx.trace(A.TraceParams.exact(0), rays)
    assert e.value.code == A.ARC_ERR_NO_WORLD
