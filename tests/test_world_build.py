"""Host world building: layout, collapse rules, BFS packing, pool packing, palette.

Pins the builder against the reference's construction rules
(crates/arcadia-core/src/chunk_tree/builder.rs:106-378, crates/arcadia-world/src/world_map.rs:158-205,
689-735, crates/arcadia-core/src/material.rs:70-80) with known answers derived from those rules."""
from __future__ import annotations

import numpy as np

import arcadia_amd as A
from oracle import oracle as O
from tests import worlds as W


def _dense_world(dense, grid=(1, 1, 1)):
    w = A.HostWorld(grid)
    w.set_chunk_dense(0, dense)
    w.pack()
    return w


def test_layout_sizes():
    assert A.NODE_DTYPE.itemsize == 16  # Node64, node64.rs:25-36
    assert A.BRICK_DTYPE.itemsize == 64  # LeafBrick64, node64.rs:116-129
    assert A.RAY_DTYPE.itemsize == 32 and A.HIT_DTYPE.itemsize == 16 and A.RESULT_DTYPE.itemsize == 80


def test_popcnt_child_slot_offset():
    """node64.rs:164-174: children 0, 5, 19, 63 of a root at base 100 sit at 100..103 — here a
    root whose four children are full uniform 64^3 cubes of materials 1..4, queried per child."""
    dense = np.zeros((256, 256, 256), np.uint8)
    for mat, b in enumerate((0, 5, 19, 63), start=1):
        x, y, z = (b & 3) * 64, ((b >> 2) & 3) * 64, (b >> 4) * 64
        dense[z:z + 64, y:y + 64, x:x + 64] = mat
    w = _dense_world(dense)
    nodes, bricks = w.chunk_tree(0)
    mask = int(nodes[0]["mask_low"]) | (int(nodes[0]["mask_high"]) << 32)
    assert mask == (1 << 0) | (1 << 5) | (1 << 19) | (1 << 63)
    assert int(nodes[0]["child_ptr"]) == 1 and len(nodes) == 5 and len(bricks) == 0
    for i in range(1, 5):  # collapsed full cubes: uniform leaves, material in lod_payload
        assert int(nodes[i]["child_ptr"]) == 0xC0000000  # IS_LEAF_FLAG | IS_UNIFORM_FLAG, base 0
        assert int(nodes[i]["lod_payload"]) >> 24 == i
    ow = O.World(*w.pools())
    for mat, b in enumerate((0, 5, 19, 63), start=1):
        x, y, z = (b & 3) * 64 + 7, ((b >> 2) & 3) * 64 + 9, (b >> 4) * 64 + 11
        assert O.query_voxel(ow, x, y, z) == mat
    assert O.query_voxel(ow, 64 + 64, 0, 0) == 0


def test_leaf_brick_pack_unpack():
    """node64.rs:177-185: voxel bit b of a 4^3 leaf holds material b + 1 in byte b&3 of word b>>2."""
    dense = np.zeros((256, 256, 256), np.uint8)
    for b in range(64):
        dense[b >> 4, (b >> 2) & 3, b & 3] = b + 1
    w = _dense_world(dense)
    nodes, bricks = w.chunk_tree(0)
    assert len(bricks) == 1
    words = bricks[0]["packed_materials"]
    for b in range(64):
        assert (int(words[b >> 2]) >> ((b & 3) * 8)) & 0xFF == b + 1
    leaf = nodes[-1]
    assert int(leaf["mask_low"]) == 0xFFFFFFFF and int(leaf["mask_high"]) == 0xFFFFFFFF
    assert int(leaf["child_ptr"]) == 0x80000000  # leaf, mixed, brick 0


def test_homogeneous_collapse_and_lookup():
    """chunk_tree.rs:26-81: stone below y=64, grass at y=64, air above. The 16 full 64^3 stone
    cubes collapse to uniform leaves at level 1; the grass layer is 16 level-1 internals, each
    with 16 level-2 internals, each with 16 uniform (non-full) level-3 leaves."""
    dense = np.zeros((256, 256, 256), np.uint8)
    dense[:, :64, :] = 2
    dense[:, 64, :] = 8
    w = _dense_world(dense)
    nodes, bricks = w.chunk_tree(0)
    assert len(bricks) == 0
    assert len(nodes) == 1 + 32 + 256 + 4096
    ow = O.World(*w.pools())
    assert O.query_voxel(ow, 10, 20, 30) == 2
    assert O.query_voxel(ow, 128, 64, 128) == 8
    assert O.query_voxel(ow, 128, 65, 128) == 0
    # BFS order: the root's 32 children come first, the 16 collapsed stone cubes (cy = 0) are
    # uniform full leaves, the 16 grass-layer children are internal
    kids = nodes[1:33]
    flags = kids["child_ptr"] & 0xC0000000
    cy = [(b >> 2) & 3 for b in range(64) if (int(nodes[0]["mask_low"]) | int(nodes[0]["mask_high"]) << 32) >> b & 1]
    for f, c in zip(flags, cy):
        assert (f == 0xC0000000) == (c == 0)


def test_round_trip_random_dense_and_sparse():
    rng = np.random.default_rng(5)
    dense = np.zeros((256, 256, 256), np.uint8)
    blob = rng.random((64, 64, 64)) < 0.3
    dense[40:104, 100:164, 7:71] = np.where(blob, rng.integers(1, 16, blob.shape), 0).astype(np.uint8)
    dense[200:232, 0:32, 200:232] = 9  # collapsible solid block
    w = _dense_world(dense)
    ow = O.World(*w.pools())
    assert np.array_equal(O.decode_dense(ow), dense)
    pts = rng.integers(0, 256, (500, 3))
    vox = [(int(x), int(y), int(z), int(rng.integers(1, 30))) for x, y, z in pts]
    pools = W.sparse_world(vox)
    expect = {}
    for x, y, z, m in vox:
        expect[(x, y, z)] = m  # last write wins
    ow = O.World(*pools)
    for (x, y, z), m in expect.items():
        assert O.query_voxel(ow, x, y, z) == m
    assert O.query_voxel(ow, 0, 0, 0) == expect.get((0, 0, 0), 0)


def test_world_pool_packing_morton_and_relocation():
    """rebuild_all_gpu_pools: slot-0 dummy node/brick, chunks packed in Morton order of their
    coordinates, empty chunks -> sentinel, child pointers relocated (world_map.rs:158-205, 705-735)."""
    grid = (2, 2, 2)
    w = A.HostWorld(grid)
    rng = np.random.default_rng(1)
    for idx in range(8):
        if idx == 5:
            continue  # stays empty
        pts = rng.integers(0, 256, (40, 3))
        w.set_chunk_sparse(idx, [(int(x), int(y), int(z), idx + 1) for x, y, z in pts])
    w.pack()
    g, roots, nodes, bricks = w.pools()
    assert roots[5] == A.native.EMPTY_CHUNK_SENTINEL
    assert nodes[0].tolist() == (0, 0, 0, 0) and not bricks[0]["packed_materials"].any()

    def morton(cx, cy, cz):
        m = 0
        for bit in range(10):
            m |= ((cx >> bit) & 1) << (3 * bit) | ((cy >> bit) & 1) << (3 * bit + 1) | ((cz >> bit) & 1) << (3 * bit + 2)
        return m

    order = sorted((morton(i % 2, (i // 2) % 2, i // 4), i) for i in range(8) if i != 5)
    base = [int(roots[i]) for _, i in order]
    assert base == sorted(base) and base[0] == 1
    ow = O.World(g, roots, nodes, bricks)
    for idx in range(8):
        if idx == 5:
            continue
        cx, cy, cz = idx % 2, (idx // 2) % 2, idx // 4
        local_nodes, _ = w.chunk_tree(idx)
        assert int(roots[idx]) + len(local_nodes) <= len(nodes)
        # every voxel of the chunk's sparse set reads back through the relocated pools
        x, y, z = (cx * 256 + 3, cy * 256 + 3, cz * 256 + 3)
        assert O.query_voxel(ow, x, y, z) in (0, idx + 1)


def test_palette_lod_payload():
    """Material::pack_lod_payload (material.rs:70-75) of deep stone (0.28, 0.30, 0.34)."""
    p = A.lod_payload(2)
    assert p >> 24 == 2
    assert p & 0xFF == int(np.float32(0.28) * np.float32(255.0))
    assert (p >> 8) & 0xFF == int(np.float32(0.30) * np.float32(255.0))
    assert (p >> 16) & 0xFF == int(np.float32(0.34) * np.float32(255.0))
    m = A.palette_material(27)  # point light: emissive
    assert m[4] > 100.0
    assert A.palette_material(0)[3] == 1.0


def test_generated_worlds_are_deterministic():
    a = W.heightmap_world(7)
    w = A.HostWorld((1, 1, 1))
    w.generate(A.gen_params(A.ARC_GEN_HEIGHTMAP, seed=7), nthreads=3)
    w.pack()
    b = w.pools()
    for x, y in zip(a[1:], b[1:]):
        assert x.tobytes() == y.tobytes()
    dense = O.decode_dense(O.World(*a))
    solid = dense > 0
    assert 0.3 < solid.mean() < 0.6
    cols = solid.sum(axis=1)  # heights per (z, x) column: terrain is a heightfield
    assert cols.min() >= 32 and cols.max() <= 200
    assert set(np.unique(dense[solid]).tolist()) <= set(range(1, 16))
