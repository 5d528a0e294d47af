"""Edit path: in-place chunk edits, partial pool sync and GPU patches.

CPU tests restate the reference's own edit tests (crates/arcadia-core/tests/edit_set.rs and the
sync tests of crates/arcadia-world/src/world_map.rs:743-1041) and check every edit against a
dense ground truth: the world decoded voxel by voxel equals a dense grid edited with the same
binary32 sphere test. GPU tests apply the queued patches to the device pools, check that the
device pools equal the host pools byte for byte, and that the traversal over the edited pools is
bit-exact against the oracle, frame after frame.
"""
from __future__ import annotations

import numpy as np
import pytest

import arcadia_amd as A
from oracle import oracle as O
from tests import worlds as W

AIR, STONE_DEEP, WATER, CRYSTAL_CYAN, SCORCHED, EMBER = 0, 2, 18, 19, 23, 24


def fresh(grid=(1, 1, 1)):
    return A.HostWorld(grid)


def solid_stone():
    w = fresh()
    w.set_chunk_dense(0, np.full((256, 256, 256), STONE_DEEP, np.uint8))
    return w


def get(w, x, y, z):
    return w.voxel(x, y, z)


def mask(n):
    return int(n["mask_low"]) | (int(n["mask_high"]) << 32)


def child_slot(n, b):
    return (int(n["child_ptr"]) & A.native.ARC_PTR_INDEX_MASK) + bin(mask(n) & ((1 << b) - 1)).count("1")


def cbit(x, y, z):
    return (z << 4) | (y << 2) | x


# ---- edit_set.rs KATs ---------------------------------------------------------------------
def test_in_place_set_voxels_air_to_solid():  # edit_set.rs:7-25
    w = fresh()
    rep = w.set_voxels_local(0, [(0, 0, 0, CRYSTAL_CYAN), (0, 0, 1, CRYSTAL_CYAN), (0, 0, 2, CRYSTAL_CYAN)])
    assert rep.added_voxels == 3 and w.chunk_info(0)["solid"] == 3
    assert [get(w, 0, 0, z) for z in range(4)] == [CRYSTAL_CYAN] * 3 + [AIR]


def test_in_place_overwrite_solid_and_clear_to_air():  # edit_set.rs:28-59
    w = fresh()
    w.set_voxels_local(0, [(10, 10, 10, CRYSTAL_CYAN), (10, 10, 11, STONE_DEEP)])
    assert w.chunk_info(0)["solid"] == 2
    assert w.set_voxels_local(0, [(10, 10, 10, STONE_DEEP)]).added_voxels == 0
    assert get(w, 10, 10, 10) == STONE_DEEP and w.chunk_info(0)["solid"] == 2
    assert w.set_voxels_local(0, [(10, 10, 10, AIR)]).added_voxels == 0
    assert w.chunk_info(0)["solid"] == 1 and get(w, 10, 10, 10) == AIR and get(w, 10, 10, 11) == STONE_DEEP
    w.set_voxels_local(0, [(10, 10, 11, AIR)])
    assert w.chunk_info(0)["solid"] == 0
    nodes, _ = w.chunk_tree(0)
    assert mask(nodes[0]) == 0


def test_in_place_across_multiple_bricks():  # edit_set.rs:62-87
    w = fresh()
    v = [(0, 0, 0, CRYSTAL_CYAN), (4, 0, 0, STONE_DEEP), (100, 150, 200, CRYSTAL_CYAN), (255, 255, 255, STONE_DEEP)]
    assert w.set_voxels_local(0, v).added_voxels == 4
    for x, y, z, m in v:
        assert get(w, x, y, z) == m


def test_setting_air_in_empty_chunk_does_not_allocate():  # edit_set.rs:90-101
    w = fresh()
    n0 = w.chunk_info(0)["n_nodes"]
    assert w.set_voxels_local(0, [(50, 50, 50, AIR), (100, 100, 100, AIR)]).added_voxels == 0
    assert w.chunk_info(0)["n_nodes"] == n0 and w.chunk_info(0)["solid"] == 0


def test_in_place_mixed_brick_transition():  # edit_set.rs:104-135
    w = fresh()
    w.set_voxels_local(0, [(0, 0, 0, CRYSTAL_CYAN), (0, 0, 1, CRYSTAL_CYAN), (0, 0, 2, CRYSTAL_CYAN)])
    assert w.chunk_info(0)["n_bricks"] == 0
    w.set_voxels_local(0, [(0, 0, 3, STONE_DEEP)])
    assert w.chunk_info(0)["n_bricks"] == 1 and w.chunk_info(0)["solid"] == 4
    assert [get(w, 0, 0, z) for z in range(4)] == [CRYSTAL_CYAN] * 3 + [STONE_DEEP]
    assert w.set_voxels_local(0, [(0, 0, 0, STONE_DEEP)]).added_voxels == 0
    assert w.chunk_info(0)["n_bricks"] == 1 and get(w, 0, 0, 0) == STONE_DEEP


def test_in_place_uncollapse_and_sculpt():  # edit_set.rs:138-162
    w = solid_stone()
    s0 = w.chunk_info(0)["solid"]
    nodes, _ = w.chunk_tree(0)
    assert int(nodes[0]["child_ptr"]) & A.native.ARC_IS_LEAF_FLAG
    assert w.set_voxels_local(0, [(128, 128, 128, CRYSTAL_CYAN)]).added_voxels == 0
    assert w.chunk_info(0)["solid"] == s0
    assert get(w, 128, 128, 128) == CRYSTAL_CYAN and get(w, 128, 128, 129) == STONE_DEEP
    w.set_voxels_local(0, [(128, 128, 129, AIR)])
    assert w.chunk_info(0)["solid"] == s0 - 1
    assert get(w, 128, 128, 129) == AIR and get(w, 128, 128, 130) == STONE_DEEP


def test_chunk_edit_report_tracking():  # edit_set.rs:187-233
    w = fresh()
    r0 = w.set_voxels_local(0, [])
    assert not r0.has_dirty_nodes() and not r0.has_dirty_bricks() and not r0.had_structural_growth
    r1 = w.set_voxels_local(0, [(0, 0, 0, CRYSTAL_CYAN), (0, 0, 1, CRYSTAL_CYAN)])
    assert r1.had_structural_growth and r1.min_dirty_node == 0 and r1.max_dirty_node >= 3
    assert not r1.has_dirty_bricks() and r1.added_voxels == 2
    r2 = w.set_voxels_local(0, [(0, 0, 2, CRYSTAL_CYAN)])
    assert not r2.had_structural_growth and r2.has_dirty_nodes() and not r2.has_dirty_bricks()
    assert r2.added_voxels == 1
    r3 = w.set_voxels_local(0, [(0, 0, 3, STONE_DEEP)])
    assert r3.had_structural_growth and r3.has_dirty_bricks()
    assert (r3.min_dirty_brick, r3.max_dirty_brick, r3.added_voxels) == (0, 0, 1)
    r4 = w.set_voxels_local(0, [(0, 0, 3, CRYSTAL_CYAN)])
    assert not r4.had_structural_growth and r4.has_dirty_bricks()
    assert (r4.min_dirty_brick, r4.max_dirty_brick, r4.added_voxels) == (0, 0, 0)


def test_in_place_remove_child_no_node_growth():  # edit_set.rs:236-261
    w = fresh()
    w.set_voxels_local(0, [(0, 0, 0, CRYSTAL_CYAN), (0, 0, 1, CRYSTAL_CYAN), (4, 0, 0, STONE_DEEP)])
    n0 = w.chunk_info(0)["n_nodes"]
    r1 = w.set_voxels_local(0, [(0, 0, 0, AIR)])
    assert not r1.had_structural_growth and w.chunk_info(0)["n_nodes"] == n0
    assert get(w, 0, 0, 0) == AIR and get(w, 0, 0, 1) == CRYSTAL_CYAN
    r2 = w.set_voxels_local(0, [(4, 0, 0, AIR)])
    assert not r2.had_structural_growth and w.chunk_info(0)["n_nodes"] == n0
    assert get(w, 4, 0, 0) == AIR and get(w, 0, 0, 1) == CRYSTAL_CYAN


def test_compact_strips_orphaned_nodes_and_preserves_voxels():  # edit_set.rs:264-301
    w = fresh()
    v = [(0, 0, 0, CRYSTAL_CYAN), (4, 4, 4, STONE_DEEP), (100, 100, 100, CRYSTAL_CYAN), (200, 200, 200, STONE_DEEP)]
    w.set_voxels_local(0, v)
    w.set_voxels_local(0, [(8, 8, 8, CRYSTAL_CYAN)])
    w.set_voxels_local(0, [(8, 8, 8, AIR)])
    before = w.chunk_info(0)
    w.compact_chunk(0)
    after = w.chunk_info(0)
    assert after["solid"] == before["solid"] and after["n_nodes"] <= before["n_nodes"]
    for x, y, z, m in v:
        assert get(w, x, y, z) == m
    assert get(w, 8, 8, 8) == AIR and get(w, 50, 50, 50) == AIR
    e = fresh()
    e.compact_chunk(0)
    assert e.chunk_info(0)["n_nodes"] == 1 and e.chunk_info(0)["n_bricks"] == 0


def test_carve_sphere_destructive():  # edit_set.rs:304-355
    w = solid_stone()
    s0 = w.chunk_info(0)["solid"]
    d, _ = w.carve_sphere_local(0, (64.0, 64.0, 64.0), 6.0, scorch_rim=True)
    assert d > 500 and w.chunk_info(0)["solid"] == s0 - d
    assert get(w, 64, 64, 64) == AIR and get(w, 65, 64, 64) == AIR
    found = any(get(w, 64 + dx, 64 + dy, 64 + dz) in (SCORCHED, EMBER)
                for dz in range(-8, 9) for dy in range(-8, 9) for dx in range(-8, 9))
    assert found
    assert w.carve_sphere_local(0, (500.0, 500.0, 500.0), 5.0)[0] == 0


def test_set_voxels_local():  # edit_set.rs:404-428
    w = fresh()
    assert w.set_voxels_local(0, [(5, 5, 5, CRYSTAL_CYAN), (5, 5, 6, CRYSTAL_CYAN),
                                  (5, 5, 7, STONE_DEEP)]).added_voxels == 3
    assert w.set_voxels_local(0, [(5, 5, 5, STONE_DEEP)]).added_voxels == 0
    assert w.chunk_info(0)["solid"] == 3 and get(w, 5, 5, 5) == STONE_DEEP


def test_empty_chunk_operations():  # edit_set.rs:431-454 (carve / clear / set parts)
    w = fresh()
    assert w.carve_sphere_local(0, (10.0, 10.0, 10.0), 5.0, scorch_rim=True)[0] == 0
    assert w.clear_voxels_local(0, [(10, 10, 10)])[0] == 0
    assert w.set_voxels_local(0, []).added_voxels == 0


def test_carve_sphere_tracked_zero_repack():  # edit_set.rs:457-482
    w = solid_stone()
    d, rep = w.carve_sphere_local(0, (64.0, 64.0, 64.0), 5.0)
    assert d > 300 and rep.had_structural_growth and rep.has_dirty_nodes() and get(w, 64, 64, 64) == AIR
    d2, rep2 = w.carve_sphere_local(0, (64.0, 64.0, 64.0), 5.5)
    assert d2 > 0 and not rep2.had_structural_growth and rep2.has_dirty_nodes()


def test_carve_crater_clears_parent_occupancy_masks():  # edit_set.rs:507-568
    w = solid_stone()
    s0 = w.chunk_info(0)["solid"]
    d, rep = w.carve_sphere_local(0, (32.0, 32.0, 32.0), 10.0)
    assert d > 0 and w.chunk_info(0)["solid"] == s0 - d and rep.has_dirty_nodes()
    for p in ((32, 32, 32), (30, 30, 30), (34, 34, 34)):
        assert get(w, *p) == AIR
    for p in ((0, 0, 0), (32, 32, 50), (50, 32, 32)):
        assert get(w, *p) == STONE_DEEP
    nodes, _ = w.chunk_tree(0)
    for b in (8, 7):
        b0, b1, b2 = cbit((b >> 4) & 3, (b >> 4) & 3, (b >> 4) & 3), cbit((b >> 2) & 3, (b >> 2) & 3, (b >> 2) & 3), \
            cbit(b & 3, b & 3, b & 3)
        assert (mask(nodes[0]) >> b0) & 1
        l1 = child_slot(nodes[0], b0)
        assert (mask(nodes[l1]) >> b1) & 1
        l2 = child_slot(nodes[l1], b1)
        assert not (mask(nodes[l2]) >> b2) & 1


# ---- world_map.rs sync KATs ----------------------------------------------------------------
def packed_world(voxels, grid=(1, 1, 1)):
    w = fresh(grid)
    if voxels:
        w.set_voxels_local(0, voxels)
    w.pack()
    w.clear_patches()
    return w


def test_sync_dirty_chunk_partial_minimal_nodes():  # world_map.rs:744-796
    w = packed_world([(0, 0, 0, STONE_DEEP), (4, 0, 0, STONE_DEEP), (8, 0, 0, STONE_DEEP), (12, 0, 0, STONE_DEEP)])
    n_initial = w.chunk_info(0)["n_nodes"]
    assert n_initial > 4
    node_base = w.chunk_info(0)["alloc"][0]
    assert node_base > 0
    rep = w.set_voxels_local(0, [(0, 0, 1, STONE_DEEP)])
    assert not rep.had_structural_growth and rep.has_dirty_nodes() and not rep.has_dirty_bricks()
    w.sync_chunk_partial(0, rep)
    ps = w.patch_list()
    assert len(ps) == 1
    p = ps[0]
    assert p["chunk_flat_idx"] == 0 and p["new_root_node_idx"] == node_base
    assert p["node_offset"] == node_base + rep.min_dirty_node
    assert len(p["nodes"]) == rep.max_dirty_node - rep.min_dirty_node + 1 <= n_initial
    assert len(p["bricks"]) == 0 and p["brick_offset"] == 0
    _, _, nodes, _ = w.pools()
    assert nodes[p["node_offset"]:p["node_offset"] + len(p["nodes"])].tobytes() == p["nodes"].tobytes()


def test_sync_dirty_chunk_partial_minimal_bricks():  # world_map.rs:798-841
    w = packed_world([(0, 0, 0, STONE_DEEP), (0, 0, 1, CRYSTAL_CYAN), (16, 0, 0, STONE_DEEP),
                      (16, 0, 1, CRYSTAL_CYAN)])
    assert w.chunk_info(0)["n_bricks"] == 2
    brick_base = w.chunk_info(0)["alloc"][2]
    rep = w.set_voxels_local(0, [(16, 0, 2, STONE_DEEP)])
    assert not rep.had_structural_growth and (rep.min_dirty_brick, rep.max_dirty_brick) == (1, 1)
    w.sync_chunk_partial(0, rep)
    ps = w.patch_list()
    assert len(ps) == 1 and len(ps[0]["bricks"]) == 1 and ps[0]["brick_offset"] == brick_base + 1
    _, bricks = w.chunk_tree(0)
    assert ps[0]["bricks"][0].tobytes() == bricks[1].tobytes()
    _, _, _, gb = w.pools()
    assert gb[brick_base + 1].tobytes() == bricks[1].tobytes()


def test_sync_dirty_chunk_partial_structural_growth_fallback():  # world_map.rs:843-862
    w = packed_world([])
    rep = w.set_voxels_local(0, [(10, 10, 10, STONE_DEEP)])
    assert rep.had_structural_growth
    w.sync_chunk_partial(0, rep)
    ps = w.patch_list()
    assert len(ps) == 1 and len(ps[0]["nodes"]) == w.chunk_info(0)["n_nodes"]


def test_sync_dirty_chunk_partial_structural_growth_within_capacity():  # world_map.rs:864-898
    w = packed_world([])
    w.set_voxels_local(0, [(0, 0, 0, STONE_DEEP), (48, 48, 48, STONE_DEEP)])
    w.sync_chunk(0)
    w.clear_patches()
    nb, ncap, _, _ = w.chunk_info(0)["alloc"]
    assert nb > 0 and ncap >= 128
    rep = w.set_voxels_local(0, [(0, 0, 4, STONE_DEEP)])
    assert rep.had_structural_growth and w.chunk_info(0)["n_nodes"] <= ncap
    w.sync_chunk_partial(0, rep)
    ps = w.patch_list()
    assert len(ps) == 1 and ps[0]["new_root_node_idx"] == nb
    assert len(ps[0]["nodes"]) == rep.max_dirty_node - rep.min_dirty_node + 1


def test_set_voxels_global_partial_sync():  # world_map.rs:900-926
    w = fresh((2, 1, 1))
    w.set_voxels_local(0, [(5, 5, 5, STONE_DEEP), (5, 5, 6, CRYSTAL_CYAN)])
    w.pack()
    w.clear_patches()
    w.set_voxels([(5, 5, 5, CRYSTAL_CYAN)])
    ps = w.patch_list()
    assert len(ps) == 1 and ps[0]["chunk_flat_idx"] == 0 and len(ps[0]["bricks"]) == 1
    assert w.voxel(5, 5, 5) == CRYSTAL_CYAN


def test_chunk_compaction_and_gpu_pool_reuse_and_patch_dedup():  # world_map.rs:928-1008
    w = packed_world([(0, 0, 0, STONE_DEEP), (0, 0, 1, CRYSTAL_CYAN), (16, 16, 16, STONE_DEEP),
                      (16, 16, 17, CRYSTAL_CYAN), (32, 32, 32, STONE_DEEP)])
    nb0, _, bb0, _ = w.chunk_info(0)["alloc"]
    _, _, n0, b0 = w.pools()
    assert nb0 > 0 and bb0 > 0
    w.clear_voxels([(32, 32, 32)])
    before = w.chunk_info(0)["n_nodes"]
    w.compact_chunk(0)
    assert w.chunk_info(0)["n_nodes"] <= before
    for i in range(5):
        w.set_voxels([(0, 0, 2 + i, STONE_DEEP)])
    _, _, n1, b1 = w.pools()
    assert len(n1) <= len(n0) + 256 and len(b1) <= len(b0) + 128
    w.clear_patches()
    w.sync_chunk(0)
    w.sync_chunk(0)
    assert sum(p["chunk_flat_idx"] == 0 for p in w.patch_list()) == 1
    w.clear_voxels([(0, 0, 0), (0, 0, 1), (16, 16, 16), (16, 16, 17), (0, 0, 2), (0, 0, 3), (0, 0, 4), (0, 0, 5),
                    (0, 0, 6)])
    w.sync_chunk(0)
    assert mask(w.chunk_tree(0)[0][0]) == 0 and w.chunk_info(0)["alloc"] == (0, 0, 0, 0)
    _, _, n2, _ = w.pools()
    w.set_voxels([(1, 1, 1, STONE_DEEP)])
    assert w.chunk_info(0)["alloc"][0] > 0
    _, _, n3, _ = w.pools()
    assert len(n3) == len(n2)  # reclaimed span reused, no tail growth


def reachable(roots, nodes):
    """Indices of the nodes and bricks the chunk roots reach (a tree walk of the packed pools)."""
    ml = nodes["mask_low"].astype(np.uint64) | (nodes["mask_high"].astype(np.uint64) << np.uint64(32))
    ptr = nodes["child_ptr"].astype(np.uint64)
    seen_n, seen_b = [], []
    frontier = np.array([r for r in roots if r != 0xFFFFFFFF], np.int64)
    while len(frontier):
        seen_n.append(frontier)
        m, p = ml[frontier], ptr[frontier]
        live = m != 0
        leaf = (p & np.uint64(0x80000000)) != 0
        uni = (p & np.uint64(0x40000000)) != 0
        base = (p & np.uint64(0x3FFFFFFF)).astype(np.int64)
        seen_b.append(base[live & leaf & ~uni])
        inner = live & ~leaf
        cnt = np.array([bin(int(x)).count("1") for x in m[inner]], np.int64)
        starts = base[inner]
        frontier = np.repeat(starts, cnt) + (np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt))
    n = np.unique(np.concatenate(seen_n)) if seen_n else np.zeros(0, np.int64)
    b = np.unique(np.concatenate(seen_b)) if seen_b else np.zeros(0, np.int64)
    return n, b


def same_reachable(roots, nodes, bricks, roots2, nodes2, bricks2):
    """Two pool images encode the same world: equal roots and equal reachable nodes / bricks (freed
    spans may differ: a full chunk sync drops the chunk's earlier partial patches, world_map.rs:302)."""
    if not np.array_equal(roots, roots2):
        return False
    n, b = reachable(roots, nodes)
    return nodes[n].tobytes() == nodes2[n].tobytes() and bricks[b].tobytes() == bricks2[b].tobytes()


# ---- dense ground truth ---------------------------------------------------------------------
def dense_carve(dense, center_m, radius_m):
    """Carve a dense [z, y, x] grid with carve_sphere_tracked's binary32 sphere test (no rim)."""
    c = (np.asarray(center_m, np.float32) * np.float32(16.0)).astype(np.float32)
    r = np.float32(radius_m) * np.float32(16.0)
    lo = np.maximum(np.floor(c - r), 0).astype(np.int64)
    hi = np.minimum(np.ceil(c + r), np.array(dense.shape[::-1]) - 1).astype(np.int64)
    z, y, x = np.mgrid[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
    dx = c[0] - (x.astype(np.float32) + np.float32(0.5))
    dy = c[1] - (y.astype(np.float32) + np.float32(0.5))
    dz = c[2] - (z.astype(np.float32) + np.float32(0.5))
    d2 = (dx * dx + dy * dy) + dz * dz
    sub = dense[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
    sub[(d2 <= r * r) & (sub != WATER)] = AIR


def edit_world(grid=(2, 2, 2), seed=3):
    w = A.HostWorld(grid)
    w.generate(A.gen_params(A.ARC_GEN_CAVES, seed=seed, n_cuboids=60, height_min=96.0, height_max=352.0,
                            wavelength=192.0))
    w.pack()
    return w


def craters(rng, grid, k):
    ext = np.array(grid, np.float64) * 16.0
    out = []
    for _ in range(k):
        c = rng.uniform(0.0, 1.0, 3) * ext
        c[1] = rng.uniform(4.0, 0.6 * ext[1])
        out.append((tuple(float(v) for v in c.astype(np.float32)), float(np.float32(rng.uniform(4, 32) / 16.0))))
    return out


def test_carves_match_dense_ground_truth():
    w = edit_world()
    dense = O.decode_dense(O.World(*w.pools()))
    rng = np.random.default_rng(5)
    for center, radius in craters(rng, (2, 2, 2), 12):
        before = int((dense != AIR).sum())
        d = w.carve_crater(center, radius)
        dense_carve(dense, center, radius)
        assert before - int((dense != AIR).sum()) == d
    got = O.decode_dense(O.World(*w.pools()))
    assert np.array_equal(got, dense)
    # the pools stay a valid world: every tree passes the upload validation walk (no GPU needed)
    for i in range(8):
        info = w.chunk_info(i)
        assert info["solid"] == int((dense[(i // 4) * 256:(i // 4 + 1) * 256, ((i // 2) % 2) * 256:((i // 2) % 2 + 1) * 256,
                                           (i % 2) * 256:(i % 2 + 1) * 256] != AIR).sum())


def test_set_voxels_match_dense_ground_truth_and_traversal_agrees_with_brute_force():
    w = edit_world(grid=(1, 1, 1), seed=4)
    dense = O.decode_dense(O.World(*w.pools()))
    rng = np.random.default_rng(9)
    n = 4000
    xyz = rng.integers(0, 256, (n, 3))
    mats = rng.choice([AIR, AIR, 3, 7, 19, 23], n)
    vox = np.concatenate([xyz, mats[:, None]], 1).astype(np.int32)
    w.set_voxels(vox)
    for (x, y, z), m in zip(xyz, mats):
        dense[z, y, x] = m
    pools = w.pools()
    assert np.array_equal(O.decode_dense(O.World(*pools)), dense)
    rays = W.random_rays(3000, (1, 1, 1), seed=2, inside=False)
    full, _ = O.trace_world(O.World(*pools), A.TraceParams.exact(), rays)
    bf = O.brute(dense, rays)
    ok = bf["tie"] == 0
    assert np.array_equal(full["hit"][ok] & 1, bf["hit"][ok])
    h = ok & (bf["hit"] == 1)
    assert np.array_equal(full["vx"][h], bf["vx"][h]) and np.array_equal(full["vz"][h], bf["vz"][h])


# ---- GPU: patches and traversal after edits ---------------------------------------------------
@pytest.mark.gpu
def test_gpu_patches_mirror_host_pools_and_trace_bit_exact(ctx):
    w = edit_world()
    pools = w.pools()
    ctx.upload(*pools)
    w.clear_patches()
    rng = np.random.default_rng(11)
    cam = A.camera_look((13.7, 30.2, 11.9), 0.6, -0.55, 65.0, 320, 180)
    prays, _ = A.primary_rays(cam, 320, 180, 250.0)
    for frame in range(4):
        for center, radius in craters(rng, (2, 2, 2), 6):
            w.carve_crater(center, radius)
        w.set_voxels(np.concatenate([rng.integers(0, 512, (300, 3)), rng.choice([AIR, 3, 19], (300, 1))], 1))
        print(f"[edit] frame {frame}: {len(w.patch_list())} patches", flush=True)
        ctx.patch_world(w)
        grid, roots, nodes, bricks = w.pools()
        droots, dnodes, dbricks = ctx.read_pools(len(roots), len(nodes), len(bricks))
        assert same_reachable(roots, nodes, bricks, droots, dnodes, dbricks)
        got = ctx.trace(A.TraceParams.exact(), prays)
        ref, _, _ = W.oracle_hits((grid, roots, nodes, bricks), A.TraceParams.exact(), prays)
        bad = W.diff_records(got, ref)
        assert len(bad) == 0, f"frame {frame}: {len(bad)} records differ"
        brays, _ = A.bounce_rays(prays, got, seed=frame, max_dist_m=26.0)
        p2 = A.TraceParams.exact(2, ignore_water=True)
        got2 = ctx.trace(p2, brays)
        ref2, _, _ = W.oracle_hits((grid, roots, nodes, bricks), p2, brays)
        assert len(W.diff_records(got2, ref2)) == 0


@pytest.mark.gpu
def test_gpu_patch_rejects_out_of_range_pointers(ctx):
    w = edit_world(grid=(1, 1, 1), seed=2)
    ctx.upload(*w.pools())
    node = np.zeros(1, A.NODE_DTYPE)
    node["mask_low"] = 1
    node["child_ptr"] = (1 << 29)  # internal node whose child lies far outside the pool
    with pytest.raises(A.ArcError):
        ctx.patch([{"chunk_flat_idx": 0, "new_root_node_idx": 1, "node_offset": 1, "nodes": node,
                    "brick_offset": 0, "bricks": np.zeros(0, A.BRICK_DTYPE)}])
    brick_leaf = np.zeros(1, A.NODE_DTYPE)
    brick_leaf["mask_low"] = 1
    brick_leaf["child_ptr"] = A.native.ARC_IS_LEAF_FLAG | ((1 << 30) - 1)
    with pytest.raises(A.ArcError):
        ctx.patch([{"chunk_flat_idx": 0, "new_root_node_idx": 1, "node_offset": 1, "nodes": brick_leaf,
                    "brick_offset": 0, "bricks": np.zeros(0, A.BRICK_DTYPE)}])


def test_patch_plan_reproduces_host_pools_frame_by_frame():
    """The patch batch of each frame, applied to last frame's pools by the same plan and scatter
    order arc_world_patch uses (host model), gives this frame's world: equal roots and equal
    reachable nodes and bricks."""
    w = edit_world()
    _, roots, nodes, bricks = w.pools()
    w.clear_patches()
    cap_n, cap_b = len(nodes) * 2, len(bricks) * 2
    dn = np.zeros(cap_n, A.NODE_DTYPE)
    db = np.zeros(cap_b, A.BRICK_DTYPE)
    dn[:len(nodes)] = nodes
    db[:len(bricks)] = bricks
    dr = roots.copy()
    rng = np.random.default_rng(11)
    for frame in range(4):
        for center, radius in craters(rng, (2, 2, 2), 6):
            w.carve_crater(center, radius)
        w.set_voxels(np.concatenate([rng.integers(0, 512, (300, 3)), rng.choice([AIR, 3, 19], (300, 1))], 1))
        assert w.apply_patches_to(dr, dn, db) > 0
        w.clear_patches()
        _, roots, nodes, bricks = w.pools()
        assert same_reachable(roots, nodes, bricks, dr, dn[:len(nodes)], db[:len(bricks)])


def test_patch_plan_last_writer_wins_and_rejects_bad_pointers():
    nodes = np.zeros(64, A.NODE_DTYPE)
    bricks = np.zeros(16, A.BRICK_DTYPE)
    roots = np.full(2, 0xFFFFFFFF, np.uint32)

    def patch(chunk, root, off, vals):
        a = np.zeros(len(vals), A.NODE_DTYPE)
        a["lod_payload"] = vals
        return a, A.native.GpuPatch(chunk, root, off, len(a), a.ctypes.data, 0, 0, None)

    keep = [patch(0, 5, 10, [1, 2, 3, 4, 5, 6]), patch(1, 7, 12, [7, 8]), patch(0, 9, 8, [9, 10, 11])]
    arr = (A.native.GpuPatch * 3)(*[k[1] for k in keep])
    rc = A.lib().arc_world_patch_host(arr, 3, 2, roots.ctypes.data, nodes.ctypes.data, 64, bricks.ctypes.data, 16)
    assert rc == 0
    assert nodes["lod_payload"][8:17].tolist() == [9, 10, 11, 2, 7, 8, 5, 6, 0]
    assert roots.tolist() == [9, 7]
    bad, p = patch(0, 1, 1, [0])
    bad["mask_low"] = 1
    bad["child_ptr"] = 60  # children at 60.. with one child: in range
    arr = (A.native.GpuPatch * 1)(p)
    assert A.lib().arc_world_patch_host(arr, 1, 2, roots.ctypes.data, nodes.ctypes.data, 64, bricks.ctypes.data,
                                        16) == 0
    bad["child_ptr"] = 64  # first child one past the pool
    assert A.lib().arc_world_patch_host(arr, 1, 2, roots.ctypes.data, nodes.ctypes.data, 64, bricks.ctypes.data,
                                        16) == A.native.ARC_ERR_INVALID
    bad["child_ptr"] = A.native.ARC_IS_LEAF_FLAG | 16  # brick one past the pool
    assert A.lib().arc_world_patch_host(arr, 1, 2, roots.ctypes.data, nodes.ctypes.data, 64, bricks.ctypes.data,
                                        16) == A.native.ARC_ERR_INVALID
