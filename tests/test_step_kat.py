"""Hand-derived known answers for the traversal outputs that only the reference's code defines:
the DDA step count, LOD stops (cell, material), contact AO and lod_rgb.

Each case is a constructed world and one ray whose reference iterations are worked out by hand
below, iteration by iteration, from the loop of trace_world_64tree_from
(shaders/common/traversal64.wgsl:302-553) and raycast_chunk_tree
(crates/arcadia-core/src/raycast_cpu.rs:77-216):
  * an iteration passes the guards (step < max_steps, t < t_limit, voxel inside the world), then
    restarts at the root on a chunk change, else at the cached ancestor of level
    max(dl, min_shift, cached_shift) where dl is the highest changed voxel bit rounded down to an
    even level (traversal64.wgsl:374-386);
  * it descends while the child bit of the current voxel is set and the node is neither a leaf nor
    at the LOD floor; an absent child ends the iteration with a box step over that cell
    (:419-436); a set bit at a leaf or at the floor is a hit (material from the brick, or from
    lod_payload >> 24 for uniform leaves and LOD stops, :439-449);
  * an empty chunk costs one iteration, a box step over its 256^3 box (:337-352).
The step count is the number of iterations that pass the guards (a hit on the first iteration
gives 1). Rays are axis-aligned with voxel-centre offsets, so every plane time is an exact binary32
value; the zero direction components become 1e-7 (world: select(d, 1e-7, |d| < 1e-7); chunk:
copysign(1e-7, d)) and their plane times are ~1e7 voxels away, so they never decide a step.

The same rays run through the oracle (CPU, always) and through the HIP kernels (marked gpu).
"""
from __future__ import annotations

import numpy as np
import pytest

import arcadia_amd as A
from arcadia_amd import native as N
from oracle import oracle as O

F = np.float32
FACE_PX, FACE_NX, FACE_PY, FACE_NY = 0, 1, 2, 3


def sparse_pools(grid, voxels_by_chunk):
    w = A.HostWorld(grid)
    for chunk, vox in voxels_by_chunk.items():
        w.set_chunk_sparse(chunk, vox)
    w.pack()
    return w.pools()


def floor_pools(height, mat, grid=(1, 1, 1)):
    dense = np.zeros((256, 256, 256), np.uint8)
    dense[:, :height, :] = mat
    w = A.HostWorld(grid)
    w.set_chunk_dense(0, dense)
    w.pack()
    return w.pools()


def ray_vox(o, d, t_max_vox):
    """One ray given in voxel units, returned in metres (world semantics)."""
    r = np.zeros(1, A.RAY_DTYPE)
    r["ox"], r["oy"], r["oz"] = (F(x) / F(16.0) for x in o)
    r["dx"], r["dy"], r["dz"] = d
    r["t_max"] = F(t_max_vox) / F(16.0)
    return r


def chunk_ray(r):
    """The same ray in voxel units (raycast_chunk_tree semantics)."""
    c = r.copy()
    for k in ("ox", "oy", "oz", "t_max"):
        c[k] = c[k] * F(16.0)
    return c


def lod_params(cls, dist=1e30, step=0xFFFFFFFF):
    p = A.TraceParams.exact(cls)
    for i in range(3):
        p.lod_dist_m[i] = 1.0e30
        p.lod_step[i] = 0xFFFFFFFF
    p.lod_dist_m[cls] = dist
    p.lod_step[cls] = step
    return p


def node_at(pools, voxel, shift):
    """The node of level `shift` on the path to `voxel` (its children are cells of 2^shift voxels),
    found by walking the packed pools with the popcount child rule; an independent reading of
    node64.rs:74-90, used for the expected lod_payload."""
    grid, roots, nodes, _ = pools
    nodes = np.asarray(nodes).view(A.NODE_DTYPE)
    c = [v >> 8 for v in voxel]
    idx = int(roots[c[0] + c[1] * grid[0] + c[2] * grid[0] * grid[1]])
    sh = 6
    while sh > shift:
        n = nodes[idx]
        b = (((voxel[2] >> sh) & 3) << 4) | (((voxel[1] >> sh) & 3) << 2) | ((voxel[0] >> sh) & 3)
        mask = int(n["mask_low"]) | (int(n["mask_high"]) << 32)
        assert mask >> b & 1 and not n["child_ptr"] & N.ARC_IS_LEAF_FLAG
        idx = (int(n["child_ptr"]) & N.ARC_PTR_INDEX_MASK) + bin(mask & ((1 << b) - 1)).count("1")
        sh -= 2
    return nodes[idx]


# ---- the cases ------------------------------------------------------------------------------------
# One material-7 voxel at (102, 21, 22) in a single chunk; rays along x through its row.
V1 = (102, 21, 22)
W1 = ("single", (1, 1, 1), {0: [(102, 21, 22, 7)]})

CASES = {
    # +x from voxel 0 (t = 0 at x = 0.5):
    #   it0 root (cell x 0..63) empty              -> step 64   to x  64, t  63.5
    #   it1 diff 64 (dl 6): root, cell 1 set, down; level-4 cell 64..79 empty  -> x  80, t  79.5
    #   it2 diff 16 (dl 4): level 4, cell 80..95 empty                        -> x  96, t  95.5
    #   it3 diff 48 (dl 4): level 4, cell 96..111 set, down; level-2 cell 96..99 empty -> x 100, t 99.5
    #   it4 diff 4 (dl 2): level 2, cell 100..103 set, down; leaf voxel 100 empty -> x 101, t 100.5
    #   it5 diff 1 (dl 0): leaf, voxel 101 empty  -> x 102, t 101.5
    #   it6 diff 3 (dl 0): leaf, voxel 102 set: hit. 7 steps, t = entry plane 102 - 0.5, face -x.
    "voxel_px": dict(world=W1, o=(0.5, 21.5, 22.5), d=(1.0, 0.0, 0.0), tmax=400.0, params=None,
                     hit=True, steps=7, t=101.5, voxel=V1, face=FACE_NX, mat=7, lod=False, chunk=True),
    # -x from voxel 250 (t = 0 at x = 250.5); a negative step exits at the cell's low plane:
    #   it0 x250 root cell 192..255 empty                         -> plane 192: x 191, t  58.5
    #   it1 x191 (diff 69, dl 6) root cell 128..191 empty         -> plane 128: x 127, t 122.5
    #   it2 x127 (diff 192, dl 6) root cell 64..127 set, level 4 cell 112..127 empty -> x 111, t 138.5
    #   it3 x111 (diff 16, dl 4) level 4 cell 96..111 set, level 2 cell 108..111 empty -> x 107, t 142.5
    #   it4 x107 (diff 4, dl 2) level 2 cell 104..107 empty        -> plane 104: x 103, t 146.5
    #   it5 x103 (diff 12, dl 2) level 2 cell 100..103 set, leaf voxel 103 empty -> x 102, t 147.5
    #   it6 x102 (diff 1, dl 0) leaf voxel 102 set: hit. 7 steps, t = entry plane 103: 250.5 - 103,
    #   face +x.
    "voxel_nx": dict(world=W1, o=(250.5, 21.5, 22.5), d=(-1.0, 0.0, 0.0), tmax=400.0, params=None,
                     hit=True, steps=7, t=147.5, voxel=V1, face=FACE_PX, mat=7, lod=False, chunk=True),
    # step cap: max_steps 6 stops voxel_px before its hit iteration (index 6): 6 steps, no hit;
    # max_steps 7 lets it hit.
    "cap6": dict(world=W1, o=(0.5, 21.5, 22.5), d=(1.0, 0.0, 0.0), tmax=400.0,
                 params=("exact", 6), hit=False, steps=6),
    "cap7": dict(world=W1, o=(0.5, 21.5, 22.5), d=(1.0, 0.0, 0.0), tmax=400.0,
                 params=("exact", 7), hit=True, steps=7, t=101.5, voxel=V1, face=FACE_NX, mat=7, lod=False),
    # t limit 90: after it2 t = 95.5 >= 90, the guard of it3 fails: 3 steps, no hit.
    "tlimit": dict(world=W1, o=(0.5, 21.5, 22.5), d=(1.0, 0.0, 0.0), tmax=90.0, params=None,
                   hit=False, steps=3),
    # Floor y < 16 of material 5 over the whole chunk: the 16^3 cells collapse to full uniform
    # nodes, and each level-4 node (a 64^3 cell) whose only children are those becomes a uniform
    # leaf with a 16-bit mask (builder.rs:277-289), so a hit resolves on a 16^3 cell.
    # Down from (40.5, 200.5, 40.5):
    #   it0 y200 root c3 empty -> 191 (t 8.5); it1 y191 root c2 empty -> 127 (72.5)
    #   it2 y127 root c1 empty -> 63 (136.5)
    #   it3 y63 root c0 set, level 4 (uniform leaf) cell y 48..63 absent -> 47 (152.5)
    #   it4 y47 (diff 16, dl 4) level 4 cell 32..47 absent -> 31 (168.5)
    #   it5 y31 (diff 48, dl 4) level 4 cell 16..31 absent -> 15 (184.5)
    #   it6 y15 (diff 16, dl 4) level 4 cell 0..15 set at a leaf: hit on the 16^3 cell, material 5.
    #   7 steps, t = entry plane y 16: 200.5 - 16, face +y, LOD flag (shift 4 > 0).
    "floor16": dict(world=("floor", 16, 5), o=(40.5, 200.5, 40.5), d=(0.0, -1.0, 0.0), tmax=400.0, params=None,
                    hit=True, steps=7, t=184.5, voxel=(40, 15, 40), face=FACE_PY, mat=5, lod=True, chunk=True,
                    rgb_node=((40, 15, 40), 4)),
    # Two chunks along x, the first empty: one iteration crosses it in a single 256 box (t 255.5),
    # then the chunk change restarts at the second chunk's root and the voxel_px walk repeats 256
    # voxels further on (7 iterations): 1 + 7 = 8 steps, t = 358 - 0.5.
    "empty_chunk_then_voxel": dict(world=("two", (2, 1, 1), {1: [(102, 21, 22, 7)]}), o=(0.5, 21.5, 22.5),
                                   d=(1.0, 0.0, 0.0), tmax=800.0, params=None, hit=True, steps=8, t=357.5,
                                   voxel=(358, 21, 22), face=FACE_NX, mat=7, lod=False),
    # The other way: from x 511.5 in the empty second chunk, through the first to the voxel:
    #   it0 empty chunk 1 -> plane 256: x 255 (t 255.5)
    #   it1 x255 (chunk change) root cell 192..255 empty -> x 191 (t 319.5)
    #   then voxel_nx from its it1 on (x 191 ... hit at 102): 6 more iterations. 8 steps,
    #   t = 511.5 - 103.
    "empty_neighbour_back": dict(world=("two", (2, 1, 1), {0: [(102, 21, 22, 7)]}), o=(511.5, 21.5, 22.5),
                                 d=(-1.0, 0.0, 0.0), tmax=800.0, params=None, hit=True, steps=8, t=408.5,
                                 voxel=V1, face=FACE_PX, mat=7, lod=False),
    # LOD floor 2 from distance (class 0, lod_dist 0 m: t > 0 from it1 on; it0 has t = 0):
    #   it0..it3 as voxel_px (level 4 > floor 2 keeps descending; it3 ends at level 2, x 100)
    #   it4 restart level max(dl 2, floor 2, cached 2) = 2: cell 100..103 set at the floor: stop.
    #   5 steps, the 4^3 cell's entry t 99.5, voxel (100, 21, 22), material = that level-2 node's
    #   lod_payload >> 24 (7: its only material), LOD flag.
    "lod2_dist": dict(world=W1, o=(0.5, 21.5, 22.5), d=(1.0, 0.0, 0.0), tmax=400.0, params=("lod", 0, 0.0, None),
                      hit=True, steps=5, t=99.5, voxel=(100, 21, 22), face=FACE_NX, mat=7, lod=True,
                      rgb_node=((100, 21, 22), 2)),
    # LOD floor 4 (class 1, secondary distance 0 m, so t > 0 * 2.2857 from it1 on):
    #   it0 root c0 empty -> 64; it1 root c1 set, level 4 (= floor, set bits stop) cell 64..79 empty
    #   -> 80; it2 level 4 cell 80..95 empty -> 96; it3 level 4 cell 96..111 set at the floor: stop.
    #   4 steps, t 95.5 (plane 96), voxel (96, 21, 22), material of the level-4 node (7).
    "lod4_dist": dict(world=W1, o=(0.5, 21.5, 22.5), d=(1.0, 0.0, 0.0), tmax=400.0, params=("lod", 1, 0.0, None),
                      hit=True, steps=4, t=95.5, voxel=(96, 21, 22), face=FACE_NX, mat=7, lod=True,
                      rgb_node=((96, 21, 22), 4)),
    # LOD floor 2 from the step count (class 2, lod_step 4: floor 2 once the iteration index > 4):
    #   it0..it4 as voxel_px (it4 ends in the leaf at x 101, cached level 0)
    #   it5 (index 5 > 4): restart level max(dl 0, floor 2, cached 0) = 2, the cached level-2 node:
    #   cell 100..103 set at the floor: stop. 6 steps, voxel (101, 21, 22); t is the entry of the
    #   4^3 cell, 99.5, behind the current t of 100.5 (exact_t = max(cell_enter, 0)).
    "lod2_steps": dict(world=W1, o=(0.5, 21.5, 22.5), d=(1.0, 0.0, 0.0), tmax=400.0, params=("lod", 2, None, 4),
                       hit=True, steps=6, t=99.5, voxel=(101, 21, 22), face=FACE_NX, mat=7, lod=True,
                       rgb_node=((101, 21, 22), 2)),
}


def build_world(spec):
    kind = spec[0]
    if kind == "single" or kind == "two":
        return sparse_pools(spec[1], spec[2])
    if kind == "floor":
        return floor_pools(spec[1], spec[2])
    raise ValueError(kind)


def params_of(spec):
    if spec is None:
        return A.TraceParams.exact(0)
    if spec[0] == "exact":
        return A.TraceParams.exact(0, max_steps=spec[1])
    _, cls, dist, step = spec
    return lod_params(cls, 1e30 if dist is None else dist, 0xFFFFFFFF if step is None else step)


_worlds = {}


def case_inputs(name):
    c = CASES[name]
    key = repr(c["world"])
    if key not in _worlds:
        _worlds[key] = build_world(c["world"])
    return _worlds[key], params_of(c["params"]), ray_vox(c["o"], c["d"], c["tmax"])


def check_record(name, h, full=None, pools=None):
    c = CASES[name]
    assert int(h["steps"]) == c["steps"], (name, int(h["steps"]))
    assert bool(h["flags"] & N.ARC_HIT_FLAG_HIT) == c["hit"], name
    if not c["hit"]:
        assert h["face"] == N.ARC_FACE_NONE
        return
    assert h["t_voxels"] == F(c["t"]), (name, float(h["t_voxels"]))
    assert (int(h["vx"]), int(h["vy"]), int(h["vz"])) == c["voxel"], name
    assert int(h["face"]) == c["face"] and int(h["mat"]) == c["mat"], name
    assert bool(h["flags"] & N.ARC_HIT_FLAG_LOD) == c["lod"], name
    if full is not None and "rgb_node" in c:
        n = node_at(pools, *c["rgb_node"])
        lp = int(n["lod_payload"])
        assert lp >> 24 == c["mat"]
        want = [F(F((lp >> s) & 0xFF) / F(255.0)) for s in (0, 8, 16)]  # unpack_rgb_lod, traversal64.wgsl:50-55
        assert list(full["lod_rgb"]) == want, (name, full["lod_rgb"], want)
        assert full["contact_ao"] == F(1.0)  # AO only for level-0 hits


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_step_kat(name):
    pools, params, ray = case_inputs(name)
    full, cnt = O.trace_world(O.World(*pools), params, ray)
    check_record(name, O.compact(full)[0], full[0], pools)
    assert int(cnt["steps"][0]) == CASES[name]["steps"]
    if CASES[name].get("chunk"):  # raycast_chunk_tree: same walk in voxel units
        fc, _ = O.trace_chunk(O.World(*pools), chunk_ray(ray))
        check_record(name, O.compact(fc)[0])


# ---- contact AO (traversal64.wgsl:60-131, 501-521) --------------------------------------------------
# A level-0 hit on voxel (41, 41, 41), local (1, 1, 1) of the leaf covering 40..43, seen from above:
# the ray (41.25, y0, 41.75) -> (0, -1, 0) hits the top face at (41.25, 42, 41.75), so the AO
# evaluation reads the air cell (1, 2, 1) in front of the face (bit 25) and, for a +y face, its
# -x / +x / -z / +z neighbours in the same leaf (bits 24, 26, 9, 41), weighted by
# 1 - fract(x) = 0.75, fract(x) = 0.25, 1 - fract(z) = 0.25, fract(z) = 0.75, each times 0.42;
# raw AO = clamp(1 - occ, 0.35, 1). Within 12 m (half the default 24 m range) the distance weight
# is 1 and contact_ao = mix(1, raw, 1) = raw.
AO_NEIGH = {"-x": ((40, 42, 41), 0.75), "+x": ((42, 42, 41), 0.25), "-z": ((41, 42, 40), 0.25),
            "+z": ((41, 42, 42), 0.75)}
AO_CASES = [(), ("-x",), ("+x", "+z"), ("-z",), ("-x", "+x", "-z", "+z")]


def ao_expected(neigh):
    occ = F(0.0)
    for k in ("-x", "+x", "-z", "+z"):  # the shader's order of the four tests
        if k in neigh:
            occ = F(occ + F(F(AO_NEIGH[k][1]) * F(0.42)))
    return F(min(max(F(F(1.0) - occ), F(0.35)), F(1.0)))


def ao_world(neigh, grid=(1, 1, 1)):
    vox = [(41, 41, 41, 9)] + [(*AO_NEIGH[k][0], 9) for k in neigh]
    return sparse_pools(grid, {0: vox})


def test_ao_closed_forms_are_distinct():
    vals = [float(ao_expected(n)) for n in AO_CASES]
    assert vals[0] == 1.0 and vals[-1] == pytest.approx(0.35) and len(set(vals)) == len(vals)


@pytest.mark.parametrize("neigh", AO_CASES, ids=lambda n: "+".join(n) or "none")
def test_oracle_contact_ao(neigh):
    pools = ao_world(neigh)
    ray = ray_vox((41.25, 100.5, 41.75), (0.0, -1.0, 0.0), 400.0)
    full, _ = O.trace_world(O.World(*pools), A.TraceParams.exact(0), ray)
    r = full[0]
    assert r["hit"] & 1 and (r["hit"] >> 16) == 0 and r["t_voxels"] == F(58.5)
    assert list(r["normal"]) == [0.0, 1.0, 0.0]
    assert r["contact_ao"] == ao_expected(neigh), (neigh, r["contact_ao"])
    # lod_bias != 0: no AO (traversal64.wgsl:501)
    f1, _ = O.trace_world(O.World(*pools), A.TraceParams.exact(1), ray)
    assert f1[0]["contact_ao"] == F(1.0)


def test_oracle_contact_ao_distance_weight():
    """At 16 m (between ao_max / 2 = 12 and ao_max = 24) the weight is
    1 - smoothstep(12, 24, 16) = 1 - (1/3)^2 (3 - 2/3) = 20/27, and contact_ao = mix(1, raw, w)."""
    neigh = ("-x", "+z")
    pools = ao_world(neigh, grid=(1, 2, 1))
    ray = ray_vox((41.25, 298.0, 41.75), (0.0, -1.0, 0.0), 600.0)  # 256 voxels above the face
    full, _ = O.trace_world(O.World(*pools), A.TraceParams.exact(0), ray)
    r = full[0]
    raw = ao_expected(neigh)
    t = F(F(F(16.0) - F(12.0)) / F(F(24.0) - F(12.0)))
    w = F(F(1.0) - F(F(t * t) * F(F(3.0) - F(F(2.0) * t))))
    want = F(F(F(1.0) * F(F(1.0) - w)) + F(raw * w))
    assert r["t_voxels"] == F(256.0) and abs(float(w) - 20.0 / 27.0) < 1e-6
    assert r["contact_ao"] == want, (r["contact_ao"], want)
    assert float(raw) < float(want) < 1.0  # raw = 0.37 near, 1 at the far end of the range


def test_oracle_contact_ao_neighbour_outside_leaf():
    """A hit on the top face of a leaf's top layer: the air cell in front lies in the next leaf, so
    the evaluation returns 1 whatever the neighbours (traversal64.wgsl:80-83)."""
    vox = [(41, 43, 41, 9), (40, 44, 41, 9), (42, 44, 41, 9)]
    pools = sparse_pools((1, 1, 1), {0: vox})
    ray = ray_vox((41.25, 100.5, 41.75), (0.0, -1.0, 0.0), 400.0)
    full, _ = O.trace_world(O.World(*pools), A.TraceParams.exact(0), ray)
    assert full[0]["hit"] & 1 and full[0]["vy"] == 43 and full[0]["contact_ao"] == F(1.0)


def test_oracle_contact_ao_side_face():
    """A -x-pointing ray on the +x face of voxel (41, 41, 41): the vertical-X branch reads the air
    cell (2, 1, 1) (bit 22) and its -y / +y / -z / +z neighbours (bits 18, 26, 6, 38) with fract(y)
    and fract(z): set here are +y (weight fract(y) = 0.625) and -z (1 - fract(z) = 0.875); voxel
    (41, 42, 41) (bit 25) touches the hit voxel but is no neighbour of the air cell."""
    vox = [(41, 41, 41, 9), (42, 42, 41, 9), (42, 41, 40, 9), (41, 42, 41, 9)]  # last: not a neighbour of the face
    pools = sparse_pools((1, 1, 1), {0: vox})
    ray = ray_vox((100.5, 41.625, 41.125), (-1.0, 0.0, 0.0), 400.0)
    full, _ = O.trace_world(O.World(*pools), A.TraceParams.exact(0), ray)
    r = full[0]
    assert r["hit"] & 1 and (r["vx"], r["vy"], r["vz"]) == (41, 41, 41) and list(r["normal"]) == [1.0, 0.0, 0.0]
    occ = F(F(F(0.625) * F(0.42)) + F(F(0.875) * F(0.42)))
    assert r["contact_ao"] == F(F(1.0) - occ)


# ---- GPU: the same rays through the kernels --------------------------------------------------------

def _gpu_trace(ctx, pools, params, rays, full):
    import torch

    ctx.upload(*pools)
    d_r = torch.from_numpy(np.ascontiguousarray(rays).view(np.uint8)).cuda()
    if full:
        d_o = torch.zeros(len(rays) * 80, dtype=torch.uint8, device="cuda")
        ctx.trace_full_device(params, d_r.data_ptr(), len(rays), d_o.data_ptr())
        torch.cuda.synchronize()
        return d_o.cpu().numpy().view(A.RESULT_DTYPE)
    return ctx.trace(params, rays)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_step_kat(ctx, name):
    pools, params, ray = case_inputs(name)
    print(f"[kat] {name}", flush=True)
    ctx.upload(*pools)
    check_record(name, ctx.trace(params, ray)[0])
    full = _gpu_trace(ctx, pools, params, ray, True)
    ref_full, _ = O.trace_world(O.World(*pools), params, ray)
    assert full.tobytes() == ref_full.tobytes(), name
    check_record(name, O.compact(full)[0], full[0], pools)
    if CASES[name].get("chunk"):
        check_record(name, ctx.trace(A.TraceParams.chunk(), chunk_ray(ray))[0])


@pytest.mark.gpu
@pytest.mark.parametrize("neigh", AO_CASES, ids=lambda n: "+".join(n) or "none")
def test_gpu_contact_ao(ctx, neigh):
    pools = ao_world(neigh)
    ray = ray_vox((41.25, 100.5, 41.75), (0.0, -1.0, 0.0), 400.0)
    print(f"[kat] contact AO {neigh}", flush=True)
    full = _gpu_trace(ctx, pools, A.TraceParams.exact(0), ray, True)
    assert full[0]["contact_ao"] == ao_expected(neigh)
    assert list(full[0]["normal"]) == [0.0, 1.0, 0.0] and full[0]["t_voxels"] == F(58.5)
