// Host edit engine: in-place voxel edits of a chunk's 64-tree and minimal GPU patches.
//
// Restates, on the host world (World / ChunkTree), the reference's edit path:
//   ChunkEditReport            crates/arcadia-core/src/chunk_tree/edit/set.rs:17-80
//   set_voxels_local_tracked   set.rs:84-138
//   compact                    set.rs:156-209
//   can_reach_leaf             set.rs:212-241
//   mutate_brick_in_place      set.rs:244-322
//   remove_child               set.rs:325-352
//   ensure_child               set.rs:355-398
//   uncollapse_if_needed       set.rs:401-424
//   apply_voxel_entries_to_leaf set.rs:427-517
//   refresh_internal_lod       set.rs:520-570
//   compute_leaf_lod           set.rs:574-618
//   carve_sphere_tracked       crates/arcadia-core/src/chunk_tree/edit/carve.rs:39-177
//   clear_voxels_local_tracked crates/arcadia-core/src/chunk_tree/edit/clear.rs:24-53
//   get_voxel / get_brick_solid_mask / intersects_aabb_local
//                              crates/arcadia-core/src/chunk_tree/tree.rs:58-275
//   sync_dirty_chunk           crates/arcadia-world/src/world_map.rs:208-315
//   sync_dirty_chunk_partial   world_map.rs:319-383
//   clear_voxels_global / set_voxels_global  world_map.rs:549-621
//   find_free_span             world_map.rs:679-686
//   carve_crater (tree part)   crates/arcadia-world/src/destruction.rs:47-155
// The tree algorithms decide node order, collapse state and LOD payloads, which the traversal's
// step counts and LOD stops depend on, so they follow the reference step by step. Where the
// reference iterates a HashMap (clear_voxels_global's chunk order), chunks go in ascending index.
// Debris particles, island detection and the emissive registry are not part of this path.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "world_internal.h"

namespace arc {
namespace {

constexpr uint32_t kMatAir = 0, kMatWater = 18, kMatScorched = 23, kMatEmber = 24;
constexpr uint32_t kChunkDim = 256;

inline uint64_t mask_of(const arc_node64& n) { return (uint64_t)n.mask_low | ((uint64_t)n.mask_high << 32); }
inline void set_mask(arc_node64& n, uint64_t m) {
  n.mask_low = (uint32_t)m;
  n.mask_high = (uint32_t)(m >> 32);
}
inline arc_node64 node_of(uint64_t mask, uint32_t ptr, uint32_t lod) {
  arc_node64 n;
  set_mask(n, mask);
  n.child_ptr = ptr;
  n.lod_payload = lod;
  return n;
}
inline arc_node64 empty_node() { return node_of(0, 0, 0); }
inline bool is_leaf(const arc_node64& n) { return (n.child_ptr & ARC_IS_LEAF_FLAG) != 0; }
inline bool is_uniform(const arc_node64& n) { return (n.child_ptr & ARC_IS_UNIFORM_FLAG) != 0; }
inline uint32_t base_ptr(const arc_node64& n) { return n.child_ptr & ARC_PTR_INDEX_MASK; }
inline uint32_t primary_mat(const arc_node64& n) { return n.lod_payload >> 24; }
inline bool has_child(const arc_node64& n, uint32_t b) { return (mask_of(n) >> b) & 1u; }
inline uint32_t child_count(const arc_node64& n) { return (uint32_t)__builtin_popcountll(mask_of(n)); }
inline uint32_t child_slot(const arc_node64& n, uint32_t b) {
  const uint64_t prefix = b >= 64 ? ~0ull : ((1ull << b) - 1ull);
  return base_ptr(n) + (uint32_t)__builtin_popcountll(mask_of(n) & prefix);
}
inline uint32_t cbit(uint32_t x, uint32_t y, uint32_t z) { return (z << 4) | (y << 2) | x; }
inline uint32_t brick_get(const arc_brick64& b, uint32_t i) { return (b.packed_materials[i >> 2] >> ((i & 3) * 8)) & 0xFFu; }
inline void brick_set(arc_brick64& b, uint32_t i, uint32_t mat) {
  const uint32_t w = i >> 2, s = (i & 3) * 8;
  b.packed_materials[w] = (b.packed_materials[w] & ~(0xFFu << s)) | ((mat & 0xFFu) << s);
}
inline int ctz64(uint64_t v) { return __builtin_ctzll(v); }

// LeafBrick64::check_uniform (node64.rs:142-156); returns -1 when mixed.
int brick_check_uniform(const arc_brick64& b, uint64_t mask) {
  if (mask == 0) return (int)kMatAir;
  const uint32_t first = brick_get(b, (uint32_t)ctz64(mask));
  for (uint64_t rem = mask & (mask - 1); rem; rem &= rem - 1)
    if (brick_get(b, (uint32_t)ctz64(rem)) != first) return -1;
  return (int)first;
}

// pack_rgb_and_mat (builder.rs:73-78): (clamp(v,0,1) * 255) as u32
inline uint32_t unorm8_trunc(float v) {
  const float c = std::fmin(std::fmax(v, 0.0f), 1.0f);
  const float s = c * 255.0f;
  return (uint32_t)s;
}

// compute_leaf_lod (set.rs:574-618)
uint32_t compute_leaf_lod(uint64_t mask, const arc_brick64& brick) {
  if (mask == 0) return 0;
  const int u = brick_check_uniform(brick, mask);
  if (u >= 0) return u == (int)kMatAir ? 0u : lod_payload_of((uint32_t)u);
  float r = 0.0f, g = 0.0f, b = 0.0f;
  uint32_t solid_mat = kMatAir, emissive_mat = kMatAir, emissive_vox = 0;
  const uint32_t count = (uint32_t)__builtin_popcountll(mask);
  for (uint64_t rem = mask; rem; rem &= rem - 1) {
    const uint32_t id = brick_get(brick, (uint32_t)ctz64(rem));
    const float* m = palette().m[id];
    const float er = m[4] * 0.25f, eg = m[5] * 0.25f, eb = m[6] * 0.25f;
    r += std::fmin(m[0] + er, 1.0f);
    g += std::fmin(m[1] + eg, 1.0f);
    b += std::fmin(m[2] + eb, 1.0f);
    if (is_emissive(id)) {
      emissive_mat = id;
      emissive_vox++;
    } else if (solid_mat == kMatAir) {
      solid_mat = id;
    }
  }
  const uint32_t rep =
      (emissive_mat != kMatAir && (emissive_vox * 2u >= count || solid_mat == kMatAir)) ? emissive_mat : solid_mat;
  const float inv = 1.0f / (float)count;
  return (rep << 24) | (unorm8_trunc(b * inv) << 16) | (unorm8_trunc(g * inv) << 8) | unorm8_trunc(r * inv);
}

}  // namespace

// ---- read queries: tree.rs ----------------------------------------------------------------
uint32_t chunk_get_voxel(const ChunkTree& t, uint32_t x, uint32_t y, uint32_t z) {
  uint32_t idx = 0;
  int shift = 6;
  for (;;) {
    const arc_node64& n = t.nodes[idx];
    const uint32_t b = cbit((x >> shift) & 3, (y >> shift) & 3, (z >> shift) & 3);
    if (!has_child(n, b)) return kMatAir;
    if (is_leaf(n)) return is_uniform(n) ? primary_mat(n) : brick_get(t.bricks[base_ptr(n)], b);
    if (shift < 2) return kMatAir;  // malformed (internal node at the voxel level)
    idx = child_slot(n, b);
    shift -= 2;
  }
}

static uint64_t chunk_brick_solid_mask(const ChunkTree& t, uint32_t bx, uint32_t by, uint32_t bz) {
  const uint32_t x = bx << 2, y = by << 2, z = bz << 2;
  uint32_t idx = 0;
  int shift = 6;
  for (;;) {
    const arc_node64& n = t.nodes[idx];
    if (is_leaf(n)) {
      if (shift > 0) {
        const uint32_t m = primary_mat(n);
        return (m != kMatAir && m != kMatWater) ? ~0ull : 0ull;
      }
      const uint64_t mask = mask_of(n);
      if (is_uniform(n)) {
        const uint32_t m = primary_mat(n);
        return (m != kMatAir && m != kMatWater) ? mask : 0ull;
      }
      const arc_brick64& br = t.bricks[base_ptr(n)];
      uint64_t solid = mask;
      for (uint64_t rem = mask; rem; rem &= rem - 1) {
        const uint32_t bit = (uint32_t)ctz64(rem);
        const uint32_t m = brick_get(br, bit);
        if (m == kMatAir || m == kMatWater) solid &= ~(1ull << bit);
      }
      return solid;
    }
    const uint32_t b = cbit((x >> shift) & 3, (y >> shift) & 3, (z >> shift) & 3);
    if (!has_child(n, b)) return 0;
    if (shift < 2) return 0;
    idx = child_slot(n, b);
    shift -= 2;
  }
}

static bool node_intersects_aabb(const ChunkTree& t, uint32_t idx, uint32_t ox, uint32_t oy, uint32_t oz, int shift,
                                 const uint32_t mn[3], const uint32_t mx[3]) {
  const arc_node64& n = t.nodes[idx];
  const uint64_t mask = mask_of(n);
  if (mask == 0) return false;
  const uint32_t size = 1u << shift;
  const uint32_t o[3] = {ox, oy, oz};
  uint32_t cmin[3], cmax[3];
  for (int a = 0; a < 3; ++a) {
    cmin[a] = (mn[a] > o[a] ? mn[a] - o[a] : 0u) >> shift;  // saturating_sub
    cmax[a] = std::min((mx[a] - o[a]) >> shift, 3u);
  }
  uint64_t q = 0;
  for (uint32_t cz = cmin[2]; cz <= cmax[2]; ++cz)
    for (uint32_t cy = cmin[1]; cy <= cmax[1]; ++cy)
      for (uint32_t cx = cmin[0]; cx <= cmax[0]; ++cx) q |= 1ull << cbit(cx, cy, cz);
  const uint64_t over = mask & q;
  if (over == 0) return false;
  if (is_leaf(n)) {
    if (is_uniform(n)) {
      const uint32_t m = primary_mat(n);
      return m != kMatAir && m != kMatWater;
    }
    const arc_brick64& br = t.bricks[base_ptr(n)];
    for (uint64_t rem = over; rem; rem &= rem - 1) {
      const uint32_t m = brick_get(br, (uint32_t)ctz64(rem));
      if (m != kMatAir && m != kMatWater) return true;
    }
    return false;
  }
  if (shift < 2) return false;
  for (uint64_t rem = over; rem; rem &= rem - 1) {
    const uint32_t b = (uint32_t)ctz64(rem);
    const uint32_t cx = b & 3, cy = (b >> 2) & 3, cz = (b >> 4) & 3;
    if (node_intersects_aabb(t, child_slot(n, b), ox + cx * size, oy + cy * size, oz + cz * size, shift - 2, mn, mx))
      return true;
  }
  return false;
}

static bool chunk_intersects_aabb(const ChunkTree& t, uint32_t x0, uint32_t y0, uint32_t z0, uint32_t x1, uint32_t y1,
                                  uint32_t z1) {
  if (mask_of(t.nodes[0]) == 0) return false;
  const uint32_t mn[3] = {x0, y0, z0}, mx[3] = {x1, y1, z1};
  return node_intersects_aabb(t, 0, 0, 0, 0, 6, mn, mx);
}

// ---- in-place mutation: set.rs ------------------------------------------------------------
namespace {

struct Editor {
  ChunkTree& t;
  EditReport& rep;

  bool empty() const { return mask_of(t.nodes[0]) == 0; }

  bool can_reach_leaf(uint32_t bx, uint32_t by, uint32_t bz) const {
    if (empty()) return false;
    const uint32_t b0 = cbit((bx >> 4) & 3, (by >> 4) & 3, (bz >> 4) & 3);
    const arc_node64& r = t.nodes[0];
    if (is_leaf(r)) return mask_of(r) != 0;
    if (!has_child(r, b0)) return false;
    const uint32_t l1 = child_slot(r, b0);
    if (is_leaf(t.nodes[l1])) return mask_of(t.nodes[l1]) != 0;
    const uint32_t b1 = cbit((bx >> 2) & 3, (by >> 2) & 3, (bz >> 2) & 3);
    if (!has_child(t.nodes[l1], b1)) return false;
    const uint32_t l2 = child_slot(t.nodes[l1], b1);
    if (is_leaf(t.nodes[l2])) return mask_of(t.nodes[l2]) != 0;
    const uint32_t b2 = cbit(bx & 3, by & 3, bz & 3);
    if (!has_child(t.nodes[l2], b2)) return false;
    const uint32_t l3 = child_slot(t.nodes[l2], b2);
    return mask_of(t.nodes[l3]) != 0;
  }

  void uncollapse_if_needed(size_t idx, uint32_t shift) {
    if (shift == 0 || !is_leaf(t.nodes[idx])) return;
    const uint32_t mat = primary_mat(t.nodes[idx]);
    const uint32_t count = child_count(t.nodes[idx]);
    if (count == 0) {
      t.nodes[idx].child_ptr = 0;
      return;
    }
    const uint32_t base = (uint32_t)t.nodes.size();
    const uint32_t lod = lod_payload_of(mat);
    t.nodes.resize(t.nodes.size() + count, node_of(~0ull, ARC_IS_LEAF_FLAG | ARC_IS_UNIFORM_FLAG, lod));
    t.nodes[idx].child_ptr = base;
  }

  size_t ensure_child(size_t parent, uint32_t b, uint32_t shift) {
    uncollapse_if_needed(parent, shift);
    if (has_child(t.nodes[parent], b)) return child_slot(t.nodes[parent], b);
    const uint32_t old_count = child_count(t.nodes[parent]);
    const uint32_t old_base = base_ptr(t.nodes[parent]);
    const uint32_t new_base = (uint32_t)t.nodes.size();
    t.nodes.resize(t.nodes.size() + old_count + 1, empty_node());
    const uint64_t prefix = (1ull << b) - 1ull;
    const uint32_t rank = (uint32_t)__builtin_popcountll(mask_of(t.nodes[parent]) & prefix);
    if (old_count > 0) {
      for (uint32_t i = 0; i < rank; ++i) t.nodes[new_base + i] = t.nodes[old_base + i];
      for (uint32_t i = rank; i < old_count; ++i) t.nodes[new_base + i + 1] = t.nodes[old_base + i];
    }
    const size_t child = (size_t)new_base + rank;
    t.nodes[child] = shift == 2 ? node_of(0, ARC_IS_LEAF_FLAG | ARC_IS_UNIFORM_FLAG, 0) : empty_node();
    t.nodes[parent].child_ptr = (t.nodes[parent].child_ptr & ~ARC_PTR_INDEX_MASK) | new_base;
    set_mask(t.nodes[parent], mask_of(t.nodes[parent]) | (1ull << b));
    return child;
  }

  void remove_child(size_t parent, uint32_t b) {
    if (!has_child(t.nodes[parent], b)) return;
    const uint32_t old_count = child_count(t.nodes[parent]);
    if (old_count <= 1) {
      const uint32_t old_base = base_ptr(t.nodes[parent]);
      if (old_count == 1 && old_base > 0 && old_base < t.nodes.size()) {
        t.nodes[old_base] = empty_node();
        rep.node(old_base);
      }
      set_mask(t.nodes[parent], 0);
      t.nodes[parent].child_ptr = 0;
      t.nodes[parent].lod_payload = 0;
      rep.node(parent);
      return;
    }
    const uint32_t old_base = base_ptr(t.nodes[parent]);
    const uint64_t prefix = (1ull << b) - 1ull;
    const uint32_t rank = (uint32_t)__builtin_popcountll(mask_of(t.nodes[parent]) & prefix);
    const uint32_t new_count = old_count - 1;
    for (uint32_t i = rank; i < new_count; ++i) t.nodes[old_base + i] = t.nodes[old_base + i + 1];
    t.nodes[old_base + new_count] = empty_node();
    set_mask(t.nodes[parent], mask_of(t.nodes[parent]) & ~(1ull << b));
    rep.node_range(old_base + rank, old_base + new_count);
    rep.node(parent);
  }

  void apply_entries_to_leaf(size_t leaf, const std::vector<std::pair<uint32_t, uint32_t>>& entries) {
    for (const auto& [b, mat] : entries) {
      const arc_node64 n = t.nodes[leaf];
      const uint64_t mask = mask_of(n);
      const uint64_t bit = 1ull << b;
      const bool was_solid = (mask & bit) != 0;
      if (is_uniform(n)) {
        if (mat == kMatAir) {
          if (was_solid) {
            set_mask(t.nodes[leaf], mask & ~bit);
            if (t.solid) t.solid--;
          }
        } else if (mask == 0) {
          set_mask(t.nodes[leaf], bit);
          t.nodes[leaf].lod_payload = mat << 24;
          t.solid++;
          rep.added++;
        } else if (mat == primary_mat(n)) {
          if (!was_solid) {
            set_mask(t.nodes[leaf], mask | bit);
            t.solid++;
            rep.added++;
          }
        } else if (was_solid && mask == bit) {
          t.nodes[leaf].lod_payload = mat << 24;
        } else {
          const uint32_t bi = (uint32_t)t.bricks.size();
          arc_brick64 br;
          std::memset(&br, 0, sizeof br);
          const uint32_t prim = primary_mat(n);
          for (uint64_t rem = mask; rem; rem &= rem - 1) brick_set(br, (uint32_t)ctz64(rem), prim);
          brick_set(br, b, mat);
          t.bricks.push_back(br);
          rep.growth = true;
          t.nodes[leaf].child_ptr = ARC_IS_LEAF_FLAG | (bi & ARC_PTR_INDEX_MASK);
          set_mask(t.nodes[leaf], mask | bit);
          if (!was_solid) {
            t.solid++;
            rep.added++;
          }
          rep.brick(bi);
        }
      } else {
        const uint32_t bi = base_ptr(n);
        brick_set(t.bricks[bi], b, mat);
        rep.brick(bi);
        if (mat == kMatAir) {
          if (was_solid) {
            set_mask(t.nodes[leaf], mask & ~bit);
            if (t.solid) t.solid--;
          }
        } else if (!was_solid) {
          set_mask(t.nodes[leaf], mask | bit);
          t.solid++;
          rep.added++;
        }
      }
    }
    const arc_node64 f = t.nodes[leaf];
    const uint64_t fm = mask_of(f);
    if (fm == 0) t.nodes[leaf].lod_payload = 0;
    else if (is_uniform(f)) t.nodes[leaf].lod_payload = lod_payload_of(primary_mat(f));
    else t.nodes[leaf].lod_payload = compute_leaf_lod(fm, t.bricks[base_ptr(f)]);
  }

  void refresh_internal_lod(size_t idx) {
    const arc_node64 n = t.nodes[idx];
    const uint32_t count = child_count(n);
    if (count == 0) {
      t.nodes[idx].lod_payload = 0;
      return;
    }
    const uint32_t base = base_ptr(n);
    uint32_t rs = 0, gs = 0, bs = 0, solid_mat = kMatAir, emissive_mat = kMatAir, emissive_kids = 0, nonempty = 0;
    for (uint32_t i = 0; i < count; ++i) {
      const arc_node64& c = t.nodes[base + i];
      if (mask_of(c) == 0) continue;
      const uint32_t l = c.lod_payload;
      rs += l & 0xFF;
      gs += (l >> 8) & 0xFF;
      bs += (l >> 16) & 0xFF;
      const uint32_t cm = l >> 24;
      if (is_emissive(cm)) {
        emissive_mat = cm;
        emissive_kids++;
      } else if (solid_mat == kMatAir) {
        solid_mat = cm;
      }
      nonempty++;
    }
    if (nonempty == 0) {
      t.nodes[idx].lod_payload = 0;
      return;
    }
    const uint32_t dom =
        (emissive_mat != kMatAir && (emissive_kids * 2u >= nonempty || solid_mat == kMatAir)) ? emissive_mat : solid_mat;
    t.nodes[idx].lod_payload = (dom << 24) | (((bs / nonempty) & 0xFF) << 16) | (((gs / nonempty) & 0xFF) << 8) |
                               ((rs / nonempty) & 0xFF);
  }

  void mutate_brick(uint32_t bx, uint32_t by, uint32_t bz, const std::vector<std::pair<uint32_t, uint32_t>>& entries) {
    const uint32_t b0 = cbit((bx >> 4) & 3, (by >> 4) & 3, (bz >> 4) & 3);
    const uint32_t b1 = cbit((bx >> 2) & 3, (by >> 2) & 3, (bz >> 2) & 3);
    const uint32_t b2 = cbit(bx & 3, by & 3, bz & 3);
    const size_t nodes_before = t.nodes.size(), bricks_before = t.bricks.size();
    size_t path[4];
    path[0] = 0;
    path[1] = ensure_child(path[0], b0, 6);
    path[2] = ensure_child(path[1], b1, 4);
    path[3] = ensure_child(path[2], b2, 2);
    const size_t nodes_after = t.nodes.size();
    if (nodes_after > nodes_before) {
      rep.growth = true;
      rep.node_range(nodes_before, nodes_after - 1);
    }
    for (size_t p : path) rep.node(p);
    const size_t leaf = path[3];
    apply_entries_to_leaf(leaf, entries);
    if (t.bricks.size() > bricks_before) rep.growth = true;
    if (mask_of(t.nodes[leaf]) == 0) {
      remove_child(path[2], b2);
      rep.node(path[2]);
      if (mask_of(t.nodes[path[2]]) == 0) {
        remove_child(path[1], b1);
        rep.node(path[1]);
        if (mask_of(t.nodes[path[1]]) == 0) {
          remove_child(path[0], b0);
          rep.node(path[0]);
          if (mask_of(t.nodes[0]) == 0) {
            t.nodes[0] = empty_node();
            t.solid = 0;
          } else {
            refresh_internal_lod(path[0]);
          }
        } else {
          refresh_internal_lod(path[1]);
          refresh_internal_lod(path[0]);
        }
      } else {
        refresh_internal_lod(path[2]);
        refresh_internal_lod(path[1]);
        refresh_internal_lod(path[0]);
      }
    } else {
      refresh_internal_lod(path[2]);
      refresh_internal_lod(path[1]);
      refresh_internal_lod(path[0]);
    }
    const size_t nodes_final = t.nodes.size();
    if (nodes_final > nodes_after) {
      rep.growth = true;
      rep.node_range(nodes_after, nodes_final - 1);
    }
  }
};

}  // namespace

EditReport chunk_set_voxels(ChunkTree& t, const uint32_t* xyzm, uint64_t n) {
  EditReport rep;
  if (n == 0) return rep;
  // group by brick, bricks in (bz, by, bx) order, entries in input order within a brick
  std::map<uint32_t, std::vector<std::pair<uint32_t, uint32_t>>> groups;
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t x = xyzm[4 * i], y = xyzm[4 * i + 1], z = xyzm[4 * i + 2], m = xyzm[4 * i + 3] & 0xFFu;
    if (x >= kChunkDim || y >= kChunkDim || z >= kChunkDim) continue;
    const uint32_t key = ((z >> 2) << 12) | ((y >> 2) << 6) | (x >> 2);
    groups[key].emplace_back(cbit(x & 3, y & 3, z & 3), m);
  }
  if (groups.empty()) return rep;
  Editor ed{t, rep};
  for (const auto& [key, entries] : groups) {
    const uint32_t bx = key & 63, by = (key >> 6) & 63, bz = key >> 12;
    bool all_air = true;
    for (const auto& e : entries) all_air = all_air && e.second == kMatAir;
    if (all_air && !ed.can_reach_leaf(bx, by, bz)) continue;
    ed.mutate_brick(bx, by, bz, entries);
  }
  if (t.solid == 0) {
    t.nodes[0] = empty_node();
    rep.node(0);
  }
  return rep;
}

uint32_t chunk_clear_voxels(ChunkTree& t, const uint32_t* xyz, uint64_t n, EditReport* out) {
  *out = EditReport{};
  if (mask_of(t.nodes[0]) == 0 || n == 0) return 0;
  std::vector<uint32_t> set;
  uint32_t removed = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    if (x >= kChunkDim || y >= kChunkDim || z >= kChunkDim) continue;
    if (chunk_get_voxel(t, x, y, z) != kMatAir) {
      removed++;
      set.insert(set.end(), {x, y, z, kMatAir});
    }
  }
  if (set.empty()) return 0;
  *out = chunk_set_voxels(t, set.data(), set.size() / 4);
  return removed;
}

void chunk_compact(ChunkTree& t) {
  if (mask_of(t.nodes[0]) == 0) {
    t.nodes.assign(1, empty_node());
    t.bricks.clear();
    t.solid = 0;
    return;
  }
  std::vector<arc_node64> nn;
  std::vector<arc_brick64> nb;
  nn.reserve(t.nodes.size());
  nb.reserve(t.bricks.size());
  std::vector<std::pair<size_t, size_t>> q;  // FIFO (new idx, old idx)
  q.reserve(t.nodes.size() / 4 + 1);
  nn.push_back(empty_node());
  q.emplace_back(0, 0);
  for (size_t head = 0; head < q.size(); ++head) {
    const auto [ni, oi] = q[head];
    const arc_node64 n = t.nodes[oi];
    if (mask_of(n) == 0) {
      nn[ni] = empty_node();
    } else if (is_leaf(n)) {
      if (is_uniform(n)) {
        nn[ni] = n;
      } else {
        const uint32_t bi = (uint32_t)nb.size();
        nb.push_back(t.bricks[base_ptr(n)]);
        arc_node64 l = n;
        l.child_ptr = (n.child_ptr & ~ARC_PTR_INDEX_MASK) | (bi & ARC_PTR_INDEX_MASK);
        nn[ni] = l;
      }
    } else {
      const uint32_t count = child_count(n);
      const uint32_t old_base = base_ptr(n), new_base = (uint32_t)nn.size();
      nn.resize(nn.size() + count, empty_node());
      arc_node64 in = n;
      in.child_ptr = (n.child_ptr & ~ARC_PTR_INDEX_MASK) | (new_base & ARC_PTR_INDEX_MASK);
      nn[ni] = in;
      for (uint32_t i = 0; i < count; ++i) q.emplace_back((size_t)new_base + i, (size_t)old_base + i);
    }
  }
  t.nodes.swap(nn);
  t.bricks.swap(nb);
}

// carve_sphere_tracked (carve.rs:39-177); glam Vec3 arithmetic restated per lane.
uint32_t chunk_carve_sphere(ChunkTree& t, const float c[3], float radius, bool scorch_rim, EditReport* out) {
  *out = EditReport{};
  if (mask_of(t.nodes[0]) == 0) return 0;
  const float outer = scorch_rim ? radius + 2.2f : radius;
  auto dist2 = [&](const float p[3]) {
    const float dx = c[0] - p[0], dy = c[1] - p[1], dz = c[2] - p[2];
    const float xx = dx * dx, yy = dy * dy, zz = dz * dz;
    return (xx + yy) + zz;
  };
  auto clamp3 = [&](const float lo[3], const float hi[3], float o[3]) {
    for (int a = 0; a < 3; ++a) o[a] = std::fmin(std::fmax(c[a], lo[a]), hi[a]);
  };
  {
    const float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {256.0f, 256.0f, 256.0f};
    float cl[3];
    clamp3(lo, hi, cl);
    if (dist2(cl) > outer * outer) return 0;
  }
  float smin[3], smax[3];
  for (int a = 0; a < 3; ++a) {
    smin[a] = std::fmax(std::floor(c[a] - outer), 0.0f);
    smax[a] = std::fmin(std::ceil(c[a] + outer), 255.0f);
  }
  if (smin[0] > smax[0] || smin[1] > smax[1] || smin[2] > smax[2]) return 0;
  auto f2u = [](float v) -> uint32_t { return v > 0.0f ? (uint32_t)v : 0u; };  // values are in [0, 255]
  const uint32_t mnx = f2u(smin[0]), mny = f2u(smin[1]), mnz = f2u(smin[2]);
  const uint32_t mxx = f2u(smax[0]), mxy = f2u(smax[1]), mxz = f2u(smax[2]);
  if (!chunk_intersects_aabb(t, mnx, mny, mnz, mxx, mxy, mxz)) return 0;
  const float inner_sq = radius * radius, outer_sq = outer * outer;
  const float er = radius + 0.95f;
  const float ember_sq = er * er;
  EditReport& rep = *out;
  Editor ed{t, rep};
  std::vector<std::pair<uint32_t, uint32_t>> entries;
  entries.reserve(64);
  uint32_t destroyed = 0;
  for (uint32_t bz = mnz >> 2; bz <= (mxz >> 2); ++bz)
    for (uint32_t by = mny >> 2; by <= (mxy >> 2); ++by)
      for (uint32_t bx = mnx >> 2; bx <= (mxx >> 2); ++bx) {
        const float bmin[3] = {(float)(bx * 4), (float)(by * 4), (float)(bz * 4)};
        const float bmax[3] = {bmin[0] + 4.0f, bmin[1] + 4.0f, bmin[2] + 4.0f};
        float cl[3];
        clamp3(bmin, bmax, cl);
        if (dist2(cl) > outer_sq) continue;
        const uint64_t solid = chunk_brick_solid_mask(t, bx, by, bz);
        if (solid == 0) continue;
        for (uint64_t rem = solid; rem; rem &= rem - 1) {
          const uint32_t b = (uint32_t)ctz64(rem);
          const uint32_t vx = (bx << 2) | (b & 3), vy = (by << 2) | ((b >> 2) & 3), vz = (bz << 2) | ((b >> 4) & 3);
          const float vc[3] = {(float)vx + 0.5f, (float)vy + 0.5f, (float)vz + 0.5f};
          const float d2 = dist2(vc);
          if (d2 <= inner_sq) {
            if (chunk_get_voxel(t, vx, vy, vz) != kMatAir) {
              destroyed++;
              entries.emplace_back(b, kMatAir);
            }
          } else if (scorch_rim && d2 <= outer_sq) {
            const uint32_t old = chunk_get_voxel(t, vx, vy, vz);
            if (old != kMatAir && old != kMatWater) {
              const uint32_t h = ((vx * 73856093u) ^ (vy * 19349663u) ^ (vz * 83492791u)) & 0xFFu;
              if (d2 <= ember_sq && h < 48) entries.emplace_back(b, kMatEmber);
              else if (h < 180) entries.emplace_back(b, kMatScorched);
            }
          }
        }
        if (!entries.empty()) {
          ed.mutate_brick(bx, by, bz, entries);
          entries.clear();
        }
      }
  if (t.solid == 0) {
    t.nodes[0] = empty_node();
    rep.node(0);
  }
  return destroyed;
}

// ---- world pools: world_map.rs ------------------------------------------------------------
static std::vector<arc_node64> relocate(const arc_node64* nodes, size_t n, uint32_t node_base, uint32_t brick_base) {
  std::vector<arc_node64> out(n);
  for (size_t i = 0; i < n; ++i) {
    const arc_node64& s = nodes[i];
    if (mask_of(s) == 0) {
      out[i] = empty_node();
      continue;
    }
    const uint32_t flags = s.child_ptr & (ARC_IS_LEAF_FLAG | ARC_IS_UNIFORM_FLAG);
    const uint32_t local = s.child_ptr & ARC_PTR_INDEX_MASK;
    uint32_t p;
    if (is_leaf(s)) p = is_uniform(s) ? flags : (flags | (brick_base + local));
    else p = flags | (node_base + local);
    out[i] = s;
    out[i].child_ptr = p;
  }
  return out;
}

static bool find_free_span(std::vector<std::pair<uint32_t, uint32_t>>& spans, uint32_t req, uint32_t* base,
                           uint32_t* cap) {
  size_t pos = spans.size();
  for (size_t i = 0; i < spans.size(); ++i)
    if (spans[i].second >= req && (uint64_t)spans[i].second <= (uint64_t)req * 4u) {
      pos = i;
      break;
    }
  if (pos == spans.size())
    for (size_t i = 0; i < spans.size(); ++i)
      if (spans[i].second >= req) {
        pos = i;
        break;
      }
  if (pos == spans.size()) return false;
  *base = spans[pos].first;
  *cap = spans[pos].second;
  spans[pos] = spans.back();  // swap_remove
  spans.pop_back();
  return true;
}

static void drop_patches_of(World& w, uint32_t idx) {
  w.pending.erase(std::remove_if(w.pending.begin(), w.pending.end(),
                                 [&](const HostPatch& p) { return p.chunk == idx; }),
                  w.pending.end());
}

void world_sync_chunk(World& w, uint32_t idx) {
  ChunkAlloc& al = w.allocs[idx];
  const ChunkAlloc old = al;
  ChunkTree& ch = w.chunks[idx];
  if (old.node_base > 0 && (ch.nodes.size() > old.node_cap || ch.bricks.size() > old.brick_cap)) chunk_compact(ch);
  if (mask_of(ch.nodes[0]) == 0) {
    if (old.node_cap > 0 && old.node_base > 0) w.free_node_spans.emplace_back(old.node_base, old.node_cap);
    if (old.brick_cap > 0 && old.brick_base > 0) w.free_brick_spans.emplace_back(old.brick_base, old.brick_cap);
    al = ChunkAlloc{0, 0, 0, 0};
    w.roots[idx] = ARC_EMPTY_CHUNK_SENTINEL;
    drop_patches_of(w, idx);
    HostPatch p;
    p.chunk = idx;
    p.root = ARC_EMPTY_CHUNK_SENTINEL;
    w.pending.push_back(std::move(p));
    return;
  }
  const uint32_t req_n = (uint32_t)ch.nodes.size(), req_b = (uint32_t)ch.bricks.size();
  uint32_t node_base;
  if (req_n <= old.node_cap && old.node_base > 0) {
    node_base = old.node_base;
  } else {
    if (old.node_cap > 0 && old.node_base > 0) w.free_node_spans.emplace_back(old.node_base, old.node_cap);
    uint32_t base, cap;
    if (!find_free_span(w.free_node_spans, req_n, &base, &cap)) {
      base = (uint32_t)w.nodes.size();
      cap = std::max(req_n + req_n / 2, 128u);
      w.nodes.resize((size_t)base + cap, empty_node());
    }
    al.node_base = base;
    al.node_cap = cap;
    node_base = base;
  }
  uint32_t brick_base;
  if (req_b <= old.brick_cap && (old.brick_base > 0 || req_b == 0)) {
    brick_base = old.brick_base;
  } else {
    if (old.brick_cap > 0 && old.brick_base > 0) w.free_brick_spans.emplace_back(old.brick_base, old.brick_cap);
    uint32_t base, cap;
    if (!find_free_span(w.free_brick_spans, req_b, &base, &cap)) {
      base = (uint32_t)w.bricks.size();
      cap = std::max(req_b + req_b / 2, 64u);
      arc_brick64 zb;
      std::memset(&zb, 0, sizeof zb);
      w.bricks.resize((size_t)base + cap, zb);
    }
    al.brick_base = base;
    al.brick_cap = cap;
    brick_base = base;
  }
  std::vector<arc_node64> rel = relocate(ch.nodes.data(), req_n, node_base, brick_base);
  std::copy(rel.begin(), rel.end(), w.nodes.begin() + node_base);
  if (req_b > 0) std::copy(ch.bricks.begin(), ch.bricks.end(), w.bricks.begin() + brick_base);
  w.roots[idx] = node_base;
  drop_patches_of(w, idx);
  HostPatch p;
  p.chunk = idx;
  p.root = node_base;
  p.node_offset = node_base;
  p.nodes = std::move(rel);
  p.brick_offset = brick_base;
  p.bricks = ch.bricks;
  w.pending.push_back(std::move(p));
}

void world_sync_chunk_partial(World& w, uint32_t idx, const EditReport& rep) {
  ChunkTree& ch = w.chunks[idx];
  if (mask_of(ch.nodes[0]) == 0) {
    world_sync_chunk(w, idx);
    return;
  }
  const ChunkAlloc al = w.allocs[idx];
  const uint32_t req_n = (uint32_t)ch.nodes.size(), req_b = (uint32_t)ch.bricks.size();
  if (req_n > al.node_cap || req_b > al.brick_cap || al.node_base == 0 || (req_b > 0 && al.brick_base == 0)) {
    world_sync_chunk(w, idx);
    return;
  }
  HostPatch p;
  p.chunk = idx;
  p.root = al.node_base;
  const bool dn = rep.min_n <= rep.max_n && rep.max_n < ch.nodes.size();
  if (dn) {
    p.nodes = relocate(ch.nodes.data() + rep.min_n, rep.max_n - rep.min_n + 1, al.node_base, al.brick_base);
    p.node_offset = al.node_base + (uint32_t)rep.min_n;
    std::copy(p.nodes.begin(), p.nodes.end(), w.nodes.begin() + p.node_offset);
  }
  const bool db = rep.min_b <= rep.max_b && req_b > 0 && rep.max_b < ch.bricks.size();
  if (db) {
    p.bricks.assign(ch.bricks.begin() + rep.min_b, ch.bricks.begin() + rep.max_b + 1);
    p.brick_offset = al.brick_base + (uint32_t)rep.min_b;
    std::copy(p.bricks.begin(), p.bricks.end(), w.bricks.begin() + p.brick_offset);
  }
  if (dn || db) {
    w.roots[idx] = al.node_base;
    w.pending.push_back(std::move(p));
  }
}

static bool chunk_of_global(const World& w, int32_t vx, int32_t vy, int32_t vz, uint32_t* idx, uint32_t local[3]) {
  const int32_t v[3] = {vx, vy, vz};
  int32_t c[3];
  for (int a = 0; a < 3; ++a) {
    c[a] = v[a] >= 0 ? v[a] / 256 : -((-v[a] + 255) / 256);  // div_euclid
    local[a] = (uint32_t)(v[a] - c[a] * 256);
    if (c[a] < 0 || (uint32_t)c[a] >= w.grid[a]) return false;
  }
  *idx = (uint32_t)c[0] + (uint32_t)c[1] * w.grid[0] + (uint32_t)c[2] * w.grid[0] * w.grid[1];
  return true;
}

uint32_t world_set_voxels(World& w, const int32_t* xyzm, uint64_t n) {
  std::map<uint32_t, std::vector<uint32_t>> per_chunk;  // ascending chunk index (set_voxels_global sorts)
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t idx, l[3];
    if (!chunk_of_global(w, xyzm[4 * i], xyzm[4 * i + 1], xyzm[4 * i + 2], &idx, l)) continue;
    per_chunk[idx].insert(per_chunk[idx].end(), {l[0], l[1], l[2], (uint32_t)xyzm[4 * i + 3] & 0xFFu});
  }
  uint32_t added = 0;
  for (auto& [idx, list] : per_chunk) {
    const EditReport rep = chunk_set_voxels(w.chunks[idx], list.data(), list.size() / 4);
    added += rep.added;
    if (w.packed) world_sync_chunk_partial(w, idx, rep);
  }
  return added;
}

uint32_t world_clear_voxels(World& w, const int32_t* xyz, uint64_t n) {
  std::map<uint32_t, std::vector<uint32_t>> per_chunk;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t idx, l[3];
    if (!chunk_of_global(w, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], &idx, l)) continue;
    per_chunk[idx].insert(per_chunk[idx].end(), {l[0], l[1], l[2]});
  }
  uint32_t removed_total = 0;
  std::vector<std::pair<uint32_t, EditReport>> modified;
  for (auto& [idx, list] : per_chunk) {
    EditReport rep;
    const uint32_t removed = chunk_clear_voxels(w.chunks[idx], list.data(), list.size() / 3, &rep);
    if (removed > 0) {
      removed_total += removed;
      modified.emplace_back(idx, rep);
    }
  }
  if (w.packed)
    for (auto& [idx, rep] : modified) world_sync_chunk_partial(w, idx, rep);
  return removed_total;
}

uint32_t world_carve_crater(World& w, const float center_m[3], float radius_m, bool scorch_rim) {
  float cv[3];
  for (int a = 0; a < 3; ++a) cv[a] = center_m[a] * 16.0f;
  const float rv = radius_m * 16.0f;
  const float outer = scorch_rim ? rv + 2.5f : rv;
  auto to_i32 = [](float v) -> int32_t {  // Rust `as i32`: saturating, NaN -> 0
    if (!(v == v)) return 0;
    if (v <= -2147483648.0f) return INT32_MIN;
    if (v >= 2147483648.0f) return INT32_MAX;
    return (int32_t)v;
  };
  auto div_e = [](int32_t v) -> int32_t { return v >= 0 ? v / 256 : -(int32_t)(((int64_t)-(int64_t)v + 255) / 256); };
  int32_t cmin[3], cmax[3];
  for (int a = 0; a < 3; ++a) {
    cmin[a] = div_e(to_i32(std::floor(cv[a] - outer)));
    cmax[a] = div_e(to_i32(std::ceil(cv[a] + outer)));
  }
  uint32_t total = 0;
  std::vector<std::pair<uint32_t, EditReport>> modified;
  for (int32_t cz = cmin[2]; cz <= cmax[2]; ++cz)
    for (int32_t cy = cmin[1]; cy <= cmax[1]; ++cy)
      for (int32_t cx = cmin[0]; cx <= cmax[0]; ++cx) {
        if (cx < 0 || cy < 0 || cz < 0 || (uint32_t)cx >= w.grid[0] || (uint32_t)cy >= w.grid[1] ||
            (uint32_t)cz >= w.grid[2])
          continue;
        const uint32_t idx = (uint32_t)cx + (uint32_t)cy * w.grid[0] + (uint32_t)cz * w.grid[0] * w.grid[1];
        ChunkTree& ch = w.chunks[idx];
        if (mask_of(ch.nodes[0]) == 0) continue;
        const float local[3] = {cv[0] - (float)(cx * 256), cv[1] - (float)(cy * 256), cv[2] - (float)(cz * 256)};
        EditReport rep;
        const uint32_t d = chunk_carve_sphere(ch, local, rv, scorch_rim, &rep);
        if (d > 0 || scorch_rim) {
          total += d;
          modified.emplace_back(idx, rep);
        }
      }
  if (w.packed)
    for (auto& [idx, rep] : modified) world_sync_chunk_partial(w, idx, rep);
  return total;
}

}  // namespace arc

// ---- C ABI (arcadia_world.h) --------------------------------------------------------------
using arc::World;

static World* W(arc_world_host* h) { return reinterpret_cast<World*>(h); }

static void to_c(const arc::EditReport& r, arc_edit_report* o) {
  if (!o) return;
  o->added_voxels = r.added;
  o->min_dirty_node = r.min_n;
  o->max_dirty_node = r.max_n;
  o->min_dirty_brick = r.min_b;
  o->max_dirty_brick = r.max_b;
  o->had_structural_growth = r.growth ? 1u : 0u;
  o->reserved[0] = o->reserved[1] = 0;
}

static arc::EditReport from_c(const arc_edit_report* o) {
  arc::EditReport r;
  r.added = o->added_voxels;
  r.min_n = o->min_dirty_node;
  r.max_n = o->max_dirty_node;
  r.min_b = o->min_dirty_brick;
  r.max_b = o->max_dirty_brick;
  r.growth = o->had_structural_growth != 0;
  return r;
}

extern "C" {

int arc_chunk_set_voxels(arc_world_host* wh, uint32_t idx, const uint32_t* xyzm, uint64_t n, arc_edit_report* rep) {
  if (!wh || (n && !xyzm)) return ARC_ERR_INVALID;
  World* w = W(wh);
  if (idx >= w->chunks.size()) return ARC_ERR_INVALID;
  try {
    to_c(arc::chunk_set_voxels(w->chunks[idx], xyzm, n), rep);
  } catch (const std::bad_alloc&) {
    return ARC_ERR_CAPACITY;
  }
  return ARC_OK;
}

int arc_chunk_clear_voxels(arc_world_host* wh, uint32_t idx, const uint32_t* xyz, uint64_t n, uint32_t* removed,
                           arc_edit_report* rep) {
  if (!wh || (n && !xyz)) return ARC_ERR_INVALID;
  World* w = W(wh);
  if (idx >= w->chunks.size()) return ARC_ERR_INVALID;
  arc::EditReport r;
  try {
    const uint32_t k = arc::chunk_clear_voxels(w->chunks[idx], xyz, n, &r);
    if (removed) *removed = k;
  } catch (const std::bad_alloc&) {
    return ARC_ERR_CAPACITY;
  }
  to_c(r, rep);
  return ARC_OK;
}

int arc_chunk_carve_sphere(arc_world_host* wh, uint32_t idx, const float center[3], float radius, int scorch_rim,
                           uint32_t* destroyed, arc_edit_report* rep) {
  if (!wh || !center) return ARC_ERR_INVALID;
  World* w = W(wh);
  if (idx >= w->chunks.size()) return ARC_ERR_INVALID;
  arc::EditReport r;
  try {
    const uint32_t d = arc::chunk_carve_sphere(w->chunks[idx], center, radius, scorch_rim != 0, &r);
    if (destroyed) *destroyed = d;
  } catch (const std::bad_alloc&) {
    return ARC_ERR_CAPACITY;
  }
  to_c(r, rep);
  return ARC_OK;
}

int arc_chunk_compact(arc_world_host* wh, uint32_t idx) {
  if (!wh) return ARC_ERR_INVALID;
  World* w = W(wh);
  if (idx >= w->chunks.size()) return ARC_ERR_INVALID;
  arc::chunk_compact(w->chunks[idx]);
  return ARC_OK;
}

int arc_chunk_info(const arc_world_host* wh, uint32_t idx, uint64_t* n_nodes, uint64_t* n_bricks, uint64_t* solid,
                   uint32_t alloc[4]) {
  if (!wh) return ARC_ERR_INVALID;
  const World* w = reinterpret_cast<const World*>(wh);
  if (idx >= w->chunks.size()) return ARC_ERR_INVALID;
  const arc::ChunkTree& c = w->chunks[idx];
  if (n_nodes) *n_nodes = c.nodes.size();
  if (n_bricks) *n_bricks = c.bricks.size();
  if (solid) *solid = c.solid;
  if (alloc) {
    const arc::ChunkAlloc a = idx < w->allocs.size() ? w->allocs[idx] : arc::ChunkAlloc{0, 0, 0, 0};
    alloc[0] = a.node_base;
    alloc[1] = a.node_cap;
    alloc[2] = a.brick_base;
    alloc[3] = a.brick_cap;
  }
  return ARC_OK;
}

uint32_t arc_world_host_voxel(const arc_world_host* wh, int32_t x, int32_t y, int32_t z) {
  if (!wh) return 0;
  const World* w = reinterpret_cast<const World*>(wh);
  uint32_t idx, l[3];
  if (!arc::chunk_of_global(*w, x, y, z, &idx, l)) return 0;
  return arc::chunk_get_voxel(w->chunks[idx], l[0], l[1], l[2]);
}

int arc_world_host_sync_chunk(arc_world_host* wh, uint32_t idx) {
  if (!wh) return ARC_ERR_INVALID;
  World* w = W(wh);
  if (!w->packed) return ARC_ERR_NO_WORLD;
  if (idx >= w->chunks.size()) return ARC_ERR_INVALID;
  arc::world_sync_chunk(*w, idx);
  return ARC_OK;
}

int arc_world_host_sync_chunk_partial(arc_world_host* wh, uint32_t idx, const arc_edit_report* rep) {
  if (!wh || !rep) return ARC_ERR_INVALID;
  World* w = W(wh);
  if (!w->packed) return ARC_ERR_NO_WORLD;
  if (idx >= w->chunks.size()) return ARC_ERR_INVALID;
  arc::world_sync_chunk_partial(*w, idx, from_c(rep));
  return ARC_OK;
}

int arc_world_host_set_voxels(arc_world_host* wh, const int32_t* xyzm, uint64_t n, uint32_t* added) {
  if (!wh || (n && !xyzm)) return ARC_ERR_INVALID;
  try {
    const uint32_t a = arc::world_set_voxels(*W(wh), xyzm, n);
    if (added) *added = a;
  } catch (const std::bad_alloc&) {
    return ARC_ERR_CAPACITY;
  }
  return ARC_OK;
}

int arc_world_host_clear_voxels(arc_world_host* wh, const int32_t* xyz, uint64_t n, uint32_t* removed) {
  if (!wh || (n && !xyz)) return ARC_ERR_INVALID;
  try {
    const uint32_t r = arc::world_clear_voxels(*W(wh), xyz, n);
    if (removed) *removed = r;
  } catch (const std::bad_alloc&) {
    return ARC_ERR_CAPACITY;
  }
  return ARC_OK;
}

int arc_world_host_carve_crater(arc_world_host* wh, const float center_m[3], float radius_m, int scorch_rim,
                                uint32_t* destroyed) {
  if (!wh || !center_m) return ARC_ERR_INVALID;
  if (!(radius_m >= 0.0f) || radius_m > 4096.0f) return ARC_ERR_INVALID;
  try {
    const uint32_t d = arc::world_carve_crater(*W(wh), center_m, radius_m, scorch_rim != 0);
    if (destroyed) *destroyed = d;
  } catch (const std::bad_alloc&) {
    return ARC_ERR_CAPACITY;
  }
  return ARC_OK;
}

int arc_world_host_pending_patches(arc_world_host* wh, const arc_gpu_patch** patches, uint64_t* n) {
  if (!wh || !patches || !n) return ARC_ERR_INVALID;
  World* w = W(wh);
  w->views.resize(w->pending.size());
  for (size_t i = 0; i < w->pending.size(); ++i) {
    const arc::HostPatch& p = w->pending[i];
    arc_gpu_patch& v = w->views[i];
    v.chunk_flat_idx = p.chunk;
    v.new_root_node_idx = p.root;
    v.node_offset = p.node_offset;
    v.n_nodes = (uint32_t)p.nodes.size();
    v.nodes = p.nodes.empty() ? nullptr : p.nodes.data();
    v.brick_offset = p.brick_offset;
    v.n_bricks = (uint32_t)p.bricks.size();
    v.bricks = p.bricks.empty() ? nullptr : p.bricks.data();
  }
  *patches = w->views.empty() ? nullptr : w->views.data();
  *n = w->views.size();
  return ARC_OK;
}

int arc_world_host_clear_patches(arc_world_host* wh) {
  if (!wh) return ARC_ERR_INVALID;
  World* w = W(wh);
  w->pending.clear();
  w->views.clear();
  return ARC_OK;
}

}  // extern "C"
