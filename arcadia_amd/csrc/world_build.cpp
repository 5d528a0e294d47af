// Host-side 64-tree construction and pool packing (C ABI in include/arcadia_world.h).
//
// Chunk trees follow ChunkTree64::build_recursive + pack_build_tree
// (crates/arcadia-core/src/chunk_tree/builder.rs:106-378): bottom-up build over 4x4x4
// children in (z, y, x) order, uniform collapse only when every present child is a fully
// solid uniform cube of one material, then breadth-first packing with contiguous sibling
// slots. The global pools follow WorldMap::rebuild_all_gpu_pools
// (crates/arcadia-world/src/world_map.rs:158-205, 689-735).
//
// Built with -ffp-contract=off: the LOD colour sums are binary32 sums in voxel order, as the
// reference computes them.
#include "world_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <new>

namespace arc {

// ---- palette: crates/arcadia-core/src/material.rs:99-204 ------------------------------------
static Palette make_palette() {
  Palette p;
  std::memset(&p, 0, sizeof p);
  auto set = [&](int i, float r, float g, float b, float rough, float er, float eg, float eb, float var) {
    p.m[i][0] = r; p.m[i][1] = g; p.m[i][2] = b; p.m[i][3] = rough;
    p.m[i][4] = er; p.m[i][5] = eg; p.m[i][6] = eb; p.m[i][7] = var;
  };
  set(0, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f);
  set(1, 0.10f, 0.11f, 0.13f, 0.90f, 0.0f, 0.0f, 0.0f, 0.03f);
  set(2, 0.28f, 0.30f, 0.34f, 0.85f, 0.0f, 0.0f, 0.0f, 0.05f);
  set(3, 0.58f, 0.54f, 0.47f, 0.80f, 0.0f, 0.0f, 0.0f, 0.06f);
  set(4, 0.46f, 0.45f, 0.40f, 0.88f, 0.0f, 0.0f, 0.0f, 0.08f);
  set(5, 0.36f, 0.40f, 0.45f, 0.72f, 0.0f, 0.0f, 0.0f, 0.05f);
  set(6, 0.48f, 0.52f, 0.57f, 0.65f, 0.0f, 0.0f, 0.0f, 0.04f);
  set(7, 0.34f, 0.23f, 0.15f, 0.92f, 0.0f, 0.0f, 0.0f, 0.06f);
  set(8, 0.22f, 0.52f, 0.18f, 0.78f, 0.0f, 0.0f, 0.0f, 0.08f);
  set(9, 0.46f, 0.58f, 0.20f, 0.78f, 0.0f, 0.0f, 0.0f, 0.08f);
  set(10, 0.16f, 0.46f, 0.22f, 0.85f, 0.0f, 0.0f, 0.0f, 0.09f);
  set(11, 0.11f, 0.32f, 0.15f, 0.88f, 0.0f, 0.0f, 0.0f, 0.06f);
  set(12, 0.68f, 0.60f, 0.46f, 0.82f, 0.0f, 0.0f, 0.0f, 0.07f);
  set(13, 0.27f, 0.18f, 0.12f, 0.88f, 0.0f, 0.0f, 0.0f, 0.06f);
  set(14, 0.48f, 0.33f, 0.20f, 0.75f, 0.0f, 0.0f, 0.0f, 0.05f);
  set(15, 0.15f, 0.44f, 0.18f, 0.70f, 0.0f, 0.0f, 0.0f, 0.10f);
  set(16, 0.76f, 0.36f, 0.11f, 0.70f, 0.0f, 0.0f, 0.0f, 0.10f);
  set(17, 0.85f, 0.48f, 0.58f, 0.72f, 0.0f, 0.0f, 0.0f, 0.08f);
  set(18, 0.06f, 0.28f, 0.38f, 0.05f, 0.0f, 0.0f, 0.0f, 0.02f);
  set(19, 0.15f, 0.85f, 1.00f, 0.15f, 0.20f, 2.40f, 3.60f, 0.05f);
  set(20, 0.95f, 0.20f, 0.80f, 0.15f, 3.20f, 0.25f, 2.50f, 0.05f);
  set(21, 1.00f, 0.68f, 0.15f, 0.15f, 3.40f, 1.75f, 0.15f, 0.05f);
  set(22, 1.00f, 0.30f, 0.04f, 0.30f, 4.20f, 1.10f, 0.08f, 0.10f);
  set(23, 0.10f, 0.09f, 0.08f, 0.95f, 0.0f, 0.0f, 0.0f, 0.05f);
  set(24, 0.95f, 0.25f, 0.04f, 0.50f, 3.20f, 0.75f, 0.05f, 0.12f);
  set(25, 0.82f, 0.80f, 0.75f, 0.35f, 0.0f, 0.0f, 0.0f, 0.03f);
  set(26, 0.85f, 0.68f, 0.24f, 0.25f, 0.0f, 0.0f, 0.0f, 0.04f);
  set(27, 1.00f, 0.97f, 0.90f, 0.05f, 116.00f, 114.50f, 111.80f, 0.02f);
  set(28, 0.92f, 0.95f, 0.98f, 0.02f, 0.0f, 0.0f, 0.0f, 0.01f);
  // Procedural fill of ids 29..255 (material.rs:189-201), binary32 with libm sinf/cosf.
  const float tau = 6.28318530717958647692f, pi = 3.14159265358979323846f;
  for (int i = 29; i < 256; ++i) {
    float t = (float)i / 255.0f;
    set(i, 0.3f + 0.4f * std::fabs(std::sin(t * tau)), 0.3f + 0.4f * std::fabs(std::cos(t * 3.0f * pi)),
        0.3f + 0.3f * t, 0.75f, 0.0f, 0.0f, 0.0f, 0.05f);
  }
  return p;
}

const Palette& palette() {
  static const Palette p = make_palette();
  return p;
}

static inline uint32_t unorm8(float v) {
  // (v.clamp(0,1) * 255.0) as u32
  float c = std::fmin(std::fmax(v, 0.0f), 1.0f);
  float s = c * 255.0f;
  return (uint32_t)s;
}

uint32_t lod_payload_of(uint32_t mat) {  // Material::pack_lod_payload, material.rs:70-75
  const float* m = palette().m[mat & 0xFF];
  return ((mat & 0xFFu) << 24) | (unorm8(m[2]) << 16) | (unorm8(m[1]) << 8) | unorm8(m[0]);
}

bool is_emissive(uint32_t mat) {  // material.rs:77-80
  const float* m = palette().m[mat & 0xFF];
  return m[4] > 0.01f || m[5] > 0.01f || m[6] > 0.01f;
}

// ---- build_recursive over a dense chunk ---------------------------------------------------
namespace {

enum Kind : uint8_t { kEmpty = 0, kUniform = 1, kMixed = 2, kInternal = 3 };

struct BNode {
  uint64_t mask;
  uint32_t lod;
  uint32_t a;      // mixed: temp brick index; internal: first child in arena
  uint16_t count;  // internal: number of children
  uint8_t kind;
  uint8_t full;    // uniform: is_fully_solid_cube
  uint8_t mat;     // uniform: material
};

struct Builder {
  const uint8_t* dense;
  std::vector<BNode> arena;
  std::vector<arc_brick64> tmp_bricks;
  uint64_t solid = 0;

  inline uint8_t at(uint32_t x, uint32_t y, uint32_t z) const {
    return dense[((size_t)z * 256u + y) * 256u + x];
  }

  BNode leaf(uint32_t ox, uint32_t oy, uint32_t oz) {  // builder.rs:139-209
    BNode n{};
    uint64_t mask = 0;
    arc_brick64 brick;
    std::memset(&brick, 0, sizeof brick);
    uint32_t first_mat = 0;
    bool all_same = true;
    float rs = 0.0f, gs = 0.0f, bs = 0.0f;
    uint32_t emissive_mat = 0, emissive_vox = 0;
    const Palette& pal = palette();
    for (uint32_t lz = 0; lz < 4; ++lz)
      for (uint32_t ly = 0; ly < 4; ++ly)
        for (uint32_t lx = 0; lx < 4; ++lx) {
          uint32_t mat = at(ox + lx, oy + ly, oz + lz);
          if (mat == 0) continue;
          uint32_t b = (lz << 4) | (ly << 2) | lx;
          mask |= 1ull << b;
          brick.packed_materials[b >> 2] |= mat << ((b & 3u) * 8u);
          if (first_mat == 0) first_mat = mat;
          else if (mat != first_mat) all_same = false;
          const float* m = pal.m[mat];
          float er = m[4] * 0.25f, eg = m[5] * 0.25f, eb = m[6] * 0.25f;
          rs += std::fmin(m[0] + er, 1.0f);
          gs += std::fmin(m[1] + eg, 1.0f);
          bs += std::fmin(m[2] + eb, 1.0f);
          if (is_emissive(mat)) {
            emissive_mat = mat;
            emissive_vox++;
          }
        }
    if (mask == 0) {
      n.kind = kEmpty;
      return n;
    }
    uint32_t count = (uint32_t)__builtin_popcountll(mask);
    solid += count;
    if (all_same) {
      n.kind = kUniform;
      n.mask = mask;
      n.mat = (uint8_t)first_mat;
      n.lod = lod_payload_of(first_mat);
      n.full = mask == ~0ull;
      return n;
    }
    float inv = 1.0f / (float)count;
    uint32_t rep = (emissive_mat != 0 && emissive_vox * 2u >= count) ? emissive_mat : first_mat;
    float r = rs * inv, g = gs * inv, b = bs * inv;
    n.kind = kMixed;
    n.mask = mask;
    n.lod = (rep << 24) | (unorm8(b) << 16) | (unorm8(g) << 8) | unorm8(r);  // pack_rgb_and_mat
    n.a = (uint32_t)tmp_bricks.size();
    tmp_bricks.push_back(brick);
    return n;
  }

  BNode build(uint32_t ox, uint32_t oy, uint32_t oz, uint32_t shift) {  // builder.rs:106-313
    if (shift == 0) return leaf(ox, oy, oz);
    const uint32_t child = 1u << shift;
    BNode kids[64];
    int nk = 0;
    uint64_t mask = 0;
    bool all_full_uniform = true;
    uint32_t common = 0, rsum = 0, gsum = 0, bsum = 0, solid_mat = 0, emissive_mat = 0, emissive_kids = 0;
    for (uint32_t cz = 0; cz < 4; ++cz)
      for (uint32_t cy = 0; cy < 4; ++cy)
        for (uint32_t cx = 0; cx < 4; ++cx) {
          BNode c = build(ox + cx * child, oy + cy * child, oz + cz * child, shift - 2);
          if (c.kind == kEmpty) continue;
          uint32_t b = (cz << 4) | (cy << 2) | cx;
          mask |= 1ull << b;
          rsum += c.lod & 0xFFu;
          gsum += (c.lod >> 8) & 0xFFu;
          bsum += (c.lod >> 16) & 0xFFu;
          uint32_t cm = c.lod >> 24;
          if (is_emissive(cm)) {
            emissive_mat = cm;
            emissive_kids++;
          } else if (solid_mat == 0) {
            solid_mat = cm;
          }
          if (c.kind == kUniform && c.full) {
            if (common == 0) common = c.mat;
            else if (common != c.mat) all_full_uniform = false;
          } else {
            all_full_uniform = false;
          }
          kids[nk++] = c;
        }
    BNode n{};
    if (mask == 0) {
      n.kind = kEmpty;
      return n;
    }
    if (all_full_uniform && common != 0) {  // :279-289
      n.kind = kUniform;
      n.mask = mask;
      n.mat = (uint8_t)common;
      n.lod = lod_payload_of(common);
      n.full = mask == ~0ull;
      // Collapsed children are dropped; their temp bricks (none: uniform) need no cleanup.
      return n;
    }
    uint32_t count = (uint32_t)nk;
    uint32_t dom = (emissive_mat != 0 && (emissive_kids * 2u >= count || solid_mat == 0)) ? emissive_mat : solid_mat;
    n.kind = kInternal;
    n.mask = mask;
    n.lod = (dom << 24) | (((bsum / count) & 0xFFu) << 16) | (((gsum / count) & 0xFFu) << 8) | ((rsum / count) & 0xFFu);
    n.a = (uint32_t)arena.size();
    n.count = (uint16_t)count;
    arena.insert(arena.end(), kids, kids + nk);
    return n;
  }
};

}  // namespace

static inline arc_node64 make_node(uint64_t mask, uint32_t ptr, uint32_t lod) {
  arc_node64 n;
  n.mask_low = (uint32_t)mask;
  n.mask_high = (uint32_t)(mask >> 32);
  n.child_ptr = ptr;
  n.lod_payload = lod;
  return n;
}

// pack_build_tree: builder.rs:318-378 (breadth-first, contiguous siblings)
static void pack_tree(const Builder& bld, const BNode& root, ChunkTree& out) {
  out.nodes.clear();
  out.bricks.clear();
  if (root.kind == kEmpty) {
    out.nodes.push_back(make_node(0, 0, 0));
    return;
  }
  out.nodes.push_back(make_node(0, 0, 0));
  std::deque<std::pair<uint32_t, BNode>> q;
  q.emplace_back(0u, root);
  while (!q.empty()) {
    auto [slot, b] = q.front();
    q.pop_front();
    switch (b.kind) {
      case kEmpty:
        out.nodes[slot] = make_node(0, 0, 0);
        break;
      case kUniform:
        out.nodes[slot] = make_node(b.mask, ARC_IS_LEAF_FLAG | ARC_IS_UNIFORM_FLAG, b.lod);
        break;
      case kMixed: {
        uint32_t bi = (uint32_t)out.bricks.size();
        out.bricks.push_back(bld.tmp_bricks[b.a]);
        out.nodes[slot] = make_node(b.mask, ARC_IS_LEAF_FLAG | bi, b.lod);
        break;
      }
      case kInternal: {
        uint32_t base = (uint32_t)out.nodes.size();
        out.nodes.resize(out.nodes.size() + b.count, make_node(0, 0, 0));
        out.nodes[slot] = make_node(b.mask, base, b.lod);
        for (uint32_t i = 0; i < b.count; ++i) q.emplace_back(base + i, bld.arena[b.a + i]);
        break;
      }
    }
  }
}

void build_chunk_dense(const uint8_t* dense, ChunkTree& out) {
  Builder bld;
  bld.dense = dense;
  BNode root = bld.build(0, 0, 0, 6);
  pack_tree(bld, root, out);
  out.solid = bld.solid;
}

// ---- rebuild_all_gpu_pools: world_map.rs:158-205, 689-735 ----------------------------------
static inline uint32_t part1by2(uint32_t n) {
  n &= 0x000003ffu;
  n = (n ^ (n << 16)) & 0xff0000ffu;
  n = (n ^ (n << 8)) & 0x0300f00fu;
  n = (n ^ (n << 4)) & 0x030c30c3u;
  n = (n ^ (n << 2)) & 0x09249249u;
  return n;
}

uint32_t morton3d(uint32_t x, uint32_t y, uint32_t z) {
  return part1by2(x) | (part1by2(y) << 1) | (part1by2(z) << 2);
}

void pack_world(World& w) {
  const uint32_t nchunks = w.grid[0] * w.grid[1] * w.grid[2];
  w.roots.assign(nchunks, ARC_EMPTY_CHUNK_SENTINEL);
  w.nodes.clear();
  w.bricks.clear();
  w.allocs.assign(nchunks, ChunkAlloc{0, 0, 0, 0});
  w.free_node_spans.clear();
  w.free_brick_spans.clear();
  w.pending.clear();
  w.views.clear();
  w.nodes.push_back(make_node(0, 0, 0));
  arc_brick64 zb;
  std::memset(&zb, 0, sizeof zb);
  w.bricks.push_back(zb);
  std::vector<std::pair<uint32_t, uint32_t>> order(nchunks);
  for (uint32_t i = 0; i < nchunks; ++i) {
    uint32_t cx = i % w.grid[0], cy = (i / w.grid[0]) % w.grid[1], cz = i / (w.grid[0] * w.grid[1]);
    order[i] = {morton3d(cx, cy, cz), i};
  }
  std::sort(order.begin(), order.end());
  size_t total_nodes = 1, total_bricks = 1;
  for (auto& c : w.chunks) {
    total_nodes += c.nodes.size();
    total_bricks += c.bricks.size();
  }
  w.nodes.reserve(total_nodes);
  w.bricks.reserve(total_bricks);
  for (auto& [code, idx] : order) {
    (void)code;
    const ChunkTree& c = w.chunks[idx];
    if (c.nodes.empty() || (c.nodes[0].mask_low == 0 && c.nodes[0].mask_high == 0)) continue;
    uint32_t nb = (uint32_t)w.nodes.size(), bb = (uint32_t)w.bricks.size();
    for (const arc_node64& n : c.nodes) {
      if (n.mask_low == 0 && n.mask_high == 0) {
        w.nodes.push_back(make_node(0, 0, 0));
        continue;
      }
      uint32_t flags = n.child_ptr & (ARC_IS_LEAF_FLAG | ARC_IS_UNIFORM_FLAG);
      uint32_t local = n.child_ptr & ARC_PTR_INDEX_MASK;
      uint32_t ptr;
      if (n.child_ptr & ARC_IS_LEAF_FLAG) ptr = (n.child_ptr & ARC_IS_UNIFORM_FLAG) ? flags : (flags | (bb + local));
      else ptr = flags | (nb + local);
      arc_node64 r = n;
      r.child_ptr = ptr;
      w.nodes.push_back(r);
    }
    w.bricks.insert(w.bricks.end(), c.bricks.begin(), c.bricks.end());
    w.roots[idx] = nb;
    w.allocs[idx] = ChunkAlloc{nb, (uint32_t)c.nodes.size(), bb, (uint32_t)c.bricks.size()};
  }
  w.packed = true;
}

}  // namespace arc

// ---- C ABI ---------------------------------------------------------------------------------
using namespace arc;

extern "C" {

int arc_world_host_create(const uint32_t grid[3], arc_world_host** out) {
  if (!grid || !out) return ARC_ERR_INVALID;
  for (int a = 0; a < 3; ++a)
    if (grid[a] == 0 || grid[a] > 256) return ARC_ERR_INVALID;
  World* w = new (std::nothrow) World();
  if (!w) return ARC_ERR_INVALID;
  std::memcpy(w->grid, grid, sizeof w->grid);
  w->chunks.resize((size_t)grid[0] * grid[1] * grid[2]);
  for (auto& c : w->chunks) c.nodes.push_back(make_node(0, 0, 0));
  *out = reinterpret_cast<arc_world_host*>(w);
  return ARC_OK;
}

void arc_world_host_destroy(arc_world_host* w) { delete reinterpret_cast<World*>(w); }

int arc_world_host_set_chunk_dense(arc_world_host* wh, uint32_t idx, const uint8_t* dense) {
  World* w = reinterpret_cast<World*>(wh);
  if (!w || !dense || idx >= w->chunks.size()) return ARC_ERR_INVALID;
  build_chunk_dense(dense, w->chunks[idx]);
  w->packed = false;
  return ARC_OK;
}

int arc_world_host_set_chunk_sparse(arc_world_host* wh, uint32_t idx, const uint32_t* v, uint64_t n) {
  World* w = reinterpret_cast<World*>(wh);
  if (!w || (!v && n) || idx >= w->chunks.size()) return ARC_ERR_INVALID;
  // from_sparse_voxels (builder.rs:422-484): last write wins per voxel, air and out-of-range
  // entries skipped; the region hint only skips empty regions, so building from the dense
  // image gives the same tree.
  std::vector<uint8_t> dense((size_t)256 * 256 * 256, 0);
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t x = v[4 * i], y = v[4 * i + 1], z = v[4 * i + 2], m = v[4 * i + 3] & 0xFFu;
    if (m == 0 || x >= 256 || y >= 256 || z >= 256) continue;
    dense[((size_t)z * 256 + y) * 256 + x] = (uint8_t)m;
  }
  build_chunk_dense(dense.data(), w->chunks[idx]);
  w->packed = false;
  return ARC_OK;
}

int arc_world_host_chunk_tree(const arc_world_host* wh, uint32_t idx, arc_node64* nodes, uint64_t* n_nodes,
                              arc_brick64* bricks, uint64_t* n_bricks) {
  const World* w = reinterpret_cast<const World*>(wh);
  if (!w || idx >= w->chunks.size() || !n_nodes || !n_bricks) return ARC_ERR_INVALID;
  const ChunkTree& c = w->chunks[idx];
  if (nodes) {
    if (*n_nodes < c.nodes.size()) return ARC_ERR_CAPACITY;
    std::memcpy(nodes, c.nodes.data(), c.nodes.size() * sizeof(arc_node64));
  }
  if (bricks) {
    if (*n_bricks < c.bricks.size()) return ARC_ERR_CAPACITY;
    std::memcpy(bricks, c.bricks.data(), c.bricks.size() * sizeof(arc_brick64));
  }
  *n_nodes = c.nodes.size();
  *n_bricks = c.bricks.size();
  return ARC_OK;
}

int arc_world_host_pack(arc_world_host* wh) {
  World* w = reinterpret_cast<World*>(wh);
  if (!w) return ARC_ERR_INVALID;
  pack_world(*w);
  return ARC_OK;
}

int arc_world_host_pools(const arc_world_host* wh, uint32_t grid[3], const uint32_t** roots,
                         const arc_node64** nodes, uint64_t* n_nodes, const arc_brick64** bricks,
                         uint64_t* n_bricks) {
  const World* w = reinterpret_cast<const World*>(wh);
  if (!w || !w->packed) return ARC_ERR_INVALID;
  if (grid) std::memcpy(grid, w->grid, sizeof w->grid);
  if (roots) *roots = w->roots.data();
  if (nodes) *nodes = w->nodes.data();
  if (n_nodes) *n_nodes = w->nodes.size();
  if (bricks) *bricks = w->bricks.data();
  if (n_bricks) *n_bricks = w->bricks.size();
  return ARC_OK;
}

int arc_palette_material(uint32_t mat, float out8[8]) {
  if (mat > 255 || !out8) return ARC_ERR_INVALID;
  std::memcpy(out8, palette().m[mat], 8 * sizeof(float));
  return ARC_OK;
}

uint32_t arc_palette_lod_payload(uint32_t mat) { return lod_payload_of(mat); }

}  // extern "C"
