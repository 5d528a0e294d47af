// Internal host-side world types shared by the world builder, generators and edit engine.
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

#include "../../include/arcadia_world.h"

namespace arc {

struct Palette {
  float m[256][8];  // albedo rgb, roughness, emission rgb, colour variance
};
const Palette& palette();
uint32_t lod_payload_of(uint32_t mat);
bool is_emissive(uint32_t mat);

// One chunk's local 64-tree (nodes[0] is the root), as ChunkTree64 holds it.
struct ChunkTree {
  std::vector<arc_node64> nodes;
  std::vector<arc_brick64> bricks;
  uint64_t solid = 0;
};

struct ChunkAlloc {
  uint32_t node_base, node_cap, brick_base, brick_cap;
};

// ChunkEditReport (crates/arcadia-core/src/chunk_tree/edit/set.rs:17-80): dirty index ranges of
// one chunk's local node / brick arrays after in-place edits.
struct EditReport {
  uint32_t added = 0;
  size_t min_n = SIZE_MAX, max_n = 0, min_b = SIZE_MAX, max_b = 0;
  bool growth = false;
  void node(size_t i) {
    min_n = std::min(min_n, i);
    max_n = std::max(max_n, i);
  }
  void node_range(size_t s, size_t e) {
    if (s <= e) {
      min_n = std::min(min_n, s);
      max_n = std::max(max_n, e);
    }
  }
  void brick(size_t i) {
    min_b = std::min(min_b, i);
    max_b = std::max(max_b, i);
  }
};

// GpuPatch (crates/arcadia-world/src/world_map.rs:20-27)
struct HostPatch {
  uint32_t chunk = 0, root = 0, node_offset = 0, brick_offset = 0;
  std::vector<arc_node64> nodes;
  std::vector<arc_brick64> bricks;
};

struct World {
  uint32_t grid[3] = {1, 1, 1};
  std::vector<ChunkTree> chunks;
  // packed pools (rebuild_all_gpu_pools) and their per-chunk allocations
  std::vector<uint32_t> roots;
  std::vector<arc_node64> nodes;
  std::vector<arc_brick64> bricks;
  std::vector<ChunkAlloc> allocs;
  std::vector<std::pair<uint32_t, uint32_t>> free_node_spans, free_brick_spans;
  std::vector<HostPatch> pending;   // queued GPU patches, in order
  std::vector<arc_gpu_patch> views; // C views of `pending` handed out by arc_world_host_pending_patches
  bool packed = false;
};

void build_chunk_dense(const uint8_t* dense, ChunkTree& out);

// edit engine (world_edit.cpp)
uint32_t chunk_get_voxel(const ChunkTree& t, uint32_t x, uint32_t y, uint32_t z);
EditReport chunk_set_voxels(ChunkTree& t, const uint32_t* xyzm, uint64_t n);
uint32_t chunk_clear_voxels(ChunkTree& t, const uint32_t* xyz, uint64_t n, EditReport* out);
uint32_t chunk_carve_sphere(ChunkTree& t, const float center[3], float radius, bool scorch_rim, EditReport* out);
void chunk_compact(ChunkTree& t);
void world_sync_chunk(World& w, uint32_t idx);
void world_sync_chunk_partial(World& w, uint32_t idx, const EditReport& rep);
uint32_t world_set_voxels(World& w, const int32_t* xyzm, uint64_t n);
uint32_t world_clear_voxels(World& w, const int32_t* xyz, uint64_t n);
uint32_t world_carve_crater(World& w, const float center_m[3], float radius_m, bool scorch_rim);
void pack_world(World& w);
uint32_t morton3d(uint32_t x, uint32_t y, uint32_t z);

}  // namespace arc
