// Internal host-side world types shared by the world builder, generators and edit engine.
#pragma once

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

struct World {
  uint32_t grid[3] = {1, 1, 1};
  std::vector<ChunkTree> chunks;
  // packed pools (rebuild_all_gpu_pools)
  std::vector<uint32_t> roots;
  std::vector<arc_node64> nodes;
  std::vector<arc_brick64> bricks;
  std::vector<ChunkAlloc> allocs;
  bool packed = false;
};

void build_chunk_dense(const uint8_t* dense, ChunkTree& out);
void pack_world(World& w);
uint32_t morton3d(uint32_t x, uint32_t y, uint32_t z);

}  // namespace arc
