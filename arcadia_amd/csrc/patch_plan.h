// Host planning of a patch batch (arc_world_patch): validation, last-writer-wins resolution of
// overlapping slices, and the staging layout the scatter kernel reads. No HIP types: the same
// plan drives the device scatter and the host model of it (arc_world_patch_host).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/arcadia_hip.h"

namespace arcplan {

// One contiguous copy of at most kPieceUnits 16-byte units into a pool.
struct Piece {
  uint64_t dst_unit;  // destination, in 16 B units of the pool (node = 1 unit, brick = 4 units)
  uint32_t src_unit;  // source, in 16 B units of the staging buffer
  uint16_t n_units;   // 1..kPieceUnits
  uint16_t pool;      // 0 = node pool, 1 = brick pool
};
static_assert(sizeof(Piece) == 16, "piece table entries are 16 B");
constexpr uint32_t kPieceUnits = 4096;

struct RootWrite {
  uint32_t chunk, root;
};

struct Plan {
  uint64_t need_nodes = 0, need_bricks = 0;  // pool sizes after growth
  std::vector<Piece> pieces;
  std::vector<RootWrite> roots;
  struct Src {
    const void* p;
    uint64_t dst, n;  // pool element index and count
    uint16_t pool;
  };
  std::vector<Src> parts;  // data in staging order
  uint64_t header_units = 0, total_units = 0;
  uint64_t max_node_written = 0, max_brick_written = 0;  // one past the last element written
};

// Checks every patch (chunk index, null slices, 30-bit ranges, child / brick pointers inside the
// grown pools) and plans the scatter. Returns ARC_OK or an ARC_ERR_* code with `err` set.
int plan_patches(const arc_gpu_patch* patches, uint64_t n, uint64_t nchunks, uint64_t node_cap, uint64_t brick_cap,
                 Plan& plan, std::string& err);
// Writes the staging image: [pieces][root writes][data], total_units * 16 bytes.
void fill_staging(const Plan& plan, void* staging);
// Host model of the scatter kernel over a staging image.
void scatter_host(const void* staging, uint32_t n_pieces, uint32_t n_roots, uint32_t* roots, void* nodes,
                  void* bricks);

}  // namespace arcplan
