// Device side of arc_world_patch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "patch_plan.h"

namespace arcdev {

using PatchPiece = arcplan::Piece;
constexpr uint32_t kPieceUnits = arcplan::kPieceUnits;

hipError_t launch_apply_patches(const PatchPiece* pieces, uint32_t n_pieces, const uint2* root_writes, uint32_t n_roots,
                                const uint4* src, void* nodes, void* bricks, uint32_t* roots, hipStream_t s);

// Relocation of snapshot chunks streamed into the node pool: nodes [begin, end) of the pool hold
// local trees; `table` has one (first node, node count, first brick, 0) entry per chunk, sorted by
// first node. Each node's pointer is rebased as rebuild_all_gpu_pools rebases it.
hipError_t launch_relocate_chunks(void* nodes, uint32_t begin, uint32_t end, const uint4* table, uint32_t n_table,
                                  hipStream_t s);

}  // namespace arcdev
