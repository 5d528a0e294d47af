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

}  // namespace arcdev
