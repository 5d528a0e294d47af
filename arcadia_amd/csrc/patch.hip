// Scatter kernel of arc_world_patch: copies staged node / brick slices into the device pools and
// writes chunk roots, in one launch. The host resolves overlapping patches beforehand (the last
// patch touching a pool index wins, as sequential buffer writes would), so pieces are disjoint and
// blocks run in any order.
#include <hip/hip_runtime.h>

#include "../../include/arcadia_hip.h"
#include "patch.h"

namespace arcdev {

__global__ __launch_bounds__(256) void k_apply_patches(const PatchPiece* __restrict__ pieces, uint32_t n_pieces,
                                                       const uint2* __restrict__ root_writes, uint32_t n_roots,
                                                       const uint4* __restrict__ src, uint4* __restrict__ nodes,
                                                       uint4* __restrict__ bricks, uint32_t* __restrict__ roots) {
  const uint32_t b = blockIdx.x;
  if (b < n_pieces) {
    const PatchPiece p = pieces[b];
    uint4* dst = (p.pool == 0 ? nodes : bricks) + p.dst_unit;
    const uint4* s = src + p.src_unit;
    for (uint32_t i = threadIdx.x; i < p.n_units; i += blockDim.x) dst[i] = s[i];
  } else {
    for (uint32_t i = threadIdx.x; i < n_roots; i += blockDim.x) {
      const uint2 r = root_writes[i];
      roots[r.x] = r.y;
    }
  }
}

hipError_t launch_apply_patches(const PatchPiece* pieces, uint32_t n_pieces, const uint2* root_writes, uint32_t n_roots,
                                const uint4* src, void* nodes, void* bricks, uint32_t* roots, hipStream_t s) {
  const uint32_t grid = n_pieces + (n_roots ? 1u : 0u);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(k_apply_patches, dim3(grid), dim3(256), 0, s, pieces, n_pieces, root_writes, n_roots, src,
                     static_cast<uint4*>(nodes), static_cast<uint4*>(bricks), roots);
  return hipGetLastError();
}

// relocate_chunk_nodes (world_map.rs:705-735): empty-mask nodes become the zero node, inner nodes
// add the chunk's first node to their child base, mixed leaves add the chunk's first brick,
// uniform leaves keep only their flags.
__global__ __launch_bounds__(256) void k_relocate_chunks(uint4* __restrict__ nodes, uint32_t begin, uint32_t end,
                                                         const uint4* __restrict__ table, uint32_t n_table) {
  const uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= end) return;
  uint32_t lo = 0u, hi = n_table;  // last entry whose first node is <= i
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (table[mid].x <= i) lo = mid;
    else hi = mid;
  }
  const uint32_t nb = table[lo].x, bb = table[lo].z;
  uint4 n = nodes[i];
  if (n.x == 0u && n.y == 0u) {
    n = make_uint4(0u, 0u, 0u, 0u);
  } else {
    const uint32_t flags = n.z & (ARC_IS_LEAF_FLAG | ARC_IS_UNIFORM_FLAG);
    const uint32_t local = n.z & ARC_PTR_INDEX_MASK;
    if (n.z & ARC_IS_LEAF_FLAG) n.z = (n.z & ARC_IS_UNIFORM_FLAG) ? flags : (flags | (bb + local));
    else n.z = flags | (nb + local);
  }
  nodes[i] = n;
}

hipError_t launch_relocate_chunks(void* nodes, uint32_t begin, uint32_t end, const uint4* table, uint32_t n_table,
                                  hipStream_t s) {
  if (end <= begin || n_table == 0) return hipSuccess;
  hipLaunchKernelGGL(k_relocate_chunks, dim3((end - begin + 255u) / 256u), dim3(256), 0, s, static_cast<uint4*>(nodes),
                     begin, end, table, n_table);
  return hipGetLastError();
}

}  // namespace arcdev
