// Device-side 64-tree DDA traversal for gfx950 (wave64).
//
// One lane carries one ray. The per-ray state machine restates trace_world_64tree_from
// (shaders/common/traversal64.wgsl:226-556) and raycast_chunk_tree
// (crates/arcadia-core/src/raycast_cpu.rs:28-218) operation for operation in binary32 with
// no contraction (the library is compiled with -ffp-contract=off and correctly rounded
// division), so hit voxel, face, t and step count equal the reference's source semantics.
//
// What differs from the WGSL text and why it cannot change a result:
//   * the reference's 2-slot chunk cache only saves chunk_roots reads: every chunk change reads
//     chunk_roots[key] directly, which yields the same root index;
//   * the ancestor restart keeps the cached Node64s (chunk root, level-1 and level-2 ancestors in
//     LDS, the deepest node in registers) and restarts at the level the reference's restart
//     conditions select; a restart is equivalent to a root descent because the cached node at
//     shift s always covers the current voxel's bits >= s+2 (the XOR `diff` tests), so results
//     are identical and only the number of node loads differs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/arcadia_hip.h"

namespace arcdev {

struct DevWorld {
  uint32_t gx, gy, gz;
  uint32_t pad;
  float ext[3];     // world extent in voxels, (float)g * 256 (binary32, computed on the host)
  float ext_in[3];  // ext - 1e-3 (binary32): the clamp bound of the start voxel
  const uint32_t* __restrict__ roots;
  const uint4* __restrict__ nodes;
  const uint32_t* __restrict__ bricks;  // 16 words per brick
};

struct DevParams {
  uint32_t lod_bias;
  uint32_t ignore_water;
  uint32_t max_steps;
  uint32_t lod_step;     // step threshold for the lod_bias class
  float lod_t;           // distance threshold (voxels) for the lod_bias class
  float lod_t_far;       // lod_bias 1: sec_lod_t * 2.2857
  float ao_max_dist;     // point_light_count.w as f32
  uint32_t flags;        // kParamLodOff
};

// DevParams::flags: the LOD floor is 0 at every step (distance thresholds +inf or NaN, step
// threshold 2^32 - 1), so the traversal skips its LOD tests.
constexpr uint32_t kParamLodOff = 1u;

__device__ __forceinline__ uint32_t f2u(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

__device__ __forceinline__ float wgsl_sign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

__device__ __forceinline__ float rust_signum(float x) {
  if (x != x) return x;
  return __builtin_signbit(x) ? -1.0f : 1.0f;
}

__device__ __forceinline__ uint32_t umin_(uint32_t a, uint32_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint32_t umax_(uint32_t a, uint32_t b) { return a > b ? a : b; }

__device__ __forceinline__ bool has_child(uint32_t ml, uint32_t mh, uint32_t b) {
  uint32_t word = b < 32u ? ml : mh;
  return ((word >> (b & 31u)) & 1u) != 0u;
}

__device__ __forceinline__ uint32_t child_offset(uint4 n, uint32_t b) {
  // base + popcount(mask & ((1<<b)-1)), split in two 32-bit halves
  uint32_t lo_mask = b >= 32u ? 0xFFFFFFFFu : ((1u << b) - 1u);
  uint32_t hi_mask = b >= 32u ? ((1u << (b - 32u)) - 1u) : 0u;
  return (n.z & ARC_PTR_INDEX_MASK) + __builtin_popcount(n.x & lo_mask) + __builtin_popcount(n.y & hi_mask);
}

__device__ __forceinline__ uint32_t pcg_hash(uint32_t input) {
  uint32_t state = input * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

__device__ __forceinline__ float clampf_(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

// eval_leaf_bitwise_ao: traversal64.wgsl:60-131
__device__ inline float eval_leaf_ao(uint32_t ml, uint32_t mh, uint32_t cx, uint32_t cy, uint32_t cz, float n0, float n1,
                                     float n2, float h0, float h1, float h2) {
  float fx = clampf_(h0 - floorf(h0), 0.0f, 1.0f);
  float fy = clampf_(h1 - floorf(h1), 0.0f, 1.0f);
  float fz = clampf_(h2 - floorf(h2), 0.0f, 1.0f);
  int nx = (int)cx + (int)n0, ny = (int)cy + (int)n1, nz = (int)cz + (int)n2;
  if (nx < 0 || nx > 3 || ny < 0 || ny > 3 || nz < 0 || nz > 3) return 1.0f;
  uint32_t bb = ((uint32_t)nz << 4) | ((uint32_t)ny << 2) | (uint32_t)nx;
  float occ = 0.0f;
  if (fabsf(n1) > 0.5f) {
    if (nx > 0 && has_child(ml, mh, bb - 1u)) occ += (1.0f - fx) * 0.42f;
    if (nx < 3 && has_child(ml, mh, bb + 1u)) occ += fx * 0.42f;
    if (nz > 0 && has_child(ml, mh, bb - 16u)) occ += (1.0f - fz) * 0.42f;
    if (nz < 3 && has_child(ml, mh, bb + 16u)) occ += fz * 0.42f;
  } else if (fabsf(n0) > 0.5f) {
    if (ny > 0 && has_child(ml, mh, bb - 4u)) occ += (1.0f - fy) * 0.42f;
    if (ny < 3 && has_child(ml, mh, bb + 4u)) occ += fy * 0.42f;
    if (nz > 0 && has_child(ml, mh, bb - 16u)) occ += (1.0f - fz) * 0.42f;
    if (nz < 3 && has_child(ml, mh, bb + 16u)) occ += fz * 0.42f;
  } else {
    if (nx > 0 && has_child(ml, mh, bb - 1u)) occ += (1.0f - fx) * 0.42f;
    if (nx < 3 && has_child(ml, mh, bb + 1u)) occ += fx * 0.42f;
    if (ny > 0 && has_child(ml, mh, bb - 4u)) occ += (1.0f - fy) * 0.42f;
    if (ny < 3 && has_child(ml, mh, bb + 4u)) occ += fy * 0.42f;
  }
  return clampf_(1.0f - occ, 0.35f, 1.0f);
}

}  // namespace arcdev
