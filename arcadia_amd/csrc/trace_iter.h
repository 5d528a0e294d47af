// One outer iteration of the traversal (the reference loop body), hit resolution and the compact
// hit record. Shared by the traversal kernels (k_trace, k_path).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "trace_device.h"
#include "trace_walk.h"

namespace arcdev {

// One outer iteration of a live lane whose guards passed (traversal64.wgsl:311-552).
// Instruction count is what bounds the kernel, so the sequence is written for few VALU ops:
//   * the child test and the child slot come from one 64-bit shift of the child mask that puts
//     bit b at the top: the sign gives has_child(b), the popcount minus one gives
//     popcount(mask & ((1 << b) - 1));
//   * the descent loop only tests and loads; the material of a hit candidate (a set bit in a
//     leaf or at the LOD floor) is resolved after the loop;
//   * cached_shift always equals the level the iteration ended on, so one field (L.cs) carries
//     both;
//   * the box step computes the clamped inside coordinate of all three axes and selects, with no
//     branches (box_step);
//   * WATER = false compiles the ignore_water tests out (ray classes that never set it).
template <uint32_t SEM, bool WATER, uint32_t BLOCK = kMaxBlock, uint32_t NCH = kLdsChunks, bool OFF32 = false>
__device__ __forceinline__ void iteration(const DevWorld& w, const DevParams& p, Lane& L, uint4* C,
                                          const RootLdsT<NCH>& R, bool lds_roots) {
  uint32_t st = L.st;
  uint4 node = L.node;
  uint32_t sh = L.cs;
  const uint32_t sidx = L.step;
  L.step = sidx + 1u;
  const uint32_t minsh = lod_min_shift<SEM>(p, L.t, sidx);
  const bool descend = restart<SEM, BLOCK, NCH, OFF32>(w, L, C, R, lds_roots, minsh, sh, node);
  const uint32_t vb = voxel_bytes(L);
  uint32_t span_mul = descend ? 1u : 4u;  // the box stepped over is span_mul << sh (empty chunk: 256)
  // Sign bit of `hw`: the voxel's child bit is set at the level the descent ended on (a hit
  // candidate: hit unless the voxel is air/water). `s` is 63 minus that child's slot.
  uint32_t hw = 0u, s = 0u;
  if (descend) {
    // descent (:390-552): at most three loads (shift 6 -> 4 -> 2 -> 0; shift 0 always stops), as
    // four unrolled tests whose exit is wave-uniform, so the compiler keeps no per-lane loop-exit
    // masks. A lane that has stopped repeats its own test with unchanged state until the wave exits.
    // The tests are sign bits combined in one word, so `go` is a single compare and the ballot
    // folds into it: bit 31 of the high word of `below` = child present; bit 31 of node.z = leaf;
    // bit 31 of minsh - sh = sh > minsh (both below 8).
#pragma unroll
    for (uint32_t pass = 0; pass < 4u; ++pass) {
      // Opaque to the compiler: left visible, it computes the next test only inside the load
      // block (the other lanes' state is unchanged) and then rebuilds `go` from a lane mask with
      // two more vector instructions per pass.
      asm volatile("" : "+v"(sh));
      s = child_shift63(vb, sh);
      const uint64_t below = (((uint64_t)node.y << 32) | node.x) << s;
      hw = (uint32_t)(below >> 32);
      if (SEM == ARC_SEM_WORLD && WATER && p.ignore_water && (node.z & ARC_IS_UNIFORM_FLAG) != 0u &&
          (node.w >> 24) == ARC_MAT_WATER) {
        hw = 0u;  // a uniform water node ignored: no child here, step over the whole cell
        span_mul = 4u;
      }
      // Level-3 nodes are leaves in every pool the upload check accepts; sh == 0 (implied by
      // sh <= minsh) only bounds the descent if a malformed node slipped in through a patch.
      // (minsh - sh) & ~node.z & hw as one three-input bit operation (truth table 0x20)
      const bool go = (int32_t)__builtin_amdgcn_bitop3_b32(minsh - sh, node.z, hw, 0x20) < 0;
      if (__builtin_amdgcn_ballot_w64(go) == 0ull) break;
      if (go) {
        node = pool_node<OFF32>(w, (node.z & ARC_PTR_INDEX_MASK) + (uint32_t)__popcll(below) - 1u);
        sh -= 2u;
        *cold_slot<BLOCK>(C, sh) = node;
      }
    }
  }
  // After the loop no lane goes on: a set child bit is at a leaf or at the LOD floor.
  const bool cand = (int32_t)hw < 0;
  bool hit = false;
  if (cand) {  // :441-538
    const uint32_t b = s ^ 63u;
    uint32_t mat = node.w >> 24;
    if ((node.z & (ARC_IS_LEAF_FLAG | ARC_IS_UNIFORM_FLAG)) == ARC_IS_LEAF_FLAG) {
      const uint32_t word = pool_brick_word<OFF32>(w, node.z & ARC_PTR_INDEX_MASK, b >> 2);
      mat = __builtin_amdgcn_ubfe(word, (b & 3u) * 8u, 8);
    }
    const bool skip_water = SEM == ARC_SEM_WORLD && WATER && p.ignore_water && mat == ARC_MAT_WATER;
    hit = !skip_water && mat != ARC_MAT_AIR;  // hit: resolved at finalize (:467-523)
    if (hit) st = fput(st, F_MAT, 8, mat);
  }
  // One way out of the iteration (a hit returns no earlier), so the lane state has one
  // definition per field where the iteration ends.
  L.node = node;
  L.cs = sh;
  if (hit) {
    st = fput(fput(fput(st, F_SHIFT, 3, sh), F_LIVE, 1, 0u), F_HIT, 1, 1u);
  } else {
    box_step(L, span_mul << sh);
    if (iteration_guards_fail<SEM>(w, p, L)) st = fput(fput(st, F_LIVE, 1, 0u), F_HIT, 1, 0u);
  }
  L.st = st;
}

// ---------------------------------------------------------------------------------------
// Finalize: hit resolution (traversal64.wgsl:467-523 / raycast_cpu.rs:162-185, 295-327)
// ---------------------------------------------------------------------------------------
struct HitInfo {
  float exact_t;
  uint32_t axis;
  float nrm_axis;
  float hx, hy, hz;  // hit point, voxel units
};

template <uint32_t SEM>
__device__ __forceinline__ HitInfo resolve_hit(const Lane& L, float rdx, float rdy, float rdz) {
  const uint32_t sh = fget(L.st, F_SHIFT, 3);
  const uint32_t span = 1u << sh;
  const uint32_t m = ~(span - 1u);
  const uint32_t bx = L.vx & m, by = L.vy & m, bz = L.vz & m;
  const float tx = plane_t(L.ix >= 0.0f ? bx : bx + span, L.ix, L.nx);
  const float ty = plane_t(L.iy >= 0.0f ? by : by + span, L.iy, L.ny);
  const float tz = plane_t(L.iz >= 0.0f ? bz : bz + span, L.iz, L.nz);
  const float cell_enter = fmaxf(fmaxf(tx, ty), tz);
  HitInfo h;
  h.exact_t = fmaxf(cell_enter, 0.0f);
  const float max_t = cell_enter;
  const float eps = 1e-4f * fmaxf(max_t * 0.01f, 1.0f);
  const float dtx = tx - max_t, dty = ty - max_t, dtz = tz - max_t;
  const float cx = fabsf(dtx) <= eps ? fabsf(rdx) : 0.0f;
  const float cy = fabsf(dty) <= eps ? fabsf(rdy) : 0.0f;
  const float cz = fabsf(dtz) <= eps ? fabsf(rdz) : 0.0f;
  uint32_t axis;
  if (cx >= cy && cx >= cz && cx > 0.0f) axis = 0;
  else if (cy >= cz && cy > 0.0f) axis = 1;
  else if (cz > 0.0f) axis = 2;
  else if (tx >= ty && tx >= tz) axis = 0;
  else if (ty >= tz) axis = 1;
  else axis = 2;
  const float dax = axis == 0 ? rdx : (axis == 1 ? rdy : rdz);
  h.axis = axis;
  h.nrm_axis = SEM == ARC_SEM_WORLD ? -wgsl_sign(dax) : -rust_signum(dax);
  const float qx = L.dx * h.exact_t, qy = L.dy * h.exact_t, qz = L.dz * h.exact_t;
  h.hx = L.ox + qx;
  h.hy = L.oy + qy;
  h.hz = L.oz + qz;
  return h;
}


// `miss_t`: t_max * 16 of the lane's ray (world semantics), kept in LDS from the refill so a miss
// record needs no second global read of the ray (a register would cost 15 VGPRs here).
template <uint32_t SEM>
__device__ __forceinline__ uint4 compact_record(const arc_ray* rays, const Lane& L, float miss_t) {
  uint4 v;
  const uint32_t steps = L.step > 0xFFFFu ? 0xFFFFu : L.step;
  if (fget(L.st, F_HIT, 1)) {
    float rdx = L.dx, rdy = L.dy, rdz = L.dz;
    if (SEM == ARC_SEM_CHUNK) {  // raw direction: the normal and tie-break use it
      const arc_ray& r = rays[L.idx];
      rdx = r.dx;
      rdy = r.dy;
      rdz = r.dz;
    }
    const HitInfo h = resolve_hit<SEM>(L, rdx, rdy, rdz);
    const uint32_t face = h.nrm_axis > 0.0f ? 2u * h.axis : (h.nrm_axis < 0.0f ? 2u * h.axis + 1u : 6u + h.axis);
    v.x = __float_as_uint(h.exact_t);
    v.y = (L.vx & 0xFFFFu) | (L.vy << 16);
    v.z = (L.vz & 0xFFFFu) | (fget(L.st, F_MAT, 8) << 16) | (face << 24);
    v.w = steps | ((ARC_HIT_FLAG_HIT | (fget(L.st, F_SHIFT, 3) > 0u ? ARC_HIT_FLAG_LOD : 0u)) << 16);
  } else {
    v.x = __float_as_uint(SEM == ARC_SEM_WORLD ? miss_t : 0.0f);
    v.y = 0u;
    v.z = (uint32_t)ARC_FACE_NONE << 24;
    v.w = steps;
  }
  return v;
}

template <uint32_t SEM>
__device__ __forceinline__ void store_compact(arc_hit* out, const arc_ray* rays, const Lane& L, float miss_t) {
  reinterpret_cast<uint4*>(out)[L.idx] = compact_record<SEM>(rays, L, miss_t);
}

}  // namespace arcdev
