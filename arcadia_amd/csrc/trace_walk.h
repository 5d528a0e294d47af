// Per-lane state and building blocks shared by the traversal walkers (k_trace, k_path,
// k_shadow): packed lane flags, the LDS cold slots of cached ancestors, chunk roots staged in LDS,
// the closed-form restart level and the exact box step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "trace_device.h"
#include "trace_launch.h"

namespace arcdev {

// packed per-lane flags
enum : uint32_t {
  F_SHIFT = 3,   // 3 bits: level of the hit cell (valid when done with a hit)
  F_HIT = 7,     // 1 bit: finished with a hit
  F_HASRAY = 8,  // 1 bit: lane owns a ray whose record is not written yet
  F_DEPTH = 9,   // 3 bits: path kernel: 0 = primary ray, d = d-th bounce
  F_MAT = 16,    // 8 bits: traversal kernels: material of the hit (valid when done with a hit)
  F_LIVE = 31,   // 1 bit: ray in flight (the sign bit: the per-trip test is one compare)
};

__device__ __forceinline__ uint32_t fget(uint32_t st, uint32_t off, uint32_t w) { return (st >> off) & ((1u << w) - 1u); }
__device__ __forceinline__ uint32_t fput(uint32_t st, uint32_t off, uint32_t w, uint32_t v) {
  const uint32_t m = ((1u << w) - 1u) << off;
  return (st & ~m) | ((v << off) & m);
}

struct Lane {
  float ox, oy, oz;  // origin, voxel units
  float dx, dy, dz;  // direction used for inside points (WORLD: dir, CHUNK: safe_dir)
  float ix, iy, iz;  // 1 / safe_dir
  float nx, ny, nz;  // -origin * inv
  uint32_t mx, my, mz;  // 1 where inv >= 0 (the box step exits through the upper plane), else 0
  float t, tl;       // t, t_limit
  uint32_t vx, vy, vz;
  uint32_t diff;     // prev_vi ^ vi of the coming iteration
  uint32_t step;     // outer iterations begun
  uint32_t st;       // packed flags
  uint32_t cs;       // cached_shift of the reference: the level the last descent ended on
  uint32_t mat;      // shadow walkers: entry marker / outcome (the traversal keeps the hit material in st)
  uint4 node;        // node of the level the last descent ended on
  uint32_t idx;
};

// Cold per-lane state in LDS, [slot][lane] so a wave's accesses are contiguous; the node cached
// at shift s sits in slot s / 2, so its address is one shift-add of s:
//   slot 3: chunk root node (shift 6),
//   slot 2: cached level-1 ancestor (shift 4), slot 1: cached level-2 ancestor (shift 2); the
//   level-3 ancestor is the lane's `node`,
//   slot 0: receives the level-3 node a descent writes like any other level (never read), so
//   the store needs no branch.
constexpr uint32_t kMaxBlock = 256u;
constexpr uint32_t kColdSlots = 4u;

__device__ __forceinline__ uint32_t rfl(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Persistent-queue claims (the queue layout is described with kQueueCounters). A wave starts on
// counter home_counter() and claims one window of `batch` items per atomic; a counter whose next
// window lies past the end stays exhausted, so the wave moves on to the next counter and is
// drained after kQueueCounters consecutive exhausted counters. `c` and `tried` are wave-uniform.
__device__ __forceinline__ uint32_t home_counter() {
  return (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (kQueueCounters - 1u);
}

__device__ __forceinline__ bool claim_window(uint32_t* queue, uint32_t& c, uint32_t& tried, uint32_t batch,
                                             uint32_t n, uint32_t& wb, uint32_t& we) {
  for (;;) {
    uint32_t got = 0u;
    if (__lane_id() == 0u) got = atomicAdd(queue + c * kQueueStride, 1u);
    got = __builtin_amdgcn_readlane(got, 0);  // lane 0's claim, into a scalar register
    const uint64_t start = ((uint64_t)got * kQueueCounters + c) * batch;
    if (start < n) {
      wb = (uint32_t)start;
      we = (uint32_t)(start + batch < n ? start + batch : n);
      return true;
    }
    if (++tried == kQueueCounters) return false;
    c = (c + 1u) & (kQueueCounters - 1u);
  }
}

// The last workgroup out resets the queue for the next launch on this stream (no memset node:
// graph-safe). `leader` is one thread of the workgroup.
__device__ __forceinline__ void queue_finish(uint32_t* queue, bool leader) {
  __syncthreads();
  if (leader) {
    __threadfence();
    const uint32_t done = atomicAdd(queue + kQueueDone, 1u);
    if (done == gridDim.x - 1u) {
      for (uint32_t c = 0; c < kQueueCounters; ++c) atomicExch(queue + c * kQueueStride, 0u);
      atomicExch(queue + kQueueDone, 0u);
      __threadfence();
    }
  }
}

// float -> u32 for a value already clamped to x >= 0 (never NaN), whose result is then clamped
// into a box below 2^32: min(x, 4294967040) (the largest float below 2^32) keeps the conversion
// defined and gives the same clamped result as the saturating WGSL u32() / Rust `as`. Plain
// arithmetic (v_min_f32 + v_cvt_u32_f32), so the compiler may speculate it into selects.
__device__ __forceinline__ uint32_t cvt_u32(float x) { return (uint32_t)fminf(x, 4294967040.0f); }

// v_cvt_u32_f32 as the hardware does it: negative values and NaN give 0, values at or above 2^32
// give 0xFFFFFFFF. Clamped afterwards into a box whose upper end stays far below 2^32, this is
// the same as cvt_u32(max(x, 0)) (one instruction instead of two; checked on MI355X for the
// special values and for 2^20 random bit patterns).
__device__ __forceinline__ uint32_t cvt_u32_sat(float x) {
  uint32_t r;
  asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// Correctly rounded 1/x. v_rcp_f32 (within 1 ulp) refined by one FMA Newton step is the
// correctly rounded reciprocal for every |x| in [2^-126, 2^126): on MI355X it matched the IEEE
// division for every binary32 of those binades. Outside them (results near or in the subnormal
// range, zero, infinities, NaN) the division runs.
__device__ __forceinline__ float rcp_exact(float x) {
  if (fabsf(x) < 0x1p126f && fabsf(x) >= 0x1p-126f) {
    const float r = __builtin_amdgcn_rcpf(x);
    const float e = fmaf(-x, r, 1.0f);
    return fmaf(e, r, r);
  }
  return 1.0f / x;
}

// min(max(x, lo), hi) for lo <= hi in one instruction (the compiler forms v_med3_u32 only for
// constant bounds). An ALU instruction, no memory access.
__device__ __forceinline__ uint32_t umed3(uint32_t x, uint32_t lo, uint32_t hi) {
  uint32_t r;
  asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(lo), "v"(hi));
  return r;
}

// max(a, b) for values that are never signaling NaNs (results of arithmetic): v_max_f32 as is.
// fmaxf also quiets its loop-carried operand first (a canonicalize, one more vector instruction
// per box step) because the compiler cannot prove it is not a signaling NaN.
__device__ __forceinline__ float fmax_nc(float a, float b) {
  float r;
  asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// t at which the ray crosses coordinate `c` on one axis: f32(c) * inv + (-o * inv).
__device__ __forceinline__ float plane_t(uint32_t c, float inv, float noi) {
  const float q = (float)c * inv;
  return q + noi;
}

// LOD min_shift (traversal64.wgsl:355-370); `sidx` is the iteration index.
// lod_bias is a launch parameter, so the class test is a scalar branch: class 1 reaches shift 4
// past lod_t_far or lod_step and shift 2 past lod_t only; classes 0 and 2 reach shift 2 past lod_t
// or lod_step and never 4. Two vector compares per step for classes 0 and 2 instead of four.
template <uint32_t SEM>
__device__ __forceinline__ uint32_t lod_min_shift(const DevParams& p, float t, uint32_t sidx) {
  if (SEM == ARC_SEM_CHUNK || (p.flags & kParamLodOff)) return 0u;
  if (p.lod_bias == 1u) return (t > p.lod_t_far || sidx > p.lod_step) ? 4u : (t > p.lod_t ? 2u : 0u);
  return (t > p.lod_t || sidx > p.lod_step) ? 2u : 0u;
}

// Guards at the top of an outer iteration (traversal64.wgsl:302-310 / raycast_cpu.rs:78-84).
template <uint32_t SEM>
__device__ __forceinline__ bool iteration_guards_fail(const DevWorld& w, const DevParams& p, const Lane& L) {
  const uint32_t max_steps = SEM == ARC_SEM_CHUNK ? 512u : p.max_steps;
  const uint32_t wx = SEM == ARC_SEM_CHUNK ? 256u : (w.gx << 8);
  const uint32_t wy = SEM == ARC_SEM_CHUNK ? 256u : (w.gy << 8);
  const uint32_t wz = SEM == ARC_SEM_CHUNK ? 256u : (w.gz << 8);
  return (L.step >= max_steps) | (L.t >= L.tl) | (L.vx >= wx) | (L.vy >= wy) | (L.vz >= wz);
}

// Restart level of an iteration without a chunk change (traversal64.wgsl:374-386). The
// reference tests c3 (diff < 4, min_shift == 0, cached_shift == 0), then c2 (diff < 16, both
// <= 2), then c1 (diff < 64, both <= 4), else restarts at the chunk root. With m =
// max(min_shift, cached_shift), both even, and dl the lowest level whose block holds both the old
// and the new voxel (0 for diff < 4, 2 below 16, 4 below 64, else 6), the first condition that
// holds is the one of level max(dl, m): that is the restart level. dl = bit length of diff,
// minus one, rounded down to even, capped at 6.
__device__ __forceinline__ uint32_t restart_shift(uint32_t diff, uint32_t minsh, uint32_t cs) {
  const uint32_t dl = umin_(6u, (31u - (uint32_t)__builtin_clz(diff | 1u)) & ~1u);
  return umax_(dl, umax_(minsh, cs));
}

// Address of the LDS slot of the cached node at shift `sh` (slot sh / 2) in a workgroup of BLOCK
// lanes.
template <uint32_t BLOCK = kMaxBlock>
__device__ __forceinline__ uint4* cold_slot(uint4* C, uint32_t sh) {
  return reinterpret_cast<uint4*>(reinterpret_cast<char*>(C) + sh * (BLOCK * 8u));
}

// Pool reads of the traversal loop. OFF32: the node pool spans at most 4 GiB (2^28 nodes) and the
// brick pool at most 4 GiB (2^26 bricks), so an element's byte offset fits 32 bits and the load
// takes the scalar-base + 32-bit vector-offset form: one shift of address arithmetic instead of
// a 64-bit multiply-add and the move that builds its high word. The launchers pick OFF32 from the
// pool capacities; without them (pools passed as bare pointers) the 64-bit form is used.
template <bool OFF32>
__device__ __forceinline__ uint4 pool_node(const DevWorld& w, uint32_t i) {
  if (!OFF32) return w.nodes[i];
  return *reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(w.nodes) + (uint32_t)(i << 4));
}

// Word `wd` (0..15) of brick `i`.
template <bool OFF32>
__device__ __forceinline__ uint32_t pool_brick_word(const DevWorld& w, uint32_t i, uint32_t wd) {
  if (!OFF32) return w.bricks[(size_t)i * 16u + wd];
  return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(w.bricks) + (uint32_t)((i << 6) | (wd << 2)));
}

// Chunk roots staged in LDS for worlds of at most N chunks: the root index and the root node of
// every chunk (20 B per chunk), read at each chunk change in place of chunk_roots[key] followed by
// nodes[root] (two dependent global loads on the longest dependency chain of a wave trip).
// kLdsChunks / kLdsChunksLarge (trace_launch.h) are the two capacities in use.
template <uint32_t N>
struct RootLdsT {
  uint4 node[N];
  uint32_t idx[N];
};
using RootLds = RootLdsT<kLdsChunks>;

// Fills `R` from the pools; returns false (block-uniform) when the world has more chunks.
template <uint32_t N>
__device__ __forceinline__ bool stage_roots(const DevWorld& w, RootLdsT<N>& R) {
  const uint32_t nc = w.gx * w.gy * w.gz;
  if (nc > N) return false;
  for (uint32_t c = threadIdx.x; c < nc; c += blockDim.x) {
    const uint32_t r = w.roots[c];
    R.idx[c] = r;
    R.node[c] = r == ARC_EMPTY_CHUNK_SENTINEL ? make_uint4(0u, 0u, 0u, 0u) : w.nodes[r];
  }
  __syncthreads();
  return true;
}

// Restart of an outer iteration (traversal64.wgsl:312-386; the shadow walkers repeat it at
// :620-669 / :830-879). On a chunk change (diff >= 256) the node is the chunk's root, read from
// chunk_roots on every change (the reference's 2-slot cache only saves those reads, the root is
// the same) or from the LDS copy; an empty chunk returns false (the caller steps over 256
// voxels). Otherwise the restart level is restart_shift() and the node the cached ancestor
// there. On return `sh` is the level the descent starts on and `node` the node at that level.
// The cached-ancestor read runs for every lane; a chunk change (restart_shift gives 6 there) then
// replaces the node it read (the previous chunk's root) with the new chunk's root. Two plain
// branches, the second skipped by whole waves, instead of an if / else.
template <uint32_t SEM, uint32_t BLOCK = kMaxBlock, uint32_t NCH = kLdsChunks, bool OFF32 = false>
__device__ __forceinline__ bool restart(const DevWorld& w, const Lane& L, uint4* C, const RootLdsT<NCH>& R,
                                        bool lds_roots, uint32_t minsh, uint32_t& sh, uint4& node) {
  const uint32_t rs = restart_shift(L.diff, minsh, sh);
  if (rs != sh) node = *cold_slot<BLOCK>(C, rs);
  sh = rs;
  bool ok = true;
  if (SEM == ARC_SEM_WORLD && L.diff >= 256u) {
    const uint32_t key = (L.vx >> 8) + __umul24(L.vy >> 8, w.gx) + __umul24(L.vz >> 8, w.gx * w.gy);
    uint32_t root;
    if (lds_roots) {
      // byte offsets by shifts (the compiler otherwise forms key * 16 as key * 4 + key * 12, a
      // quarter-rate multiply)
      const uint32_t k4 = key << 2;
      root = *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(R.idx) + k4);
      node = *reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(R.node) + (k4 << 2));
    } else {
      root = *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(w.roots) + (key << 2));
      if (root != ARC_EMPTY_CHUNK_SENTINEL) node = pool_node<OFF32>(w, root);
    }
    ok = root != ARC_EMPTY_CHUNK_SENTINEL;
    if (ok) *cold_slot<BLOCK>(C, 6u) = node;
  }
  return ok;
}

// step_dda_box (traversal64.wgsl:152-179 / raycast_cpu.rs:245-292) over the box of `span`
// voxels that holds the lane's voxel: t moves to the box exit, the voxel to the first one beyond
// it (exit axis) or to the clamped inside point (other axes), and `diff` records what changed.
// All three axes are computed and selected, with no branches.
__device__ __forceinline__ void box_step(Lane& L, uint32_t span) {
  const uint32_t last = span - 1u;
  // v & ~last as one three-input bit operation (truth table 0x30), so -span is not formed
  const uint32_t bx = __builtin_amdgcn_bitop3_b32(L.vx, last, 0u, 0x30), by = __builtin_amdgcn_bitop3_b32(L.vy, last, 0u, 0x30),
                 bz = __builtin_amdgcn_bitop3_b32(L.vz, last, 0u, 0x30);
  // exit plane: b + span going up (m = 1), b going down (m = 0); span <= 256 fits the 24-bit
  // multiply-add
  const uint32_t ex = __umul24(span, L.mx) + bx, ey = __umul24(span, L.my) + by, ez = __umul24(span, L.mz) + bz;
  const float tx = plane_t(ex, L.ix, L.nx), ty = plane_t(ey, L.iy, L.ny), tz = plane_t(ez, L.iz, L.nz);
  const float t_exit = fminf(fminf(tx, ty), tz);
  const float next_t = fmax_nc(L.t, t_exit);
  const bool is_x = tx <= ty && tx <= tz;
  const bool is_y = !is_x && (ty <= tz);
  const float qx = L.dx * next_t, qy = L.dy * next_t, qz = L.dz * next_t;
  // the inside point, clamped into the box (bx <= bx + last: boxes stay below 2^32 - 256)
  const uint32_t cx = umed3(cvt_u32_sat(L.ox + qx), bx, bx + last);
  const uint32_t cy = umed3(cvt_u32_sat(L.oy + qy), by, by + last);
  const uint32_t cz = umed3(cvt_u32_sat(L.oz + qz), bz, bz + last);
  // first voxel beyond the box on the exit axis: ex (= b + span) going up, b - 1 (= ex - 1) going
  // down, i.e. ex + m - 1
  const uint32_t fx = ex + L.mx - 1u, fy = ey + L.my - 1u, fz = ez + L.mz - 1u;
  const uint32_t nvx = is_x ? fx : cx;
  const uint32_t nvy = is_y ? fy : cy;
  const uint32_t nvz = (is_x || is_y) ? cz : fz;
  // `diff` is read only through its highest set bit (restart_shift, the chunk-change test), and
  // (nvx ^ bx) | (nvy ^ by) | (nvz ^ bz) has the same highest bit as the XOR of the old and new
  // voxels: on the other axes the new voxel stays inside the box (old and new share the bits at
  // and above log2(span) with b), and on the exit axis the old voxel and b differ only below
  // log2(span) while f ^ b has a bit at or above it. Written with b instead of the old voxel,
  // the old coordinates are dead once b is formed, so the new ones can take their registers.
  // (a ^ b) | c as one three-input bit operation (truth table 0xBE)
  L.diff = __builtin_amdgcn_bitop3_b32(nvz, bz, __builtin_amdgcn_bitop3_b32(nvy, by, nvx ^ bx, 0xBE), 0xBE);
  L.vx = nvx;
  L.vy = nvy;
  L.vz = nvz;
  L.t = next_t;
}

// Child slot b of the voxel at shift `sh` (traversal64.wgsl:391-394): (z << 4) | (y << 2) | x of
// the voxel's 2-bit digits at that level.
__device__ __forceinline__ uint32_t child_bit(const Lane& L, uint32_t sh) {
  return (__builtin_amdgcn_ubfe(L.vz, sh, 2) << 4) | (__builtin_amdgcn_ubfe(L.vy, sh, 2) << 2) |
         __builtin_amdgcn_ubfe(L.vx, sh, 2);
}

// In-chunk voxel coordinates, one byte each: (x & 255) | (y & 255) << 8 | (z & 255) << 16, by two
// byte permutes (selector 0x0C gives a zero byte).
__device__ __forceinline__ uint32_t voxel_bytes(const Lane& L) {
  const uint32_t xy = __builtin_amdgcn_perm(L.vy, L.vx, 0x0C0C0400u);
  return __builtin_amdgcn_perm(L.vz, xy, 0x0C040100u);
}

// 63 - child_bit(L, sh), from vb = voxel_bytes(L): the complemented digits (3 - x, 3 - y, 3 - z),
// one per byte, summed as (3 - x) + 4 (3 - y) + 16 (3 - z) by one 24-bit multiply with 0x1041
// (2^12 + 2^6 + 1): the sum lands on bits 12..17 and the other partial products on bits 0-1, 6-9,
// 20-23 and 28-29, with no carries between them. Four instructions where child_bit and the
// subtraction take six.
__device__ __forceinline__ uint32_t child_shift63(uint32_t vb, uint32_t sh) {
  const uint32_t d = ~(vb >> sh) & 0x030303u;
  return __builtin_amdgcn_ubfe(__umul24(d, 0x1041u), 12, 6);
}

// Marker a shadow walker's `mat` holds once its setup found the world entry (before that it
// holds ARC_SHADOW_MISS_WORLD; after an in-loop return, the outcome).
constexpr uint32_t kEntered = 0xFFu;

// Ray setup: traversal64.wgsl:234-300 (WORLD) / raycast_cpu.rs:34-76 (CHUNK), then the guards
// of iteration 0. SHADOW: the setup of trace_shadow_vis / _di (:576-605 / :786-815), which
// starts at the world entry (t_start unused) and misses when the entry lies at or beyond
// max_dist (world_hit.x >= max_t_v) rather than when max(entry, 0) does.
template <uint32_t SEM, bool SHADOW = false>
__device__ __forceinline__ void ray_init(const DevWorld& w, const DevParams& p, const arc_ray& r, Lane& L,
                                         uint4* C) {
  L.step = 0u;
  L.st = fput(0u, F_HASRAY, 1, 1u);  // not live until the setup succeeds
  if (SEM == ARC_SEM_WORLD) {
    const float max_t_v = r.t_max * 16.0f;
    const float start_t_v = SHADOW ? 0.0f : fmaxf(r.t_start * 16.0f, 0.0f);
    L.ox = r.ox * 16.0f;
    L.oy = r.oy * 16.0f;
    L.oz = r.oz * 16.0f;
    L.dx = r.dx;
    L.dy = r.dy;
    L.dz = r.dz;
    const float sx = fabsf(r.dx) < 1e-7f ? 1e-7f : r.dx;
    const float sy = fabsf(r.dy) < 1e-7f ? 1e-7f : r.dy;
    const float sz = fabsf(r.dz) < 1e-7f ? 1e-7f : r.dz;
    L.ix = rcp_exact(sx);
    L.iy = rcp_exact(sy);
    L.iz = rcp_exact(sz);
    const float wx = w.ext[0], wy = w.ext[1], wz = w.ext[2];
    // intersect_aabb (traversal64.wgsl:134-142)
    const float tnx = ((L.ix > 0.0f ? 0.0f : wx) - L.ox) * L.ix;
    const float tny = ((L.iy > 0.0f ? 0.0f : wy) - L.oy) * L.iy;
    const float tnz = ((L.iz > 0.0f ? 0.0f : wz) - L.oz) * L.iz;
    const float txx = ((L.ix > 0.0f ? wx : 0.0f) - L.ox) * L.ix;
    const float txy = ((L.iy > 0.0f ? wy : 0.0f) - L.oy) * L.iy;
    const float txz = ((L.iz > 0.0f ? wz : 0.0f) - L.oz) * L.iz;
    const float wh_x = fmaxf(fmaxf(tnx, tny), tnz);
    const float wh_y = fminf(fminf(txx, txy), txz);
    const float t_enter = fmaxf(wh_x, start_t_v);
    if (wh_y < t_enter || (SHADOW ? wh_x : t_enter) >= max_t_v) return;
    if (SHADOW) L.mat = kEntered;
    L.t = t_enter;
    L.tl = fminf(wh_y, max_t_v);
    const float tt = L.t + 1e-4f;
    const float qx = L.dx * tt, qy = L.dy * tt, qz = L.dz * tt;
    L.vx = cvt_u32(clampf_(L.ox + qx, 0.0f, w.ext_in[0]));
    L.vy = cvt_u32(clampf_(L.oy + qy, 0.0f, w.ext_in[1]));
    L.vz = cvt_u32(clampf_(L.oz + qz, 0.0f, w.ext_in[2]));
  } else {
    const uint32_t root = w.roots[0];
    if (root == ARC_EMPTY_CHUNK_SENTINEL) return;
    const uint4 rn = w.nodes[root];
    if (rn.x == 0u && rn.y == 0u) return;  // tree.is_empty()
    L.ox = r.ox;
    L.oy = r.oy;
    L.oz = r.oz;
    L.dx = fabsf(r.dx) > 1e-7f ? r.dx : copysignf(1e-7f, r.dx);
    L.dy = fabsf(r.dy) > 1e-7f ? r.dy : copysignf(1e-7f, r.dy);
    L.dz = fabsf(r.dz) > 1e-7f ? r.dz : copysignf(1e-7f, r.dz);
    L.ix = rcp_exact(L.dx);
    L.iy = rcp_exact(L.dy);
    L.iz = rcp_exact(L.dz);
    // intersect_aabb (raycast_cpu.rs:222-234)
    const float ax0 = (0.0f - L.ox) * L.ix, ax1 = (256.0f - L.ox) * L.ix;
    const float ay0 = (0.0f - L.oy) * L.iy, ay1 = (256.0f - L.oy) * L.iy;
    const float az0 = (0.0f - L.oz) * L.iz, az1 = (256.0f - L.oz) * L.iz;
    const float tmin = fmaxf(fmaxf(fminf(ax0, ax1), fminf(ay0, ay1)), fminf(az0, az1));
    const float tmax = fminf(fminf(fmaxf(ax0, ax1), fmaxf(ay0, ay1)), fmaxf(az0, az1));
    if (!(tmax >= fmaxf(tmin, 0.0f))) return;
    const float t_enter = fmaxf(tmin, 0.0f);
    if (tmax < t_enter || t_enter >= r.t_max) return;
    L.t = t_enter;
    L.tl = fminf(tmax, r.t_max);
    const float tt = L.t + 1e-4f;
    const float qx = L.dx * tt, qy = L.dy * tt, qz = L.dz * tt;
    L.vx = cvt_u32(clampf_(L.ox + qx, 0.0f, 256.0f - 1e-3f));
    L.vy = cvt_u32(clampf_(L.oy + qy, 0.0f, 256.0f - 1e-3f));
    L.vz = cvt_u32(clampf_(L.oz + qz, 0.0f, 256.0f - 1e-3f));
    *cold_slot<kMaxBlock>(C, 6u) = rn;  // chunk root for every root restart (k_trace's workgroups)
    L.node = rn;  // `node` == cached ancestor at cached_shift (6): iteration 0 restarts here
  }
  L.nx = (-L.ox) * L.ix;
  L.ny = (-L.oy) * L.iy;
  L.nz = (-L.oz) * L.iz;
  L.mx = L.ix >= 0.0f ? 1u : 0u;
  L.my = L.iy >= 0.0f ? 1u : 0u;
  L.mz = L.iz >= 0.0f ? 1u : 0u;
  L.diff = 0xFFFFFFFFu;  // prev_vi starts at u32::MAX
  if (iteration_guards_fail<SEM>(w, p, L)) return;
  L.cs = 6u;
  L.st = fput(L.st, F_LIVE, 1, 1u);
}

// Device view of the pools (host side of a launch).
inline DevWorld make_world(const uint32_t grid[3], const uint32_t* roots, const arc_node64* nodes,
                           const arc_brick64* bricks) {
  DevWorld w;
  w.gx = grid[0];
  w.gy = grid[1];
  w.gz = grid[2];
  w.pad = 0;
  for (int a = 0; a < 3; ++a) {
    w.ext[a] = (float)grid[a] * 256.0f;
    w.ext_in[a] = w.ext[a] - 1e-3f;
  }
  w.roots = roots;
  w.nodes = reinterpret_cast<const uint4*>(nodes);
  w.bricks = reinterpret_cast<const uint32_t*>(bricks);
  return w;
}

}  // namespace arcdev
