// Shadow any-hit walkers for gfx950: trace_shadow_vis (shaders/common/traversal64.wgsl:571-773)
// and trace_shadow_vis_di (:781-989) on the persistent ray queue of the traversal kernel.
//
// The walkers share the world traversal's outer loop (guards, chunk change, ancestor restart,
// descent, exact box step: trace_walk.h) and differ from it in what a descent tests:
//   * min_shift follows the walker's own schedule (SUN: 2 beyond 48 voxels or 6 steps, 4 beyond
//     192 voxels or 16 steps; DI: the same levels at 48 / 160 voxels and 6 / 14 steps, only while
//     more than 32 / 64 voxels remain to the target);
//   * uniform water nodes are always stepped over (4 << shift), whatever ignore_water says;
//   * a set child bit at a leaf or at the LOD floor ends the walk (DI: also at any uniform node)
//     unless the cell's material is water; SUN takes the node's LOD material, DI reads the brick
//     at voxel level; DI counts an emissive cell within 16 voxels of max_dist as unoccluded;
//   * after the loop, SUN treats a walk cut by the step budget below 32 m height as occluded,
//     DI one that stopped short of max_dist by more than 2 voxels.
// The lane's `mat` carries the outcome code (ARC_SHADOW_*) once the walk returned in the loop.
#include <hip/hip_runtime.h>

#include "trace_launch.h"
#include "trace_walk.h"

namespace arcdev {

__device__ __forceinline__ bool is_emissive_mat(uint32_t m) {  // traversal64.wgsl:775-777
  return (m >= 19u && m <= 22u) || m == 24u || m == 27u;
}

// One outer iteration of a live shadow walker whose guards passed.
template <uint32_t MODE>
__device__ __forceinline__ void shadow_iteration(const DevWorld& w, const DevParams& p, Lane& L, uint4* C,
                                                 const RootLds& R, bool lds_roots) {
  uint32_t st = L.st;
  uint4 node = L.node;
  uint32_t sh = L.cs;
  const uint32_t sidx = L.step;
  L.step = sidx + 1u;
  uint32_t minsh;
  if (MODE == ARC_SHADOW_SUN) {  // :648-654
    minsh = (L.t > 48.0f || sidx > 6u) ? 2u : 0u;
    if (L.t > 192.0f || sidx > 16u) minsh = 4u;
  } else {  // :858-865
    const float rem = L.tl - L.t;
    minsh = (L.t > 48.0f && rem > 32.0f && sidx > 6u) ? 2u : 0u;
    if (L.t > 160.0f && rem > 64.0f && sidx > 14u) minsh = 4u;
  }
  const bool descend = restart<ARC_SEM_WORLD>(w, L, C, R, lds_roots, minsh, sh, node);
  uint32_t span_mul = descend ? 1u : 4u;  // the box stepped over is span_mul << sh
  bool cand = false;                      // the walk ends here unless the cell is water
  if (descend) {
    // :672-755 / :882-971, in the traversal's form: at most three loads, four unrolled tests with
    // a wave-uniform exit; a lane that has stopped repeats its own test until the wave exits.
    const uint32_t stop_flags = MODE == ARC_SHADOW_SUN ? ARC_IS_LEAF_FLAG : (ARC_IS_LEAF_FLAG | ARC_IS_UNIFORM_FLAG);
#pragma unroll
    for (uint32_t pass = 0; pass < 4u; ++pass) {
      const uint32_t b = child_bit(L, sh);
      const bool water = (node.z & ARC_IS_UNIFORM_FLAG) != 0u && (node.w >> 24) == ARC_MAT_WATER;  // fast-skip
      const uint64_t below = (((uint64_t)node.y << 32) | node.x) << (63u - b);
      const bool has = (int64_t)below < 0 && !water;  // child at b
      const bool stop = (node.z & stop_flags) != 0u || sh <= minsh;
      if (water) span_mul = 4u;
      cand = has && stop;
      const bool go = has && !stop;
      if (__builtin_amdgcn_ballot_w64(go) == 0ull) break;
      if (go) {
        node = w.nodes[(node.z & ARC_PTR_INDEX_MASK) + (uint32_t)__popcll(below) - 1u];
        sh -= 2u;
        *cold_slot(C, sh) = node;
      }
    }
  }
  bool hit = false;
  if (cand) {
    uint32_t mat = node.w >> 24;
    if (MODE == ARC_SHADOW_DI && sh == 0u && (node.z & (ARC_IS_LEAF_FLAG | ARC_IS_UNIFORM_FLAG)) == ARC_IS_LEAF_FLAG) {
      const uint32_t b = child_bit(L, 0u);
      const uint32_t word = w.bricks[(size_t)(node.z & ARC_PTR_INDEX_MASK) * 16u + (b >> 2)];
      mat = __builtin_amdgcn_ubfe(word, (b & 3u) * 8u, 8);
    }
    hit = mat != ARC_MAT_WATER;  // water cells are stepped over (span 1 << sh)
    if (hit) {
      const bool lit = MODE == ARC_SHADOW_DI && is_emissive_mat(mat) && L.t >= L.tl - 16.0f;
      L.mat = lit ? ARC_SHADOW_EMITTER : ARC_SHADOW_BLOCKED;
    }
  }
  // one way out, as in the traversal's iteration
  L.node = node;
  L.cs = sh;
  if (hit) {
    st = fput(fput(st, F_LIVE, 1, 0u), F_HIT, 1, 1u);
  } else {
    box_step(L, span_mul << sh);
    if (iteration_guards_fail<ARC_SEM_WORLD>(w, p, L)) st = fput(fput(st, F_LIVE, 1, 0u), F_HIT, 1, 0u);
  }
  L.st = st;
}

// Visibility and record of a finished walker.
template <uint32_t MODE>
__device__ __forceinline__ void shadow_finish(float* __restrict__ vis, arc_shadow_rec* __restrict__ rec,
                                              const DevParams& p, const Lane& L) {
  uint32_t outcome;
  float t = L.t;
  if (fget(L.st, F_HIT, 1)) {
    outcome = L.mat;
  } else if (L.mat != kEntered) {
    outcome = ARC_SHADOW_MISS_WORLD;
    t = 0.0f;
  } else {
    // The loop tests the step budget, then t >= t_limit, then the world bounds (:610-617); only
    // leaving the world sets t = t_limit.
    if (!(L.step >= p.max_steps) && !(t >= L.tl)) t = L.tl;
    if (MODE == ARC_SHADOW_SUN) {  // :762-772
      if (t < L.tl) {
        const float q = L.dy * t;
        const float py = L.oy + q;
        outcome = py * 0.0625f < 32.0f ? ARC_SHADOW_BUDGET_DARK : ARC_SHADOW_BUDGET_LIT;
      } else {
        outcome = ARC_SHADOW_CLEAR;
      }
    } else {  // :976-984
      outcome = t >= L.tl - 2.0f ? (t >= L.tl ? ARC_SHADOW_CLEAR : ARC_SHADOW_BUDGET_LIT) : ARC_SHADOW_BUDGET_DARK;
    }
  }
  const bool dark = outcome == ARC_SHADOW_BLOCKED || outcome == ARC_SHADOW_BUDGET_DARK;
  vis[L.idx] = dark ? 0.0f : 1.0f;
  if (rec) {
    const bool miss = outcome == ARC_SHADOW_MISS_WORLD;
    uint4 v;
    v.x = __float_as_uint(t);
    v.y = L.step;
    v.z = miss ? 0u : ((L.vx & 0xFFFFu) | (L.vy << 16));
    v.w = (miss ? 0u : (L.vz & 0xFFFFu)) | (outcome << 16);
    reinterpret_cast<uint4*>(rec)[L.idx] = v;
  }
}

// Persistent walker kernel: the ray queue and refill of k_trace (`queue`: the walkers' own queue
// region, zero between launches).
template <uint32_t MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_shadow(
    DevWorld w, DevParams p, const arc_ray* __restrict__ rays, float* __restrict__ vis, arc_shadow_rec* __restrict__ rec,
    uint32_t n, uint32_t* __restrict__ queue, uint32_t refill, uint32_t batch) {
  __shared__ uint4 s_cold[kColdSlots * kMaxBlock];
  __shared__ RootLds s_roots;
  uint4* const C = s_cold + threadIdx.x;
  const bool lds_roots = stage_roots(w, s_roots);
  const uint32_t lane = __lane_id();
  Lane L;
  L.st = 0u;
  L.step = 0u;
  L.idx = 0u;
  L.mat = 0u;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
  uint32_t wb = 0u, we = 0u;  // wave-uniform: window of rays being handed out
  uint32_t qc = home_counter(), qtried = 0u;
  bool drained = false;
  // The refill block runs when nidle >= thr (wave-uniform): thr = refill until the queue is
  // drained, then 64, when the wave leaves; an ordinary trip costs one ballot and one scalar
  // compare beyond the iteration.
  uint32_t thr = refill;
  for (;;) {
    // the iteration first, then the refill block (this order keeps the lane state in its loop
    // registers without per-trip copies); lanes started there iterate from the next trip
    if (fget(L.st, F_LIVE, 1)) shadow_iteration<MODE>(w, p, L, C, s_roots, lds_roots);
    const bool idle_lane = fget(L.st, F_LIVE, 1) == 0u;
    const uint64_t idle = __ballot(idle_lane);
    const uint32_t nidle = (uint32_t)__popcll(idle);
    if (nidle >= thr) {  // uniform: a ballot count against a uniform threshold
      if (rfl((uint32_t)drained)) break;
      if (idle_lane && fget(L.st, F_HASRAY, 1)) {
        shadow_finish<MODE>(vis, rec, p, L);
        L.st = fput(L.st, F_HASRAY, 1, 0u);
      }
      const uint32_t rank = (uint32_t)__popcll(idle & lt_mask);
      uint32_t assigned = 0u;
      uint32_t my = 0xFFFFFFFFu;
      while (assigned < nidle) {
        if (wb >= we && !claim_window(queue, qc, qtried, batch, n, wb, we)) {
          drained = true;
          break;
        }
        const uint32_t take = umin_(nidle - assigned, we - wb);
        if (idle_lane && rank >= assigned && rank < assigned + take) my = wb + (rank - assigned);
        wb += take;
        assigned += take;
      }
      if (my != 0xFFFFFFFFu) {
        L.idx = my;
        L.mat = ARC_SHADOW_MISS_WORLD;  // until the setup finds the world entry
        const arc_ray r = rays[my];
        ray_init<ARC_SEM_WORLD, true>(w, p, r, L, C);  // sets mat = kEntered past the entry test
      }
      thr = rfl(drained ? 64u : refill);
    }
    // One back-edge: the lanes refilled above begin at the top of the next trip. (A `continue`
    // back to the ballot made a second back-edge, and the register allocator then copied the
    // whole lane state on every wave trip.)
  }
  if (fget(L.st, F_HASRAY, 1)) shadow_finish<MODE>(vis, rec, p, L);
  queue_finish(queue, threadIdx.x == 0u);
}

hipError_t shadow_occupancy(uint32_t block_threads, uint32_t* blocks_per_cu) {
  int nb = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &nb, reinterpret_cast<const void*>(&k_shadow<ARC_SHADOW_SUN>), (int)block_threads, 0);
  *blocks_per_cu = nb > 0 ? (uint32_t)nb : 1u;
  return e;
}

hipError_t launch_shadow(const ShadowLaunchArgs& a, hipStream_t stream) {
  DevWorld w = make_world(a.grid, a.roots, a.nodes, a.bricks);
  DevParams p{};
  p.max_steps = a.mode == ARC_SHADOW_DI ? (a.max_steps < 32u ? 32u : (a.max_steps > 48u ? 48u : a.max_steps))
                                        : a.max_steps;  // :593 / :803
  if (a.mode == ARC_SHADOW_DI)
    hipLaunchKernelGGL(k_shadow<ARC_SHADOW_DI>, dim3(a.blocks), dim3(a.block_threads), 0, stream, w, p, a.rays, a.vis,
                       a.rec, (uint32_t)a.n, a.queue, a.refill, a.batch);
  else
    hipLaunchKernelGGL(k_shadow<ARC_SHADOW_SUN>, dim3(a.blocks), dim3(a.block_threads), 0, stream, w, p, a.rays, a.vis,
                       a.rec, (uint32_t)a.n, a.queue, a.refill, a.batch);
  return hipGetLastError();
}

}  // namespace arcdev
