// 64-tree DDA traversal kernels for gfx950 and their launchers.
//
// Execution model
// ---------------
// A persistent ray queue feeds one ray per lane. Each trip of a wave through its loop runs one
// outer iteration of trace_world_64tree_from (traversal64.wgsl:302-553) for every live lane:
// choose the restart node, descend at most four levels, then either record a hit or take one
// exact DDA box step. A wave's iteration is a chain of dependent node loads and a long exact float
// sequence that the wave issues one instruction at a time (vector, scalar and branch alike), so
// the kernel is shaped to cut the instructions a wave issues:
//   * finished lanes are not resolved in place: the hit record (cell entry planes, tie-break
//     normal, hash, AO) is computed once per refill for all finished lanes together, not in
//     every wave trip in which some lane of 64 hits;
//   * cold per-lane state (cached ancestors, chunk root node) lives in LDS
//     indexed by level: a descent writes one slot, a restart reads one slot;
//   * the lane keeps `diff` (XOR of the last step's voxels) instead of prev_vi, and keeps the
//     node of its deepest level, which is the restart node whenever the restart level equals
//     cached_shift (every level-3 restart).
// Results are written by ray index, so neither the schedule nor these choices can change an
// output: the cached-ancestor restart is equivalent to a root descent (the cached node at shift
// s covers the current voxel's bits >= s+2 whenever the reference uses it).
//
// Queue: a wave claims windows of `batch` consecutive rays with one atomic and refills its idle
// lanes from the window. The claims go to kQueueCounters counters that deal out interleaved
// windows (claim_window), so all waves of all XCDs still sweep the ray array together and the
// rays in flight form one coherent band of screen tiles. Measured on MI355X: 8 contiguous per-XCD
// segments ran 65% slower than one shared sweep. The last workgroup to finish resets the
// counters, which keeps a launch self-contained (no memset node: graph-safe).
#include <hip/hip_runtime.h>

#include <mutex>

#include "raymath.h"
#include "trace_device.h"
#include "trace_launch.h"
#include "trace_iter.h"
#include "trace_walk.h"

namespace arcdev {


template <uint32_t SEM>
__device__ __forceinline__ void store_full(arc_trace_result* out, const arc_ray* rays, const DevParams& p,
                                           const Lane& L, const uint4* C) {
  const arc_ray r = rays[L.idx];
  float t_vox, hp0, hp1, hp2, n0 = 0.0f, n1 = 0.0f, n2 = 0.0f, rgb0 = 0.0f, rgb1 = 0.0f, rgb2 = 0.0f, ao = 1.0f;
  uint32_t vh = 0u, hitw = 0u, mat = 0u;
  const bool hit = fget(L.st, F_HIT, 1) != 0u;
  if (hit) {
    const HitInfo h = resolve_hit<SEM>(L, r.dx, r.dy, r.dz);
    const uint32_t sh = fget(L.st, F_SHIFT, 3);
    const uint4 node = L.node;  // the node the hit resolved in
    t_vox = h.exact_t;
    n0 = h.axis == 0 ? h.nrm_axis : 0.0f;
    n1 = h.axis == 1 ? h.nrm_axis : 0.0f;
    n2 = h.axis == 2 ? h.nrm_axis : 0.0f;
    mat = fget(L.st, F_MAT, 8);
    hitw = 1u | (h.axis << 8) | (sh << 16);
    if (SEM == ARC_SEM_WORLD) {
      hp0 = h.hx * 0.0625f;
      hp1 = h.hy * 0.0625f;
      hp2 = h.hz * 0.0625f;
      const uint32_t lp = node.w;
      rgb0 = (float)(lp & 0xFFu) / 255.0f;
      rgb1 = (float)((lp >> 8) & 0xFFu) / 255.0f;
      rgb2 = (float)((lp >> 16) & 0xFFu) / 255.0f;
      vh = pcg_hash((L.vx * 73856093u) ^ (L.vy * 19349663u) ^ (L.vz * 83492791u));
      if (sh == 0u && p.lod_bias == 0u) {  // contact AO, traversal64.wgsl:501-521
        const float hit_dist_m = L.t * 0.0625f;
        const float raw = p.ao_max_dist;
        const float ao_max = raw < 1.0f ? 24.0f : raw;
        if (hit_dist_m < ao_max) {
          const float raw_ao =
              eval_leaf_ao(node.x, node.y, L.vx & 3u, L.vy & 3u, L.vz & 3u, n0, n1, n2, h.hx, h.hy, h.hz);
          const float ao_min = ao_max * 0.5f;
          const float den = ao_max - ao_min;
          const float num = hit_dist_m - ao_min;
          const float tt = clampf_(num / den, 0.0f, 1.0f);
          const float t2 = tt * tt;
          const float tw = 2.0f * tt;
          const float sm = t2 * (3.0f - tw);
          const float wgt = 1.0f - sm;
          const float omw = 1.0f - wgt;
          const float a0 = 1.0f * omw;
          const float a1 = raw_ao * wgt;
          ao = a0 + a1;
        }
      }
    } else {
      hp0 = h.hx;
      hp1 = h.hy;
      hp2 = h.hz;
    }
  } else if (SEM == ARC_SEM_WORLD) {
    t_vox = r.t_max * 16.0f;
    const float q0 = r.dx * r.t_max, q1 = r.dy * r.t_max, q2 = r.dz * r.t_max;
    hp0 = r.ox + q0;
    hp1 = r.oy + q1;
    hp2 = r.oz + q2;
    n1 = 1.0f;
  } else {
    t_vox = 0.0f;
    hp0 = hp1 = hp2 = 0.0f;
  }
  uint4* q = reinterpret_cast<uint4*>(out + L.idx);
  q[0] = make_uint4(__float_as_uint(t_vox), __float_as_uint(hp0), __float_as_uint(hp1), __float_as_uint(hp2));
  q[1] = make_uint4(__float_as_uint(n0), __float_as_uint(n1), __float_as_uint(n2), __float_as_uint(rgb0));
  q[2] = make_uint4(__float_as_uint(rgb1), __float_as_uint(rgb2), __float_as_uint(ao), mat);
  q[3] = make_uint4(vh, L.step, hitw, hit ? L.vx : 0u);
  q[4] = make_uint4(hit ? L.vy : 0u, hit ? L.vz : 0u, 0u, 0u);
}

template <bool FULL>
struct OutPtr;
template <>
struct OutPtr<false> {
  using T = arc_hit;
};
template <>
struct OutPtr<true> {
  using T = arc_trace_result;
};

template <uint32_t SEM, bool FULL>
__device__ __forceinline__ void finalize(typename OutPtr<FULL>::T* out, const arc_ray* rays, const DevParams& p,
                                         const Lane& L, const uint4* C, float miss_t) {
  if (FULL) store_full<SEM>(reinterpret_cast<arc_trace_result*>(out), rays, p, L, C);
  else store_compact<SEM>(reinterpret_cast<arc_hit*>(out), rays, L, miss_t);
}

// Persistent ray-queue kernel.
// `queue`: one queue region (claim counters and finished workgroups, zero between launches).
// Occupancy target: the compact variants fit 64 VGPRs (8 waves/SIMD) without spilling; the
// full-record variants carry more live values through finalize and keep the default budget.
template <uint32_t SEM, bool FULL, bool WATER, bool OFF32>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FULL ? 1 : 6, 8))) void k_trace(
    DevWorld w, DevParams p, const arc_ray* __restrict__ rays,
                                               typename OutPtr<FULL>::T* __restrict__ out, uint32_t n,
                                               uint32_t* __restrict__ queue, uint32_t refill, uint32_t batch) {
  __shared__ uint4 s_cold[kColdSlots * kMaxBlock];
  __shared__ RootLds s_roots;
  uint4* const C = s_cold + threadIdx.x;
  const bool lds_roots = SEM == ARC_SEM_WORLD && stage_roots(w, s_roots);
  const uint32_t lane = __lane_id();
  Lane L;
  L.st = 0u;
  L.step = 0u;
  L.idx = 0u;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
  uint32_t wb = 0u, we = 0u;  // wave-uniform: window of rays being handed out
  uint32_t qc = home_counter(), qtried = 0u;
  __shared__ float s_miss[kMaxBlock];  // t_max * 16 of each lane's ray (compact world records)
  bool drained = false;
  // A trip runs the iteration of the live lanes, then the refill block when nidle >= thr
  // (wave-uniform): thr = refill until the queue is drained, then 64, when the wave leaves (its
  // finished lanes are written after the loop). An ordinary trip costs one ballot and one scalar
  // compare beyond the iteration; the exit is taken under a readfirstlane'd condition, so the
  // compiler sees it uniform (an exit it took for divergent kept per-lane exit masks and copied
  // every loop-carried lane register twice per trip). The iteration comes first in the loop body:
  // with the refill block first, the lane state merged from the refill and the iteration was
  // copied back into its loop registers on every trip (8 vector moves); this order needs 3.
  uint32_t thr = refill;
  for (;;) {
    if (fget(L.st, F_LIVE, 1)) iteration<SEM, WATER, kMaxBlock, kLdsChunks, OFF32>(w, p, L, C, s_roots, lds_roots);
    const bool idle_lane = fget(L.st, F_LIVE, 1) == 0u;
    const uint64_t idle = __ballot(idle_lane);
    const uint32_t nidle = (uint32_t)__popcll(idle);
    if (nidle >= thr) {  // uniform: a ballot count against a uniform threshold
      if (rfl((uint32_t)drained)) break;
      // finalize finished lanes together, then refill every idle lane
      if (idle_lane && fget(L.st, F_HASRAY, 1)) {
        finalize<SEM, FULL>(out, rays, p, L, C, s_miss[threadIdx.x]);
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
        const arc_ray r = rays[my];
        if (!FULL && SEM == ARC_SEM_WORLD) s_miss[threadIdx.x] = r.t_max * 16.0f;
        ray_init<SEM>(w, p, r, L, C);
      }
      thr = rfl(drained ? 64u : refill);
    }
    // One back-edge: the lanes refilled here begin their iteration at the top of the next trip.
    // (A `continue` back to the ballot made a second back-edge, and the register allocator then
    // copied the whole lane state on every wave trip.)
  }
  if (fget(L.st, F_HASRAY, 1)) finalize<SEM, FULL>(out, rays, p, L, C, s_miss[threadIdx.x]);
  queue_finish(queue, C == s_cold);  // thread 0 (C = s_cold + threadIdx.x)
}

// ---------------------------------------------------------------------------------------
// Path kernel: batched views with in-lane bounces
// ---------------------------------------------------------------------------------------
// One persistent launch traces every pixel slot of a batch of views and, for each primary ray
// that hits, up to `n_bounces` diffuse bounces generated in the lane from the finished hit
// (arcmath::bounce_ray, bit-identical to arc_gen_bounce_rays_dense). A queue item is one pixel
// slot: views are laid out one after another in padded 8x8-tile order (view_first[v] = first slot
// of view v). Records: out[d * n_items + slot] = the compact hit of depth d (0 = primary); slots
// without a ray at depth d (padding pixels, bounces of misses) get the record of the empty ray
// (t_voxels -16, no hit, 0 steps), the same the traversal writes for a ray with t_max = -1.

__device__ __forceinline__ DevParams sel_params(bool b, const DevParams& pb, const DevParams& p0) {
  DevParams p;
  p.lod_bias = b ? pb.lod_bias : p0.lod_bias;
  p.ignore_water = b ? pb.ignore_water : p0.ignore_water;
  p.max_steps = b ? pb.max_steps : p0.max_steps;
  p.lod_step = b ? pb.lod_step : p0.lod_step;
  p.lod_t = b ? pb.lod_t : p0.lod_t;
  p.lod_t_far = b ? pb.lod_t_far : p0.lod_t_far;
  p.ao_max_dist = b ? pb.ao_max_dist : p0.ao_max_dist;
  p.flags = p0.flags & pb.flags;  // LOD tests skipped only when neither class has them (uniform)
  return p;
}

__device__ __forceinline__ uint4 empty_record() {
  return make_uint4(__float_as_uint(-16.0f), 0u, (uint32_t)ARC_FACE_NONE << 24, 0u);
}

struct PathCtl {
  const ViewDev* views;
  const uint32_t* view_first;  // n_views + 1 entries, view_first[n_views] = n_items
  uint32_t n_views, n_items, n_bounces;
  float bounce_tmax;
  uint64_t seed;
};

// Starts slot `item`: its primary ray, or the empty records of a padding slot.
__device__ __forceinline__ void path_start(const DevWorld& w, const DevParams& p0, const PathCtl& pc, uint32_t item,
                                           uint4* out, Lane& L, float& tmax, uint4* C) {
  uint32_t lo = 0u, hi = pc.n_views;  // last view with view_first <= item
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pc.view_first[mid] <= item) lo = mid;
    else hi = mid;
  }
  const ViewDev v = pc.views[lo];
  const uint32_t local = item - pc.view_first[lo];
  const uint32_t tile = local >> 6, j = local & 63u;
  const uint32_t px = (tile % v.tiles_x) * 8u + (j & 7u), py = (tile / v.tiles_x) * 8u + (j >> 3);
  L.idx = item;
  if (px >= v.width || py >= v.height) {
    for (uint32_t d = 0; d <= pc.n_bounces; ++d) out[(size_t)d * pc.n_items + item] = empty_record();
    L.st = 0u;
    return;
  }
  float dir[3];
  arcmath::primary_dir(v.cam, px, py, v.width, v.height, dir);
  arc_ray r;
  r.ox = v.cam.eye[0];
  r.oy = v.cam.eye[1];
  r.oz = v.cam.eye[2];
  r.t_start = 0.0f;
  r.dx = dir[0];
  r.dy = dir[1];
  r.dz = dir[2];
  r.t_max = v.max_dist_m;
  tmax = v.max_dist_m;
  ray_init<ARC_SEM_WORLD>(w, p0, r, L, C);
}

// Writes the finished ray's record, then starts its bounce or closes the slot.
// `frames`: arcmath::bounce_frame of the nine face codes (the workgroup's LDS table).
__device__ __forceinline__ void path_finish(const DevWorld& w, const DevParams& pb, const PathCtl& pc, uint4* out,
                                            Lane& L, float& tmax, uint4* C, const arcmath::BounceFrame* frames) {
  const uint32_t depth = fget(L.st, F_DEPTH, 3);
  const uint32_t steps = L.step > 0xFFFFu ? 0xFFFFu : L.step;
  const bool hit = fget(L.st, F_HIT, 1) != 0u;
  uint4 v;
  uint32_t face = ARC_FACE_NONE;
  float exact_t = 0.0f;
  if (hit) {
    const HitInfo h = resolve_hit<ARC_SEM_WORLD>(L, L.dx, L.dy, L.dz);
    face = h.nrm_axis > 0.0f ? 2u * h.axis : (h.nrm_axis < 0.0f ? 2u * h.axis + 1u : 6u + h.axis);
    exact_t = h.exact_t;
    v.x = __float_as_uint(h.exact_t);
    v.y = (L.vx & 0xFFFFu) | (L.vy << 16);
    v.z = (L.vz & 0xFFFFu) | (fget(L.st, F_MAT, 8) << 16) | (face << 24);
    v.w = steps | ((ARC_HIT_FLAG_HIT | (fget(L.st, F_SHIFT, 3) > 0u ? ARC_HIT_FLAG_LOD : 0u)) << 16);
  } else {
    v.x = __float_as_uint(tmax * 16.0f);
    v.y = 0u;
    v.z = (uint32_t)ARC_FACE_NONE << 24;
    v.w = steps;
  }
  const uint32_t item = L.idx;
  out[(size_t)depth * pc.n_items + item] = v;
  if (hit && depth < pc.n_bounces) {
    const float ov[3] = {L.ox, L.oy, L.oz}, dr[3] = {L.dx, L.dy, L.dz};
    arc_ray r;
    arcmath::bounce_ray_framed(ov, dr, exact_t, face, frames[face], pc.seed, (uint64_t)depth * pc.n_items + item,
                               pc.bounce_tmax, &r);
    ray_init<ARC_SEM_WORLD>(w, pb, r, L, C);
    L.st = fput(L.st, F_DEPTH, 3, depth + 1u);
    tmax = pc.bounce_tmax;
  } else {
    for (uint32_t d = depth + 1u; d <= pc.n_bounces; ++d) out[(size_t)d * pc.n_items + item] = empty_record();
    L.st = fput(L.st, F_HASRAY, 1, 0u);
  }
}

// `queue`: one queue region (claim counters and finished workgroups, zero between launches).
// Two shapes: workgroups of 256 lanes with the roots of worlds up to 256 chunks in LDS (21 KB per
// workgroup), or one workgroup of 1024 lanes per CU holding the roots of worlds up to 4096 chunks
// (80 KB) beside its 64 KB of cold slots; both keep the 4 waves per SIMD the kernel's registers
// allow.
// WATER = false compiles the ignore_water tests out (neither ray class sets it). UNI: the classes
// trace alike apart from ignore_water (params_uniform), so the iteration reads the primary class's
// scalar parameters and only the water flag is per lane.
template <uint32_t BLOCK, uint32_t NCH, bool WATER, bool OFF32, bool UNI>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_path(
    DevWorld w, DevParams p0, DevParams pb, PathCtl pc, uint4* __restrict__ out, uint32_t* __restrict__ queue,
    uint32_t refill, uint32_t batch) {
  __shared__ uint4 s_cold[kColdSlots * BLOCK];
  __shared__ RootLdsT<NCH> s_roots;
  __shared__ arcmath::BounceFrame s_frame[9];  // the bounce lobe's frame per face code
  uint4* const C = s_cold + threadIdx.x;
  if (threadIdx.x < 9u) s_frame[threadIdx.x] = arcmath::bounce_frame(threadIdx.x);
  __syncthreads();
  const bool lds_roots = stage_roots(w, s_roots);
  const uint32_t lane = __lane_id();
  Lane L;
  L.st = 0u;
  L.step = 0u;
  L.idx = 0u;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
  float tmax = 0.0f;
  uint32_t wb = 0u, we = 0u;
  uint32_t qc = home_counter(), qtried = 0u;
  bool drained = false;
  // The finish / refill block runs when nidle >= thr (wave-uniform): thr = refill while slots are
  // left to claim; once they are drained, one more than the idle lanes that hold nothing, so the
  // block runs as soon as a ray finishes (its bounce starts at once). An ordinary trip costs one
  // ballot and one scalar compare beyond the iteration.
  uint32_t thr = refill;
  for (;;) {
    // the iteration first, then the finish / refill block (this order keeps the lane state in its
    // loop registers without per-trip copies); lanes started there iterate from the next trip
    if (fget(L.st, F_LIVE, 1)) {
      const bool b = fget(L.st, F_DEPTH, 3) != 0u;
      if (UNI) {
        DevParams pl = p0;
        if (WATER) pl.ignore_water = b ? pb.ignore_water : p0.ignore_water;
        iteration<ARC_SEM_WORLD, WATER, BLOCK, NCH, OFF32>(w, pl, L, C, s_roots, lds_roots);
      } else {
        iteration<ARC_SEM_WORLD, WATER, BLOCK, NCH, OFF32>(w, sel_params(b, pb, p0), L, C, s_roots, lds_roots);
      }
    }
    const bool idle_lane = fget(L.st, F_LIVE, 1) == 0u;
    const uint32_t nidle = (uint32_t)__popcll(__ballot(idle_lane));
    if (nidle >= thr) {  // uniform: a ballot count against a uniform threshold
      const uint64_t pending = __ballot(idle_lane && fget(L.st, F_HASRAY, 1) != 0u);
      // drained, nothing pending, every lane idle: done (uniform exit)
      if (rfl((uint32_t)(drained && nidle == 64u && pending == 0ull))) break;
      if (idle_lane && fget(L.st, F_HASRAY, 1)) path_finish(w, pb, pc, out, L, tmax, C, s_frame);
      if (!drained && nidle >= refill) {
        const bool needs = fget(L.st, F_LIVE, 1) == 0u && fget(L.st, F_HASRAY, 1) == 0u;
        const uint64_t nm = __ballot(needs);
        const uint32_t nneed = (uint32_t)__popcll(nm);
        const uint32_t rank = (uint32_t)__popcll(nm & lt_mask);
        uint32_t assigned = 0u, my = 0xFFFFFFFFu;
        while (assigned < nneed) {
          if (wb >= we && !claim_window(queue, qc, qtried, batch, pc.n_items, wb, we)) {
            drained = true;
            break;
          }
          const uint32_t take = umin_(nneed - assigned, we - wb);
          if (needs && rank >= assigned && rank < assigned + take) my = wb + (rank - assigned);
          wb += take;
          assigned += take;
        }
        if (my != 0xFFFFFFFFu) path_start(w, p0, pc, my, out, L, tmax, C);
      }
      const bool live = fget(L.st, F_LIVE, 1) != 0u;
      const uint32_t nfree = (uint32_t)__popcll(__ballot(!live && fget(L.st, F_HASRAY, 1) == 0u));
      thr = rfl(drained ? umin_(nfree + 1u, 64u) : refill);  // all 64 free: the next trip exits
    }
  }
  queue_finish(queue, C == s_cold);  // thread 0 (C = s_cold + threadIdx.x)
}

// ---------------------------------------------------------------------------------------
// Primary rays and their diffuse bounces in one launch
// ---------------------------------------------------------------------------------------
// A ray set and one cosine bounce off each of its hits, traced by one persistent launch. Two
// launches (primary, then bounce) each end in a drain: once a queue runs dry, most waves idle
// while the last long rays finish, ~0.15 ms of a 0.38 ms launch on the 1080p full-frame world.
// Here the waves claim primary rays until that queue runs dry, then bounce indices, so the
// primaries' drain is filled with bounce work and only the last bounces drain.
//
// A bounce needs its primary's hit. The lane that finishes primary i publishes one 64-bit
// message msg[i] with a single agent-scope atomic store: 0 = not finished (the array is zero
// between launches), kMsgMiss = no hit (the bounce slot gets the empty-ray record), else
// (1 + face) << 32 | bits of t_voxels, which with ray i itself (read-only input) is all
// arcmath::bounce_ray needs. The message is the whole payload, so no ordering between stores and
// no fence is involved. Bounce index windows are claimed in index order after the primary
// windows, so the primaries they need finished long before, except near the end; a lane whose
// message is not there yet keeps the index (F_PEND) and polls it at later refills. The consumer
// clears msg[i], so every launch leaves the array zero.
//
// Waits are bounded: a lane waits only for a primary that has been claimed, i.e. that a
// resident wave is tracing or has finished and publishes at its next refill (once the primary
// windows are drained, a finished primary publishes at once).
//
// UNI: both ray classes trace with the same parameters apart from ignore_water (LOD off in both,
// same step cap), so the iteration reads the primary class's scalar parameters and only the water
// flag is per lane; otherwise every field is selected per lane as in the path kernel.
constexpr uint32_t F_PEND = 12;  // 1 bit: lane holds a bounce index whose primary has not published
constexpr uint64_t kMsgMiss = 0xF00000000ull;

__device__ __forceinline__ void msg_store(uint64_t* m, uint64_t v) {
  __hip_atomic_store(m, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t msg_load(uint64_t* m) {
  return __hip_atomic_load(m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounce index i's message `m` (non-zero) taken: the slot's empty record, or the bounce ray set
// up in the lane. `frames`: arcmath::bounce_frame of the nine face codes.
__device__ __forceinline__ void bounce_take(const DevWorld& w, const DevParams& pb, const arc_ray& pr, uint64_t m,
                                            uint64_t seed, float bounce_tmax, uint4* out_b, Lane& L, uint4* C,
                                            const arcmath::BounceFrame* frames) {
  const uint32_t i = L.idx;
  if (m == kMsgMiss) {
    out_b[i] = empty_record();
    L.st = 0u;
    return;
  }
  const float ov[3] = {pr.ox * 16.0f, pr.oy * 16.0f, pr.oz * 16.0f}, dr[3] = {pr.dx, pr.dy, pr.dz};
  const uint32_t face = (uint32_t)(m >> 32) - 1u;
  arc_ray r;
  arcmath::bounce_ray_framed(ov, dr, __uint_as_float((uint32_t)m), face, frames[face], seed, i, bounce_tmax, &r);
  ray_init<ARC_SEM_WORLD>(w, pb, r, L, C);
  L.st = fput(L.st, F_DEPTH, 3, 1u);
}

// The last workgroup out resets both queues (primary windows, bounce windows).
__device__ __forceinline__ void queue_finish2(uint32_t* queue, uint32_t* queue_b, bool leader) {
  __syncthreads();
  if (leader) {
    __threadfence();
    const uint32_t done = atomicAdd(queue + kQueueDone, 1u);
    if (done == gridDim.x - 1u) {
      for (uint32_t c = 0; c < kQueueCounters; ++c) {
        atomicExch(queue + c * kQueueStride, 0u);
        atomicExch(queue_b + c * kQueueStride, 0u);
      }
      atomicExch(queue + kQueueDone, 0u);
      __threadfence();
    }
  }
}

// out: 2n compact records, [0, n) the primaries, [n, 2n) the bounce of primary i at n + i.
// queue: two queue regions (primary windows, then bounce windows); msg: n words, zero.
template <bool WATER, bool OFF32, bool UNI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_trace_bounce(
    DevWorld w, DevParams p0, DevParams pb, const arc_ray* __restrict__ rays, uint4* __restrict__ out, uint32_t n,
    uint64_t* __restrict__ msg, uint32_t* __restrict__ queue, uint64_t seed, float bounce_tmax, uint32_t refill,
    uint32_t batch) {
  __shared__ uint4 s_cold[kColdSlots * kMaxBlock];
  __shared__ RootLds s_roots;
  __shared__ float s_miss[kMaxBlock];  // t_max * 16 of each lane's primary ray
  __shared__ arcmath::BounceFrame s_frame[9];  // the bounce lobe's frame per face code
  uint4* const C = s_cold + threadIdx.x;
  uint4* const out_b = out + n;
  uint32_t* const queue_b = queue + kQueueRegionWords;
  if (threadIdx.x < 9u) s_frame[threadIdx.x] = arcmath::bounce_frame(threadIdx.x);
  __syncthreads();
  const bool lds_roots = stage_roots(w, s_roots);
  const uint32_t lane = __lane_id();
  Lane L;
  L.st = 0u;
  L.step = 0u;
  L.idx = 0u;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
  const float miss_b = bounce_tmax * 16.0f;
  uint32_t wb = 0u, we = 0u;  // wave-uniform: window being handed out (primary until drained_p)
  uint32_t qc = home_counter(), qtried = 0u, qc_b = qc, qtried_b = 0u;
  bool drained_p = false, drained_b = false;
  constexpr uint32_t kHold = (1u << F_HASRAY) | (1u << F_PEND);
  // Primary phase: every lane holds a primary ray or nothing, so the loop is k_trace's (finalize
  // also publishes the message). It ends once the primary windows run dry, after the lanes of that
  // refill are set up; the lanes' state carries over into the loop below, which also claims
  // bounce indices. With one loop for both phases the primary half of the launch ran the bounce
  // bookkeeping too (+23% scalar instructions).
  {
    for (;;) {
      if (fget(L.st, F_LIVE, 1)) iteration<ARC_SEM_WORLD, WATER, kMaxBlock, kLdsChunks, OFF32>(
          w, p0, L, C, s_roots, lds_roots);
      const bool idle_lane = fget(L.st, F_LIVE, 1) == 0u;
      const uint64_t idle = __ballot(idle_lane);
      const uint32_t nidle = (uint32_t)__popcll(idle);
      if (nidle >= refill) {  // uniform: a ballot count against a uniform threshold
        if (idle_lane && fget(L.st, F_HASRAY, 1)) {
          const uint4 v = compact_record<ARC_SEM_WORLD>(rays, L, s_miss[threadIdx.x]);
          out[L.idx] = v;
          msg_store(msg + L.idx, (v.w & ((uint32_t)ARC_HIT_FLAG_HIT << 16)) != 0u
                                     ? ((uint64_t)(1u + (v.z >> 24)) << 32) | v.x
                                     : kMsgMiss);
          L.st = 0u;
        }
        const uint32_t rank = (uint32_t)__popcll(idle & lt_mask);
        uint32_t assigned = 0u, my = 0xFFFFFFFFu;
        while (assigned < nidle) {
          if (wb >= we && !claim_window(queue, qc, qtried, batch, n, wb, we)) {
            drained_p = true;
            break;
          }
          const uint32_t take = umin_(nidle - assigned, we - wb);
          if (idle_lane && rank >= assigned && rank < assigned + take) my = wb + (rank - assigned);
          wb += take;
          assigned += take;
        }
        if (my != 0xFFFFFFFFu) {
          L.idx = my;
          const arc_ray r = rays[my];
          s_miss[threadIdx.x] = r.t_max * 16.0f;
          ray_init<ARC_SEM_WORLD>(w, p0, r, L, C);
        }
        if (rfl((uint32_t)drained_p)) break;
      }
    }
  }
  // The refill block runs when nidle >= thr (wave-uniform): thr = refill while claims go on, 64
  // once nothing is left to claim, and 0 (every trip) while a lane waits for a message or, with
  // nothing left to claim, while the wave still traces a primary (published at once when it
  // finishes: a bounce lane elsewhere may wait for it). In that mode the block's work (publish,
  // claims, polling) runs when enough lanes are idle, when all are, or on every 8th entry. An
  // ordinary trip costs one ballot and one scalar compare beyond the iteration.
  uint32_t thr = refill, tick = 0u;
  for (;;) {
    // the iteration first, then the refill block (this order keeps the lane state in its loop
    // registers without per-trip copies); lanes started there iterate from the next trip
    if (fget(L.st, F_LIVE, 1)) {
      const bool b = fget(L.st, F_DEPTH, 3) != 0u;
      if (UNI) {
        DevParams pl = p0;
        if (WATER) pl.ignore_water = b ? pb.ignore_water : p0.ignore_water;
        iteration<ARC_SEM_WORLD, WATER, kMaxBlock, kLdsChunks, OFF32>(w, pl, L, C, s_roots, lds_roots);
      } else {
        iteration<ARC_SEM_WORLD, WATER, kMaxBlock, kLdsChunks, OFF32>(w, sel_params(b, pb, p0), L, C, s_roots,
                                                                      lds_roots);
      }
    }
    const bool idle_lane = fget(L.st, F_LIVE, 1) == 0u;  // pending lanes are not live
    const uint32_t nidle = (uint32_t)__popcll(__ballot(idle_lane));
    if (nidle >= thr) {  // uniform: a ballot count against a uniform threshold
      const uint64_t hold = __ballot(idle_lane && (L.st & kHold) != 0u);
      // nothing left to claim, every lane idle and nothing held: done (uniform exit)
      if (rfl((uint32_t)(drained_b && nidle == 64u && hold == 0ull))) break;
      tick = rfl(tick + 1u);
      if (rfl((uint32_t)(nidle >= refill || nidle == 64u || (tick & 7u) == 0u))) {
        if (idle_lane && fget(L.st, F_HASRAY, 1)) {
          const uint32_t depth = fget(L.st, F_DEPTH, 3);
          const uint4 v = compact_record<ARC_SEM_WORLD>(rays, L, depth ? miss_b : s_miss[threadIdx.x]);
          out[(size_t)depth * n + L.idx] = v;
          if (depth == 0u)
            msg_store(msg + L.idx, (v.w & ((uint32_t)ARC_HIT_FLAG_HIT << 16)) != 0u
                                       ? ((uint64_t)(1u + (v.z >> 24)) << 32) | v.x
                                       : kMsgMiss);
          L.st = 0u;
        }
        if (!drained_b && nidle >= refill) {
          const bool needs = idle_lane && (L.st & kHold) == 0u;
          const uint64_t nm = __ballot(needs);
          const uint32_t nneed = (uint32_t)__popcll(nm);
          const uint32_t rank = (uint32_t)__popcll(nm & lt_mask);
          uint32_t assigned = 0u, my = 0xFFFFFFFFu;
          bool my_b = false;
          while (assigned < nneed) {
            if (wb >= we) {
              if (!drained_p && !claim_window(queue, qc, qtried, batch, n, wb, we)) drained_p = true;
              if (drained_p && !claim_window(queue_b, qc_b, qtried_b, batch, n, wb, we)) {
                drained_b = true;
                break;
              }
            }
            const uint32_t take = umin_(nneed - assigned, we - wb);
            if (needs && rank >= assigned && rank < assigned + take) {
              my = wb + (rank - assigned);
              my_b = drained_p;
            }
            wb += take;
            assigned += take;
          }
          if (my != 0xFFFFFFFFu) {
            L.idx = my;
            if (my_b) {
              L.st = 1u << F_PEND;
            } else {
              const arc_ray r = rays[my];
              s_miss[threadIdx.x] = r.t_max * 16.0f;
              ray_init<ARC_SEM_WORLD>(w, p0, r, L, C);
            }
          }
        }
        if (fget(L.st, F_PEND, 1)) {
          // the parent ray is loaded beside the message (one memory round trip instead of two)
          const arc_ray pr = rays[L.idx];
          const uint64_t m = msg_load(msg + L.idx);
          if (m != 0ull) {
            msg_store(msg + L.idx, 0ull);
            bounce_take(w, pb, pr, m, seed, bounce_tmax, out_b, L, C, s_frame);
          }
        }
      }
      const bool live = fget(L.st, F_LIVE, 1) != 0u;
      const bool pend = __ballot(fget(L.st, F_PEND, 1) != 0u) != 0ull;
      const bool prim = drained_b && __ballot((L.st & (1u << F_HASRAY)) != 0u && fget(L.st, F_DEPTH, 3) == 0u) != 0ull;
      thr = rfl((pend || prim) ? 0u : (drained_b ? 64u : refill));
      // nothing to trace and a message still missing: poll more slowly (~500 cycles)
      if (rfl((uint32_t)(pend && __ballot(live) == 0ull))) __builtin_amdgcn_s_sleep(8);
    }
  }
  queue_finish2(queue, queue_b, C == s_cold);  // thread 0 (C = s_cold + threadIdx.x)
}

}  // namespace arcdev

// ---------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------
namespace arcdev {

static DevParams make_params(const arc_trace_params& t) {
  DevParams p;
  p.lod_bias = t.lod_bias;
  p.ignore_water = t.ignore_water;
  const uint32_t cls = t.lod_bias == 0u ? 0u : (t.lod_bias == 1u ? 1u : 2u);
  p.max_steps = t.max_steps[cls];
  p.lod_step = t.lod_step[cls];
  p.lod_t = t.lod_dist_m[cls] * 16.0f;
  p.lod_t_far = (t.lod_dist_m[1] * 16.0f) * 2.2857f;
  p.ao_max_dist = t.ao_max_dist_bits;
  // LOD never engages: t > lod_t (and t > lod_t_far for class 1) is false for +inf or NaN
  // thresholds, sidx > lod_step is false for 2^32 - 1
  const bool far_off = !(p.lod_t < INFINITY) && (cls != 1u || !(p.lod_t_far < INFINITY));
  p.flags = (far_off && p.lod_step == 0xFFFFFFFFu) ? kParamLodOff : 0u;
  return p;
}

template <uint32_t SEM, bool FULL>
static const void* kernel_ptr() {
  return reinterpret_cast<const void*>(&k_trace<SEM, FULL, SEM == ARC_SEM_WORLD, true>);
}

hipError_t kernel_occupancy(uint32_t semantics, bool full, uint32_t block_threads, uint32_t* blocks_per_cu) {
  int nb = 0;
  const void* f = semantics == ARC_SEM_CHUNK ? (full ? kernel_ptr<ARC_SEM_CHUNK, true>() : kernel_ptr<ARC_SEM_CHUNK, false>())
                                             : (full ? kernel_ptr<ARC_SEM_WORLD, true>() : kernel_ptr<ARC_SEM_WORLD, false>());
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, (int)block_threads, 0);
  *blocks_per_cu = nb > 0 ? (uint32_t)nb : 1u;
  return e;
}

template <uint32_t SEM, bool FULL, bool WATER>
static void launch_one(const LaunchArgs& a, const DevWorld& w, const DevParams& p, hipStream_t stream) {
  auto out = reinterpret_cast<typename OutPtr<FULL>::T*>(a.out);
  if (a.off32)
    hipLaunchKernelGGL((k_trace<SEM, FULL, WATER, true>), dim3(a.blocks), dim3(a.block_threads), 0, stream, w, p,
                       a.rays, out, (uint32_t)a.n, a.queue, a.refill, a.batch);
  else
    hipLaunchKernelGGL((k_trace<SEM, FULL, WATER, false>), dim3(a.blocks), dim3(a.block_threads), 0, stream, w, p,
                       a.rays, out, (uint32_t)a.n, a.queue, a.refill, a.batch);
}

hipError_t launch_trace(const LaunchArgs& a, hipStream_t stream) {
  const DevWorld w = make_world(a.grid, a.roots, a.nodes, a.bricks);
  const DevParams p = make_params(a.params);
  if (a.params.semantics == ARC_SEM_CHUNK) {
    if (a.full) launch_one<ARC_SEM_CHUNK, true, false>(a, w, p, stream);
    else launch_one<ARC_SEM_CHUNK, false, false>(a, w, p, stream);
  } else if (p.ignore_water) {
    if (a.full) launch_one<ARC_SEM_WORLD, true, true>(a, w, p, stream);
    else launch_one<ARC_SEM_WORLD, false, true>(a, w, p, stream);
  } else {
    if (a.full) launch_one<ARC_SEM_WORLD, true, false>(a, w, p, stream);
    else launch_one<ARC_SEM_WORLD, false, false>(a, w, p, stream);
  }
  return hipGetLastError();
}

static const void* path_kernel(uint32_t block_threads) {
  return block_threads == kPathLargeBlock
             ? reinterpret_cast<const void*>(&k_path<kPathLargeBlock, kLdsChunksLarge, true, true, false>)
             : reinterpret_cast<const void*>(&k_path<kMaxBlock, kLdsChunks, true, true, false>);
}

static bool params_uniform(const DevParams& p0, const DevParams& pb);

template <uint32_t BLOCK, uint32_t NCH, bool WATER>
static void launch_path_as(const PathLaunchArgs& a, const DevWorld& w, const PathCtl& pc, hipStream_t stream) {
  const DevParams p0 = make_params(a.primary), pb = make_params(a.bounce);
  uint4* const out = reinterpret_cast<uint4*>(a.out);
  const bool uni = params_uniform(p0, pb);
  if (a.off32 && uni)
    hipLaunchKernelGGL((k_path<BLOCK, NCH, WATER, true, true>), dim3(a.blocks), dim3(a.block_threads), 0, stream, w,
                       p0, pb, pc, out, a.queue, a.refill, a.batch);
  else if (a.off32)
    hipLaunchKernelGGL((k_path<BLOCK, NCH, WATER, true, false>), dim3(a.blocks), dim3(a.block_threads), 0, stream, w,
                       p0, pb, pc, out, a.queue, a.refill, a.batch);
  else if (uni)
    hipLaunchKernelGGL((k_path<BLOCK, NCH, WATER, false, true>), dim3(a.blocks), dim3(a.block_threads), 0, stream, w,
                       p0, pb, pc, out, a.queue, a.refill, a.batch);
  else
    hipLaunchKernelGGL((k_path<BLOCK, NCH, WATER, false, false>), dim3(a.blocks), dim3(a.block_threads), 0, stream, w,
                       p0, pb, pc, out, a.queue, a.refill, a.batch);
}

hipError_t path_occupancy(uint32_t block_threads, uint32_t* blocks_per_cu) {
  int nb = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, path_kernel(block_threads), (int)block_threads, 0);
  *blocks_per_cu = nb > 0 ? (uint32_t)nb : 1u;
  return e;
}

hipError_t launch_path(const PathLaunchArgs& a, hipStream_t stream) {
  const DevWorld w = make_world(a.grid, a.roots, a.nodes, a.bricks);
  PathCtl pc;
  pc.views = a.views;
  pc.view_first = a.view_first;
  pc.n_views = a.n_views;
  pc.n_items = a.n_items;
  pc.n_bounces = a.n_bounces;
  pc.bounce_tmax = a.bounce_max_dist_m;
  pc.seed = a.seed;
  const bool water = a.primary.ignore_water || a.bounce.ignore_water;
  const bool large = a.block_threads == kPathLargeBlock;
  if (large && water) launch_path_as<kPathLargeBlock, kLdsChunksLarge, true>(a, w, pc, stream);
  else if (large) launch_path_as<kPathLargeBlock, kLdsChunksLarge, false>(a, w, pc, stream);
  else if (water) launch_path_as<kMaxBlock, kLdsChunks, true>(a, w, pc, stream);
  else launch_path_as<kMaxBlock, kLdsChunks, false>(a, w, pc, stream);
  return hipGetLastError();
}

// The two classes trace alike apart from ignore_water: LOD never engages in either and the step
// caps agree (lod_bias then only selects the LOD rule, which is off).
static bool params_uniform(const DevParams& p0, const DevParams& pb) {
  return (p0.flags & kParamLodOff) && (pb.flags & kParamLodOff) && p0.max_steps == pb.max_steps;
}

hipError_t bounce_occupancy(uint32_t block_threads, uint32_t* blocks_per_cu) {
  int nb = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &nb, reinterpret_cast<const void*>(&k_trace_bounce<true, true, false>), (int)block_threads, 0);
  *blocks_per_cu = nb > 0 ? (uint32_t)nb : 1u;
  return e;
}

template <bool WATER, bool OFF32>
static void launch_bounce_as(const BounceLaunchArgs& a, const DevWorld& w, const DevParams& p0, const DevParams& pb,
                             hipStream_t stream) {
  uint4* const out = reinterpret_cast<uint4*>(a.out);
  if (params_uniform(p0, pb))
    hipLaunchKernelGGL((k_trace_bounce<WATER, OFF32, true>), dim3(a.blocks), dim3(a.block_threads), 0, stream, w, p0,
                       pb, a.rays, out, (uint32_t)a.n, a.msg, a.queue, a.seed, a.bounce_max_dist_m, a.refill, a.batch);
  else
    hipLaunchKernelGGL((k_trace_bounce<WATER, OFF32, false>), dim3(a.blocks), dim3(a.block_threads), 0, stream, w,
                       p0, pb, a.rays, out, (uint32_t)a.n, a.msg, a.queue, a.seed, a.bounce_max_dist_m, a.refill,
                       a.batch);
}

hipError_t launch_trace_bounce(const BounceLaunchArgs& a, hipStream_t stream) {
  const DevWorld w = make_world(a.grid, a.roots, a.nodes, a.bricks);
  const DevParams p0 = make_params(a.primary), pb = make_params(a.bounce);
  const bool water = a.primary.ignore_water || a.bounce.ignore_water;
  if (water && a.off32) launch_bounce_as<true, true>(a, w, p0, pb, stream);
  else if (water) launch_bounce_as<true, false>(a, w, p0, pb, stream);
  else if (a.off32) launch_bounce_as<false, true>(a, w, p0, pb, stream);
  else launch_bounce_as<false, false>(a, w, p0, pb, stream);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Queue scratch of the allocation-free entry point (arc_trace_pools_noalloc)
// ---------------------------------------------------------------------------------------
// One queue region (kQueueRegionWords words, 32 KiB) in static device memory: each GPU the library
// runs on holds its own copy until the process ends. The array is zero when the code object loads,
// and the last workgroup of every launch returns the counters to zero, so consecutive calls need
// no clearing. Calls that overlap on one GPU (two streams) share the counters and race.
// In host code `scratch` names a host-side stand-in, never a device address: the device address
// comes from hipGetSymbolAddress, once per device, with that device current.
static __device__ __align__(256) uint32_t scratch[kQueueRegionWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};
uint32_t g_scratch_cu[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(uint32_t** queue, uint32_t* n_cu) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    int cu = 0;
    e = hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    if (p == nullptr || cu <= 0) return hipErrorInvalidValue;
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
    g_scratch_cu[dev] = (uint32_t)cu;
  }
  *queue = g_scratch_ptr[dev];
  *n_cu = g_scratch_cu[dev];
  return hipSuccess;
}

}  // namespace arcdev

