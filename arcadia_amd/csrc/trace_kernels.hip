// 64-tree DDA traversal kernels for gfx950 and their launchers.
//
// Execution model
// ---------------
// A persistent ray queue feeds one ray per lane. Each trip of a wave through its loop runs one
// outer iteration of trace_world_64tree_from (traversal64.wgsl:302-553) for every live lane:
// choose the restart node, descend at most four levels, then either record a hit or take one
// exact DDA box step. The traversal is VALU-issue bound (pointer chasing with a long
// per-step float sequence), so the kernel is shaped to cut the instructions a wave issues:
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
// Queue: a wave claims windows of `batch` consecutive rays with one atomic on a single counter
// and refills its idle lanes from the window, so all waves of all XCDs sweep the ray array
// together and the rays in flight form one coherent band of screen tiles. Measured on MI355X:
// per-XCD counters (the array cut into stripes of 1..4096 windows dealt round-robin to the XCDs,
// each XCD draining its own stripes first) ran 15-25% slower at every stripe size, and 8
// contiguous per-XCD segments 65% slower. The last workgroup to finish resets the counters,
// which keeps a launch self-contained (no memset node: graph-safe).
#include <hip/hip_runtime.h>

#include "raymath.h"
#include "trace_device.h"
#include "trace_launch.h"

namespace arcdev {


// packed per-lane flags
enum : uint32_t {
  F_CSHIFT = 0,  // 3 bits: cached_shift of the reference (level the last descent ended on)
  F_SHIFT = 3,   // 3 bits: level of the hit cell (valid when done with a hit)
  F_LIVE = 6,    // 1 bit: ray in flight
  F_HIT = 7,     // 1 bit: finished with a hit
  F_HASRAY = 8,  // 1 bit: lane owns a ray whose record is not written yet
  F_DEPTH = 9,   // 3 bits: path kernel: 0 = primary ray, d = d-th bounce
  F_DESC = 12,   // 1 bit: descent parked at the pass cap, continues at the F_CSHIFT level
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
  float t, tl;       // t, t_limit
  uint32_t vx, vy, vz;
  uint32_t diff;     // prev_vi ^ vi of the coming iteration
  uint32_t step;     // outer iterations begun
  uint32_t st;       // packed flags
  uint32_t mat;      // material of the hit
  uint4 node;        // node of the level the last descent ended on
  uint32_t idx;
};

// Cold per-lane state in LDS, [slot][lane] so a wave's accesses are contiguous:
//   slot 0: chunk root node (after a hit: the node the hit resolved in),
//   slot 1: cached level-1 ancestor, slot 2: cached level-2 ancestor (slot of shift s:
//   (6 - s) / 2; the level-3 ancestor is the lane's `node`).
constexpr uint32_t kMaxBlock = 256u;
constexpr uint32_t kColdSlots = 3u;

__device__ __forceinline__ uint32_t rfl(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// float -> u32 for a value already clamped to x >= 0 (never NaN), whose result is then clamped
// into a box below 2^32: min(x, 4294967040) (the largest float below 2^32) keeps the conversion
// defined and gives the same clamped result as the saturating WGSL u32() / Rust `as`. Plain
// arithmetic (v_min_f32 + v_cvt_u32_f32), so the compiler may speculate it into selects.
__device__ __forceinline__ uint32_t cvt_u32(float x) { return (uint32_t)fminf(x, 4294967040.0f); }

// t at which the ray crosses coordinate `c` on one axis: f32(c) * inv + (-o * inv).
__device__ __forceinline__ float plane_t(uint32_t c, float inv, float noi) {
  const float q = (float)c * inv;
  return q + noi;
}

// LOD min_shift (traversal64.wgsl:355-370); `sidx` is the iteration index.
template <uint32_t SEM>
__device__ __forceinline__ uint32_t lod_min_shift(const DevParams& p, float t, uint32_t sidx) {
  if (SEM == ARC_SEM_CHUNK) return 0u;
  uint32_t m = 0u;
  if (p.lod_bias == 0u) {
    if (t > p.lod_t || sidx > p.lod_step) m = 2u;
  } else if (p.lod_bias == 1u) {
    if (t > p.lod_t_far || sidx > p.lod_step) m = 4u;
    else if (t > p.lod_t) m = 2u;
  } else {
    if (t > p.lod_t || sidx > p.lod_step) m = 2u;
  }
  return m;
}

// Guards at the top of an outer iteration (traversal64.wgsl:302-310 / raycast_cpu.rs:78-84).
template <uint32_t SEM>
__device__ __forceinline__ bool iteration_guards_fail(const DevWorld& w, const DevParams& p, const Lane& L) {
  const uint32_t max_steps = SEM == ARC_SEM_CHUNK ? 512u : p.max_steps;
  const uint32_t wx = SEM == ARC_SEM_CHUNK ? 256u : (w.gx << 8);
  const uint32_t wy = SEM == ARC_SEM_CHUNK ? 256u : (w.gy << 8);
  const uint32_t wz = SEM == ARC_SEM_CHUNK ? 256u : (w.gz << 8);
  return L.step >= max_steps || L.t >= L.tl || L.vx >= wx || L.vy >= wy || L.vz >= wz;
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

// Address of the LDS slot of the cached node at shift `sh` (slot (6 - sh) / 2).
__device__ __forceinline__ uint4* cold_slot(uint4* C, uint32_t sh) {
  return reinterpret_cast<uint4*>(reinterpret_cast<char*>(C) + (6u - sh) * (kMaxBlock * 8u));
}

// Chunk roots staged in LDS for worlds of at most kLdsChunks chunks: the root index and the root
// node of every chunk, read at each chunk change in place of chunk_roots[key] followed by
// nodes[root] (two dependent global loads on the longest dependency chain of a wave trip).
constexpr uint32_t kLdsChunks = 256u;
struct RootLds {
  uint4 node[kLdsChunks];
  uint32_t idx[kLdsChunks];
};

// Fills `R` from the pools; returns false (block-uniform) when the world has more chunks.
__device__ __forceinline__ bool stage_roots(const DevWorld& w, RootLds& R) {
  const uint32_t nc = w.gx * w.gy * w.gz;
  if (nc > kLdsChunks) return false;
  for (uint32_t c = threadIdx.x; c < nc; c += blockDim.x) {
    const uint32_t r = w.roots[c];
    R.idx[c] = r;
    R.node[c] = r == ARC_EMPTY_CHUNK_SENTINEL ? make_uint4(0u, 0u, 0u, 0u) : w.nodes[r];
  }
  __syncthreads();
  return true;
}

// Descent passes (child-bit tests) one wave trip may run for a lane before it parks the descent
// and continues it on the next trip; 4 = never park (a descent has at most four levels).
constexpr uint32_t kMaxPasses = 4u;

// One outer iteration of a live lane whose guards passed (traversal64.wgsl:311-552).
// Instruction count is what bounds the kernel, so the sequence is written for few VALU ops:
//   * the child test and the child slot come from one 64-bit shift of the child mask that puts
//     bit b at the top: the sign gives has_child(b), the popcount minus one gives
//     popcount(mask & ((1 << b) - 1));
//   * the descent loop only tests and loads; the material of a hit candidate (a set bit in a
//     leaf or at the LOD floor) is resolved after the loop;
//   * cached_shift always equals the level the iteration ended on, so one field carries both;
//   * the chunk root is read from the chunk_roots array on every chunk change (the array is a
//     few KB and stays in cache), which gives the same root as the reference's 2-slot cache;
//   * the box step computes the clamped inside coordinate of all three axes and selects, with no
//     branches;
//   * WATER = false compiles the ignore_water tests out (ray classes that never set it).
template <uint32_t SEM, bool WATER>
__device__ __forceinline__ void iteration(const DevWorld& w, const DevParams& p, Lane& L, uint4* C, const RootLds& R,
                                          bool lds_roots) {
  uint32_t st = L.st;
  uint4 node = L.node;
  uint32_t sh = fget(st, F_CSHIFT, 3);
  uint32_t span_mul = 1u;  // the box stepped over is span_mul << sh
  bool descend = true;
  uint32_t minsh;
  if (kMaxPasses < 4u && fget(st, F_DESC, 1)) {  // continue a parked descent
    minsh = lod_min_shift<SEM>(p, L.t, L.step - 1u);
  } else {
    const uint32_t sidx = L.step;
    L.step = sidx + 1u;
    minsh = lod_min_shift<SEM>(p, L.t, sidx);
    const uint32_t diff = L.diff;
    if (SEM == ARC_SEM_WORLD && diff >= 256u) {  // chunk change (:316-335)
      const uint32_t key = (L.vx >> 8) + __umul24(L.vy >> 8, w.gx) + __umul24(L.vz >> 8, w.gx * w.gy);
      uint32_t root;
      if (lds_roots) {
        root = R.idx[key];
        node = R.node[key];
      } else {
        root = w.roots[key];
        if (root != ARC_EMPTY_CHUNK_SENTINEL) node = w.nodes[root];
      }
      sh = 6u;
      if (root == ARC_EMPTY_CHUNK_SENTINEL) {
        descend = false;  // step over the empty chunk (:337-352)
        span_mul = 4u;
      } else {
        C[0] = node;
      }
    } else {  // ancestor restart (:374-386)
      const uint32_t rs = restart_shift(diff, minsh, sh);
      if (rs != sh) node = *cold_slot(C, rs);
      sh = rs;
    }
  }
  bool cand = false;  // set bit at a leaf or at the LOD floor: hit unless the voxel is air/water
  if (descend) {
    uint32_t loads = 0u;
    for (;;) {  // descent (:390-552), at most four levels
      const uint32_t b = (__builtin_amdgcn_ubfe(L.vz, sh, 2) << 4) | (__builtin_amdgcn_ubfe(L.vy, sh, 2) << 2) |
                         __builtin_amdgcn_ubfe(L.vx, sh, 2);
      if (SEM == ARC_SEM_WORLD && WATER && p.ignore_water && (node.z & ARC_IS_UNIFORM_FLAG) != 0u &&
          (node.w >> 24) == ARC_MAT_WATER) {
        span_mul = 4u;
        break;
      }
      const uint64_t below = (((uint64_t)node.y << 32) | node.x) << (63u - b);
      if ((int64_t)below >= 0) break;  // no child at b
      // Level-3 nodes are leaves in every pool the upload check accepts; sh == 0 (implied by
      // sh <= minsh) only bounds the descent if a malformed node slipped in through a patch.
      if ((node.z & ARC_IS_LEAF_FLAG) != 0u || sh <= minsh) {
        cand = true;
        break;
      }
      node = w.nodes[(node.z & ARC_PTR_INDEX_MASK) + (uint32_t)__popcll(below) - 1u];
      sh -= 2u;
      if (sh != 0u) *cold_slot(C, sh) = node;
      if (kMaxPasses < 4u && ++loads == kMaxPasses) {  // park: the next trip tests `node`
        L.node = node;
        L.st = fput(fput(st, F_CSHIFT, 3, sh), F_DESC, 1, 1u);
        return;
      }
    }
  }
  if (cand) {  // :441-538
    const uint32_t b = (__builtin_amdgcn_ubfe(L.vz, sh, 2) << 4) | (__builtin_amdgcn_ubfe(L.vy, sh, 2) << 2) |
                       __builtin_amdgcn_ubfe(L.vx, sh, 2);
    uint32_t mat = node.w >> 24;
    if ((node.z & (ARC_IS_LEAF_FLAG | ARC_IS_UNIFORM_FLAG)) == ARC_IS_LEAF_FLAG) {
      const uint32_t word = w.bricks[(size_t)(node.z & ARC_PTR_INDEX_MASK) * 16u + (b >> 2)];
      mat = __builtin_amdgcn_ubfe(word, (b & 3u) * 8u, 8);
    }
    const bool skip_water = SEM == ARC_SEM_WORLD && WATER && p.ignore_water && mat == ARC_MAT_WATER;
    if (!skip_water && mat != ARC_MAT_AIR) {  // hit: resolved at finalize (:467-523)
      L.mat = mat;
      L.node = node;
      L.st = fput(fput(fput(fput(fput(st, F_CSHIFT, 3, sh), F_SHIFT, 3, sh), F_LIVE, 1, 0u), F_HIT, 1, 1u),
                  F_DESC, 1, 0u);
      return;
    }
  }
  L.node = node;
  // step_dda_box (traversal64.wgsl:152-179 / raycast_cpu.rs:245-292)
  const uint32_t span = span_mul << sh;
  const uint32_t mk = 0u - span;
  const uint32_t bx = L.vx & mk, by = L.vy & mk, bz = L.vz & mk;
  const bool px = L.ix >= 0.0f, py = L.iy >= 0.0f, pz = L.iz >= 0.0f;
  const uint32_t ex = px ? bx + span : bx, ey = py ? by + span : by, ez = pz ? bz + span : bz;
  const float tx = plane_t(ex, L.ix, L.nx), ty = plane_t(ey, L.iy, L.ny), tz = plane_t(ez, L.iz, L.nz);
  const float t_exit = fminf(fminf(tx, ty), tz);
  const float next_t = fmaxf(L.t, t_exit);
  const bool is_x = tx <= ty && tx <= tz;
  const bool is_y = !is_x && (ty <= tz);
  const float qx = L.dx * next_t, qy = L.dy * next_t, qz = L.dz * next_t;
  const float ipx = fmaxf(L.ox + qx, 0.0f), ipy = fmaxf(L.oy + qy, 0.0f), ipz = fmaxf(L.oz + qz, 0.0f);
  const uint32_t last = span - 1u;
  const uint32_t cx = umin_(umax_(cvt_u32(ipx), bx), bx + last);
  const uint32_t cy = umin_(umax_(cvt_u32(ipy), by), by + last);
  const uint32_t cz = umin_(umax_(cvt_u32(ipz), bz), bz + last);
  const uint32_t fx = px ? ex : bx - 1u, fy = py ? ey : by - 1u, fz = pz ? ez : bz - 1u;
  const uint32_t nvx = is_x ? fx : cx;
  const uint32_t nvy = is_y ? fy : cy;
  const uint32_t nvz = (is_x || is_y) ? cz : fz;
  L.diff = (nvx ^ L.vx) | (nvy ^ L.vy) | (nvz ^ L.vz);
  L.vx = nvx;
  L.vy = nvy;
  L.vz = nvz;
  L.t = next_t;
  st = fput(fput(st, F_CSHIFT, 3, sh), F_DESC, 1, 0u);
  if (iteration_guards_fail<SEM>(w, p, L)) st = fput(fput(st, F_LIVE, 1, 0u), F_HIT, 1, 0u);
  L.st = st;
}

// Ray setup: traversal64.wgsl:234-300 (WORLD) / raycast_cpu.rs:34-76 (CHUNK), then the guards
// of iteration 0.
template <uint32_t SEM>
__device__ __forceinline__ void ray_init(const DevWorld& w, const DevParams& p, const arc_ray& r, Lane& L,
                                         uint4* C) {
  L.step = 0u;
  L.st = fput(0u, F_HASRAY, 1, 1u);  // not live until the setup succeeds
  if (SEM == ARC_SEM_WORLD) {
    const float max_t_v = r.t_max * 16.0f;
    const float start_t_v = fmaxf(r.t_start * 16.0f, 0.0f);
    L.ox = r.ox * 16.0f;
    L.oy = r.oy * 16.0f;
    L.oz = r.oz * 16.0f;
    L.dx = r.dx;
    L.dy = r.dy;
    L.dz = r.dz;
    const float sx = fabsf(r.dx) < 1e-7f ? 1e-7f : r.dx;
    const float sy = fabsf(r.dy) < 1e-7f ? 1e-7f : r.dy;
    const float sz = fabsf(r.dz) < 1e-7f ? 1e-7f : r.dz;
    L.ix = 1.0f / sx;
    L.iy = 1.0f / sy;
    L.iz = 1.0f / sz;
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
    if (wh_y < t_enter || t_enter >= max_t_v) return;
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
    L.ix = 1.0f / L.dx;
    L.iy = 1.0f / L.dy;
    L.iz = 1.0f / L.dz;
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
    C[0] = rn;    // chunk root for every root restart
    L.node = rn;  // `node` == cached ancestor at cached_shift (6): iteration 0 restarts here
  }
  L.nx = (-L.ox) * L.ix;
  L.ny = (-L.oy) * L.iy;
  L.nz = (-L.oz) * L.iz;
  L.diff = 0xFFFFFFFFu;  // prev_vi starts at u32::MAX
  if (iteration_guards_fail<SEM>(w, p, L)) return;
  L.st = fput(fput(L.st, F_CSHIFT, 3, 6u), F_LIVE, 1, 1u);
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

template <uint32_t SEM>
__device__ __forceinline__ void store_compact(arc_hit* out, const arc_ray* rays, const Lane& L) {
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
    v.z = (L.vz & 0xFFFFu) | ((L.mat & 0xFFu) << 16) | (face << 24);
    v.w = steps | ((ARC_HIT_FLAG_HIT | (fget(L.st, F_SHIFT, 3) > 0u ? ARC_HIT_FLAG_LOD : 0u)) << 16);
  } else {
    float miss_t = 0.0f;
    if (SEM == ARC_SEM_WORLD) miss_t = rays[L.idx].t_max * 16.0f;
    v.x = __float_as_uint(miss_t);
    v.y = 0u;
    v.z = (uint32_t)ARC_FACE_NONE << 24;
    v.w = steps;
  }
  reinterpret_cast<uint4*>(out)[L.idx] = v;
}

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
    mat = L.mat;
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
                                         const Lane& L, const uint4* C) {
  if (FULL) store_full<SEM>(reinterpret_cast<arc_trace_result*>(out), rays, p, L, C);
  else store_compact<SEM>(reinterpret_cast<arc_hit*>(out), rays, L);
}

// The last workgroup out resets the counters for the next launch on this stream.
__device__ __forceinline__ void queue_finish(uint32_t* queue, bool leader) {
  __syncthreads();
  if (leader) {
    __threadfence();
    const uint32_t done = atomicAdd(queue + 1, 1u);
    if (done == gridDim.x - 1u) {
      for (uint32_t s = 0; s < kQueueWords; ++s) atomicExch(queue + s, 0u);
      __threadfence();
    }
  }
}

// Persistent ray-queue kernel.
// queue[0]: rays claimed; queue[1]: finished workgroups (both zero between launches).
// Occupancy target: the compact variants fit 64 VGPRs (8 waves/SIMD) without spilling; the
// full-record variants carry more live values through finalize and keep the default budget.
template <uint32_t SEM, bool FULL, bool WATER>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FULL ? 1 : 8, 8))) void k_trace(
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
  bool drained = false;
  for (;;) {
    const bool idle_lane = fget(L.st, F_LIVE, 1) == 0u;
    const uint64_t idle = __ballot(idle_lane);
    const uint32_t nidle = (uint32_t)__popcll(idle);
    if (drained) {
      if (nidle == 64u) break;
    } else if (nidle >= refill) {
      // finalize finished lanes together, then refill every idle lane
      if (idle_lane && fget(L.st, F_HASRAY, 1)) {
        finalize<SEM, FULL>(out, rays, p, L, C);
        L.st = fput(L.st, F_HASRAY, 1, 0u);
      }
      const uint32_t rank = (uint32_t)__popcll(idle & lt_mask);
      uint32_t assigned = 0u;
      uint32_t my = 0xFFFFFFFFu;
      while (assigned < nidle) {
        if (wb >= we) {
          uint32_t got = 0u;
          if (lane == 0) got = atomicAdd(queue, batch);
          got = rfl(__shfl(got, 0));
          const uint32_t lo = umin_(got, n);
          if (lo >= n) {
            drained = true;
            break;
          }
          wb = lo;
          we = umin_(lo + batch, n);
        }
        const uint32_t take = umin_(nidle - assigned, we - wb);
        if (idle_lane && rank >= assigned && rank < assigned + take) my = wb + (rank - assigned);
        wb += take;
        assigned += take;
      }
      if (my != 0xFFFFFFFFu) {
        L.idx = my;
        const arc_ray r = rays[my];
        ray_init<SEM>(w, p, r, L, C);
      }
      continue;  // re-ballot: freshly initialised lanes may already be done
    }
    if (!idle_lane) iteration<SEM, WATER>(w, p, L, C, s_roots, lds_roots);
  }
  if (fget(L.st, F_HASRAY, 1)) finalize<SEM, FULL>(out, rays, p, L, C);
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
  p.pad = 0;
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
__device__ __forceinline__ void path_finish(const DevWorld& w, const DevParams& pb, const PathCtl& pc, uint4* out,
                                            Lane& L, float& tmax, uint4* C) {
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
    v.z = (L.vz & 0xFFFFu) | ((L.mat & 0xFFu) << 16) | (face << 24);
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
    arcmath::bounce_ray(ov, dr, exact_t, face, pc.seed, (uint64_t)depth * pc.n_items + item, pc.bounce_tmax, &r);
    ray_init<ARC_SEM_WORLD>(w, pb, r, L, C);
    L.st = fput(L.st, F_DEPTH, 3, depth + 1u);
    tmax = pc.bounce_tmax;
  } else {
    for (uint32_t d = depth + 1u; d <= pc.n_bounces; ++d) out[(size_t)d * pc.n_items + item] = empty_record();
    L.st = fput(L.st, F_HASRAY, 1, 0u);
  }
}

// queue[0]: slots claimed; queue[1]: finished workgroups (both zero between launches).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_path(
    DevWorld w, DevParams p0, DevParams pb, PathCtl pc, uint4* __restrict__ out, uint32_t* __restrict__ queue,
    uint32_t refill, uint32_t batch) {
  __shared__ uint4 s_cold[kColdSlots * kMaxBlock];
  __shared__ RootLds s_roots;
  uint4* const C = s_cold + threadIdx.x;
  const bool lds_roots = stage_roots(w, s_roots);
  const uint32_t lane = __lane_id();
  Lane L;
  L.st = 0u;
  L.step = 0u;
  L.idx = 0u;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
  float tmax = 0.0f;
  uint32_t wb = 0u, we = 0u;
  bool drained = false;
  for (;;) {
    const bool idle_lane = fget(L.st, F_LIVE, 1) == 0u;
    const uint32_t nidle = (uint32_t)__popcll(__ballot(idle_lane));
    const bool pending = __ballot(idle_lane && fget(L.st, F_HASRAY, 1)) != 0ull;
    if (drained ? pending : nidle >= refill) {
      if (idle_lane && fget(L.st, F_HASRAY, 1)) path_finish(w, pb, pc, out, L, tmax, C);
      if (!drained) {
        const bool needs = fget(L.st, F_LIVE, 1) == 0u && fget(L.st, F_HASRAY, 1) == 0u;
        const uint64_t nm = __ballot(needs);
        const uint32_t nneed = (uint32_t)__popcll(nm);
        const uint32_t rank = (uint32_t)__popcll(nm & lt_mask);
        uint32_t assigned = 0u, my = 0xFFFFFFFFu;
        while (assigned < nneed) {
          if (wb >= we) {
            uint32_t got = 0u;
            if (lane == 0) got = atomicAdd(queue, batch);
            got = rfl(__shfl(got, 0));
            if (got >= pc.n_items) {
              drained = true;
              break;
            }
            wb = got;
            we = umin_(got + batch, pc.n_items);
          }
          const uint32_t take = umin_(nneed - assigned, we - wb);
          if (needs && rank >= assigned && rank < assigned + take) my = wb + (rank - assigned);
          wb += take;
          assigned += take;
        }
        if (my != 0xFFFFFFFFu) path_start(w, p0, pc, my, out, L, tmax, C);
      }
      continue;
    }
    if (drained && nidle == 64u) break;
    if (!idle_lane)
      iteration<ARC_SEM_WORLD, true>(w, sel_params(fget(L.st, F_DEPTH, 3) != 0u, pb, p0), L, C, s_roots, lds_roots);
  }
  queue_finish(queue, C == s_cold);  // thread 0 (C = s_cold + threadIdx.x)
}

}  // namespace arcdev

// ---------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------
namespace arcdev {

static DevWorld make_world(const uint32_t grid[3], const uint32_t* roots, const arc_node64* nodes,
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
  p.pad = 0;
  return p;
}

template <uint32_t SEM, bool FULL>
static const void* kernel_ptr() {
  return reinterpret_cast<const void*>(&k_trace<SEM, FULL, SEM == ARC_SEM_WORLD>);
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
  hipLaunchKernelGGL((k_trace<SEM, FULL, WATER>), dim3(a.blocks), dim3(a.block_threads), 0, stream, w, p, a.rays,
                     reinterpret_cast<typename OutPtr<FULL>::T*>(a.out), (uint32_t)a.n, a.queue, a.refill, a.batch);
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

hipError_t path_occupancy(uint32_t block_threads, uint32_t* blocks_per_cu) {
  int nb = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&k_path),
                                                               (int)block_threads, 0);
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
  hipLaunchKernelGGL(k_path, dim3(a.blocks), dim3(a.block_threads), 0, stream, w, make_params(a.primary),
                     make_params(a.bounce), pc, reinterpret_cast<uint4*>(a.out), a.queue, a.refill, a.batch);
  return hipGetLastError();
}

}  // namespace arcdev
