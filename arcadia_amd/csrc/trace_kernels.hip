// 64-tree DDA traversal kernels for gfx950 and their launchers.
//
// Launch model: a persistent ray queue. The grid holds blocks_per_cu x n_cu workgroups;
// every wave64 loops { refill idle lanes from a global atomic ray counter when at least
// `refill` lanes are idle (one atomic per wave, lanes ranked with mbcnt over the ballot of
// idle lanes); advance each live lane by one outer DDA iteration } until the queue is drained
// and no lane is live. Results are written by ray index, so the schedule never changes an
// output. The last workgroup to finish resets the queue counters, which keeps a launch
// self-contained (no memset node, safe under HIP graph capture).
#include <hip/hip_runtime.h>

#include "trace_device.h"
#include "trace_launch.h"

namespace arcdev {

template <uint32_t SEM>
struct Sem;

// ---------------------------------------------------------------------------------------
// Results
// ---------------------------------------------------------------------------------------
struct Outcome {
  bool hit;
  float t_voxels;
  uint32_t vi[3];
  uint32_t mat;
  uint32_t axis;
  float nrm_axis;  // value of the normal on `axis` (-sign / -signum of dir)
  uint32_t shift;
  uint32_t steps;
  // full-record extras
  float hp[3];
  uint32_t lod_payload;
  float contact_ao;
};

__device__ __forceinline__ void store_compact(arc_hit* out, uint32_t idx, const Outcome& o) {
  uint4 v;
  v.x = __float_as_uint(o.t_voxels);
  uint32_t steps = o.steps > 0xFFFFu ? 0xFFFFu : o.steps;
  if (o.hit) {
    uint32_t face = o.nrm_axis > 0.0f ? 2u * o.axis : (o.nrm_axis < 0.0f ? 2u * o.axis + 1u : 6u + o.axis);
    v.y = (o.vi[0] & 0xFFFFu) | (o.vi[1] << 16);
    v.z = (o.vi[2] & 0xFFFFu) | ((o.mat & 0xFFu) << 16) | (face << 24);
    v.w = steps | ((ARC_HIT_FLAG_HIT | (o.shift > 0 ? ARC_HIT_FLAG_LOD : 0u)) << 16);
  } else {
    v.y = 0u;
    v.z = (uint32_t)ARC_FACE_NONE << 24;
    v.w = steps;
  }
  reinterpret_cast<uint4*>(out)[idx] = v;
}

// ---------------------------------------------------------------------------------------
// World semantics: trace_world_64tree_from (traversal64.wgsl:226-556)
// ---------------------------------------------------------------------------------------
template <>
struct Sem<ARC_SEM_WORLD> {
  // Returns false when the ray is already finished (miss before the loop).
  static __device__ __forceinline__ bool init(const DevWorld& w, const arc_ray& r, RayState& s, Outcome& o) {
    const float o3[3] = {r.ox, r.oy, r.oz};
    const float d3[3] = {r.dx, r.dy, r.dz};
    const uint32_t g[3] = {w.gx, w.gy, w.gz};
    o.hit = false;
    o.t_voxels = r.t_max * 16.0f;
    o.steps = 0;
    o.contact_ao = 1.0f;
    o.lod_payload = 0;
    o.mat = 0;
    o.shift = 0;
    o.axis = 1;
    o.nrm_axis = 1.0f;
    for (int a = 0; a < 3; ++a) {
      float m = d3[a] * r.t_max;
      o.hp[a] = o3[a] + m;  // res.hit_pos_meters = origin_m + dir * max_dist_m
      o.vi[a] = 0;
    }
    const float max_t_v = r.t_max * 16.0f;
    const float start_t_v = fmaxf(r.t_start * 16.0f, 0.0f);
    float tminv[3], tmaxv[3];
    for (int a = 0; a < 3; ++a) {
      s.vo[a] = o3[a] * 16.0f;
      s.dir[a] = d3[a];
      s.rdir[a] = d3[a];
      float safe = fabsf(d3[a]) < 1e-7f ? 1e-7f : d3[a];
      s.inv[a] = 1.0f / safe;
      float wmax = (float)g[a] * 256.0f;
      float eb = s.inv[a] > 0.0f ? 0.0f : wmax;
      float xb = s.inv[a] > 0.0f ? wmax : 0.0f;
      float e0 = eb - s.vo[a], x0 = xb - s.vo[a];
      tminv[a] = e0 * s.inv[a];
      tmaxv[a] = x0 * s.inv[a];
    }
    float wh_x = fmaxf(fmaxf(tminv[0], tminv[1]), tminv[2]);
    float wh_y = fminf(fminf(tmaxv[0], tmaxv[1]), tmaxv[2]);
    float t_enter = fmaxf(wh_x, start_t_v);
    if (wh_y < t_enter || t_enter >= max_t_v) return false;
    s.t = t_enter;
    s.t_limit = fminf(wh_y, max_t_v);
    for (int a = 0; a < 3; ++a) {
      float nv = -s.vo[a];
      s.noi[a] = nv * s.inv[a];
      float wmax = (float)g[a] * 256.0f;
      float tt = s.t + 1e-4f;
      float dm = s.dir[a] * tt;
      float sp = clampf_(s.vo[a] + dm, 0.0f, wmax - 1e-3f);
      s.vi[a] = f2u(sp);
      s.prev[a] = 0xFFFFFFFFu;
    }
    s.mtf_k0 = s.mtf_k1 = 0xFFFFFFFFu;
    s.mtf_r0 = s.mtf_r1 = ARC_EMPTY_CHUNK_SENTINEL;
    s.cached_shift = 6u;
    s.step = 0;
    s.l1 = s.l2 = s.l3 = make_uint4(0, 0, 0, 0);
    return true;
  }
};

// ---------------------------------------------------------------------------------------
// Chunk semantics: raycast_chunk_tree (raycast_cpu.rs:28-218) on chunk 0
// ---------------------------------------------------------------------------------------
template <>
struct Sem<ARC_SEM_CHUNK> {
  static __device__ __forceinline__ bool init(const DevWorld& w, const arc_ray& r, RayState& s, Outcome& o) {
    const float o3[3] = {r.ox, r.oy, r.oz};
    const float d3[3] = {r.dx, r.dy, r.dz};
    o.hit = false;
    o.t_voxels = 0.0f;
    o.steps = 0;
    o.contact_ao = 1.0f;
    o.lod_payload = 0;
    o.mat = 0;
    o.shift = 0;
    o.axis = 0;
    o.nrm_axis = 0.0f;
    for (int a = 0; a < 3; ++a) {
      o.hp[a] = 0.0f;
      o.vi[a] = 0;
    }
    const uint32_t root = w.roots[0];
    if (root == ARC_EMPTY_CHUNK_SENTINEL) return false;
    const uint4 rn = w.nodes[root];
    if (rn.x == 0u && rn.y == 0u) return false;  // tree.is_empty()
    float tmn[3], tmx[3];
    for (int a = 0; a < 3; ++a) {
      s.vo[a] = o3[a];
      s.rdir[a] = d3[a];
      float safe = fabsf(d3[a]) > 1e-7f ? d3[a] : copysignf(1e-7f, d3[a]);
      s.dir[a] = safe;
      s.inv[a] = 1.0f / safe;
      float a0 = 0.0f - o3[a], a1 = 256.0f - o3[a];
      float t0 = a0 * s.inv[a], t1 = a1 * s.inv[a];
      tmn[a] = fminf(t0, t1);
      tmx[a] = fmaxf(t0, t1);
    }
    float tmin = fmaxf(fmaxf(tmn[0], tmn[1]), tmn[2]);
    float tmax = fminf(fminf(tmx[0], tmx[1]), tmx[2]);
    if (!(tmax >= fmaxf(tmin, 0.0f))) return false;
    float t_enter = fmaxf(tmin, 0.0f);
    if (tmax < t_enter || t_enter >= r.t_max) return false;
    s.t = t_enter;
    s.t_limit = fminf(tmax, r.t_max);
    for (int a = 0; a < 3; ++a) {
      float nv = -o3[a];
      s.noi[a] = nv * s.inv[a];
      float tt = s.t + 1e-4f;
      float dm = s.dir[a] * tt;
      float sp = clampf_(o3[a] + dm, 0.0f, 256.0f - 1e-3f);
      s.vi[a] = f2u(sp);
      s.prev[a] = 0xFFFFFFFFu;
    }
    s.root_cpu = root;
    s.mtf_k0 = s.mtf_k1 = 0xFFFFFFFFu;
    s.mtf_r0 = s.mtf_r1 = root;
    s.cached_shift = 6u;
    s.step = 0;
    s.l1 = s.l2 = s.l3 = make_uint4(0, 0, 0, 0);
    return true;
  }
};

// One outer DDA iteration. Returns true when the ray finished (o filled).
template <uint32_t SEM, bool FULL>
__device__ __forceinline__ bool dda_iteration(const DevWorld& w, const DevParams& p, RayState& s, Outcome& o) {
  const uint32_t max_steps = SEM == ARC_SEM_CHUNK ? 512u : p.max_steps;
  const uint32_t wmx = SEM == ARC_SEM_CHUNK ? 256u : (w.gx << 8);
  const uint32_t wmy = SEM == ARC_SEM_CHUNK ? 256u : (w.gy << 8);
  const uint32_t wmz = SEM == ARC_SEM_CHUNK ? 256u : (w.gz << 8);
  if (s.step >= max_steps || s.t >= s.t_limit || s.vi[0] >= wmx || s.vi[1] >= wmy || s.vi[2] >= wmz) {
    o.steps = s.step;
    return true;
  }
  const uint32_t step = s.step;
  s.step = step + 1u;
  const uint32_t diff = (s.prev[0] ^ s.vi[0]) | (s.prev[1] ^ s.vi[1]) | (s.prev[2] ^ s.vi[2]);
  s.prev[0] = s.vi[0];
  s.prev[1] = s.vi[1];
  s.prev[2] = s.vi[2];

  uint32_t span;  // edge of the box to step over when no hit
  uint32_t root_idx;
  uint32_t min_shift = 0u;
  bool hit = false;
  uint4 node = make_uint4(0, 0, 0, 0);
  uint32_t shift = 6u, b = 0u, mat = 0u;
  if (SEM == ARC_SEM_WORLD) {
    root_idx = s.mtf_r0;
    if (diff >= 256u) {  // traversal64.wgsl:316-335
      const uint32_t key = (s.vi[0] >> 8) + (s.vi[1] >> 8) * w.gx + (s.vi[2] >> 8) * w.gx * w.gy;
      s.cached_shift = 6u;
      if (key == s.mtf_k1) {
        root_idx = s.mtf_r1;
      } else {
        root_idx = w.roots[key];
      }
      s.mtf_k1 = s.mtf_k0;
      s.mtf_r1 = s.mtf_r0;
      s.mtf_k0 = key;
      s.mtf_r0 = root_idx;
    }
  } else {
    root_idx = s.root_cpu;
  }
  if (SEM == ARC_SEM_WORLD && root_idx == ARC_EMPTY_CHUNK_SENTINEL) {
    span = 256u;
  } else {
    if (SEM == ARC_SEM_WORLD) {  // :355-370
      if (p.lod_bias == 0u) {
        if (s.t > p.lod_t || step > p.lod_step) min_shift = 2u;
      } else if (p.lod_bias == 1u) {
        if (s.t > p.lod_t_far || step > p.lod_step) min_shift = 4u;
        else if (s.t > p.lod_t) min_shift = 2u;
      } else {
        if (s.t > p.lod_t || step > p.lod_step) min_shift = 2u;
      }
    }
    if (diff < 4u && min_shift == 0u && s.cached_shift == 0u) {
      node = s.l3;
      shift = 0u;
    } else if (diff < 16u && min_shift <= 2u && s.cached_shift <= 2u) {
      node = s.l2;
      shift = 2u;
    } else if (diff < 64u && min_shift <= 4u && s.cached_shift <= 4u) {
      node = s.l1;
      shift = 4u;
    } else {
      node = w.nodes[root_idx];
      shift = 6u;
    }
    s.cached_shift = shift;
    const uint32_t lx = s.vi[0] & 255u, ly = s.vi[1] & 255u, lz = s.vi[2] & 255u;
    for (;;) {  // :390-552, at most 4 levels
      const uint32_t cx = (lx >> shift) & 3u, cy = (ly >> shift) & 3u, cz = (lz >> shift) & 3u;
      b = (cz << 4) | (cy << 2) | cx;
      if (SEM == ARC_SEM_WORLD && p.ignore_water && (node.z & ARC_IS_UNIFORM_FLAG) != 0u &&
          ((node.w >> 24) & 0xFFu) == ARC_MAT_WATER) {
        span = 4u << shift;
        break;
      }
      if (!has_child(node.x, node.y, b)) {
        span = 1u << shift;
        break;
      }
      // Level-3 nodes are leaves in every pool the builder or the upload check accepts; the
      // shift test only bounds the descent if a patch ever slipped a malformed node in.
      const bool is_leaf = (node.z & ARC_IS_LEAF_FLAG) != 0u || shift == 0u;
      if (is_leaf || shift <= min_shift) {
        mat = (node.w >> 24) & 0xFFu;
        if (is_leaf && (node.z & ARC_IS_UNIFORM_FLAG) == 0u) {
          const uint32_t word = w.bricks[(size_t)(node.z & ARC_PTR_INDEX_MASK) * 16u + (b >> 2)];
          mat = (word >> ((b & 3u) * 8u)) & 0xFFu;
        }
        span = 1u << shift;
        if (SEM == ARC_SEM_WORLD && p.ignore_water && mat == ARC_MAT_WATER) break;
        if (mat != ARC_MAT_AIR) hit = true;
        break;
      }
      const uint32_t next = child_offset(node, b);
      node = w.nodes[next];
      shift -= 2u;
      s.cached_shift = shift;
      if (shift == 4u) s.l1 = node;
      else if (shift == 2u) s.l2 = node;
      else if (shift == 0u) s.l3 = node;
    }
  }

  const uint32_t bmin0 = s.vi[0] & ~(span - 1u), bmin1 = s.vi[1] & ~(span - 1u), bmin2 = s.vi[2] & ~(span - 1u);
  const uint32_t bmin[3] = {bmin0, bmin1, bmin2};
  if (hit) {  // :467-523 / raycast_cpu.rs:162-185
    float tm[3];
    for (int a = 0; a < 3; ++a) {
      const bool pos = s.inv[a] >= 0.0f;
      const uint32_t ec = pos ? bmin[a] : bmin[a] + span;
      float m = (float)ec * s.inv[a];
      tm[a] = m + s.noi[a];
    }
    const float cell_enter = fmaxf(fmaxf(tm[0], tm[1]), tm[2]);
    const float exact_t = fmaxf(cell_enter, 0.0f);
    const float max_t = cell_enter;
    const float eps = 1e-4f * fmaxf(max_t * 0.01f, 1.0f);
    float cand[3];
    for (int a = 0; a < 3; ++a) {
      float dd = tm[a] - max_t;
      cand[a] = fabsf(dd) <= eps ? fabsf(s.rdir[a]) : 0.0f;
    }
    uint32_t axis;
    if (cand[0] >= cand[1] && cand[0] >= cand[2] && cand[0] > 0.0f) axis = 0;
    else if (cand[1] >= cand[2] && cand[1] > 0.0f) axis = 1;
    else if (cand[2] > 0.0f) axis = 2;
    else if (tm[0] >= tm[1] && tm[0] >= tm[2]) axis = 0;
    else if (tm[1] >= tm[2]) axis = 1;
    else axis = 2;
    const float dax = axis == 0 ? s.rdir[0] : (axis == 1 ? s.rdir[1] : s.rdir[2]);
    o.hit = true;
    o.t_voxels = exact_t;
    o.vi[0] = s.vi[0];
    o.vi[1] = s.vi[1];
    o.vi[2] = s.vi[2];
    o.mat = mat;
    o.axis = axis;
    o.nrm_axis = SEM == ARC_SEM_WORLD ? -wgsl_sign(dax) : -rust_signum(dax);
    o.shift = shift;
    o.steps = s.step;
    if (FULL) {
      float hpv[3];
      for (int a = 0; a < 3; ++a) {
        float m = s.dir[a] * exact_t;
        hpv[a] = s.vo[a] + m;
        o.hp[a] = SEM == ARC_SEM_WORLD ? hpv[a] * 0.0625f : hpv[a];
      }
      if (SEM == ARC_SEM_WORLD) {
        o.lod_payload = node.w;
        o.contact_ao = 1.0f;
        if (shift == 0u && p.lod_bias == 0u) {  // :501-521
          const float hit_dist_m = s.t * 0.0625f;
          const float raw = p.ao_max_dist;
          const float ao_max = raw < 1.0f ? 24.0f : raw;
          if (hit_dist_m < ao_max) {
            float n3[3] = {0.0f, 0.0f, 0.0f};
            n3[axis] = o.nrm_axis;
            const uint32_t cx = s.vi[0] & 3u, cy = s.vi[1] & 3u, cz = s.vi[2] & 3u;
            const float raw_ao = eval_leaf_ao(node.x, node.y, cx, cy, cz, n3[0], n3[1], n3[2], hpv[0], hpv[1], hpv[2]);
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
            o.contact_ao = a0 + a1;
          }
        }
      }
    }
    return true;
  }

  // step_dda_box (traversal64.wgsl:152-179 / raycast_cpu.rs:245-292)
  float tmx[3];
  uint32_t exitc[3];
  for (int a = 0; a < 3; ++a) {
    const bool pos = s.inv[a] >= 0.0f;
    exitc[a] = pos ? bmin[a] + span : bmin[a];
    float m = (float)exitc[a] * s.inv[a];
    tmx[a] = m + s.noi[a];
  }
  const float t_exit = fminf(fminf(tmx[0], tmx[1]), tmx[2]);
  const float next_t = fmaxf(s.t, t_exit);
  const bool is_x = tmx[0] <= tmx[1] && tmx[0] <= tmx[2];
  const bool is_y = !is_x && (tmx[1] <= tmx[2]);
  const bool is_z = !is_x && !is_y;
  const bool is_exit[3] = {is_x, is_y, is_z};
  for (int a = 0; a < 3; ++a) {
    const bool pos = s.inv[a] >= 0.0f;
    const uint32_t outside = exitc[a] - (pos ? 0u : 1u);
    float m = s.dir[a] * next_t;
    const float ip = fmaxf(s.vo[a] + m, 0.0f);
    const uint32_t inside = umin_(umax_(f2u(ip), bmin[a]), bmin[a] + span - 1u);
    s.vi[a] = is_exit[a] ? outside : inside;
  }
  s.t = next_t;
  return false;
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

template <uint32_t SEM>
__device__ __forceinline__ void store_full(arc_trace_result* out, uint32_t idx, const Outcome& o) {
  float nrm[3];
  if (o.hit) {
    nrm[0] = nrm[1] = nrm[2] = 0.0f;
    nrm[o.axis] = o.nrm_axis;
  } else if (SEM == ARC_SEM_WORLD) {
    nrm[0] = 0.0f;
    nrm[1] = 1.0f;
    nrm[2] = 0.0f;
  } else {
    nrm[0] = nrm[1] = nrm[2] = 0.0f;
  }
  float rgb[3] = {0.0f, 0.0f, 0.0f};
  uint32_t vh = 0u;
  if (o.hit && SEM == ARC_SEM_WORLD) {
    rgb[0] = (float)(o.lod_payload & 0xFFu) / 255.0f;
    rgb[1] = (float)((o.lod_payload >> 8) & 0xFFu) / 255.0f;
    rgb[2] = (float)((o.lod_payload >> 16) & 0xFFu) / 255.0f;
    vh = pcg_hash((o.vi[0] * 73856093u) ^ (o.vi[1] * 19349663u) ^ (o.vi[2] * 83492791u));
  }
  uint4* q = reinterpret_cast<uint4*>(out + idx);
  q[0] = make_uint4(__float_as_uint(o.t_voxels), __float_as_uint(o.hp[0]), __float_as_uint(o.hp[1]),
                    __float_as_uint(o.hp[2]));
  q[1] = make_uint4(__float_as_uint(nrm[0]), __float_as_uint(nrm[1]), __float_as_uint(nrm[2]), __float_as_uint(rgb[0]));
  q[2] = make_uint4(__float_as_uint(rgb[1]), __float_as_uint(rgb[2]), __float_as_uint(o.contact_ao), o.mat);
  const uint32_t hitw = o.hit ? (1u | (o.axis << 8) | (o.shift << 16)) : 0u;
  q[3] = make_uint4(vh, o.steps, hitw, o.hit ? o.vi[0] : 0u);
  q[4] = make_uint4(o.hit ? o.vi[1] : 0u, o.hit ? o.vi[2] : 0u, 0u, 0u);
}

// Persistent ray-queue kernel.
template <uint32_t SEM, bool FULL>
__global__ __launch_bounds__(256) void k_trace(DevWorld w, DevParams p, const arc_ray* __restrict__ rays,
                                               typename OutPtr<FULL>::T* __restrict__ out, uint32_t n,
                                               uint32_t* __restrict__ queue, uint32_t refill, uint32_t batch) {
  (void)batch;
  const uint32_t lane = __lane_id();
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
  RayState s;
  Outcome o;
  uint32_t idx = 0;
  bool live = false;
  bool drained = false;
  for (;;) {
    const uint64_t idle = __ballot(!live);
    const uint32_t nidle = __popcll(idle);
    if (!drained && (nidle >= refill || nidle == 64u)) {
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(queue, nidle);
      base = __shfl(base, 0);
      if (base + nidle >= n) drained = true;
      if (!live) {
        const uint32_t my = base + (uint32_t)__popcll(idle & lt_mask);
        if (my < n) {
          idx = my;
          const arc_ray r = rays[my];
          live = Sem<SEM>::init(w, r, s, o);
          if (!live) {
            if (FULL) store_full<SEM>(reinterpret_cast<arc_trace_result*>(out), my, o);
            else store_compact(reinterpret_cast<arc_hit*>(out), my, o);
          }
        }
      }
    }
    if (drained && __ballot(live) == 0ull) break;
    if (live) {
      if (dda_iteration<SEM, FULL>(w, p, s, o)) {
        live = false;
        if (FULL) store_full<SEM>(reinterpret_cast<arc_trace_result*>(out), idx, o);
        else store_compact(reinterpret_cast<arc_hit*>(out), idx, o);
      }
    }
  }
  // Last workgroup out resets the queue for the next launch on this stream.
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const uint32_t done = atomicAdd(queue + 1, 1u);
    if (done == gridDim.x - 1u) {
      atomicExch(queue, 0u);
      atomicExch(queue + 1, 0u);
      __threadfence();
    }
  }
}

}  // namespace arcdev

// ---------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------
namespace arcdev {

hipError_t launch_trace(const LaunchArgs& a, hipStream_t stream) {
  DevWorld w;
  w.gx = a.grid[0];
  w.gy = a.grid[1];
  w.gz = a.grid[2];
  w.pad = 0;
  w.roots = a.roots;
  w.nodes = reinterpret_cast<const uint4*>(a.nodes);
  w.bricks = reinterpret_cast<const uint32_t*>(a.bricks);
  DevParams p;
  p.lod_bias = a.params.lod_bias;
  p.ignore_water = a.params.ignore_water;
  const uint32_t cls = a.params.lod_bias == 0u ? 0u : (a.params.lod_bias == 1u ? 1u : 2u);
  p.max_steps = a.params.max_steps[cls];
  p.lod_step = a.params.lod_step[cls];
  p.lod_t = a.params.lod_dist_m[cls] * 16.0f;
  p.lod_t_far = (a.params.lod_dist_m[1] * 16.0f) * 2.2857f;
  p.ao_max_dist = a.params.ao_max_dist_bits;
  p.pad = 0;
  const dim3 grid(a.blocks), block(a.block_threads);
  const uint32_t n = (uint32_t)a.n;
  if (a.params.semantics == ARC_SEM_CHUNK) {
    if (a.full)
      hipLaunchKernelGGL((k_trace<ARC_SEM_CHUNK, true>), grid, block, 0, stream, w, p, a.rays,
                         reinterpret_cast<arc_trace_result*>(a.out), n, a.queue, a.refill, a.batch);
    else
      hipLaunchKernelGGL((k_trace<ARC_SEM_CHUNK, false>), grid, block, 0, stream, w, p, a.rays,
                         reinterpret_cast<arc_hit*>(a.out), n, a.queue, a.refill, a.batch);
  } else {
    if (a.full)
      hipLaunchKernelGGL((k_trace<ARC_SEM_WORLD, true>), grid, block, 0, stream, w, p, a.rays,
                         reinterpret_cast<arc_trace_result*>(a.out), n, a.queue, a.refill, a.batch);
    else
      hipLaunchKernelGGL((k_trace<ARC_SEM_WORLD, false>), grid, block, 0, stream, w, p, a.rays,
                         reinterpret_cast<arc_hit*>(a.out), n, a.queue, a.refill, a.batch);
  }
  return hipGetLastError();
}

}  // namespace arcdev
