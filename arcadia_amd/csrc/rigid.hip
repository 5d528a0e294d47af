// Rigid voxel bodies and point queries for gfx950: trace_rigid_bodies / trace_single_rigid_body
// (shaders/common/traversal64.wgsl:1005-1244) and query_voxel_world_meters (:182-220).
//
// One thread per ray: a ray visits at most 32 bodies, each behind a bounding-sphere test, and a
// body's walk is at most 64 steps of the exact 64-tree DDA in the body's local frame (no LOD, no
// water rules), so rays are short and a plain grid beats the persistent queue. The instances
// are staged in LDS (7 KB) and read by every thread.
//
// Arithmetic follows the WGSL text in binary32 with no contraction. Where WGSL leaves the order
// open: matrix-vector products are ((c0 x + c1 y) + c2 z) + c3 w per component, dot products
// x, y, z left to right, normalize(v) = v / sqrt(dot(v, v)), sign(0) = 0.
// Body trees are not reachable from the chunk roots that the upload check walks, so every node
// and brick index is checked against the pool sizes; a tree that leaves the pools traces as empty.
#include <hip/hip_runtime.h>

#include "../../include/arcadia_hip.h"
#include "trace_device.h"
#include "trace_launch.h"

namespace arcdev {

struct RigidWorld {
  const uint4* __restrict__ nodes;
  const uint32_t* __restrict__ bricks;
  uint64_t n_nodes, n_bricks;
};

__device__ __forceinline__ void mat_vec(const float* m, float x, float y, float z, float wc, float out[3]) {
  for (int r = 0; r < 3; ++r) {
    const float a = m[0 + r] * x;
    const float b = m[4 + r] * y;
    const float c = m[8 + r] * z;
    const float d = m[12 + r] * wc;
    out[r] = ((a + b) + c) + d;
  }
}

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
  const float x = ax * bx, y = ay * by, z = az * bz;
  return (x + y) + z;
}

// RigidTraceResult of a miss (:1011-1019, :1210-1218).
__device__ __forceinline__ void rigid_miss(const float o[3], const float d[3], float max_dist_m, arc_rigid_hit& r) {
  r.t_voxels = max_dist_m * 16.0f;
  for (int a = 0; a < 3; ++a) {
    const float q = d[a] * max_dist_m;
    r.hit_pos_m[a] = o[a] + q;
    r.prev_hit_pos_m[a] = r.hit_pos_m[a];
    r.normal[a] = a == 1 ? 1.0f : 0.0f;
  }
  r.contact_ao = 1.0f;
  r.mat = ARC_MAT_AIR;
  r.voxel_hash = 0u;
  r.hit = 0u;
  r.body = 0xFFFFFFFFu;
  r.reserved = 0u;
}

// Local frame of a body (:1047-1054): ray origin in local voxels and local direction.
__device__ __forceinline__ void body_frame(const arc_rigid_body& rb, const float o[3], const float d[3], float vo[3],
                                           float ld[3]) {
  float lo[3];
  mat_vec(rb.world_to_local, o[0], o[1], o[2], 1.0f, lo);
  mat_vec(rb.world_to_local, d[0], d[1], d[2], 0.0f, ld);
  for (int a = 0; a < 3; ++a) vo[a] = lo[a] * 16.0f;
}

// What the body loop keeps of a hit; the world-space fields of the record are rebuilt from it
// for the nearest body only (rigid_record), with the same arithmetic as the walk.
struct RigidCand {
  float t;          // exact_t, local voxels
  float ao;         // contact_ao
  uint32_t mat;
  uint32_t hash;
  uint32_t axis;    // axis of the local face normal -sign(ld[axis])
};

// trace_single_rigid_body (:1005-1202) up to the hit: true and `c` filled when the body is hit.
__device__ __forceinline__ bool rigid_single(const RigidWorld& w, const arc_rigid_body& rb, const float o[3],
                                             const float d[3], float max_dist_m, RigidCand& c) {
  const uint32_t root_idx = __float_as_uint(rb.local_bounds_and_root[3]);
  if (root_idx == ARC_EMPTY_CHUNK_SENTINEL || root_idx >= w.n_nodes) return false;
  // 1. sphere pre-reject (:1028-1045)
  const float ocx = o[0] - rb.world_sphere[0], ocy = o[1] - rb.world_sphere[1], ocz = o[2] - rb.world_sphere[2];
  const float b_proj = dot3(ocx, ocy, ocz, d[0], d[1], d[2]);
  const float r2 = rb.world_sphere[3] * rb.world_sphere[3];
  const float c_dist = dot3(ocx, ocy, ocz, ocx, ocy, ocz) - r2;
  if (c_dist > 0.0f && b_proj > 0.0f) return false;
  const float disc = b_proj * b_proj - c_dist;
  if (disc < 0.0f) return false;
  const float near_m = fmaxf(-b_proj - sqrtf(disc), 0.0f);
  if (near_m >= max_dist_m) return false;
  // 2. local frame (:1047-1054)
  float vo[3], ld[3];
  body_frame(rb, o, d, vo, ld);
  float lmax[3], inv[3], noi[3];
  bool pos[3];
  uint32_t lmaxi[3];
  for (int a = 0; a < 3; ++a) {
    lmax[a] = rb.local_bounds_and_root[a] * 16.0f;
    const float safe = fabsf(ld[a]) < 1e-7f ? 1e-7f : ld[a];
    inv[a] = 1.0f / safe;
  }
  const float max_t_v = max_dist_m * 16.0f;
  float tn[3], tx[3];
  for (int a = 0; a < 3; ++a) {  // intersect_aabb :134-142
    const float eb = inv[a] > 0.0f ? 0.0f : lmax[a];
    const float xb = inv[a] > 0.0f ? lmax[a] : 0.0f;
    tn[a] = (eb - vo[a]) * inv[a];
    tx[a] = (xb - vo[a]) * inv[a];
  }
  const float bx = fmaxf(fmaxf(tn[0], tn[1]), tn[2]);
  const float by = fminf(fminf(tx[0], tx[1]), tx[2]);
  const float t_enter = fmaxf(bx, 0.0f);
  if (by < t_enter || t_enter >= max_t_v) return false;
  float t = t_enter;
  const float t_limit = fminf(by, max_t_v);
  uint32_t vi[3];
  for (int a = 0; a < 3; ++a) {
    pos[a] = inv[a] >= 0.0f;
    noi[a] = -vo[a] * inv[a];
    lmaxi[a] = f2u(lmax[a] + 0.5f);
    const float q = ld[a] * (t + 1e-4f);
    vi[a] = f2u(clampf_(vo[a] + q, 0.0f, lmax[a] - 1e-3f));
  }
  uint32_t prev[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
  uint32_t cached_shift = 6u;
  uint4 l1 = make_uint4(0, 0, 0, 0), l2 = l1, l3 = l1;
  for (uint32_t step = 0; step < 64u; ++step) {  // :1077-1197
    if (t >= t_limit) break;
    if (vi[0] >= lmaxi[0] || vi[1] >= lmaxi[1] || vi[2] >= lmaxi[2]) break;
    const uint32_t diff = (prev[0] ^ vi[0]) | (prev[1] ^ vi[1]) | (prev[2] ^ vi[2]);
    prev[0] = vi[0];
    prev[1] = vi[1];
    prev[2] = vi[2];
    uint4 node;
    uint32_t shift;
    if (diff < 4u && cached_shift == 0u) {
      node = l3;
      shift = 0u;
    } else if (diff < 16u && cached_shift <= 2u) {
      node = l2;
      shift = 2u;
    } else if (diff < 64u && cached_shift <= 4u) {
      node = l1;
      shift = 4u;
    } else {
      node = w.nodes[root_idx];
      shift = 6u;
    }
    cached_shift = shift;
    const uint32_t lx = vi[0] & 255u, ly = vi[1] & 255u, lz = vi[2] & 255u;
    for (;;) {
      const uint32_t cx = (lx >> shift) & 3u, cy = (ly >> shift) & 3u, cz = (lz >> shift) & 3u;
      const uint32_t b = (cz << 4) | (cy << 2) | cx;
      const uint32_t span = 1u << shift;
      bool step_box = false;
      if (!has_child(node.x, node.y, b)) {
        step_box = true;
      } else if ((node.z & ARC_IS_LEAF_FLAG) != 0u) {
        uint32_t mat = (node.w >> 24) & 0xFFu;
        if ((node.z & ARC_IS_UNIFORM_FLAG) == 0u) {
          const uint32_t bi = node.z & ARC_PTR_INDEX_MASK;
          mat = bi < w.n_bricks ? (w.bricks[(size_t)bi * 16u + (b >> 2)] >> ((b & 3u) * 8u)) & 0xFFu : ARC_MAT_AIR;
        }
        if (mat != ARC_MAT_AIR) {  // :1143-1182
          float tm[3];
          for (int a = 0; a < 3; ++a) {
            const uint32_t cmin = vi[a] & ~(span - 1u);
            const uint32_t ec = pos[a] ? cmin : cmin + span;
            const float m = (float)ec * inv[a];
            tm[a] = m + noi[a];
          }
          const float cell_enter = fmaxf(fmaxf(tm[0], tm[1]), tm[2]);
          const float exact_t = fmaxf(cell_enter, 0.0f);
          const uint32_t axis = (tm[0] >= tm[1] && tm[0] >= tm[2]) ? 0u : (tm[1] >= tm[2] ? 1u : 2u);
          c.t = exact_t;
          c.mat = mat;
          c.axis = axis;
          c.hash = pcg_hash((vi[0] * 73856093u) ^ (vi[1] * 19349663u) ^ (vi[2] * 83492791u));
          c.ao = 1.0f;
          if (shift == 0u) {
            float hv[3];
            for (int a = 0; a < 3; ++a) {
              const float q = ld[a] * exact_t;
              hv[a] = vo[a] + q;
            }
            const float ln0 = axis == 0u ? -wgsl_sign(ld[0]) : 0.0f;
            const float ln1 = axis == 1u ? -wgsl_sign(ld[1]) : 0.0f;
            const float ln2 = axis == 2u ? -wgsl_sign(ld[2]) : 0.0f;
            c.ao = eval_leaf_ao(node.x, node.y, cx, cy, cz, ln0, ln1, ln2, hv[0], hv[1], hv[2]);
          }
          return true;
        }
        step_box = true;
      }
      if (step_box) {  // step_dda_box (:152-179) over the cell
        float te[3];
        uint32_t bmin[3], ex[3];
        for (int a = 0; a < 3; ++a) {
          bmin[a] = vi[a] & ~(span - 1u);
          ex[a] = pos[a] ? bmin[a] + span : bmin[a];
          const float m = (float)ex[a] * inv[a];
          te[a] = m + noi[a];
        }
        const float t_exit = fminf(fminf(te[0], te[1]), te[2]);
        const float next_t = fmaxf(t, t_exit);
        const bool is_x = te[0] <= te[1] && te[0] <= te[2];
        const bool is_y = !is_x && te[1] <= te[2];
        const bool is_exit[3] = {is_x, is_y, !is_x && !is_y};
        for (int a = 0; a < 3; ++a) {
          const uint32_t outside = ex[a] - (pos[a] ? 0u : 1u);
          const float q = ld[a] * next_t;
          const float ip = fmaxf(vo[a] + q, 0.0f);
          const uint32_t cc = f2u(ip);
          const uint32_t hi = bmin[a] + span - 1u;
          const uint32_t inside = cc < bmin[a] ? bmin[a] : (cc > hi ? hi : cc);
          vi[a] = is_exit[a] ? outside : inside;
        }
        t = next_t;
        break;
      }
      const uint32_t next = child_offset(node, b);
      if (next >= w.n_nodes) return false;  // tree leaves the pool: traced as empty
      node = w.nodes[next];
      shift -= 2u;
      cached_shift = shift;
      if (shift == 4u) l1 = node;
      else if (shift == 2u) l2 = node;
      else if (shift == 0u) l3 = node;
    }
  }
  return false;
}

// The RigidTraceResult of a hit (:1151-1181) from the nearest body's candidate: the local hit
// point, its transforms and the world normal, computed as the walk of that body computes them.
__device__ __forceinline__ void rigid_record(const arc_rigid_body& rb, uint32_t rb_idx, const float o[3],
                                             const float d[3], const RigidCand& c, arc_rigid_hit& res) {
  float vo[3], ld[3];
  body_frame(rb, o, d, vo, ld);
  float hm[3];
  for (int a = 0; a < 3; ++a) {
    const float q = ld[a] * c.t;
    hm[a] = (vo[a] + q) * 0.0625f;
  }
  res.hit = 1u;
  res.t_voxels = c.t;
  mat_vec(rb.local_to_world, hm[0], hm[1], hm[2], 1.0f, res.hit_pos_m);
  mat_vec(rb.prev_local_to_world, hm[0], hm[1], hm[2], 1.0f, res.prev_hit_pos_m);
  const float ln[3] = {c.axis == 0u ? -wgsl_sign(ld[0]) : 0.0f, c.axis == 1u ? -wgsl_sign(ld[1]) : 0.0f,
                       c.axis == 2u ? -wgsl_sign(ld[2]) : 0.0f};
  float wn[3];
  mat_vec(rb.local_to_world, ln[0], ln[1], ln[2], 0.0f, wn);
  const float len = sqrtf(dot3(wn[0], wn[1], wn[2], wn[0], wn[1], wn[2]));
  for (int a = 0; a < 3; ++a) res.normal[a] = wn[a] / len;
  res.mat = c.mat;
  res.voxel_hash = c.hash;
  res.contact_ao = c.ao;
  res.body = rb_idx;
}

__global__ __launch_bounds__(256) void k_rigid(RigidWorld w, const arc_rigid_body* __restrict__ bodies, uint32_t n_rb,
                                               const arc_ray* __restrict__ rays, uint32_t n,
                                               arc_rigid_hit* __restrict__ out) {
  __shared__ arc_rigid_body s_rb[ARC_MAX_RIGID_BODIES];
  for (uint32_t i = threadIdx.x; i < n_rb * (uint32_t)(sizeof(arc_rigid_body) / 4); i += blockDim.x)
    reinterpret_cast<float*>(s_rb)[i] = reinterpret_cast<const float*>(bodies)[i];
  __syncthreads();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const arc_ray r = rays[i];
  const float o[3] = {r.ox, r.oy, r.oz}, d[3] = {r.dx, r.dy, r.dz};
  // trace_rigid_bodies (:1204-1244): the nearest hit over the bodies, each walk cut at the
  // nearest hit so far
  RigidCand best;
  best.t = r.t_max * 16.0f;
  uint32_t best_b = 0xFFFFFFFFu;
  float cur_max_m = r.t_max;
  for (uint32_t b = 0; b < n_rb; ++b) {
    RigidCand c;
    if (rigid_single(w, s_rb[b], o, d, cur_max_m, c) && c.t < best.t) {
      best = c;
      best_b = b;
      cur_max_m = c.t * 0.0625f;
    }
  }
  arc_rigid_hit res;
  rigid_miss(o, d, r.t_max, res);
  if (best_b != 0xFFFFFFFFu) rigid_record(s_rb[best_b], best_b, o, d, best, res);
  out[i] = res;
}

__global__ __launch_bounds__(256) void k_query_voxels(DevWorld w, uint64_t n_nodes, uint64_t n_bricks,
                                                      const float* __restrict__ pts, uint32_t n,
                                                      uint32_t* __restrict__ mat) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t vi[3];
  for (int a = 0; a < 3; ++a) {
    const float vp = pts[3 * (size_t)i + a] * 16.0f;
    if (vp < 0.0f || vp >= w.ext[a]) {  // NaN passes both tests, as in the WGSL, and becomes voxel 0
      mat[i] = ARC_MAT_AIR;
      return;
    }
    vi[a] = f2u(vp);
  }
  uint32_t out = ARC_MAT_AIR;
  const uint32_t key = (vi[0] >> 8) + (vi[1] >> 8) * w.gx + (vi[2] >> 8) * w.gx * w.gy;
  uint32_t node_idx = w.roots[key];
  if (node_idx != ARC_EMPTY_CHUNK_SENTINEL) {
    uint32_t shift = 6u;
    for (uint32_t level = 0; level < 4u && node_idx < n_nodes; ++level) {  // :200-219
      const uint4 node = w.nodes[node_idx];
      const uint32_t b = ((((vi[2] & 255u) >> shift) & 3u) << 4) | ((((vi[1] & 255u) >> shift) & 3u) << 2) |
                         (((vi[0] & 255u) >> shift) & 3u);
      if (!has_child(node.x, node.y, b)) break;
      if ((node.z & ARC_IS_LEAF_FLAG) != 0u) {
        if ((node.z & ARC_IS_UNIFORM_FLAG) != 0u) {
          out = (node.w >> 24) & 0xFFu;
        } else {
          const uint32_t bi = node.z & ARC_PTR_INDEX_MASK;
          if (bi < n_bricks) out = (w.bricks[(size_t)bi * 16u + (b >> 2)] >> ((b & 3u) * 8u)) & 0xFFu;
        }
        break;
      }
      node_idx = child_offset(node, b);
      shift -= 2u;
    }
  }
  mat[i] = out;
}

hipError_t launch_rigid(const RigidLaunchArgs& a, hipStream_t stream) {
  RigidWorld w;
  w.nodes = reinterpret_cast<const uint4*>(a.nodes);
  w.bricks = reinterpret_cast<const uint32_t*>(a.bricks);
  w.n_nodes = a.n_nodes;
  w.n_bricks = a.n_bricks;
  const uint32_t n = (uint32_t)a.n;
  hipLaunchKernelGGL(k_rigid, dim3((n + 255) / 256), dim3(256), 0, stream, w, a.bodies, a.n_bodies, a.rays, n, a.out);
  return hipGetLastError();
}

hipError_t launch_query(const QueryLaunchArgs& a, hipStream_t stream) {
  DevWorld w{};
  w.gx = a.grid[0];
  w.gy = a.grid[1];
  w.gz = a.grid[2];
  for (int k = 0; k < 3; ++k) w.ext[k] = (float)a.grid[k] * 256.0f;
  w.roots = a.roots;
  w.nodes = reinterpret_cast<const uint4*>(a.nodes);
  w.bricks = reinterpret_cast<const uint32_t*>(a.bricks);
  const uint32_t n = (uint32_t)a.n;
  hipLaunchKernelGGL(k_query_voxels, dim3((n + 255) / 256), dim3(256), 0, stream, w, a.n_nodes, a.n_bricks, a.points,
                     n, a.mat);
  return hipGetLastError();
}

}  // namespace arcdev
