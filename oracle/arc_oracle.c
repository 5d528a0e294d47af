/*
 * CPU oracle for the 64-tree DDA traversal — TEST INFRASTRUCTURE ONLY (see arc_oracle.h).
 *
 * Build: gcc -O3 -ffp-contract=off -fno-fast-math (oracle/Makefile). Every float operation
 * below is one IEEE-754 binary32 operation in source order, with no contraction, which is
 * how Rust evaluates raycast_cpu.rs and how the WGSL source reads once FMA fusion is ruled
 * out. min/max return the non-NaN operand (fminf/fmaxf), float->u32 saturates with NaN->0
 * (Rust `as`, WGSL u32()).
 */
#include "arc_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---- helpers ---------------------------------------------------------------------------- */

static inline uint32_t f2u(float x) {
  if (!(x > 0.0f)) return 0u; /* NaN and <= 0 */
  if (x >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

static inline float wgsl_sign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

/* Rust f32::signum: +1 for +0.0 and positive, -1 for -0.0 and negative (NaN -> NaN). */
static inline float rust_signum(float x) {
  if (x != x) return x;
  return signbit(x) ? -1.0f : 1.0f;
}

static inline float clampf_(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

static inline uint32_t clampu_(uint32_t v, uint32_t lo, uint32_t hi) {
  uint32_t a = v > lo ? v : lo;
  return a < hi ? a : hi;
}

/* popcount child slot: node64.rs:86-90; traversal64.wgsl:32-40 */
static inline uint32_t child_offset(const arc_node64* n, uint32_t b) {
  uint64_t mask = (uint64_t)n->mask_low | ((uint64_t)n->mask_high << 32);
  uint64_t prefix = (b >= 64) ? ~0ull : ((1ull << b) - 1ull);
  return (n->child_ptr & ARC_PTR_INDEX_MASK) + (uint32_t)__builtin_popcountll(mask & prefix);
}

/* traversal64.wgsl:25-29 */
static inline int has_child(const arc_node64* n, uint32_t b) {
  uint32_t word = b < 32u ? n->mask_low : n->mask_high;
  return ((word >> (b & 31u)) & 1u) != 0u;
}

/* traversal64.wgsl:43-47 */
static inline uint32_t brick_mat(const orc_world* w, uint32_t brick_idx, uint32_t b) {
  uint32_t word = w->bricks[brick_idx].packed_materials[b >> 2];
  return (word >> ((b & 3u) * 8u)) & 0xFFu;
}

/* types.wgsl:258-262 */
static inline uint32_t pcg_hash(uint32_t input) {
  uint32_t state = input * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

/* Normal code of the compact hit record. */
static inline uint8_t face_code(const float nrm[3]) {
  for (int a = 0; a < 3; ++a) {
    if (nrm[a] > 0.0f) return (uint8_t)(2 * a);
    if (nrm[a] < 0.0f) return (uint8_t)(2 * a + 1);
  }
  return 255u;
}

typedef struct step_res {
  float t;
  uint32_t vi[3];
} step_res;

/* step_dda_box: traversal64.wgsl:152-179 (GPU, `dir` for the inside point) and
 * raycast_cpu.rs:245-292 (CPU, `safe_dir`); the caller passes the direction to use. */
static inline step_res step_dda_box(const float vo[3], const float dir[3], const float inv[3],
                                    const float noi[3], const int pos[3], const uint32_t bmin[3],
                                    const uint32_t bmax[3], float cur_t) {
  uint32_t exitc[3];
  float tm[3];
  for (int a = 0; a < 3; ++a) {
    exitc[a] = pos[a] ? bmax[a] : bmin[a];
    float m = (float)exitc[a] * inv[a];
    tm[a] = m + noi[a];
  }
  float t_exit = fminf(fminf(tm[0], tm[1]), tm[2]);
  float next_t = fmaxf(cur_t, t_exit);
  int is_x = tm[0] <= tm[1] && tm[0] <= tm[2];
  int is_y = !is_x && (tm[1] <= tm[2]);
  int is_z = !is_x && !is_y;
  int is_exit[3] = {is_x, is_y, is_z};
  step_res r;
  r.t = next_t;
  for (int a = 0; a < 3; ++a) {
    uint32_t outside = exitc[a] - (pos[a] ? 0u : 1u); /* wrapping */
    float dn = dir[a] * next_t;
    float ip = fmaxf(vo[a] + dn, 0.0f);
    uint32_t hi = bmax[a] - 1u; /* box_max - 1 (box_max >= 1 always) */
    uint32_t inside = clampu_(f2u(ip), bmin[a], hi);
    r.vi[a] = is_exit[a] ? outside : inside;
  }
  return r;
}

/* eval_leaf_bitwise_ao: traversal64.wgsl:60-131 */
static float eval_leaf_ao(uint32_t ml, uint32_t mh, uint32_t cx, uint32_t cy, uint32_t cz,
                          const float nrm[3], const float hp[3]) {
  arc_node64 n = {ml, mh, 0, 0};
  float fx = clampf_(hp[0] - floorf(hp[0]), 0.0f, 1.0f);
  float fy = clampf_(hp[1] - floorf(hp[1]), 0.0f, 1.0f);
  float fz = clampf_(hp[2] - floorf(hp[2]), 0.0f, 1.0f);
  int nx = (int)cx + (int)nrm[0];
  int ny = (int)cy + (int)nrm[1];
  int nz = (int)cz + (int)nrm[2];
  if (nx < 0 || nx > 3 || ny < 0 || ny > 3 || nz < 0 || nz > 3) return 1.0f;
  uint32_t bb = ((uint32_t)nz << 4) | ((uint32_t)ny << 2) | (uint32_t)nx;
  float occ = 0.0f;
  if (fabsf(nrm[1]) > 0.5f) {
    if (nx > 0 && has_child(&n, bb - 1u)) occ += (1.0f - fx) * 0.42f;
    if (nx < 3 && has_child(&n, bb + 1u)) occ += fx * 0.42f;
    if (nz > 0 && has_child(&n, bb - 16u)) occ += (1.0f - fz) * 0.42f;
    if (nz < 3 && has_child(&n, bb + 16u)) occ += fz * 0.42f;
  } else if (fabsf(nrm[0]) > 0.5f) {
    if (ny > 0 && has_child(&n, bb - 4u)) occ += (1.0f - fy) * 0.42f;
    if (ny < 3 && has_child(&n, bb + 4u)) occ += fy * 0.42f;
    if (nz > 0 && has_child(&n, bb - 16u)) occ += (1.0f - fz) * 0.42f;
    if (nz < 3 && has_child(&n, bb + 16u)) occ += fz * 0.42f;
  } else {
    if (nx > 0 && has_child(&n, bb - 1u)) occ += (1.0f - fx) * 0.42f;
    if (nx < 3 && has_child(&n, bb + 1u)) occ += fx * 0.42f;
    if (ny > 0 && has_child(&n, bb - 4u)) occ += (1.0f - fy) * 0.42f;
    if (ny < 3 && has_child(&n, bb + 4u)) occ += fy * 0.42f;
  }
  return clampf_(1.0f - occ, 0.35f, 1.0f);
}

/* smoothstep / mix as SPIR-V GLSL.std.450 define them: t*t*(3-2t), x*(1-a)+y*a. */
static inline float smoothstep_(float e0, float e1, float x) {
  float t = clampf_((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
static inline float mix_(float x, float y, float a) { return x * (1.0f - a) + y * a; }

static void result_init_world(arc_trace_result* r, const float o[3], const float d[3],
                              float max_dist_m) {
  memset(r, 0, sizeof(*r));
  r->t_voxels = max_dist_m * 16.0f;
  for (int a = 0; a < 3; ++a) r->hit_pos_m[a] = o[a] + d[a] * max_dist_m;
  r->normal[0] = 0.0f;
  r->normal[1] = 1.0f;
  r->normal[2] = 0.0f;
  r->contact_ao = 1.0f;
}

/* ---- trace_world_64tree_from: traversal64.wgsl:226-556 ---------------------------------- */
static void trace_world_one(const orc_world* w, const arc_trace_params* p, const arc_ray* ray,
                            arc_trace_result* res, orc_counters* cnt) {
  const float o[3] = {ray->ox, ray->oy, ray->oz};
  const float dir[3] = {ray->dx, ray->dy, ray->dz};
  const float start_t_m = ray->t_start, max_dist_m = ray->t_max;
  orc_counters c;
  memset(&c, 0, sizeof c);
  result_init_world(res, o, dir, max_dist_m); /* :234-242 */

  float vo[3], safe[3], inv[3], wmax[3], noi[3];
  int pos[3];
  uint32_t wmaxi[3];
  for (int a = 0; a < 3; ++a) vo[a] = o[a] * 16.0f; /* :244 */
  const float max_t_v = max_dist_m * 16.0f;         /* :245 */
  const float start_t_v = fmaxf(start_t_m * 16.0f, 0.0f);
  for (int a = 0; a < 3; ++a) {
    safe[a] = fabsf(dir[a]) < 1e-7f ? 1e-7f : dir[a]; /* :248 */
    inv[a] = 1.0f / safe[a];
    wmax[a] = (float)w->grid[a] * 256.0f;
  }
  /* intersect_aabb :134-142 */
  float tminv[3], tmaxv[3];
  for (int a = 0; a < 3; ++a) {
    float eb = inv[a] > 0.0f ? 0.0f : wmax[a];
    float xb = inv[a] > 0.0f ? wmax[a] : 0.0f;
    tminv[a] = (eb - vo[a]) * inv[a];
    tmaxv[a] = (xb - vo[a]) * inv[a];
  }
  float wh_x = fmaxf(fmaxf(tminv[0], tminv[1]), tminv[2]);
  float wh_y = fminf(fminf(tmaxv[0], tmaxv[1]), tmaxv[2]);
  float t_enter = fmaxf(wh_x, start_t_v); /* :255 */
  if (wh_y < t_enter || t_enter >= max_t_v) {
    if (cnt) *cnt = c;
    return;
  }
  float t = t_enter;
  const float t_limit = fminf(wh_y, max_t_v);
  uint32_t max_steps = p->max_steps[0];
  if (p->lod_bias == 1u) max_steps = p->max_steps[1];
  else if (p->lod_bias >= 2u) max_steps = p->max_steps[2];
  const float prim_lod_t = p->lod_dist_m[0] * 16.0f;
  const uint32_t prim_lod_step = p->lod_step[0];
  const float sec_lod_t = p->lod_dist_m[1] * 16.0f;
  const uint32_t sec_lod_step = p->lod_step[1];
  const float gi_lod_t = p->lod_dist_m[2] * 16.0f;
  const uint32_t gi_lod_step = p->lod_step[2];
  for (int a = 0; a < 3; ++a) {
    pos[a] = inv[a] >= 0.0f;        /* :277 */
    noi[a] = -vo[a] * inv[a];       /* :278 */
    wmaxi[a] = w->grid[a] << 8;     /* :279 */
  }
  uint32_t vi[3];
  for (int a = 0; a < 3; ++a) { /* :281-282 */
    float sp = clampf_(vo[a] + dir[a] * (t + 1e-4f), 0.0f, wmax[a] - 1e-3f);
    vi[a] = f2u(sp);
  }
  uint32_t mtf_c0[3] = {~0u, ~0u, ~0u}, mtf_c1[3] = {~0u, ~0u, ~0u};
  uint32_t mtf_root0 = ARC_EMPTY_CHUNK_SENTINEL, mtf_root1 = ARC_EMPTY_CHUNK_SENTINEL;
  uint32_t prev[3] = {~0u, ~0u, ~0u};
  uint32_t cached_shift = 6u;
  arc_node64 l1 = {0, 0, 0, 0}, l2 = {0, 0, 0, 0}, l3 = {0, 0, 0, 0};
  const uint32_t gx = w->grid[0], gy = w->grid[1];

  for (uint32_t step = 0; step < max_steps; ++step) { /* :302 */
    if (t >= t_limit) break;
    if (vi[0] >= wmaxi[0] || vi[1] >= wmaxi[1] || vi[2] >= wmaxi[2]) {
      t = t_limit;
      break;
    }
    c.steps++;
    uint32_t diff = (prev[0] ^ vi[0]) | (prev[1] ^ vi[1]) | (prev[2] ^ vi[2]);
    prev[0] = vi[0];
    prev[1] = vi[1];
    prev[2] = vi[2];
    uint32_t root_idx = mtf_root0;
    if (diff >= 256u) { /* :316-335 */
      uint32_t cc[3] = {vi[0] >> 8, vi[1] >> 8, vi[2] >> 8};
      cached_shift = 6u;
      if (cc[0] == mtf_c1[0] && cc[1] == mtf_c1[1] && cc[2] == mtf_c1[2]) {
        root_idx = mtf_root1;
        memcpy(mtf_c1, mtf_c0, sizeof mtf_c1);
        mtf_root1 = mtf_root0;
        memcpy(mtf_c0, cc, sizeof cc);
        mtf_root0 = root_idx;
      } else {
        uint32_t c_idx = cc[0] + cc[1] * gx + cc[2] * gx * gy;
        root_idx = w->roots[c_idx];
        c.root_loads++;
        memcpy(mtf_c1, mtf_c0, sizeof mtf_c1);
        mtf_root1 = mtf_root0;
        memcpy(mtf_c0, cc, sizeof cc);
        mtf_root0 = root_idx;
      }
    }
    if (root_idx == ARC_EMPTY_CHUNK_SENTINEL) { /* :337-352 */
      uint32_t bmin[3], bmax[3];
      for (int a = 0; a < 3; ++a) {
        bmin[a] = vi[a] & ~255u;
        bmax[a] = bmin[a] + 256u;
      }
      step_res st = step_dda_box(vo, dir, inv, noi, pos, bmin, bmax, t);
      t = st.t;
      memcpy(vi, st.vi, sizeof vi);
      c.empty_chunk_steps++;
      continue;
    }
    uint32_t min_shift = 0u; /* :355-370 */
    if (p->lod_bias == 0u) {
      if (t > prim_lod_t || step > prim_lod_step) min_shift = 2u;
    } else if (p->lod_bias == 1u) {
      if (t > (sec_lod_t * 2.2857f) || step > sec_lod_step) min_shift = 4u;
      else if (t > sec_lod_t) min_shift = 2u;
    } else {
      if (t > gi_lod_t || step > gi_lod_step) min_shift = 2u;
    }
    arc_node64 node;
    uint32_t shift;
    if (diff < 4u && min_shift == 0u && cached_shift == 0u) { /* :374-386 */
      node = l3;
      shift = 0u;
    } else if (diff < 16u && min_shift <= 2u && cached_shift <= 2u) {
      node = l2;
      shift = 2u;
    } else if (diff < 64u && min_shift <= 4u && cached_shift <= 4u) {
      node = l1;
      shift = 4u;
    } else {
      node = w->nodes[root_idx];
      c.node_loads++;
      c.level_loads[0]++;
      shift = 6u;
    }
    c.restarts[(6u - shift) / 2u]++;
    cached_shift = shift;
    const uint32_t l[3] = {vi[0] & 255u, vi[1] & 255u, vi[2] & 255u};
    for (;;) { /* :390-552 */
      uint32_t cx = (l[0] >> shift) & 3u, cy = (l[1] >> shift) & 3u, cz = (l[2] >> shift) & 3u;
      uint32_t b = (cz << 4) | (cy << 2) | cx;
      if (p->ignore_water && (node.child_ptr & ARC_IS_UNIFORM_FLAG) != 0u) { /* :398-417 */
        uint32_t uni = (node.lod_payload >> 24) & 0xFFu;
        if (uni == ARC_MAT_WATER) {
          uint32_t span = 4u << shift, bmin[3], bmax[3];
          for (int a = 0; a < 3; ++a) {
            bmin[a] = vi[a] & ~(span - 1u);
            bmax[a] = bmin[a] + span;
          }
          step_res st = step_dda_box(vo, dir, inv, noi, pos, bmin, bmax, t);
          t = st.t;
          memcpy(vi, st.vi, sizeof vi);
          break;
        }
      }
      uint32_t span = 1u << shift, cmin[3], cmax[3];
      for (int a = 0; a < 3; ++a) {
        cmin[a] = vi[a] & ~(span - 1u);
        cmax[a] = cmin[a] + span;
      }
      if (!has_child(&node, b)) { /* :419-436 */
        step_res st = step_dda_box(vo, dir, inv, noi, pos, cmin, cmax, t);
        t = st.t;
        memcpy(vi, st.vi, sizeof vi);
        break;
      }
      int is_leaf = (node.child_ptr & ARC_IS_LEAF_FLAG) != 0u;
      int stop_lod = shift <= min_shift;
      if (is_leaf || stop_lod) { /* :441-538 */
        uint32_t mat = (node.lod_payload >> 24) & 0xFFu;
        if (is_leaf && (node.child_ptr & ARC_IS_UNIFORM_FLAG) == 0u) {
          mat = brick_mat(w, node.child_ptr & ARC_PTR_INDEX_MASK, b);
          c.brick_loads++;
        }
        if (p->ignore_water && mat == ARC_MAT_WATER) {
          step_res st = step_dda_box(vo, dir, inv, noi, pos, cmin, cmax, t);
          t = st.t;
          memcpy(vi, st.vi, sizeof vi);
          break;
        }
        if (mat != ARC_MAT_AIR) { /* :467-523 */
          float tm[3];
          for (int a = 0; a < 3; ++a) {
            uint32_t ec = pos[a] ? cmin[a] : cmax[a];
            float m = (float)ec * inv[a];
            tm[a] = m + noi[a];
          }
          float cell_enter = fmaxf(fmaxf(tm[0], tm[1]), tm[2]);
          float exact_t = fmaxf(cell_enter, 0.0f);
          float hp[3];
          for (int a = 0; a < 3; ++a) hp[a] = vo[a] + dir[a] * exact_t;
          res->hit = 1u;
          res->t_voxels = exact_t;
          for (int a = 0; a < 3; ++a) res->hit_pos_m[a] = hp[a] * 0.0625f;
          float max_t = fmaxf(fmaxf(tm[0], tm[1]), tm[2]);
          float eps = 1e-4f * fmaxf(max_t * 0.01f, 1.0f);
          float cand[3];
          for (int a = 0; a < 3; ++a) cand[a] = fabsf(tm[a] - max_t) <= eps ? fabsf(dir[a]) : 0.0f;
          float nrm[3] = {0.0f, 0.0f, 0.0f};
          if (cand[0] >= cand[1] && cand[0] >= cand[2] && cand[0] > 0.0f) nrm[0] = -wgsl_sign(dir[0]);
          else if (cand[1] >= cand[2] && cand[1] > 0.0f) nrm[1] = -wgsl_sign(dir[1]);
          else if (cand[2] > 0.0f) nrm[2] = -wgsl_sign(dir[2]);
          else if (tm[0] >= tm[1] && tm[0] >= tm[2]) nrm[0] = -wgsl_sign(dir[0]);
          else if (tm[1] >= tm[2]) nrm[1] = -wgsl_sign(dir[1]);
          else nrm[2] = -wgsl_sign(dir[2]);
          /* which axis was chosen (needed when its sign is zero) */
          int axis;
          if (cand[0] >= cand[1] && cand[0] >= cand[2] && cand[0] > 0.0f) axis = 0;
          else if (cand[1] >= cand[2] && cand[1] > 0.0f) axis = 1;
          else if (cand[2] > 0.0f) axis = 2;
          else if (tm[0] >= tm[1] && tm[0] >= tm[2]) axis = 0;
          else if (tm[1] >= tm[2]) axis = 1;
          else axis = 2;
          memcpy(res->normal, nrm, sizeof nrm);
          res->mat = mat;
          res->lod_rgb[0] = (float)(node.lod_payload & 0xFFu) / 255.0f;
          res->lod_rgb[1] = (float)((node.lod_payload >> 8) & 0xFFu) / 255.0f;
          res->lod_rgb[2] = (float)((node.lod_payload >> 16) & 0xFFu) / 255.0f;
          res->voxel_hash = pcg_hash((vi[0] * 73856093u) ^ (vi[1] * 19349663u) ^ (vi[2] * 83492791u));
          if (shift == 0u && p->lod_bias == 0u) { /* :501-521 */
            float hit_dist_m = t * 0.0625f;
            float raw = p->ao_max_dist_bits;
            float ao_max = raw < 1.0f ? 24.0f : raw;
            if (hit_dist_m < ao_max) {
              float raw_ao = eval_leaf_ao(node.mask_low, node.mask_high, cx, cy, cz, nrm, hp);
              float ao_min = ao_max * 0.5f;
              float wgt = 1.0f - smoothstep_(ao_min, ao_max, hit_dist_m);
              res->contact_ao = mix_(1.0f, raw_ao, wgt);
            } else {
              res->contact_ao = 1.0f;
            }
          }
          res->vx = vi[0];
          res->vy = vi[1];
          res->vz = vi[2];
          res->steps = c.steps;
          /* stash axis in the upper bits of `hit` for face coding (1 | axis<<8 | shift<<16) */
          res->hit = 1u | ((uint32_t)axis << 8) | (shift << 16);
          if (cnt) *cnt = c;
          return;
        }
        step_res st = step_dda_box(vo, dir, inv, noi, pos, cmin, cmax, t);
        t = st.t;
        memcpy(vi, st.vi, sizeof vi);
        break;
      }
      uint32_t next = child_offset(&node, b); /* :540-551 */
      node = w->nodes[next];
      c.node_loads++;
      shift -= 2u;
      c.level_loads[(6u - shift) / 2u]++;
      cached_shift = shift;
      if (shift == 4u) l1 = node;
      else if (shift == 2u) l2 = node;
      else if (shift == 0u) l3 = node;
    }
  }
  res->steps = c.steps;
  if (cnt) *cnt = c;
}

/* ---- shadow any-hit walkers: trace_shadow_vis (traversal64.wgsl:571-773) and
 * trace_shadow_vis_di (:781-989) ------------------------------------------------------------ */
static inline int is_emissive_mat(uint32_t m) { /* :775-777 */
  return (m >= 19u && m <= 22u) || m == 24u || m == 27u;
}

static void shadow_one(const orc_world* w, uint32_t mode, uint32_t ray_steps_w, const arc_ray* ray,
                       float* vis, arc_shadow_rec* rec) {
  const float o[3] = {ray->ox, ray->oy, ray->oz};
  const float dir[3] = {ray->dx, ray->dy, ray->dz};
  float vo[3], safe[3], inv[3], wmax[3], noi[3];
  int pos[3];
  uint32_t wmaxi[3];
  for (int a = 0; a < 3; ++a) vo[a] = o[a] * 16.0f; /* :576 / :786 */
  const float max_t_v = ray->t_max * 16.0f;
  for (int a = 0; a < 3; ++a) {
    safe[a] = fabsf(dir[a]) < 1e-7f ? 1e-7f : dir[a]; /* :579 */
    inv[a] = 1.0f / safe[a];
    wmax[a] = (float)w->grid[a] * 256.0f;
  }
  float tminv[3], tmaxv[3];
  for (int a = 0; a < 3; ++a) { /* intersect_aabb :134-142 */
    float eb = inv[a] > 0.0f ? 0.0f : wmax[a];
    float xb = inv[a] > 0.0f ? wmax[a] : 0.0f;
    tminv[a] = (eb - vo[a]) * inv[a];
    tmaxv[a] = (xb - vo[a]) * inv[a];
  }
  const float wh_x = fmaxf(fmaxf(tminv[0], tminv[1]), tminv[2]);
  const float wh_y = fminf(fminf(tmaxv[0], tmaxv[1]), tmaxv[2]);
  memset(rec, 0, sizeof *rec);
  if (wh_y < fmaxf(wh_x, 0.0f) || wh_x >= max_t_v) { /* :587-589 */
    *vis = 1.0f;
    rec->outcome = ARC_SHADOW_MISS_WORLD;
    return;
  }
  const float t_enter = fmaxf(wh_x, 0.0f);
  float t = t_enter;
  const float t_limit = fminf(wh_y, max_t_v);
  uint32_t max_steps = ray_steps_w; /* :593 */
  if (mode == ARC_SHADOW_DI) max_steps = max_steps < 32u ? 32u : (max_steps > 48u ? 48u : max_steps); /* :803 */
  for (int a = 0; a < 3; ++a) {
    pos[a] = inv[a] >= 0.0f;
    noi[a] = -vo[a] * inv[a];
    wmaxi[a] = w->grid[a] << 8;
  }
  uint32_t vi[3];
  for (int a = 0; a < 3; ++a) { /* :599-600 */
    float sp = clampf_(vo[a] + dir[a] * (t + 1e-4f), 0.0f, wmax[a] - 1e-3f);
    vi[a] = f2u(sp);
  }
  uint32_t prev[3] = {~0u, ~0u, ~0u};
  uint32_t cached_shift = 6u;
  arc_node64 l1 = {0, 0, 0, 0}, l2 = {0, 0, 0, 0}, l3 = {0, 0, 0, 0};
  uint32_t cached_root = ARC_EMPTY_CHUNK_SENTINEL;
  const uint32_t gx = w->grid[0], gy = w->grid[1];
  uint32_t steps = 0u;
  uint32_t step;
  for (step = 0; step < max_steps; ++step) { /* :610 */
    if (t >= t_limit) break;
    if (vi[0] >= wmaxi[0] || vi[1] >= wmaxi[1] || vi[2] >= wmaxi[2]) {
      t = t_limit;
      break;
    }
    steps++;
    const uint32_t diff = (prev[0] ^ vi[0]) | (prev[1] ^ vi[1]) | (prev[2] ^ vi[2]);
    memcpy(prev, vi, sizeof prev);
    uint32_t root_idx = cached_root;
    if (diff >= 256u) { /* :624-630 */
      cached_shift = 6u;
      const uint32_t c_idx = (vi[0] >> 8) + (vi[1] >> 8) * gx + (vi[2] >> 8) * gx * gy;
      root_idx = w->roots[c_idx];
      cached_root = root_idx;
    }
    if (root_idx == ARC_EMPTY_CHUNK_SENTINEL) { /* :632-646 */
      uint32_t bmin[3], bmax[3];
      for (int a = 0; a < 3; ++a) {
        bmin[a] = vi[a] & ~255u;
        bmax[a] = bmin[a] + 256u;
      }
      step_res st = step_dda_box(vo, dir, inv, noi, pos, bmin, bmax, t);
      t = st.t;
      memcpy(vi, st.vi, sizeof vi);
      continue;
    }
    uint32_t min_shift = 0u;
    if (mode == ARC_SHADOW_SUN) { /* :648-654 */
      if (t > 48.0f || step > 6u) min_shift = 2u;
      if (t > 192.0f || step > 16u) min_shift = 4u;
    } else { /* :858-865 */
      const float dist_rem = t_limit - t;
      if (t > 48.0f && dist_rem > 32.0f && step > 6u) min_shift = 2u;
      if (t > 160.0f && dist_rem > 64.0f && step > 14u) min_shift = 4u;
    }
    arc_node64 node;
    uint32_t shift;
    if (diff < 4u && min_shift == 0u && cached_shift == 0u) { /* :656-669 */
      node = l3;
      shift = 0u;
    } else if (diff < 16u && min_shift <= 2u && cached_shift <= 2u) {
      node = l2;
      shift = 2u;
    } else if (diff < 64u && min_shift <= 4u && cached_shift <= 4u) {
      node = l1;
      shift = 4u;
    } else {
      node = w->nodes[root_idx];
      shift = 6u;
    }
    cached_shift = shift;
    const uint32_t l[3] = {vi[0] & 255u, vi[1] & 255u, vi[2] & 255u};
    for (;;) { /* :672-755 / :882-971 */
      const uint32_t b = (((l[2] >> shift) & 3u) << 4) | (((l[1] >> shift) & 3u) << 2) | ((l[0] >> shift) & 3u);
      uint32_t span = 0u;
      if ((node.child_ptr & ARC_IS_UNIFORM_FLAG) != 0u && ((node.lod_payload >> 24) & 0xFFu) == ARC_MAT_WATER) {
        span = 4u << shift; /* water fast-skip */
      } else if (!has_child(&node, b)) {
        span = 1u << shift;
      } else {
        const int is_leaf = (node.child_ptr & ARC_IS_LEAF_FLAG) != 0u;
        const int is_uniform = (node.child_ptr & ARC_IS_UNIFORM_FLAG) != 0u;
        const int stop_lod = shift <= min_shift;
        const int stop = mode == ARC_SHADOW_SUN ? (is_leaf || stop_lod) : (is_leaf || is_uniform || stop_lod);
        if (stop) {
          uint32_t mat = (node.lod_payload >> 24) & 0xFFu;
          if (mode == ARC_SHADOW_DI && is_leaf && !is_uniform && shift == 0u)
            mat = brick_mat(w, node.child_ptr & ARC_PTR_INDEX_MASK, b);
          if (mat == ARC_MAT_WATER) {
            span = 1u << shift;
          } else {
            int lit = 0;
            if (mode == ARC_SHADOW_DI && is_emissive_mat(mat) && t >= t_limit - 16.0f) lit = 1; /* :960-962 */
            *vis = lit ? 1.0f : 0.0f;
            rec->t_voxels = t;
            rec->steps = steps;
            rec->vx = (uint16_t)vi[0];
            rec->vy = (uint16_t)vi[1];
            rec->vz = (uint16_t)vi[2];
            rec->outcome = lit ? ARC_SHADOW_EMITTER : ARC_SHADOW_BLOCKED;
            return;
          }
        }
      }
      if (span != 0u) {
        uint32_t bmin[3], bmax[3];
        for (int a = 0; a < 3; ++a) {
          bmin[a] = vi[a] & ~(span - 1u);
          bmax[a] = bmin[a] + span;
        }
        step_res st = step_dda_box(vo, dir, inv, noi, pos, bmin, bmax, t);
        t = st.t;
        memcpy(vi, st.vi, sizeof vi);
        break;
      }
      node = w->nodes[child_offset(&node, b)];
      shift -= 2u;
      cached_shift = shift;
      if (shift == 4u) l1 = node;
      else if (shift == 2u) l2 = node;
      else if (shift == 0u) l3 = node;
    }
  }
  rec->t_voxels = t;
  rec->steps = steps;
  rec->vx = (uint16_t)vi[0];
  rec->vy = (uint16_t)vi[1];
  rec->vz = (uint16_t)vi[2];
  const int budget = step >= max_steps;
  if (mode == ARC_SHADOW_SUN) { /* :762-772 */
    int dark = 0;
    if (t < t_limit) {
      const float q = dir[1] * t;
      const float py = vo[1] + q;
      if (py * 0.0625f < 32.0f) dark = 1;
    }
    *vis = dark ? 0.0f : 1.0f;
    rec->outcome = dark ? ARC_SHADOW_BUDGET_DARK : (t < t_limit ? ARC_SHADOW_BUDGET_LIT : ARC_SHADOW_CLEAR);
  } else { /* :976-984 */
    const int lit = t >= t_limit - 2.0f;
    *vis = lit ? 1.0f : 0.0f;
    rec->outcome = lit ? (t >= t_limit ? ARC_SHADOW_CLEAR : ARC_SHADOW_BUDGET_LIT) : ARC_SHADOW_BUDGET_DARK;
  }
  (void)budget;
}

/* ---- rigid bodies: trace_single_rigid_body / trace_rigid_bodies (traversal64.wgsl:1005-1244)
 * Matrix-vector products are evaluated column by column, ((c0 x + c1 y) + c2 z) + c3 w, and dot
 * products x, y, z left to right; normalize(v) = v / sqrt(dot(v, v)); sign(0) = 0. ----------- */
static void mat_vec(const float m[16], const float v[4], float out[3]) {
  for (int r = 0; r < 3; ++r) {
    float a = m[0 + r] * v[0];
    float b = m[4 + r] * v[1];
    float c = m[8 + r] * v[2];
    float d = m[12 + r] * v[3];
    out[r] = ((a + b) + c) + d;
  }
}

static inline float dot3(const float a[3], const float b[3]) {
  float x = a[0] * b[0], y = a[1] * b[1], z = a[2] * b[2];
  return (x + y) + z;
}

static inline float sign_wgsl(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

static void rigid_miss(const float o[3], const float d[3], float max_dist_m, arc_rigid_hit* res) {
  memset(res, 0, sizeof *res);
  res->t_voxels = max_dist_m * 16.0f;
  for (int a = 0; a < 3; ++a) {
    const float q = d[a] * max_dist_m;
    res->hit_pos_m[a] = o[a] + q;
    res->prev_hit_pos_m[a] = res->hit_pos_m[a];
  }
  res->normal[1] = 1.0f;
  res->mat = ARC_MAT_AIR;
  res->contact_ao = 1.0f;
  res->body = 0xFFFFFFFFu;
}

static void rigid_single(const orc_world* w, const arc_rigid_body* rb, uint32_t rb_idx, const float o[3],
                         const float d[3], float max_dist_m, arc_rigid_hit* res) {
  rigid_miss(o, d, max_dist_m, res);
  uint32_t root_idx;
  memcpy(&root_idx, &rb->local_bounds_and_root[3], 4);
  if (root_idx == ARC_EMPTY_CHUNK_SENTINEL || root_idx >= w->n_nodes) return;
  /* 1. sphere pre-reject (:1028-1045) */
  float oc[3];
  for (int a = 0; a < 3; ++a) oc[a] = o[a] - rb->world_sphere[a];
  const float b_proj = dot3(oc, d);
  const float r2 = rb->world_sphere[3] * rb->world_sphere[3];
  const float c_dist = dot3(oc, oc) - r2;
  if (c_dist > 0.0f && b_proj > 0.0f) return;
  const float disc = b_proj * b_proj - c_dist;
  if (disc < 0.0f) return;
  const float near_m = fmaxf(-b_proj - sqrtf(disc), 0.0f);
  if (near_m >= max_dist_m) return;
  /* 2. local frame (:1047-1054) */
  const float o4[4] = {o[0], o[1], o[2], 1.0f}, d4[4] = {d[0], d[1], d[2], 0.0f};
  float lo[3], ld[3];
  mat_vec(rb->world_to_local, o4, lo);
  mat_vec(rb->world_to_local, d4, ld);
  float vo[3], lmax[3], safe[3], inv[3], noi[3];
  int pos[3];
  for (int a = 0; a < 3; ++a) {
    vo[a] = lo[a] * 16.0f;
    lmax[a] = rb->local_bounds_and_root[a] * 16.0f;
  }
  const float max_t_v = max_dist_m * 16.0f;
  for (int a = 0; a < 3; ++a) {
    safe[a] = fabsf(ld[a]) < 1e-7f ? 1e-7f : ld[a];
    inv[a] = 1.0f / safe[a];
  }
  float tminv[3], tmaxv[3];
  for (int a = 0; a < 3; ++a) { /* intersect_aabb :134-142 */
    const float eb = inv[a] > 0.0f ? 0.0f : lmax[a];
    const float xb = inv[a] > 0.0f ? lmax[a] : 0.0f;
    tminv[a] = (eb - vo[a]) * inv[a];
    tmaxv[a] = (xb - vo[a]) * inv[a];
  }
  const float bx = fmaxf(fmaxf(tminv[0], tminv[1]), tminv[2]);
  const float by = fminf(fminf(tmaxv[0], tmaxv[1]), tmaxv[2]);
  const float t_enter = fmaxf(bx, 0.0f);
  if (by < t_enter || t_enter >= max_t_v) return;
  float t = t_enter;
  const float t_limit = fminf(by, max_t_v);
  uint32_t lmaxi[3];
  for (int a = 0; a < 3; ++a) {
    pos[a] = inv[a] >= 0.0f;
    noi[a] = -vo[a] * inv[a];
    lmaxi[a] = f2u(lmax[a] + 0.5f);
  }
  uint32_t vi[3];
  for (int a = 0; a < 3; ++a) vi[a] = f2u(clampf_(vo[a] + ld[a] * (t + 1e-4f), 0.0f, lmax[a] - 1e-3f));
  uint32_t prev[3] = {~0u, ~0u, ~0u};
  uint32_t cached_shift = 6u;
  arc_node64 l1 = {0, 0, 0, 0}, l2 = {0, 0, 0, 0}, l3 = {0, 0, 0, 0};
  for (uint32_t step = 0; step < 64u; ++step) { /* :1077 */
    if (t >= t_limit) break;
    if (vi[0] >= lmaxi[0] || vi[1] >= lmaxi[1] || vi[2] >= lmaxi[2]) break; /* t = t_limit */
    const uint32_t diff = (prev[0] ^ vi[0]) | (prev[1] ^ vi[1]) | (prev[2] ^ vi[2]);
    memcpy(prev, vi, sizeof prev);
    arc_node64 node;
    uint32_t shift;
    if (diff < 4u && cached_shift == 0u) { /* :1089-1102 */
      node = l3;
      shift = 0u;
    } else if (diff < 16u && cached_shift <= 2u) {
      node = l2;
      shift = 2u;
    } else if (diff < 64u && cached_shift <= 4u) {
      node = l1;
      shift = 4u;
    } else {
      node = w->nodes[root_idx];
      shift = 6u;
    }
    cached_shift = shift;
    const uint32_t l[3] = {vi[0] & 255u, vi[1] & 255u, vi[2] & 255u};
    for (;;) { /* :1106-1190 */
      const uint32_t cx = (l[0] >> shift) & 3u, cy = (l[1] >> shift) & 3u, cz = (l[2] >> shift) & 3u;
      const uint32_t b = (cz << 4) | (cy << 2) | cx;
      const uint32_t span = 1u << shift;
      uint32_t cmin[3], cmax[3];
      for (int a = 0; a < 3; ++a) {
        cmin[a] = vi[a] & ~(span - 1u);
        cmax[a] = cmin[a] + span;
      }
      int step_box = 0;
      if (!has_child(&node, b)) {
        step_box = 1;
      } else if ((node.child_ptr & ARC_IS_LEAF_FLAG) != 0u) {
        uint32_t mat = (node.lod_payload >> 24) & 0xFFu;
        if ((node.child_ptr & ARC_IS_UNIFORM_FLAG) == 0u) {
          const uint32_t bi = node.child_ptr & ARC_PTR_INDEX_MASK;
          mat = bi < w->n_bricks ? brick_mat(w, bi, b) : ARC_MAT_AIR;
        }
        if (mat != ARC_MAT_AIR) { /* :1143-1182 */
          float tm[3];
          for (int a = 0; a < 3; ++a) {
            const uint32_t ec = pos[a] ? cmin[a] : cmax[a];
            const float m = (float)ec * inv[a];
            tm[a] = m + noi[a];
          }
          const float cell_enter = fmaxf(fmaxf(tm[0], tm[1]), tm[2]);
          const float exact_t = fmaxf(cell_enter, 0.0f);
          float hv[4], hm[4];
          for (int a = 0; a < 3; ++a) {
            const float q = ld[a] * exact_t;
            hv[a] = vo[a] + q;
            hm[a] = hv[a] * 0.0625f;
          }
          hm[3] = 1.0f;
          res->hit = 1u;
          res->t_voxels = exact_t;
          mat_vec(rb->local_to_world, hm, res->hit_pos_m);
          mat_vec(rb->prev_local_to_world, hm, res->prev_hit_pos_m);
          float ln[4] = {0.0f, 1.0f, 0.0f, 0.0f};
          if (tm[0] >= tm[1] && tm[0] >= tm[2]) {
            ln[0] = -sign_wgsl(ld[0]);
            ln[1] = 0.0f;
          } else if (tm[1] >= tm[2]) {
            ln[1] = -sign_wgsl(ld[1]);
          } else {
            ln[1] = 0.0f;
            ln[2] = -sign_wgsl(ld[2]);
          }
          float wn[3];
          mat_vec(rb->local_to_world, ln, wn);
          const float len = sqrtf(dot3(wn, wn));
          for (int a = 0; a < 3; ++a) res->normal[a] = wn[a] / len;
          res->mat = mat;
          res->voxel_hash = pcg_hash((vi[0] * 73856093u) ^ (vi[1] * 19349663u) ^ (vi[2] * 83492791u));
          res->contact_ao = 1.0f;
          if (shift == 0u) {
            const float nv[3] = {ln[0], ln[1], ln[2]};
            res->contact_ao = eval_leaf_ao(node.mask_low, node.mask_high, cx, cy, cz, nv, hv);
          }
          res->body = rb_idx;
          return;
        }
        step_box = 1;
      }
      if (step_box) {
        step_res st = step_dda_box(vo, ld, inv, noi, pos, cmin, cmax, t);
        t = st.t;
        memcpy(vi, st.vi, sizeof vi);
        break;
      }
      const uint32_t next = child_offset(&node, b);
      if (next >= w->n_nodes) return; /* tree leaves the pool: traced as empty */
      node = w->nodes[next];
      shift -= 2u;
      cached_shift = shift;
      if (shift == 4u) l1 = node;
      else if (shift == 2u) l2 = node;
      else if (shift == 0u) l3 = node;
    }
  }
}

static void rigid_all(const orc_world* w, const arc_rigid_body* rbs, uint32_t n_rb, const arc_ray* ray,
                      arc_rigid_hit* best) {
  const float o[3] = {ray->ox, ray->oy, ray->oz}, d[3] = {ray->dx, ray->dy, ray->dz};
  rigid_miss(o, d, ray->t_max, best);
  float cur_max_m = ray->t_max;
  const uint32_t count = n_rb < ARC_MAX_RIGID_BODIES ? n_rb : ARC_MAX_RIGID_BODIES;
  for (uint32_t i = 0; i < count; ++i) { /* :1232-1241 */
    arc_rigid_hit cand;
    rigid_single(w, &rbs[i], i, o, d, cur_max_m, &cand);
    if (cand.hit && cand.t_voxels < best->t_voxels) {
      *best = cand;
      cur_max_m = cand.t_voxels * 0.0625f;
    }
  }
}

void orc_trace_rigid(const orc_world* w, const arc_rigid_body* bodies, uint32_t n_bodies, const arc_ray* rays,
                     uint64_t n, arc_rigid_hit* out) {
  for (uint64_t i = 0; i < n; ++i) rigid_all(w, bodies, n_bodies, &rays[i], &out[i]);
}

/* Single-ray forms of the traversals, for restatements that trace inline (restir_oracle.c). */
void orc_trace_world_ray(const orc_world* w, const arc_trace_params* p, const arc_ray* ray, arc_trace_result* res) {
  orc_counters c;
  trace_world_one(w, p, ray, res, &c);
}
float orc_shadow_ray(const orc_world* w, uint32_t mode, uint32_t ray_steps_w, const arc_ray* ray) {
  float vis;
  arc_shadow_rec rec;
  shadow_one(w, mode, ray_steps_w, ray, &vis, &rec);
  return vis;
}
void orc_rigid_ray(const orc_world* w, const arc_rigid_body* bodies, uint32_t n_bodies, const arc_ray* ray,
                   arc_rigid_hit* out) {
  rigid_all(w, bodies, n_bodies, ray, out);
}

/* query_voxel_world_meters (traversal64.wgsl:182-220) */
uint32_t orc_query_point(const orc_world* w, const float p[3]) {
  float vp[3];
  for (int a = 0; a < 3; ++a) {
    vp[a] = p[a] * 16.0f;
    const float wm = (float)w->grid[a] * 256.0f;
    if (vp[a] < 0.0f || vp[a] >= wm) return ARC_MAT_AIR;
  }
  return orc_query_voxel(w, f2u(vp[0]), f2u(vp[1]), f2u(vp[2]));
}

/* ---- raycast_chunk_tree: raycast_cpu.rs:28-218 ------------------------------------------ */
/* Operates on chunk 0 of the pools (root = roots[0]); pool indices are global, which is the
 * relocated image of the chunk's local nodes (world_map.rs:705-735). */
static void trace_chunk_one(const orc_world* w, uint32_t root, const float o[3], const float d[3],
                            float max_dist, arc_trace_result* res, orc_counters* cnt) {
  orc_counters c;
  memset(&c, 0, sizeof c);
  memset(res, 0, sizeof(*res));
  res->contact_ao = 1.0f;
  if (root == ARC_EMPTY_CHUNK_SENTINEL ||
      (w->nodes[root].mask_low == 0u && w->nodes[root].mask_high == 0u)) { /* :34-36 */
    if (cnt) *cnt = c;
    return;
  }
  float safe[3], inv[3], noi[3];
  int pos[3];
  for (int a = 0; a < 3; ++a) {
    safe[a] = fabsf(d[a]) > 1e-7f ? d[a] : copysignf(1e-7f, d[a]); /* :38-54 */
    inv[a] = 1.0f / safe[a];
  }
  /* intersect_aabb :222-234 */
  float t0[3], t1[3], tmn[3], tmx[3];
  for (int a = 0; a < 3; ++a) {
    t0[a] = (0.0f - o[a]) * inv[a];
    t1[a] = (256.0f - o[a]) * inv[a];
    tmn[a] = fminf(t0[a], t1[a]);
    tmx[a] = fmaxf(t0[a], t1[a]);
  }
  float tmin = fmaxf(fmaxf(tmn[0], tmn[1]), tmn[2]);
  float tmax = fminf(fminf(tmx[0], tmx[1]), tmx[2]);
  if (!(tmax >= fmaxf(tmin, 0.0f))) {
    if (cnt) *cnt = c;
    return;
  }
  float t_enter = fmaxf(tmin, 0.0f); /* :59-63 */
  float t_limit = fminf(tmax, max_dist);
  if (tmax < t_enter || t_enter >= max_dist) {
    if (cnt) *cnt = c;
    return;
  }
  for (int a = 0; a < 3; ++a) {
    pos[a] = inv[a] >= 0.0f;
    noi[a] = -o[a] * inv[a];
  }
  float t = t_enter;
  uint32_t vi[3];
  for (int a = 0; a < 3; ++a) { /* :69-70 */
    float sp = clampf_(o[a] + safe[a] * (t + 1e-4f), 0.0f, 256.0f - 1e-3f);
    vi[a] = f2u(sp);
  }
  uint32_t prev[3] = {~0u, ~0u, ~0u};
  uint32_t cached_shift = 6u;
  arc_node64 l1 = {0, 0, 0, 0}, l2 = {0, 0, 0, 0}, l3 = {0, 0, 0, 0};
  for (int step = 0; step < 512; ++step) { /* :78 */
    if (t >= t_limit) break;
    if (vi[0] >= 256u || vi[1] >= 256u || vi[2] >= 256u) break;
    c.steps++;
    uint32_t diff = (prev[0] ^ vi[0]) | (prev[1] ^ vi[1]) | (prev[2] ^ vi[2]);
    memcpy(prev, vi, sizeof prev);
    arc_node64 node;
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
      node = w->nodes[root];
      c.node_loads++;
      shift = 6u;
    }
    cached_shift = shift;
    const uint32_t l[3] = {vi[0] & 255u, vi[1] & 255u, vi[2] & 255u};
    for (;;) {
      uint32_t cx = (l[0] >> shift) & 3u, cy = (l[1] >> shift) & 3u, cz = (l[2] >> shift) & 3u;
      uint32_t b = (cz << 4) | (cy << 2) | cx;
      uint32_t span = 1u << shift, cmin[3], cmax[3];
      for (int a = 0; a < 3; ++a) {
        cmin[a] = vi[a] & ~(span - 1u);
        cmax[a] = cmin[a] + span;
      }
      if (!has_child(&node, b)) {
        step_res st = step_dda_box(o, safe, inv, noi, pos, cmin, cmax, t);
        t = st.t;
        memcpy(vi, st.vi, sizeof vi);
        break;
      }
      if (node.child_ptr & ARC_IS_LEAF_FLAG) {
        uint32_t mat;
        if (node.child_ptr & ARC_IS_UNIFORM_FLAG) {
          mat = (node.lod_payload >> 24) & 0xFFu;
        } else {
          mat = brick_mat(w, node.child_ptr & ARC_PTR_INDEX_MASK, b);
          c.brick_loads++;
        }
        if (mat != ARC_MAT_AIR) {
          float tm[3];
          for (int a = 0; a < 3; ++a) {
            uint32_t ec = pos[a] ? cmin[a] : cmax[a];
            float m = (float)ec * inv[a];
            tm[a] = m + noi[a];
          }
          float cell_enter = fmaxf(fmaxf(tm[0], tm[1]), tm[2]);
          float exact_t = fmaxf(cell_enter, 0.0f);
          /* resolve_entry_normal :295-327 (Rust signum: never zero) */
          float max_t = fmaxf(fmaxf(tm[0], tm[1]), tm[2]);
          float eps = 1e-4f * fmaxf(max_t * 0.01f, 1.0f);
          float cand[3];
          for (int a = 0; a < 3; ++a) cand[a] = fabsf(tm[a] - max_t) <= eps ? fabsf(d[a]) : 0.0f;
          int axis;
          if (cand[0] >= cand[1] && cand[0] >= cand[2] && cand[0] > 0.0f) axis = 0;
          else if (cand[1] >= cand[2] && cand[1] > 0.0f) axis = 1;
          else if (cand[2] > 0.0f) axis = 2;
          else if (tm[0] >= tm[1] && tm[0] >= tm[2]) axis = 0;
          else if (tm[1] >= tm[2]) axis = 1;
          else axis = 2;
          res->normal[0] = res->normal[1] = res->normal[2] = 0.0f;
          res->normal[axis] = -rust_signum(d[axis]);
          res->t_voxels = exact_t;
          for (int a = 0; a < 3; ++a) res->hit_pos_m[a] = o[a] + safe[a] * exact_t;
          res->mat = mat;
          res->vx = vi[0];
          res->vy = vi[1];
          res->vz = vi[2];
          res->steps = c.steps;
          res->hit = 1u | ((uint32_t)axis << 8) | (shift << 16);
          if (cnt) *cnt = c;
          return;
        }
        step_res st = step_dda_box(o, safe, inv, noi, pos, cmin, cmax, t);
        t = st.t;
        memcpy(vi, st.vi, sizeof vi);
        break;
      }
      uint32_t next = child_offset(&node, b);
      node = w->nodes[next];
      c.node_loads++;
      shift -= 2u;
      cached_shift = shift;
      if (shift == 4u) l1 = node;
      else if (shift == 2u) l2 = node;
      else if (shift == 0u) l3 = node;
    }
  }
  res->steps = c.steps;
  if (cnt) *cnt = c;
}

/* ---- WorldMap::raycast_meters: world_map.rs:464-545 --------------------------------------
 * glam 0.29.3 (Cargo.lock), Vec3::normalize_or_zero = normalize_or(ZERO): rcp = 1.0 / length(),
 * length = sqrt(dot) with dot = (x*x + y*y) + z*z; the result is self * rcp when rcp is finite
 * and > 0, else zero (then the reference returns None). floor() of a positive value (p/256) as
 * floorf. */
static void trace_meters_one(const orc_world* w, const arc_ray* ray, arc_trace_result* res) {
  memset(res, 0, sizeof(*res));
  res->contact_ao = 1.0f;
  float d0[3] = {ray->dx, ray->dy, ray->dz};
  float dd = (d0[0] * d0[0] + d0[1] * d0[1]) + d0[2] * d0[2];
  float rl = 1.0f / sqrtf(dd);
  float d[3];
  if (!(rl > 0.0f && rl < INFINITY)) return; /* normalize_or_zero -> zero -> None */
  for (int a = 0; a < 3; ++a) d[a] = d0[a] * rl;
  if (d[0] == 0.0f && d[1] == 0.0f && d[2] == 0.0f) return;
  float vo[3] = {ray->ox * 16.0f, ray->oy * 16.0f, ray->oz * 16.0f};
  float max_v = ray->t_max * 16.0f;
  float inv[3];
  for (int a = 0; a < 3; ++a) inv[a] = fabsf(d[a]) > 1e-7f ? 1.0f / d[a] : copysignf(INFINITY, d[a]);
  float wmax[3];
  for (int a = 0; a < 3; ++a) wmax[a] = (float)(w->grid[a] * 256u);
  float t0[3], t1[3], tmn[3], tmx[3];
  for (int a = 0; a < 3; ++a) {
    t0[a] = (0.0f - vo[a]) * inv[a];
    t1[a] = (wmax[a] - vo[a]) * inv[a];
    tmn[a] = fminf(t0[a], t1[a]);
    tmx[a] = fmaxf(t0[a], t1[a]);
  }
  float tmin = fmaxf(fmaxf(tmn[0], tmn[1]), tmn[2]);
  float tw_exit = fminf(fminf(tmx[0], tmx[1]), tmx[2]);
  if (!(tw_exit >= fmaxf(tmin, 0.0f))) return;
  float t = tmin;
  if (t >= max_v || tw_exit <= 0.0f) return;
  t = fmaxf(t, 0.0f);
  uint32_t total_steps = 0;
  for (int hop = 0; hop < 128; ++hop) {
    if (t >= fminf(tw_exit, max_v)) break;
    float pp[3];
    int cc[3];
    for (int a = 0; a < 3; ++a) {
      pp[a] = clampf_(vo[a] + d[a] * (t + 1e-2f), 0.0f, wmax[a] - 1e-3f);
      cc[a] = (int)floorf(pp[a] / 256.0f);
    }
    float cmin[3], cmax[3];
    for (int a = 0; a < 3; ++a) {
      cmin[a] = (float)cc[a] * 256.0f;
      cmax[a] = cmin[a] + 256.0f;
    }
    if (cc[0] >= 0 && cc[1] >= 0 && cc[2] >= 0 && cc[0] < (int)w->grid[0] &&
        cc[1] < (int)w->grid[1] && cc[2] < (int)w->grid[2]) {
      uint32_t idx = (uint32_t)cc[0] + (uint32_t)cc[1] * w->grid[0] +
                     (uint32_t)cc[2] * w->grid[0] * w->grid[1];
      uint32_t root = w->roots[idx];
      if (root != ARC_EMPTY_CHUNK_SENTINEL) {
        float lo[3] = {vo[0] - cmin[0], vo[1] - cmin[1], vo[2] - cmin[2]};
        arc_trace_result hr;
        orc_counters hc;
        trace_chunk_one(w, root, lo, d, max_v, &hr, &hc);
        total_steps += hc.steps;
        if (hr.hit) {
          *res = hr;
          res->vx = hr.vx + (uint32_t)cc[0] * 256u;
          res->vy = hr.vy + (uint32_t)cc[1] * 256u;
          res->vz = hr.vz + (uint32_t)cc[2] * 256u;
          float dist_m = hr.t_voxels * 0.0625f;
          res->t_voxels = hr.t_voxels; /* voxel units, like every record of this oracle */
          for (int a = 0; a < 3; ++a) {
            float ray_o[3] = {ray->ox, ray->oy, ray->oz};
            res->hit_pos_m[a] = ray_o[a] + d[a] * dist_m;
          }
          res->steps = total_steps;
          return;
        }
      }
    }
    /* chunk box exit */
    float c0[3], c1[3], cn[3], cx[3];
    for (int a = 0; a < 3; ++a) {
      c0[a] = (cmin[a] - vo[a]) * inv[a];
      c1[a] = (cmax[a] - vo[a]) * inv[a];
      cn[a] = fminf(c0[a], c1[a]);
      cx[a] = fmaxf(c0[a], c1[a]);
    }
    float ctmin = fmaxf(fmaxf(cn[0], cn[1]), cn[2]);
    float ctmax = fminf(fminf(cx[0], cx[1]), cx[2]);
    if (ctmax >= fmaxf(ctmin, 0.0f)) {
      t = fmaxf(ctmax, t + 0.05f);
    } else {
      break;
    }
  }
  res->steps = total_steps;
}

/* ---- thread fan-out ---------------------------------------------------------------------- */
typedef struct job {
  int kind;
  const orc_world* w;
  const arc_trace_params* p;
  const arc_ray* rays;
  uint64_t n;
  arc_trace_result* out;
  orc_counters* cnt;
  uint64_t next;
  pthread_mutex_t mu;
  const arc_shadow_params* sp; /* kind 3 */
  float* vis;
  arc_shadow_rec* srec;
} job;

static void* worker(void* arg) {
  job* j = (job*)arg;
  const uint64_t grain = 256;
  for (;;) {
    pthread_mutex_lock(&j->mu);
    uint64_t s = j->next;
    j->next += grain;
    pthread_mutex_unlock(&j->mu);
    if (s >= j->n) break;
    uint64_t e = s + grain < j->n ? s + grain : j->n;
    for (uint64_t i = s; i < e; ++i) {
      orc_counters* c = j->cnt ? &j->cnt[i] : NULL;
      if (j->kind == 0) {
        trace_world_one(j->w, j->p, &j->rays[i], &j->out[i], c);
      } else if (j->kind == 1) {
        const float o[3] = {j->rays[i].ox, j->rays[i].oy, j->rays[i].oz};
        const float d[3] = {j->rays[i].dx, j->rays[i].dy, j->rays[i].dz};
        trace_chunk_one(j->w, j->w->roots[0], o, d, j->rays[i].t_max, &j->out[i], c);
      } else if (j->kind == 2) {
        trace_meters_one(j->w, &j->rays[i], &j->out[i]);
      } else {
        shadow_one(j->w, j->sp->mode, j->sp->max_steps, &j->rays[i], &j->vis[i], &j->srec[i]);
      }
    }
  }
  return NULL;
}

static void run_job(job* j, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  j->next = 0;
  pthread_mutex_init(&j->mu, NULL);
  if (nthreads == 1) {
    worker(j);
  } else {
    pthread_t th[256];
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, worker, j);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
  }
  pthread_mutex_destroy(&j->mu);
}

void orc_trace_world(const orc_world* w, const arc_trace_params* p, const arc_ray* rays,
                     uint64_t n, arc_trace_result* out, orc_counters* counters, int nthreads) {
  job j = {0, w, p, rays, n, out, counters, 0, PTHREAD_MUTEX_INITIALIZER};
  run_job(&j, nthreads);
}

void orc_trace_chunk(const orc_world* w, const arc_ray* rays, uint64_t n, arc_trace_result* out,
                     orc_counters* counters, int nthreads) {
  job j = {1, w, NULL, rays, n, out, counters, 0, PTHREAD_MUTEX_INITIALIZER};
  run_job(&j, nthreads);
}

void orc_trace_meters(const orc_world* w, const arc_ray* rays, uint64_t n, arc_trace_result* out,
                      int nthreads) {
  job j = {2, w, NULL, rays, n, out, NULL, 0, PTHREAD_MUTEX_INITIALIZER};
  run_job(&j, nthreads);
}

void orc_trace_shadow(const orc_world* w, const arc_shadow_params* sp, const arc_ray* rays, uint64_t n, float* vis,
                      arc_shadow_rec* rec, int nthreads) {
  job j = {3, w, NULL, rays, n, NULL, NULL, 0, PTHREAD_MUTEX_INITIALIZER, sp, vis, rec};
  run_job(&j, nthreads);
}

void orc_compact(const arc_trace_result* full, uint64_t n, int world_units, arc_hit* out) {
  (void)world_units;
  for (uint64_t i = 0; i < n; ++i) {
    const arc_trace_result* r = &full[i];
    arc_hit h;
    memset(&h, 0, sizeof h);
    h.t_voxels = r->t_voxels;
    h.steps = (uint16_t)(r->steps > 0xFFFFu ? 0xFFFFu : r->steps);
    if (r->hit & 1u) {
      uint32_t axis = (r->hit >> 8) & 0xFFu;
      uint32_t shift = (r->hit >> 16) & 0xFFu;
      h.vx = (uint16_t)r->vx;
      h.vy = (uint16_t)r->vy;
      h.vz = (uint16_t)r->vz;
      h.mat = (uint8_t)r->mat;
      uint8_t fc = face_code(r->normal);
      h.face = fc == 255u ? (uint8_t)(6u + axis) : fc;
      h.flags = (uint16_t)(ARC_HIT_FLAG_HIT | (shift > 0 ? ARC_HIT_FLAG_LOD : 0));
    } else {
      h.face = (uint8_t)ARC_FACE_NONE;
    }
    out[i] = h;
  }
}

/* ---- point query: traversal64.wgsl:182-220 (integer part) -------------------------------- */
uint32_t orc_query_voxel(const orc_world* w, uint32_t x, uint32_t y, uint32_t z) {
  uint32_t cx = x >> 8, cy = y >> 8, cz = z >> 8;
  if (cx >= w->grid[0] || cy >= w->grid[1] || cz >= w->grid[2]) return 0u;
  uint32_t root = w->roots[cx + cy * w->grid[0] + cz * w->grid[0] * w->grid[1]];
  if (root == ARC_EMPTY_CHUNK_SENTINEL) return 0u;
  uint32_t l[3] = {x & 255u, y & 255u, z & 255u};
  uint32_t idx = root, shift = 6u;
  for (int level = 0; level < 4; ++level) {
    const arc_node64* n = &w->nodes[idx];
    uint32_t b = (((l[2] >> shift) & 3u) << 4) | (((l[1] >> shift) & 3u) << 2) | ((l[0] >> shift) & 3u);
    if (!has_child(n, b)) return 0u;
    if (n->child_ptr & ARC_IS_LEAF_FLAG) {
      if (n->child_ptr & ARC_IS_UNIFORM_FLAG) return (n->lod_payload >> 24) & 0xFFu;
      return brick_mat(w, n->child_ptr & ARC_PTR_INDEX_MASK, b);
    }
    idx = child_offset(n, b);
    shift -= 2u;
  }
  return 0u;
}

typedef struct dec_job {
  const orc_world* w;
  uint8_t* dense;
  uint32_t z0, z1;
} dec_job;

static void* dec_worker(void* arg) {
  dec_job* j = (dec_job*)arg;
  uint64_t X = (uint64_t)j->w->grid[0] * 256u, Y = (uint64_t)j->w->grid[1] * 256u;
  for (uint32_t z = j->z0; z < j->z1; ++z)
    for (uint32_t y = 0; y < Y; ++y)
      for (uint32_t x = 0; x < X; ++x)
        j->dense[((uint64_t)z * Y + y) * X + x] = (uint8_t)orc_query_voxel(j->w, x, y, z);
  return NULL;
}

void orc_decode_dense(const orc_world* w, uint8_t* dense, int nthreads) {
  uint32_t Z = w->grid[2] * 256u;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 64) nthreads = 64;
  pthread_t th[64];
  dec_job jobs[64];
  uint32_t per = (Z + (uint32_t)nthreads - 1u) / (uint32_t)nthreads;
  for (int i = 0; i < nthreads; ++i) {
    jobs[i].w = w;
    jobs[i].dense = dense;
    jobs[i].z0 = (uint32_t)i * per < Z ? (uint32_t)i * per : Z;
    jobs[i].z1 = jobs[i].z0 + per < Z ? jobs[i].z0 + per : Z;
    pthread_create(&th[i], NULL, dec_worker, &jobs[i]);
  }
  for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
}

/* ---- independent dense walker ----------------------------------------------------------- */
typedef struct brute_job {
  const uint8_t* dense;
  const uint32_t* dims;
  const arc_ray* rays;
  uint64_t n;
  int world_units;
  double max_t_scale;
  orc_brute_hit* out;
  uint64_t next;
  pthread_mutex_t mu;
} brute_job;

static void brute_one(const brute_job* j, const arc_ray* r, orc_brute_hit* h) {
  const double tie_rel = 2e-6;
  double s = j->world_units ? 16.0 : 1.0;
  double o[3] = {r->ox * s, r->oy * s, r->oz * s};
  double d[3] = {r->dx, r->dy, r->dz};
  double tstart = j->world_units ? (r->t_start > 0 ? r->t_start * 16.0 : 0.0) : 0.0;
  double tmax_ray = r->t_max * s * j->max_t_scale;
  memset(h, 0, sizeof(*h));
  h->face = -1;
  /* box entry/exit */
  double tin = -1e300, tout = 1e300;
  int in_axis = -1;
  for (int a = 0; a < 3; ++a) {
    double dim = (double)j->dims[a];
    if (d[a] == 0.0) {
      if (o[a] < 0.0 || o[a] >= dim) return;
      continue;
    }
    double ta = (0.0 - o[a]) / d[a], tb = (dim - o[a]) / d[a];
    double lo = ta < tb ? ta : tb, hi = ta < tb ? tb : ta;
    if (lo > tin) {
      tin = lo;
      in_axis = a;
    }
    if (hi < tout) tout = hi;
  }
  double t = tin > tstart ? tin : tstart;
  if (t > tin) in_axis = -1; /* starts inside: no entry face */
  if (tout < t || t >= tmax_ray) return;
  double limit = tout < tmax_ray ? tout : tmax_ray;
  int v[3], step[3];
  double tnext[3], tdelta[3];
  for (int a = 0; a < 3; ++a) {
    double p = o[a] + d[a] * t;
    int vi = (int)floor(p);
    if (p == floor(p) && d[a] < 0) vi -= 1; /* exactly on a plane: the ray moves into the lower cell */
    if (in_axis == a) vi = d[a] > 0 ? 0 : (int)j->dims[a] - 1;
    if (vi < 0) vi = 0;
    if (vi >= (int)j->dims[a]) vi = (int)j->dims[a] - 1;
    v[a] = vi;
    if (d[a] > 0) {
      step[a] = 1;
      tnext[a] = ((double)(vi + 1) - o[a]) / d[a];
      tdelta[a] = 1.0 / d[a];
    } else if (d[a] < 0) {
      step[a] = -1;
      tnext[a] = ((double)vi - o[a]) / d[a];
      tdelta[a] = -1.0 / d[a];
    } else {
      step[a] = 0;
      tnext[a] = 1e300;
      tdelta[a] = 1e300;
    }
  }
  int face_axis = in_axis; /* -1: the ray starts inside, first voxel has no entry face */
  int tie = 0;
  /* entry tie: two axes enter together */
  if (in_axis >= 0) {
    for (int a = 0; a < 3; ++a) {
      if (a == in_axis || d[a] == 0.0) continue;
      double dim = (double)j->dims[a];
      double ta = (0.0 - o[a]) / d[a], tb = (dim - o[a]) / d[a];
      double lo = ta < tb ? ta : tb;
      if (fabs(lo - tin) <= tie_rel * (fabs(tin) > 1.0 ? fabs(tin) : 1.0)) tie = 1;
    }
  }
  /* a start point within tie distance of a voxel boundary is ambiguous too */
  for (int a = 0; a < 3; ++a) {
    double p = o[a] + d[a] * t;
    double fr = p - floor(p);
    if (a != in_axis && fr != 0.0 && (fr < 1e-4 || fr > 1.0 - 1e-4)) tie = 1;
  }
  uint64_t X = j->dims[0], Y = j->dims[1];
  for (uint64_t it = 0; it < 4u * (X + Y + j->dims[2]) + 8u; ++it) {
    uint8_t m = j->dense[((uint64_t)v[2] * Y + (uint64_t)v[1]) * X + (uint64_t)v[0]];
    if (m != 0) {
      h->hit = 1;
      h->t = t;
      h->vx = v[0];
      h->vy = v[1];
      h->vz = v[2];
      h->mat = m;
      h->face = face_axis < 0 ? -1 : 2 * face_axis + (d[face_axis] > 0 ? 1 : 0);
      h->tie = tie || face_axis < 0 || (fabs(t - limit) <= tie_rel * (limit > 1.0 ? limit : 1.0));
      return;
    }
    int a = 0;
    if (tnext[1] < tnext[a]) a = 1;
    if (tnext[2] < tnext[a]) a = 2;
    double tn = tnext[a];
    /* near-tie among axes: the path may legitimately differ at float precision */
    for (int b = 0; b < 3; ++b)
      if (b != a && fabs(tnext[b] - tn) <= tie_rel * (tn > 1.0 ? tn : 1.0)) tie = 1;
    if (tn >= limit) {
      if (fabs(tn - limit) <= tie_rel * (limit > 1.0 ? limit : 1.0)) tie = 1;
      break;
    }
    t = tn;
    v[a] += step[a];
    tnext[a] += tdelta[a];
    face_axis = a;
    if (v[a] < 0 || v[a] >= (int)j->dims[a]) break;
  }
  h->hit = 0;
  h->tie = tie;
}

static void* brute_worker(void* arg) {
  brute_job* j = (brute_job*)arg;
  for (;;) {
    pthread_mutex_lock(&j->mu);
    uint64_t s = j->next;
    j->next += 256;
    pthread_mutex_unlock(&j->mu);
    if (s >= j->n) break;
    uint64_t e = s + 256 < j->n ? s + 256 : j->n;
    for (uint64_t i = s; i < e; ++i) brute_one(j, &j->rays[i], &j->out[i]);
  }
  return NULL;
}

void orc_brute_dense(const uint8_t* dense, const uint32_t dims[3], const arc_ray* rays, uint64_t n,
                     int world_units, double max_t_scale, orc_brute_hit* out, int nthreads) {
  brute_job j;
  j.dense = dense;
  j.dims = dims;
  j.rays = rays;
  j.n = n;
  j.world_units = world_units;
  j.max_t_scale = max_t_scale;
  j.out = out;
  j.next = 0;
  pthread_mutex_init(&j.mu, NULL);
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 64) nthreads = 64;
  pthread_t th[64];
  for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, brute_worker, &j);
  for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
  pthread_mutex_destroy(&j.mu);
}

/* ---- primary G-buffer: shaders/primary_gbuffer.wgsl:93-312 (opaque path) -------------------- */
uint16_t orc_f32_to_f16(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  uint32_t sign = (u >> 16) & 0x8000u;
  uint32_t e = (u >> 23) & 0xFFu;
  uint32_t m = u & 0x7FFFFFu;
  if (e == 0xFFu) return (uint16_t)(sign | 0x7C00u | (m ? 0x200u | (m >> 13) : 0u)); /* inf / NaN */
  int32_t he = (int32_t)e - 127 + 15;
  if (he >= 31) return (uint16_t)(sign | 0x7C00u); /* overflow -> inf */
  if (he <= 0) {                                   /* subnormal half (or zero) */
    if (he < -10) return (uint16_t)sign;
    uint32_t mant = m | 0x800000u;
    uint32_t shift = (uint32_t)(14 - he);
    uint32_t h = mant >> shift;
    uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (h & 1u))) h++;
    return (uint16_t)(sign | h);
  }
  uint32_t h = ((uint32_t)he << 10) | (m >> 13);
  uint32_t rem = m & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++; /* may carry into the exponent: correct */
  return (uint16_t)(sign | h);
}

static inline uint32_t pk_snorm16(float e) {
  float c = fminf(1.0f, fmaxf(-1.0f, e));
  float s = 32767.0f * c;
  return (uint32_t)(int32_t)floorf(0.5f + s) & 0xFFFFu;
}
static inline uint32_t pk_unorm8(float e) {
  float c = fminf(1.0f, fmaxf(0.0f, e));
  float s = 255.0f * c;
  return (uint32_t)floorf(0.5f + s);
}
/* oct_encode (types.wgsl:112-119) + pack_norm_motion (:310-313) */
static void pk_norm_motion(const float n[3], float mx, float my, uint32_t out[2]) {
  float l1 = fabsf(n[0]) + fabsf(n[1]);
  l1 = l1 + fabsf(n[2]);
  float inv = 1.0f / fmaxf(l1, 1e-6f);
  float px = n[0] * inv, py = n[1] * inv;
  if (n[2] < 0.0f) {
    float wx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
    float wy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
    px = wx;
    py = wy;
  }
  out[0] = pk_snorm16(px) | (pk_snorm16(py) << 16);
  out[1] = (uint32_t)orc_f32_to_f16(mx) | ((uint32_t)orc_f32_to_f16(my) << 16);
}

typedef struct gb_job {
  const orc_world* w;
  const arc_trace_params* p;
  const arc_gbuffer_params* g;
  const float* pal;
  uint32_t *dm, *nm, *aa;
  uint32_t n;
  uint32_t next;
  pthread_mutex_t mu;
} gb_job;

static void gb_one(gb_job* j, uint32_t i) {
  const arc_gbuffer_params* g = j->g;
  uint32_t tiles_x = (g->width + 7u) / 8u, tile = i / 64u, k = i % 64u;
  uint32_t px = (tile % tiles_x) * 8u + (k % 8u), py = (tile / tiles_x) * 8u + (k / 8u);
  if (px >= g->width || py >= g->height) return;
  /* primary_gbuffer.wgsl:103-115 */
  float ux = ((float)px + 0.5f) + g->jitter[0];
  float uy = ((float)py + 0.5f) + g->jitter[1];
  float uvx = ux / (float)g->width, uvy = uy / (float)g->height;
  float ndx = uvx * 2.0f - 1.0f, ndy = 1.0f - uvy * 2.0f;
  float sx = (ndx * g->aspect) * g->tan_half_fov, sy = ndy * g->tan_half_fov;
  float v[3];
  for (int a = 0; a < 3; ++a) {
    float rr = g->right[a] * sx, uu = g->up[a] * sy;
    v[a] = (g->forward[a] + rr) + uu;
  }
  float dd = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
  float len = sqrtf(dd);
  arc_ray ray = {g->eye[0], g->eye[1], g->eye[2], 0.0f, v[0] / len, v[1] / len, v[2] / len, g->max_dist_m};
  arc_trace_result r;
  trace_world_one(j->w, j->p, &ray, &r, NULL);
  uint32_t pix = py * g->width + px;
  if (!(r.hit & 1u)) {
    float d3[3] = {ray.dx, ray.dy, ray.dz};
    uint32_t nmv[2];
    pk_norm_motion(d3, 0.0f, 0.0f, nmv);
    float far = 1e5f;
    memcpy(&j->dm[2 * pix], &far, 4);
    j->dm[2 * pix + 1] = 0u;
    j->nm[2 * pix] = nmv[0];
    j->nm[2 * pix + 1] = nmv[1];
    j->aa[pix] = pk_unorm8(1.0f) << 24;
    return;
  }
  float hit_dist = r.t_voxels * 0.0625f;
  uint32_t mat = r.mat & 0xFFu;
  const float* m = j->pal + 8u * mat;
  float tt = clampf_((hit_dist - 18.0f) / (48.0f - 18.0f), 0.0f, 1.0f);
  float sm = tt * tt * (3.0f - 2.0f * tt);
  float wgt = 1.0f - sm;
  float v_rand = (((float)(r.voxel_hash & 0xFFu) / 255.0f) * 2.0f - 1.0f) * wgt;
  float kk = 1.0f + v_rand * m[7];
  float alb[3];
  for (int a = 0; a < 3; ++a) alb[a] = clampf_(m[a] * kk, 0.01f, 1.0f);
  float pu = ((float)px + 0.5f) / (float)g->width, pv = ((float)py + 0.5f) / (float)g->height;
  float clip[4];
  for (int row = 0; row < 4; ++row) {
    float a0 = g->prev_view_proj[0 * 4 + row] * r.hit_pos_m[0];
    float a1 = g->prev_view_proj[1 * 4 + row] * r.hit_pos_m[1];
    float a2 = g->prev_view_proj[2 * 4 + row] * r.hit_pos_m[2];
    float a3 = g->prev_view_proj[3 * 4 + row] * 1.0f;
    clip[row] = ((a0 + a1) + a2) + a3;
  }
  float mx = -2.0f, my = -2.0f;
  if (clip[3] > 0.001f) {
    float nx = clip[0] / clip[3], ny = clip[1] / clip[3];
    mx = (nx * 0.5f + 0.5f) - pu;
    my = (0.5f - ny * 0.5f) - pv;
  }
  memcpy(&j->dm[2 * pix], &hit_dist, 4);
  j->dm[2 * pix + 1] = mat & 0xFFFFu;
  uint32_t nmv[2];
  pk_norm_motion(r.normal, mx, my, nmv);
  j->nm[2 * pix] = nmv[0];
  j->nm[2 * pix + 1] = nmv[1];
  float ao = clampf_(r.contact_ao, 0.0f, 1.0f);
  j->aa[pix] = pk_unorm8(alb[0]) | (pk_unorm8(alb[1]) << 8) | (pk_unorm8(alb[2]) << 16) | (pk_unorm8(ao) << 24);
}

static void* gb_worker(void* arg) {
  gb_job* j = (gb_job*)arg;
  for (;;) {
    pthread_mutex_lock(&j->mu);
    uint32_t s = j->next;
    j->next += 256u;
    pthread_mutex_unlock(&j->mu);
    if (s >= j->n) break;
    uint32_t e = s + 256u < j->n ? s + 256u : j->n;
    for (uint32_t i = s; i < e; ++i) gb_one(j, i);
  }
  return NULL;
}

void orc_primary_gbuffer(const orc_world* w, const arc_trace_params* p, const arc_gbuffer_params* g,
                         const float* palette, uint32_t* depth_mat, uint32_t* norm_motion,
                         uint32_t* albedo_ao, int nthreads) {
  gb_job j;
  j.w = w;
  j.p = p;
  j.g = g;
  j.pal = palette;
  j.dm = depth_mat;
  j.nm = norm_motion;
  j.aa = albedo_ao;
  j.n = ((g->width + 7u) / 8u) * ((g->height + 7u) / 8u) * 64u;
  j.next = 0;
  pthread_mutex_init(&j.mu, NULL);
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 64) nthreads = 64;
  pthread_t th[64];
  for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, gb_worker, &j);
  for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
  pthread_mutex_destroy(&j.mu);
}

/* ---- SHARC cache: shaders/common/sharc.wgsl, sharc_update.wgsl:129-169, sharc_resolve.wgsl,
 * sharc_query.wgsl. Sequential: deposits claim slots in input order. ----------------------- */
float orc_f16_to_f32(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1Fu, m = h & 0x3FFu, u;
  if (e == 0) {
    if (m == 0) {
      u = sign;
    } else { /* subnormal: m * 2^-24 */
      float f = (float)m * (1.0f / 16777216.0f);
      memcpy(&u, &f, 4);
      u |= sign;
    }
  } else if (e == 31) {
    u = sign | 0x7F800000u | (m << 13);
  } else {
    u = sign | ((e - 15 + 127) << 23) | (m << 13);
  }
  float f;
  memcpy(&f, &u, 4);
  return f;
}

static uint32_t sh_lod(float d) { /* floor(log2(max(d,4)*0.25)) clamped 0..3, exact */
  float y = fmaxf(d, 4.0f) * 0.25f;
  int e = 0;
  frexpf(y, &e); /* y = m * 2^e, m in [0.5, 1) -> floor(log2 y) = e - 1 */
  e -= 1;
  return (uint32_t)(e < 0 ? 0 : (e > 3 ? 3 : e));
}

static uint32_t sh_norm_id(const float n[3]) {
  float ax = fabsf(n[0]), ay = fabsf(n[1]), az = fabsf(n[2]);
  if (ax > ay && ax > az) return n[0] > 0.0f ? 0u : 1u;
  if (ay > az) return n[1] > 0.0f ? 2u : 3u;
  return n[2] > 0.0f ? 4u : 5u;
}

static void sh_cross(const float a[3], const float b[3], float o[3]) {
  float p0 = a[1] * b[2], q0 = a[2] * b[1];
  float p1 = a[2] * b[0], q1 = a[0] * b[2];
  float p2 = a[0] * b[1], q2 = a[1] * b[0];
  o[0] = p0 - q0;
  o[1] = p1 - q1;
  o[2] = p2 - q2;
}

void orc_sharc_make_key(const float pos[3], const float n[3], const float cam[3], float jx, float jy,
                        uint32_t bucket_mask, uint32_t* bucket, uint32_t* key) {
  float d[3], a[3];
  for (int k = 0; k < 3; ++k) d[k] = pos[k] - cam[k];
  float dd = d[0] * d[0] + d[1] * d[1];
  dd = dd + d[2] * d[2];
  uint32_t lod = sh_lod(sqrtf(dd));
  float s = 0.25f * (float)(1u << lod);
  for (int k = 0; k < 3; ++k) a[k] = pos[k] - n[k] * 0.02f;
  if (jx != 0.0f || jy != 0.0f) {
    float up[3] = {1.0f, 0.0f, 0.0f};
    if (fabsf(n[2]) < 0.999f) {
      up[0] = 0.0f;
      up[2] = 1.0f;
    }
    float t[3], b[3];
    sh_cross(up, n, t);
    float tl = t[0] * t[0] + t[1] * t[1];
    tl = sqrtf(tl + t[2] * t[2]);
    for (int k = 0; k < 3; ++k) t[k] = t[k] / tl;
    sh_cross(n, t, b);
    float sc = 0.42f * s;
    for (int k = 0; k < 3; ++k) {
      float w = t[k] * jx + b[k] * jy;
      a[k] = a[k] + w * sc;
    }
  }
  uint32_t c[3];
  for (int k = 0; k < 3; ++k) c[k] = (uint32_t)((int32_t)floorf(a[k] / s) + 16384);
  uint32_t nid = sh_norm_id(n);
  uint32_t bk = pcg_hash(c[0] * 73856093u ^ c[1] * 19349663u ^ c[2] * 83492791u ^ nid * 2654435761u ^ lod * 38069u) &
                bucket_mask;
  *bucket = bk;
  *key = pcg_hash(bk ^ (c[0] * 1013904223u) ^ (c[1] * 1664525u) ^ (c[2] * 2891336453u) ^ (nid << 20) ^ (lod << 24)) |
         0x80000000u;
}

void orc_sharc_deposit(uint32_t log2_buckets, uint32_t* keys, uint32_t* accum, const arc_sharc_sample* s, uint64_t n,
                       const float cam[3]) {
  uint32_t mask = (1u << log2_buckets) - 1u;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t bk, key, slot = 0xFFFFFFFFu;
    orc_sharc_make_key(s[i].pos, s[i].normal, cam, 0.0f, 0.0f, mask, &bk, &key);
    for (uint32_t w = 0; w < 4 && slot == 0xFFFFFFFFu; ++w)
      if (keys[bk * 4 + w] == key) slot = bk * 4 + w;
    for (uint32_t w = 0; w < 4 && slot == 0xFFFFFFFFu; ++w)
      if (keys[bk * 4 + w] == 0u) {
        keys[bk * 4 + w] = key;
        slot = bk * 4 + w;
      }
    if (slot == 0xFFFFFFFFu) continue;
    for (int c = 0; c < 3; ++c) {
      float v = clampf_(s[i].radiance[c], 0.0f, 32.0f) * 1024.0f;
      accum[slot * 4 + c] += f2u(v);
    }
    accum[slot * 4 + 3] += 1u;
  }
}

void orc_sharc_resolve(uint32_t n_slots, uint32_t* keys, uint32_t* accum, uint32_t* resolved) {
  for (uint32_t slot = 0; slot < n_slots; ++slot) {
    uint32_t key = keys[slot], *rs = resolved + 4 * slot, *ac = accum + 4 * slot;
    if (key == 0u) {
      rs[0] = rs[1] = rs[2] = rs[3] = 0u;
      continue;
    }
    int prev_valid = rs[0] == key;
    float rad[3] = {0.0f, 0.0f, 0.0f};
    uint32_t hist = 0, age = 0;
    if (prev_valid) {
      rad[0] = orc_f16_to_f32((uint16_t)(rs[1] & 0xFFFFu));
      rad[1] = orc_f16_to_f32((uint16_t)(rs[1] >> 16));
      rad[2] = orc_f16_to_f32((uint16_t)(rs[2] & 0xFFFFu));
      hist = (rs[2] >> 16) & 0xFFu;
      age = (rs[2] >> 24) & 0xFFu;
    }
    if (ac[3] > 0u) {
      float inv = 1.0f / ((float)ac[3] * 1024.0f);
      hist = hist + 1u < 255u ? hist + 1u : 255u;
      float alpha = 1.0f / fminf((float)hist, 16.0f);
      for (int c = 0; c < 3; ++c) {
        float smp = (float)ac[c] * inv;
        rad[c] = mix_(rad[c], smp, alpha);
      }
      age = 0;
      ac[0] = ac[1] = ac[2] = ac[3] = 0u;
    } else {
      age = age + 1u;
      if (!prev_valid || age > 64u) {
        keys[slot] = 0u;
        rs[0] = rs[1] = rs[2] = rs[3] = 0u;
        continue;
      }
    }
    float c0 = clampf_(rad[0], 0.0f, 32.0f), c1 = clampf_(rad[1], 0.0f, 32.0f), c2 = clampf_(rad[2], 0.0f, 32.0f);
    rs[0] = key;
    rs[1] = (uint32_t)orc_f32_to_f16(c0) | ((uint32_t)orc_f32_to_f16(c1) << 16);
    rs[2] = (uint32_t)orc_f32_to_f16(c2) | ((hist & 0xFFu) << 16) | ((age & 0xFFu) << 24);
    rs[3] = 0u;
  }
}

static void sh_lookup(const uint32_t* resolved, uint32_t bk, uint32_t key, float out[4]) {
  out[0] = out[1] = out[2] = out[3] = 0.0f;
  for (uint32_t w = 0; w < 4; ++w) {
    const uint32_t* rs = resolved + 4 * (bk * 4 + w);
    if (rs[0] == key) {
      out[0] = orc_f16_to_f32((uint16_t)(rs[1] & 0xFFFFu));
      out[1] = orc_f16_to_f32((uint16_t)(rs[1] >> 16));
      out[2] = orc_f16_to_f32((uint16_t)(rs[2] & 0xFFFFu));
      out[3] = 1.0f;
      return;
    }
  }
}

void orc_sharc_query(uint32_t log2_buckets, const uint32_t* resolved, const arc_sharc_probe* q, uint64_t n,
                     const float cam[3], float* out) {
  uint32_t mask = (1u << log2_buckets) - 1u;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t bk, key;
    float* o = out + 4 * i;
    if (q[i].jitter[0] != 0.0f || q[i].jitter[1] != 0.0f) {
      orc_sharc_make_key(q[i].pos, q[i].normal, cam, q[i].jitter[0], q[i].jitter[1], mask, &bk, &key);
      sh_lookup(resolved, bk, key, o);
      if (o[3] > 0.5f) continue;
    }
    orc_sharc_make_key(q[i].pos, q[i].normal, cam, 0.0f, 0.0f, mask, &bk, &key);
    sh_lookup(resolved, bk, key, o);
  }
}
