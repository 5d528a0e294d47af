/*
 * CPU oracle of the ReSTIR GI / DI passes — TEST INFRASTRUCTURE ONLY.
 *
 * temporal_main (shaders/lighting_restir_gi.wgsl:62-455) and spatial_main (:457-762) restated one
 * pixel at a time, in the shader's own shape: each secondary ray is traced inline, when the shader
 * traces it (the bounce-sun shadow ray only on a SHARC miss, the rigid-body walk only with bodies,
 * the DI visibility ray only when not skipped). Helpers follow shaders/common/types.wgsl and
 * shaders/common/di.wgsl. Binary32 arithmetic in source order, no contraction; where WGSL leaves it
 * open: dot = x + y + z left to right, length = sqrt(dot), normalize(v) = v / length(v),
 * mix(a, b, t) = a (1 - t) + b t, round = to nearest even, sin / cos = the Cody-Waite + Taylor
 * polynomial the product's ray generators use, f32 -> f16 to nearest even, conversions to integers
 * truncate (float -> u32 saturating).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <string.h>

#include "arc_oracle.h"

typedef struct {
  float x, y, z;
} v3f;

static v3f V(float x, float y, float z) {
  v3f r = {x, y, z};
  return r;
}
static v3f vadd(v3f a, v3f b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3f vsub(v3f a, v3f b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3f vscale(v3f a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static v3f vmulv(v3f a, v3f b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static v3f vdivs(v3f a, float s) { return V(a.x / s, a.y / s, a.z / s); }
static float vdot(v3f a, v3f b) {
  float x = a.x * b.x, y = a.y * b.y, z = a.z * b.z;
  return (x + y) + z;
}
static float vlen(v3f a) { return sqrtf(vdot(a, a)); }
static v3f vnorm(v3f a) { return vdivs(a, vlen(a)); }
static v3f vcross(v3f a, v3f b) {
  float x0 = a.y * b.z, x1 = a.z * b.y, y0 = a.z * b.x, y1 = a.x * b.z, z0 = a.x * b.y, z1 = a.y * b.x;
  return V(x0 - x1, y0 - y1, z0 - z1);
}
static v3f vload(const float* p) { return V(p[0], p[1], p[2]); }
static float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
static float mixf(float a, float b, float t) {
  float u = a * (1.0f - t), v = b * t;
  return u + v;
}
static v3f vmix(v3f a, v3f b, float t) { return V(mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)); }
static float fractf(float x) { return x - floorf(x); }
static uint32_t f2u(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}
static uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
static uint32_t pcg(uint32_t input) {
  uint32_t state = input * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}
/* rand_f32 (types.wgsl:264-267) */
static float rnd(uint32_t* seed) {
  *seed = pcg(*seed);
  return (float)*seed / 4294967295.0f;
}
/* luma (types.wgsl:183-185) */
static float luma(v3f c) { return vdot(c, V(0.2126f, 0.7152f, 0.0722f)); }

/* sin and cos of 0 <= x < 8 (Cody-Waite by pi/2 in two parts, Taylor polynomials). */
static void sincos8(float x, float* s, float* c) {
  const float kf = floorf(x * 0.636619772f + 0.5f);
  const int k = (int)kf;
  const float a = kf * 1.5703125f, b = kf * 4.83826794897e-4f;
  const float r = (x - a) - b;
  const float r2 = r * r;
  float ps = 2.75573188e-06f;
  ps = ps * r2 + -1.98412701e-04f;
  ps = ps * r2 + 8.33333377e-03f;
  ps = ps * r2 + -1.66666672e-01f;
  const float sr = r + (r * r2) * ps;
  float pc = 2.48015876e-05f;
  pc = pc * r2 + -1.38888892e-03f;
  pc = pc * r2 + 4.16666679e-02f;
  pc = pc * r2 + -0.5f;
  const float cr = 1.0f + r2 * pc;
  switch (k & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}

static float h2f(uint32_t b) { return orc_f16_to_f32((uint16_t)(b & 0xFFFFu)); }
static uint32_t f2h(float x) { return (uint32_t)orc_f32_to_f16(x); }
static uint32_t pack2x16float(float a, float b) { return f2h(a) | (f2h(b) << 16); }

/* oct_encode / oct_decode (types.wgsl:108-130) */
static void oct_wrap(float px, float py, float* ox, float* oy) {
  *ox = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
  *oy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
}
static void oct_encode(v3f n, float* px, float* py) {
  float l1 = (fabsf(n.x) + fabsf(n.y)) + fabsf(n.z);
  float inv = 1.0f / fmaxf(l1, 1e-6f);
  float x = n.x * inv, y = n.y * inv;
  if (n.z < 0.0f) oct_wrap(x, y, &x, &y);
  *px = x;
  *py = y;
}
static v3f oct_decode(float px, float py) {
  v3f n = V(px, py, (1.0f - fabsf(px)) - fabsf(py));
  if (n.z < 0.0f) oct_wrap(px, py, &n.x, &n.y);
  return vnorm(n);
}
/* unpack_normal_oct (types.wgsl:313-316) with unpack2x16snorm */
static float snorm16(uint32_t b) { return clampf((float)(int16_t)(uint16_t)(b & 0xFFFFu) / 32767.0f, -1.0f, 1.0f); }
static v3f unpack_normal_oct(uint32_t p) { return oct_decode(snorm16(p), snorm16(p >> 16)); }

/* GiReservoir packing (types.wgsl:132-196) */
static uint32_t pack_norm_m_flags(v3f n, uint32_t m, uint32_t age, int sky) {
  float ox, oy;
  oct_encode(n, &ox, &oy);
  int ix = (int)rintf(ox * 127.0f), iy = (int)rintf(oy * 127.0f);
  ix = (ix < -128 ? -128 : (ix > 127 ? 127 : ix)) + 128;
  iy = (iy < -128 ? -128 : (iy > 127 ? 127 : iy)) + 128;
  return ((uint32_t)ix & 0xFFu) | (((uint32_t)iy & 0xFFu) << 8) | ((umin(m, 16u) & 0xFFu) << 16) |
         ((umin(age, 127u) & 0x7Fu) << 24) | (sky ? 0x80000000u : 0u);
}
static v3f unpack_normal(uint32_t f) {
  float ux = (float)(f & 0xFFu), uy = (float)((f >> 8) & 0xFFu);
  return oct_decode((ux - 128.0f) / 127.0f, (uy - 128.0f) / 127.0f);
}
static uint32_t unpack_m(uint32_t f) { return (f >> 16) & 0xFFu; }
static uint32_t unpack_sample_age(uint32_t f) { return (f >> 24) & 0x7Fu; }
static int unpack_is_sky(uint32_t f) { return (f & 0x80000000u) != 0u; }
static void pack_lo_w(v3f lo, float w, uint32_t* rg, uint32_t* bw) {
  float r = clampf(lo.x, 0.0f, 65500.0f), g = clampf(lo.y, 0.0f, 65500.0f), b = clampf(lo.z, 0.0f, 65500.0f);
  *rg = pack2x16float(r, g);
  *bw = pack2x16float(b, fminf(w, 65500.0f));
}
static v3f unpack_lo(uint32_t rg, uint32_t bw) { return V(h2f(rg), h2f(rg >> 16), h2f(bw)); }
static float unpack_w(uint32_t bw) { return h2f(bw >> 16); }
static arc_gi_reservoir create_empty_reservoir(void) {
  arc_gi_reservoir r;
  memset(&r, 0, sizeof r);
  r.primary_depth = 1e5f;
  return r;
}
static arc_di_reservoir create_empty_di_reservoir(void) {
  arc_di_reservoir r;
  r.light_idx_seed = 0xFFFFFFFFu;
  r.w_sum = 0.0f;
  r.m = 0.0f;
  r.w = 0.0f;
  return r;
}

/* interleaved_gradient_noise (types.wgsl:483-497) */
static void interleaved_gradient_noise(float px, float py, uint32_t frame, float* u1, float* u2) {
  float f = (float)frame;
  uint32_t ux = f2u(px), uy = f2u(py);
  float ign1 = fractf(52.9829189f * fractf(0.06711056f * px + 0.00583715f * py));
  float ign2 = fractf(52.9829189f * fractf(0.00583715f * px - 0.06711056f * py + 0.5f));
  uint32_t h1 = pcg(ux * 1973u + uy * 9277u + frame * 26699u);
  uint32_t h2 = pcg(h1 ^ 0x9E3779B9u);
  float r1 = (float)(h1 & 0xFFFFu) / 65536.0f;
  float r2 = (float)(h2 & 0xFFFFu) / 65536.0f;
  *u1 = fractf(ign1 + f * 0.754877666f + r1);
  *u2 = fractf(ign2 + f * 0.569840291f + r2);
}

/* eval_sky_radiance (types.wgsl:500-523); pow by repeated products */
static v3f eval_sky_radiance(v3f dir, v3f sun) {
  float sun_elev = sun.y;
  float day_t = clampf((sun_elev + 0.05f) / 0.30f, 0.0f, 1.0f);
  float twilight_t = clampf(1.0f - fabsf(sun_elev - 0.02f) / 0.22f, 0.0f, 1.0f);
  float zenith = clampf(dir.y, 0.0f, 1.0f);
  float om = 1.0f - zenith;
  float horizon = (om * om) * om;
  v3f sky_z = vmix(V(0.0015f, 0.0025f, 0.0055f), V(0.08f, 0.22f, 0.55f), day_t);
  v3f sky_h = vmix(V(0.0035f, 0.0065f, 0.0120f), V(0.38f, 0.50f, 0.68f), day_t);
  sky_h = vmix(sky_h, V(0.65f, 0.25f, 0.10f), twilight_t * 0.75f);
  v3f sky = vmix(sky_z, sky_h, horizon);
  float sun_cos = fmaxf(vdot(dir, sun), 0.0f);
  if (sun_elev > -0.12f) {
    float x2 = sun_cos * sun_cos, x4 = x2 * x2, x8 = x4 * x4;
    float p14 = (x8 * x4) * x2;
    sky = vadd(sky, vscale(vscale(vscale(V(0.85f, 0.52f, 0.24f), p14), 0.35f), fmaxf(day_t, twilight_t)));
  }
  return sky;
}

/* make_basis / sample_cosine_hemisphere (types.wgsl:531-545) */
static void make_basis(v3f n, v3f* t, v3f* b) {
  v3f up = fabsf(n.z) < 0.999f ? V(0.0f, 0.0f, 1.0f) : V(1.0f, 0.0f, 0.0f);
  *t = vnorm(vcross(up, n));
  *b = vcross(n, *t);
}
static v3f sample_cosine_hemisphere(float u1, float u2, v3f n, v3f t, v3f b) {
  float r = sqrtf(u1), phi = 6.2831853f * u2, z = sqrtf(fmaxf(0.0f, 1.0f - u1));
  float s, c;
  sincos8(phi, &s, &c);
  return vadd(vadd(vscale(t, r * c), vscale(b, r * s)), vscale(n, z));
}

/* ---- di.wgsl:42-115 ---------------------------------------------------------------------- */
static void di_unpack_normal_and_eta(uint32_t np, v3f* normal, float* eta) {
  float nx = (float)(np & 1023u) / 1023.0f * 2.0f - 1.0f;
  float ny = (float)((np >> 10) & 1023u) / 1023.0f * 2.0f - 1.0f;
  float nz = (float)((np >> 20) & 1023u) / 1023.0f * 2.0f - 1.0f;
  v3f n = V(nx, ny, nz);
  float len = vlen(n);
  *eta = clampf(len, 0.0f, 1.0f);
  *normal = len > 0.05f ? vdivs(n, fmaxf(len, 1e-4f)) : V(0.0f, 1.0f, 0.0f);
}
static v3f di_emitter_surface_pos(v3f centroid, v3f normal, float radius, float eta, uint32_t seed) {
  if (radius <= 0.04f) return centroid;
  float u1 = (float)(seed & 0xFFu) / 255.0f, u2 = (float)((seed >> 8) & 0xFFu) / 255.0f;
  float max_r = fminf(radius * 0.20f, 0.08f);
  float s, c;
  if (eta >= 0.35f) {
    v3f t, b;
    make_basis(normal, &t, &b);
    float r = sqrtf(u1) * max_r;
    sincos8(6.2831853f * u2, &s, &c);
    return vadd(centroid, vadd(vscale(t, c * r), vscale(b, s * r)));
  }
  float cos_theta = u2 * 2.0f - 1.0f;
  float sin_theta = sqrtf(fmaxf(0.0f, 1.0f - cos_theta * cos_theta));
  sincos8(6.2831853f * u1, &s, &c);
  return vadd(centroid, vscale(V(c * sin_theta, cos_theta, s * sin_theta), max_r));
}
static v3f di_eval_unshadowed(const arc_emissive_source* em, uint32_t seed, v3f x_v, v3f n_v) {
  v3f un_n;
  float eta;
  di_unpack_normal_and_eta(em->normal_packed, &un_n, &eta);
  v3f em_pos = di_emitter_surface_pos(vload(em->pos), un_n, em->radius, eta, seed);
  v3f to_l = vsub(em_pos, x_v);
  float d2 = vdot(to_l, to_l);
  if (d2 < 1e-6f || d2 > 1296.0f) return V(0.0f, 0.0f, 0.0f);
  float d = sqrtf(d2);
  v3f l_dir = vdivs(to_l, d);
  float cos_theta_v = fmaxf(vdot(n_v, l_dir), 0.0f);
  if (cos_theta_v <= 0.0f) return V(0.0f, 0.0f, 0.0f);
  float dir_factor = (1.0f - eta) + eta * fmaxf(vdot(un_n, V(-l_dir.x, -l_dir.y, -l_dir.z)), 0.0f);
  float range_fade = clampf(1.0f - d2 / 1296.0f, 0.0f, 1.0f);
  float geom = (cos_theta_v * dir_factor * 0.31830989f * range_fade * range_fade) / fmaxf(d2, 0.20f);
  return vscale(vload(em->intensity), geom);
}

/* ---- helpers of the passes ----------------------------------------------------------------- */
typedef struct {
  const orc_world* w;
  const orc_restir_in* in;
  arc_gi_reservoir* gi_out;
  arc_di_reservoir* di_out;
  uint32_t first, last;
} job;

static v3f reconstruct_ray_dir(const arc_restir_params* p, uint32_t px, uint32_t py) { /* types.wgsl:334-345 */
  float uvx = ((float)px + 0.5f + p->cam.jitter[0]) / (float)p->width;
  float uvy = ((float)py + 0.5f + p->cam.jitter[1]) / (float)p->height;
  float ndx = uvx * 2.0f - 1.0f, ndy = 1.0f - uvy * 2.0f;
  v3f f = vload(p->cam.forward), r = vload(p->cam.right), u = vload(p->cam.up);
  return vnorm(vadd(vadd(f, vscale(r, ndx * p->cam.aspect * p->cam.tan_half_fov)), vscale(u, ndy * p->cam.tan_half_fov)));
}

static arc_ray mkray(v3f o, v3f d, float t_max) {
  arc_ray r;
  r.ox = o.x;
  r.oy = o.y;
  r.oz = o.z;
  r.t_start = 0.0f;
  r.dx = d.x;
  r.dy = d.y;
  r.dz = d.z;
  r.t_max = t_max;
  return r;
}

/* trace_world_64tree(origin, dir, max, lod_bias, ignore_water) on the given traversal parameters */
static void trace_world(const orc_world* w, const arc_trace_params* base, uint32_t lod_bias, v3f o, v3f d,
                        float max_d, arc_trace_result* res) {
  arc_trace_params tp = *base;
  tp.semantics = ARC_SEM_WORLD;
  tp.lod_bias = lod_bias;
  tp.ignore_water = 1u;
  arc_ray r = mkray(o, d, max_d);
  orc_trace_world_ray(w, &tp, &r, res);
}
static float trace_shadow_vis(const orc_world* w, const arc_restir_params* p, v3f o, v3f d, float max_d) {
  arc_ray r = mkray(o, d, max_d);
  return orc_shadow_ray(w, ARC_SHADOW_SUN, p->shadow_max_steps, &r);
}
static float trace_shadow_vis_di(const orc_world* w, const arc_restir_params* p, v3f o, v3f d, float max_d) {
  arc_ray r = mkray(o, d, max_d);
  return orc_shadow_ray(w, ARC_SHADOW_DI, p->shadow_max_steps, &r);
}

/* probe_emissive_macro_candidate (lighting_restir_gi.wgsl:19-56) */
static float probe_emissive_macro_candidate(const orc_world* w, const float* palette, v3f pos_m, v3f cand_dir,
                                            float probe_dist_m) {
  v3f vp = vscale(vadd(pos_m, vscale(cand_dir, probe_dist_m)), 16.0f);
  float wmx = (float)w->grid[0] * 256.0f, wmy = (float)w->grid[1] * 256.0f, wmz = (float)w->grid[2] * 256.0f;
  if (vp.x < 0.0f || vp.y < 0.0f || vp.z < 0.0f || vp.x >= wmx || vp.y >= wmy || vp.z >= wmz) return 0.0f;
  uint32_t vi[3] = {f2u(vp.x), f2u(vp.y), f2u(vp.z)};
  uint32_t c_idx = (vi[0] >> 8) + (vi[1] >> 8) * w->grid[0] + (vi[2] >> 8) * w->grid[0] * w->grid[1];
  uint32_t root_idx = w->roots[c_idx];
  if (root_idx == ARC_EMPTY_CHUNK_SENTINEL) return 0.0f;
  uint32_t l[3] = {vi[0] & 255u, vi[1] & 255u, vi[2] & 255u};
  const arc_node64* root = &w->nodes[root_idx];
  uint32_t b0 = (((l[2] >> 6) & 3u) << 4) | (((l[1] >> 6) & 3u) << 2) | ((l[0] >> 6) & 3u);
  uint64_t m0 = (uint64_t)root->mask_low | ((uint64_t)root->mask_high << 32);
  if (!((m0 >> b0) & 1u)) return 0.0f;
  uint32_t m_id;
  if (root->child_ptr & ARC_IS_LEAF_FLAG) {
    m_id = (root->lod_payload >> 24) & 0xFFu;
  } else {
    uint32_t l1_idx = (root->child_ptr & ARC_PTR_INDEX_MASK) +
                      (uint32_t)__builtin_popcountll(m0 & ((b0 == 0u) ? 0ull : ((1ull << b0) - 1ull)));
    const arc_node64* l1 = &w->nodes[l1_idx];
    uint32_t b1 = (((l[2] >> 4) & 3u) << 4) | (((l[1] >> 4) & 3u) << 2) | ((l[0] >> 4) & 3u);
    uint64_t m1 = (uint64_t)l1->mask_low | ((uint64_t)l1->mask_high << 32);
    if (!((m1 >> b1) & 1u)) return 0.0f;
    m_id = (l1->lod_payload >> 24) & 0xFFu;
  }
  const float* m = palette + 8u * m_id;
  return vdot(V(m[4], m[5], m[6]), V(0.333f, 0.333f, 0.333f));
}

/* query_sharc (sharc_query.wgsl:20-29) over the given table */
static void query_sharc(const orc_restir_in* in, v3f pos, v3f n, const float cam[3], float jx, float jy, float out[4]) {
  out[0] = out[1] = out[2] = out[3] = 0.0f;
  if (!in->sharc_resolved) return;
  arc_sharc_probe q;
  q.pos[0] = pos.x;
  q.pos[1] = pos.y;
  q.pos[2] = pos.z;
  q.normal[0] = n.x;
  q.normal[1] = n.y;
  q.normal[2] = n.z;
  q.jitter[0] = jx;
  q.jitter[1] = jy;
  orc_sharc_query(in->sharc_log2_buckets, in->sharc_resolved, &q, 1, cam, out);
}

/* ---- temporal_main (lighting_restir_gi.wgsl:62-455) -------------------------------------------- */
static void temporal_pixel(const job* j, uint32_t gx, uint32_t gy) {
  const orc_world* w = j->w;
  const orc_restir_in* in = j->in;
  const arc_restir_params* u = in->p;
  const uint32_t W = u->width, H = u->height;
  const uint32_t p_idx = gy * W + gx;
  const float hit_dist = ((const float*)in->depth_mat)[2 * p_idx];
  if (hit_dist >= 1e4f) {
    j->gi_out[p_idx] = create_empty_reservoir();
    j->di_out[p_idx] = create_empty_di_reservoir();
    return;
  }
  v3f ray_dir = reconstruct_ray_dir(u, gx, gy);
  v3f hit_pos = vadd(vload(u->cam.eye), vscale(ray_dir, hit_dist));
  uint32_t nm0 = in->norm_motion[2 * p_idx], nm1 = in->norm_motion[2 * p_idx + 1];
  v3f normal = unpack_normal_oct(nm0);
  float mot_x = h2f(nm1), mot_y = h2f(nm1 >> 16);
  float puvx = ((float)gx + 0.5f) / (float)W, puvy = ((float)gy + 0.5f) / (float)H;
  float prev_uvx = puvx + mot_x, prev_uvy = puvy + mot_y;
  uint32_t frame_idx = u->frame_index;
  float xi0, xi1;
  interleaved_gradient_noise((float)gx, (float)gy, frame_idx, &xi0, &xi1);
  v3f sun_dir = vload(u->sun_dir);
  float sun_power = u->sun_intensity;
  float n_dot_l = fmaxf(vdot(normal, sun_dir), 0.0f);

  int prev_valid = 0;
  arc_gi_reservoir prev_res;
  memset(&prev_res, 0, sizeof prev_res); /* WGSL zero-initialises the var */
  uint32_t prev_idx = 0u;
  if (prev_uvx >= 0.0f && prev_uvx < 1.0f && prev_uvy >= 0.0f && prev_uvy < 1.0f) {
    uint32_t prev_px = umin(f2u(prev_uvx * (float)W), W - 1u);
    uint32_t prev_py = umin(f2u(prev_uvy * (float)H), H - 1u);
    prev_idx = prev_py * W + prev_px;
    prev_res = in->gi_in[prev_idx];
    uint32_t prev_m = unpack_m(prev_res.norm_m_flags), prev_age = unpack_sample_age(prev_res.norm_m_flags);
    if (prev_m > 0u && prev_age <= 12u && prev_res.primary_depth < 1e4f) {
      float prev_cam_dist = vlen(vsub(hit_pos, vload(u->prev_eye)));
      float depth_tol = fmaxf(0.12f, hit_dist * 0.05f);
      if (fabsf(prev_cam_dist - prev_res.primary_depth) < depth_tol) prev_valid = 1;
    }
  }
  uint32_t denoiser_taps = u->denoiser_taps;
  int checkerboard_active = denoiser_taps <= 12u;
  uint32_t checkerboard_slot = (gx & 1u) ^ ((gy & 1u) << 1);
  uint32_t frame_slot = frame_idx & 3u;
  int skip = checkerboard_active && prev_valid && (((checkerboard_slot ^ frame_slot) & 1u) != 0u);

  float sun_vis = 0.0f, shadow_m = 1.0f, accum_sun_shadow = 0.0f;
  float psx = h2f(prev_res.shadow_vis_m), psm = h2f(prev_res.shadow_vis_m >> 16);
  if (skip) {
    accum_sun_shadow = psx;
    shadow_m = fminf(psm, 16.0f);
  } else {
    if (n_dot_l > 0.001f && sun_power > 0.01f) {
      float j1 = xi0 * 2.0f - 1.0f, j2 = xi1 * 2.0f - 1.0f;
      v3f sun_up = fabsf(sun_dir.y) < 0.9f ? V(0.0f, 1.0f, 0.0f) : V(1.0f, 0.0f, 0.0f);
      v3f s_tan = vnorm(vcross(sun_up, sun_dir));
      v3f s_bitan = vcross(sun_dir, s_tan);
      v3f jittered_sun = vnorm(vadd(sun_dir, vscale(vadd(vscale(s_tan, j1), vscale(s_bitan, j2)), u->shadow_jitter)));
      v3f shadow_origin = vadd(hit_pos, vscale(normal, 0.08f));
      float shadow_max_d = u->shadow_max_dist;
      if (u->shadow_mode == 0u) {
        arc_trace_result st;
        trace_world(w, in->sun_trace, 1u, shadow_origin, jittered_sun, shadow_max_d, &st);
        if (!(st.hit & 1u)) {
          v3f end_p = vadd(shadow_origin, vscale(jittered_sun, shadow_max_d));
          if (end_p.y >= 32.0f || jittered_sun.y < 0.0f) sun_vis = 1.0f;
        }
      } else {
        sun_vis = trace_shadow_vis(w, u, shadow_origin, jittered_sun, shadow_max_d);
      }
    }
    if (prev_valid) {
      float prev_shd_m = fminf(psm, 16.0f);
      shadow_m = prev_shd_m + 1.0f;
      accum_sun_shadow = (psx * prev_shd_m + sun_vis) / shadow_m;
    } else {
      accum_sun_shadow = sun_vis;
      shadow_m = 1.0f;
    }
  }
  if (skip) {
    arc_gi_reservoir out_res = prev_res;
    out_res.primary_depth = hit_dist;
    out_res.shadow_vis_m = pack2x16float(accum_sun_shadow, shadow_m);
    uint32_t f = prev_res.norm_m_flags;
    out_res.norm_m_flags = pack_norm_m_flags(unpack_normal(f), unpack_m(f), unpack_sample_age(f) + 1u, unpack_is_sky(f));
    j->gi_out[p_idx] = out_res;
    j->di_out[p_idx] = prev_valid ? in->di_in[prev_idx] : create_empty_di_reservoir();
    return;
  }

  /* 2. initial candidate */
  v3f smooth_hemi_n = vnorm(vadd(normal, V(0.0f, 0.22f, 0.0f)));
  v3f bt, bb;
  make_basis(smooth_hemi_n, &bt, &bb);
  v3f bounce_dir = sample_cosine_hemisphere(xi0, xi1, smooth_hemi_n, bt, bb);
  v3f cand_dir = sample_cosine_hemisphere(fractf(xi0 + 0.5f), fractf(xi1 + 0.381966f), smooth_hemi_n, bt, bb);
  float em_near = probe_emissive_macro_candidate(w, in->palette, hit_pos, cand_dir, 1.25f);
  float em_mid = probe_emissive_macro_candidate(w, in->palette, hit_pos, cand_dir, 3.0f);
  int guided_pick = fmaxf(em_near, em_mid) > 0.25f;
  if (guided_pick) bounce_dir = cand_dir;
  v3f gi_origin = vadd(hit_pos, vscale(normal, 0.10f));
  arc_trace_result gi;
  trace_world(w, in->gi_trace, 2u, gi_origin, bounce_dir, u->gi_max_dist, &gi);
  int gi_hit = (gi.hit & 1u) != 0u;
  float gi_t = gi.t_voxels;
  v3f gi_pos = vload(gi.hit_pos_m), gi_n = vload(gi.normal), gi_lod = vload(gi.lod_rgb);
  uint32_t gi_mat = gi.mat;
  float rb_gi_max = gi_hit ? gi.t_voxels * 0.0625f : u->gi_max_dist;
  if (in->n_bodies > 0u) {
    arc_rigid_hit rb;
    arc_ray rr = mkray(gi_origin, bounce_dir, rb_gi_max);
    orc_rigid_ray(w, in->bodies, in->n_bodies, &rr, &rb);
    if (rb.hit) {
      gi_hit = 1;
      gi_t = rb.t_voxels;
      gi_pos = vload(rb.hit_pos_m);
      gi_n = vload(rb.normal);
      gi_mat = rb.mat;
      gi_lod = vload(in->palette + 8u * (rb.mat & 0xFFu));
    }
  }
  v3f cand_pos = vadd(hit_pos, vscale(bounce_dir, 500.0f));
  v3f cand_norm = V(-bounce_dir.x, -bounce_dir.y, -bounce_dir.z);
  int cand_is_sky = 1;
  v3f cand_lo = V(0.0f, 0.0f, 0.0f);
  if (gi_hit) {
    cand_pos = gi_pos;
    cand_norm = gi_n;
    cand_is_sky = 0;
    const float* hm = in->palette + 8u * (gi_mat & 0xFFu);
    v3f hit_albedo = vmix(V(hm[0], hm[1], hm[2]), gi_lod, 0.55f);
    v3f hit_emission = V(hm[4], hm[5], hm[6]);
    float gi_dist_m = gi_t * 0.0625f;
    float em_atten = 1.0f / (1.0f + gi_dist_m * 0.06f + gi_dist_m * gi_dist_m * 0.025f);
    float guided_weight = guided_pick ? 0.55f : 1.0f;
    float di_dedup = u->n_emissive > 0u ? 0.25f : 1.0f;
    cand_lo = vadd(cand_lo, vscale(vscale(vscale(hit_emission, em_atten), guided_weight), di_dedup));
    float jx, jy, sh[4];
    interleaved_gradient_noise((float)gx + 101.0f, (float)gy + 57.0f, frame_idx + 7919u, &jx, &jy);
    query_sharc(in, gi_pos, gi_n, u->cam.eye, jx * 2.0f - 1.0f, jy * 2.0f - 1.0f, sh);
    if (sh[3] > 0.5f) {
      cand_lo = vadd(cand_lo, V(sh[0], sh[1], sh[2]));
    } else {
      float b_ndl = fmaxf(vdot(gi_n, sun_dir), 0.0f);
      v3f bounce_sun = V(0.0f, 0.0f, 0.0f);
      if (b_ndl > 0.001f && sun_power > 0.01f) {
        float bvis = trace_shadow_vis(w, u, vadd(gi_pos, vscale(gi_n, 0.10f)), sun_dir, 32.0f);
        if (bvis > 0.001f) bounce_sun = vscale(vscale(vscale(vscale(V(1.0f, 0.90f, 0.75f), sun_power), b_ndl), bvis), 0.55f);
      }
      cand_lo = vadd(cand_lo, vmulv(hit_albedo, bounce_sun));
    }
    float ao = clampf(gi_dist_m / 0.40f, 0.25f, 1.0f);
    cand_lo = vscale(cand_lo, ao);
  } else {
    v3f end_p = vadd(gi_origin, vscale(bounce_dir, u->gi_max_dist));
    if (bounce_dir.y > 0.15f && end_p.y >= 32.0f) {
      float sky_escape = clampf((bounce_dir.y - 0.15f) / 0.35f, 0.0f, 1.0f);
      cand_lo = vadd(cand_lo, vscale(vscale(eval_sky_radiance(bounce_dir, sun_dir), 0.45f), sky_escape));
    }
  }
  float gi_peak = fmaxf(cand_lo.x, fmaxf(cand_lo.y, cand_lo.z));
  if (gi_peak > 0.001f) cand_lo = vdivs(cand_lo, 1.0f + gi_peak * 0.16f);

  /* 3. temporal RIS */
  float cand_cos = fmaxf(0.0f, vdot(smooth_hemi_n, bounce_dir));
  float cand_p_hat = luma(cand_lo) * (cand_cos / 3.14159265f);
  float cand_w = cand_p_hat > 1e-5f ? luma(cand_lo) : 0.0f;
  float w_sum = cand_w;
  v3f selected_pos = cand_pos, selected_norm = cand_norm, selected_lo = cand_lo;
  float selected_p_hat = cand_p_hat;
  uint32_t selected_age = 0u, m_total = 1u;
  int selected_is_sky = cand_is_sky;
  if (prev_valid) {
    uint32_t max_m = u->target_valid ? 8u : 12u;
    uint32_t prev_m = umin(unpack_m(prev_res.norm_m_flags), max_m);
    float prev_w = unpack_w(prev_res.lo_b_w);
    v3f prev_lo = unpack_lo(prev_res.lo_rg, prev_res.lo_b_w);
    v3f prev_norm = unpack_normal(prev_res.norm_m_flags);
    uint32_t prev_age = unpack_sample_age(prev_res.norm_m_flags);
    int prev_is_sky = unpack_is_sky(prev_res.norm_m_flags);
    v3f to_prev = vsub(vload(prev_res.sample_pos), hit_pos);
    float d_prev = vlen(to_prev);
    float prev_p_hat = 0.0f;
    if (d_prev > 1e-4f) {
      v3f dir_prev = vdivs(to_prev, d_prev);
      float prev_cos = fmaxf(0.0f, vdot(smooth_hemi_n, dir_prev));
      prev_p_hat = luma(prev_lo) * (prev_cos / 3.14159265f);
    }
    uint32_t effective_prev_m = umin(prev_m, max_m - 1u);
    float prev_weight = prev_p_hat > 1e-5f ? prev_p_hat * prev_w * (float)effective_prev_m : 0.0f;
    w_sum += prev_weight;
    m_total = effective_prev_m + 1u;
    uint32_t seed = pcg(gx + pcg(gy + pcg(frame_idx + 0x9E3779B9u)));
    int pick_prev = (rnd(&seed) * w_sum) < prev_weight;
    if (pick_prev) {
      selected_pos = vload(prev_res.sample_pos);
      selected_norm = prev_norm;
      selected_lo = prev_lo;
      selected_p_hat = prev_p_hat;
      selected_age = prev_age + 1u;
      selected_is_sky = prev_is_sky;
    }
    float blend_t = clampf(prev_weight / fmaxf(w_sum, 1e-5f), 0.0f, 1.0f);
    v3f mixed_lo = vmix(cand_lo, prev_lo, blend_t);
    float winner_luma = luma(selected_lo), mixed_luma = luma(mixed_lo);
    if (mixed_luma > 1e-5f && winner_luma > 1e-5f) selected_lo = vscale(mixed_lo, winner_luma / mixed_luma);
  }
  float final_w = (selected_p_hat > 1e-5f && w_sum > 0.0f)
                      ? clampf(w_sum / ((float)m_total * selected_p_hat), 0.0f, 12.0f)
                      : 0.0f;
  arc_gi_reservoir out_res;
  out_res.sample_pos[0] = selected_pos.x;
  out_res.sample_pos[1] = selected_pos.y;
  out_res.sample_pos[2] = selected_pos.z;
  out_res.norm_m_flags = pack_norm_m_flags(selected_norm, m_total, selected_age, selected_is_sky);
  pack_lo_w(selected_lo, final_w, &out_res.lo_rg, &out_res.lo_b_w);
  out_res.primary_depth = hit_dist;
  out_res.shadow_vis_m = pack2x16float(accum_sun_shadow, shadow_m);
  j->gi_out[p_idx] = out_res;

  /* 4. ReSTIR DI */
  arc_di_reservoir out_di = create_empty_di_reservoir();
  uint32_t count = u->n_emissive;
  if (count > 0u && u->total_flux > 0.0f) {
    arc_di_reservoir di_cand = create_empty_di_reservoir();
    uint32_t di_rng = pcg(gx * 1973u + gy * 9277u + frame_idx * 26699u + 0xA341316Cu);
    arc_di_reservoir prev_di = create_empty_di_reservoir();
    uint32_t prev_di_idx = 0xFFFFu, prev_di_seed = 0u;
    if (prev_valid) {
      prev_di = in->di_in[prev_idx];
      prev_di_idx = prev_di.light_idx_seed & 0xFFFFu;
      prev_di_seed = prev_di.light_idx_seed >> 16;
    }
    float avg_flux = fmaxf(1e-5f, u->total_flux / (float)count);
    for (uint32_t a = 0u; a < 4u; ++a) {
      uint32_t cand_idx = 0u;
      int alias = 0;
      if (a < 2u) {
        alias = 1;
      } else if (a == 2u) {
        cand_idx = pcg((gx >> 2) * 1973u + (gy >> 2) * 9277u + frame_idx * 26699u) % count;
      } else if (prev_valid && prev_di_idx < count) {
        int offset = (int)(rnd(&di_rng) * 33.0f) - 16;
        int ci = (int)prev_di_idx + offset;
        ci = ci < 0 ? 0 : (ci > (int)count - 1 ? (int)count - 1 : ci);
        cand_idx = (uint32_t)ci;
      } else {
        alias = 1;
      }
      if (alias) {
        float u1 = rnd(&di_rng);
        float u2 = rnd(&di_rng);
        uint32_t col = umin(f2u(u1 * (float)count), count - 1u);
        arc_alias_entry ae = in->alias[col];
        cand_idx = u2 < ae.prob ? col : ae.alias_idx;
      }
      if (cand_idx < count) {
        uint32_t cand_seed = pcg(di_rng + a * 7919u + 0x517CC1B7u) & 0xFFFFu;
        const arc_emissive_source* em = &in->sources[cand_idx];
        float p_hat = luma(di_eval_unshadowed(em, cand_seed, hit_pos, normal));
        if (p_hat > 1e-5f) {
          float em_flux = fmaxf(1e-5f, luma(vload(em->intensity)));
          float q_rel = clampf(em_flux / avg_flux, 0.25f, 4.0f);
          float cand_weight = p_hat / (a < 2u ? q_rel : 1.0f);
          di_cand.w_sum += cand_weight;
          di_cand.m += 1.0f;
          if (rnd(&di_rng) * di_cand.w_sum < cand_weight) di_cand.light_idx_seed = cand_idx | (cand_seed << 16);
        }
      }
    }
    uint32_t cwi = di_cand.light_idx_seed & 0xFFFFu, cws = di_cand.light_idx_seed >> 16;
    float cand_winner_p_hat = 0.0f;
    if (cwi < count) cand_winner_p_hat = luma(di_eval_unshadowed(&in->sources[cwi], cws, hit_pos, normal));
    float di_w_sum = di_cand.w_sum, di_m_total = di_cand.m;
    uint32_t selected_di_idx_seed = di_cand.light_idx_seed;
    float selected_di_p_hat = cand_winner_p_hat;
    if (prev_valid && prev_di.m > 0.0f && prev_di.w > 0.0f && prev_di_idx < count) {
      float prev_p_hat = luma(di_eval_unshadowed(&in->sources[prev_di_idx], prev_di_seed, hit_pos, normal));
      if (prev_p_hat > 1e-5f) {
        float prev_m_clamped = fminf(prev_di.m, 8.0f);
        float prev_weight = prev_p_hat * prev_di.w * prev_m_clamped;
        di_w_sum += prev_weight;
        di_m_total += prev_m_clamped;
        if (rnd(&di_rng) * di_w_sum < prev_weight) {
          selected_di_idx_seed = prev_di.light_idx_seed;
          selected_di_p_hat = prev_p_hat;
        }
      }
    }
    float di_final_w = (selected_di_p_hat > 1e-5f && di_w_sum > 0.0f)
                           ? di_w_sum / (fmaxf(di_m_total, 1.0f) * selected_di_p_hat)
                           : 0.0f;
    out_di.light_idx_seed = selected_di_idx_seed;
    out_di.w_sum = di_w_sum;
    out_di.m = fminf(di_m_total, 8.0f);
    out_di.w = clampf(di_final_w, 0.0f, 2.5f);
  }
  j->di_out[p_idx] = out_di;
}

/* ---- spatial_main (lighting_restir_gi.wgsl:457-762) -------------------------------------------- */
static void spatial_pixel(const job* j, uint32_t gx, uint32_t gy) {
  const orc_world* w = j->w;
  const orc_restir_in* in = j->in;
  const arc_restir_params* u = in->p;
  const uint32_t W = u->width, H = u->height;
  const uint32_t p_idx = gy * W + gx;
  const float* depth = (const float*)in->depth_mat;
  const float hit_dist = depth[2 * p_idx];
  if (hit_dist >= 1e4f) {
    j->gi_out[p_idx] = create_empty_reservoir();
    j->di_out[p_idx] = create_empty_di_reservoir();
    return;
  }
  v3f ray_dir = reconstruct_ray_dir(u, gx, gy);
  v3f hit_pos = vadd(vload(u->cam.eye), vscale(ray_dir, hit_dist));
  v3f normal = unpack_normal_oct(in->norm_motion[2 * p_idx]);
  arc_gi_reservoir center_res = in->gi_in[p_idx];
  uint32_t center_m = unpack_m(center_res.norm_m_flags);
  arc_di_reservoir center_di = in->di_in[p_idx];
  uint32_t center_di_idx = center_di.light_idx_seed & 0xFFFFu, center_di_seed = center_di.light_idx_seed >> 16;
  if (center_m == 0u && center_di.m == 0.0f) {
    j->gi_out[p_idx] = center_res;
    j->di_out[p_idx] = center_di;
    return;
  }
  v3f smooth_hemi_n = vnorm(vadd(normal, V(0.0f, 0.22f, 0.0f)));
  v3f center_lo = unpack_lo(center_res.lo_rg, center_res.lo_b_w);
  float center_w = unpack_w(center_res.lo_b_w);
  v3f center_to_s = vsub(vload(center_res.sample_pos), hit_pos);
  float center_d = vlen(center_to_s);
  v3f center_dir = smooth_hemi_n;
  if (center_d > 1e-4f) center_dir = vdivs(center_to_s, center_d);
  float center_cos = fmaxf(0.0f, vdot(smooth_hemi_n, center_dir));
  float center_p_hat = luma(center_lo) * (center_cos / 3.14159265f);
  float w_sum = center_p_hat > 1e-5f ? center_p_hat * center_w * (float)center_m : 0.0f;
  v3f accum_lo = vscale(center_lo, w_sum);
  float csx = h2f(center_res.shadow_vis_m), csm = h2f(center_res.shadow_vis_m >> 16);
  float shd_sum = csx * 2.0f, shd_w_sum = 2.0f;
  uint32_t m_sum = center_m;
  arc_gi_reservoir selected_res = center_res;
  float selected_p_hat = center_p_hat;
  int neighbor_won = 0;
  v3f won_neighbor_dir = center_dir;
  float won_neighbor_dist = center_d;
  uint32_t count = u->n_emissive;
  float di_w_sum = 0.0f, di_m_sum = 0.0f;
  uint32_t selected_di_idx_seed = center_di.light_idx_seed;
  float selected_di_p_hat = 0.0f;
  uint32_t selected_di_idx = center_di_idx, selected_di_seed = center_di_seed;
  if (count > 0u && center_di.m > 0.0f && center_di.w > 0.0f && center_di_idx < count) {
    selected_di_p_hat = luma(di_eval_unshadowed(&in->sources[center_di_idx], center_di_seed, hit_pos, normal));
    if (selected_di_p_hat > 1e-5f) {
      di_w_sum = selected_di_p_hat * center_di.w * center_di.m;
      di_m_sum = center_di.m;
    }
  }
  uint32_t frame_idx = u->frame_index;
  uint32_t seed = pcg(gx + pcg(gy + pcg(frame_idx + 0x85EBCA6Bu)));
  uint32_t denoiser_taps = u->denoiser_taps;
  uint32_t num_taps = denoiser_taps <= 12u ? 2u : 3u;
  float angle = rnd(&seed) * 6.2831853f;
  float sa, ca;
  sincos8(angle, &sa, &ca);
  for (uint32_t i = 0u; i < num_taps; ++i) {
    float box = 3.0f, boy = 1.0f;
    if (i == 1u) {
      box = -3.5f;
      boy = 3.0f;
    } else if (i == 2u) {
      box = 0.5f;
      boy = -5.0f;
    }
    int rx = (int)rintf(box * ca - boy * sa);
    int ry = (int)rintf(box * sa + boy * ca);
    int qx = (int)gx + rx, qy = (int)gy + ry;
    if (qx < 0 || qy < 0 || qx >= (int)W || qy >= (int)H) continue;
    uint32_t q_idx = (uint32_t)qy * W + (uint32_t)qx;
    float q_dist = depth[2 * q_idx];
    if (q_dist >= 1e4f) continue;
    if (fabsf(hit_dist - q_dist) / fmaxf(hit_dist, 0.001f) > 0.10f) continue;
    v3f q_norm = unpack_normal_oct(in->norm_motion[2 * q_idx]);
    if (vdot(normal, q_norm) < 0.80f) continue;
    if (center_m > 0u) {
      arc_gi_reservoir q_res = in->gi_in[q_idx];
      float qsx = h2f(q_res.shadow_vis_m), qsm = h2f(q_res.shadow_vis_m >> 16);
      if (qsm > 0.0f) {
        float shd_tap_w = fminf(qsm, 4.0f) * 0.35f;
        shd_sum += qsx * shd_tap_w;
        shd_w_sum += shd_tap_w;
      }
      uint32_t q_m = unpack_m(q_res.norm_m_flags);
      if (q_m > 0u) {
        uint32_t q_m_clamped = umin(q_m, 4u);
        m_sum += q_m_clamped;
        if (unpack_sample_age(q_res.norm_m_flags) <= 4u) {
          v3f q_lo = unpack_lo(q_res.lo_rg, q_res.lo_b_w);
          float q_w = unpack_w(q_res.lo_b_w);
          v3f q_n_s = unpack_normal(q_res.norm_m_flags);
          float jacobian = 1.0f, p_hat_target = 0.0f, dist_v_to_s = 0.0f;
          v3f dir_v_to_s = V(0.0f, 0.0f, 0.0f);
          if (unpack_is_sky(q_res.norm_m_flags)) {
            v3f to_s = vsub(vload(q_res.sample_pos), hit_pos);
            dist_v_to_s = vlen(to_s);
            if (dist_v_to_s >= 1e-4f) {
              dir_v_to_s = vdivs(to_s, dist_v_to_s);
              float cos_theta_v = fmaxf(0.0f, vdot(smooth_hemi_n, dir_v_to_s));
              p_hat_target = luma(q_lo) * (cos_theta_v / 3.14159265f);
            }
          } else {
            v3f q_pos = vadd(vload(u->cam.eye), vscale(reconstruct_ray_dir(u, (uint32_t)qx, (uint32_t)qy), q_dist));
            v3f v_new = vsub(hit_pos, vload(q_res.sample_pos));
            v3f v_old = vsub(q_pos, vload(q_res.sample_pos));
            float d_new = vlen(v_new), d_old = vlen(v_old);
            if (d_new >= 0.25f && d_old >= 0.25f) {
              v3f hat_v_new = vdivs(v_new, d_new), hat_v_old = vdivs(v_old, d_old);
              float cos_s_new = vdot(q_n_s, hat_v_new), cos_s_old = vdot(q_n_s, hat_v_old);
              if (cos_s_new > 0.05f && cos_s_old > 0.05f) {
                float j_raw = (cos_s_new / cos_s_old) * ((d_old * d_old) / fmaxf(0.0625f, d_new * d_new));
                jacobian = clampf(j_raw, 0.2f, 2.0f);
                dir_v_to_s = V(-hat_v_new.x, -hat_v_new.y, -hat_v_new.z);
                dist_v_to_s = d_new;
                float cos_theta_v = fmaxf(0.0f, vdot(smooth_hemi_n, dir_v_to_s));
                p_hat_target = luma(q_lo) * (cos_theta_v / 3.14159265f);
              }
            }
          }
          if (p_hat_target > 1e-5f) {
            float w_q = p_hat_target * jacobian * q_w * (float)q_m_clamped;
            w_sum += w_q;
            accum_lo = vadd(accum_lo, vscale(q_lo, w_q));
            if ((rnd(&seed) * w_sum) < w_q) {
              selected_res = q_res;
              selected_p_hat = p_hat_target;
              neighbor_won = 1;
              won_neighbor_dir = dir_v_to_s;
              won_neighbor_dist = dist_v_to_s;
            }
          }
        }
      }
    }
    if (count > 0u) {
      arc_di_reservoir q_di = in->di_in[q_idx];
      uint32_t q_di_idx = q_di.light_idx_seed & 0xFFFFu, q_di_seed = q_di.light_idx_seed >> 16;
      float q_di_m = fminf(q_di.m, 16.0f);
      if (q_di_m > 0.0f && q_di.w > 0.0f && q_di_idx < count) {
        float p_hat_target = luma(di_eval_unshadowed(&in->sources[q_di_idx], q_di_seed, hit_pos, normal));
        if (p_hat_target > 1e-5f) {
          float w_q = p_hat_target * q_di.w * q_di_m;
          di_w_sum += w_q;
          di_m_sum += q_di_m;
          if (rnd(&seed) * di_w_sum < w_q) {
            selected_di_idx_seed = q_di.light_idx_seed;
            selected_di_p_hat = p_hat_target;
            selected_di_idx = q_di_idx;
            selected_di_seed = q_di_seed;
          }
        }
      }
    }
  }
  if (center_m > 0u) {
    if (neighbor_won && !unpack_is_sky(selected_res.norm_m_flags)) {
      v3f vis_origin = vadd(hit_pos, vscale(normal, 0.08f));
      float shadow_d = fmaxf(0.05f, won_neighbor_dist - 0.12f);
      float vis = trace_shadow_vis(w, u, vis_origin, won_neighbor_dir, shadow_d);
      if (vis < 0.1f) {
        selected_res = center_res;
        selected_p_hat = center_p_hat;
        w_sum = center_p_hat > 1e-5f ? center_p_hat * center_w * (float)center_m : 0.0f;
        accum_lo = vscale(center_lo, w_sum);
        m_sum = center_m;
      }
    }
    uint32_t max_m = u->target_valid ? 8u : 12u;
    uint32_t out_m = umin(m_sum, max_m);
    float final_w = (selected_p_hat > 1e-5f && w_sum > 0.0f)
                        ? clampf(w_sum / ((float)(m_sum > 1u ? m_sum : 1u) * selected_p_hat), 0.0f, 12.0f)
                        : 0.0f;
    v3f sel_lo = unpack_lo(selected_res.lo_rg, selected_res.lo_b_w);
    float sel_luma = luma(sel_lo);
    v3f avg_lo = vdivs(accum_lo, fmaxf(w_sum, 1e-5f));
    float avg_luma = luma(avg_lo);
    if (avg_luma > 1e-5f && sel_luma > 1e-5f) sel_lo = vscale(avg_lo, sel_luma / avg_luma);
    uint32_t lo_rg, lo_b_w;
    pack_lo_w(sel_lo, final_w, &lo_rg, &lo_b_w);
    uint32_t f = selected_res.norm_m_flags;
    arc_gi_reservoir out_res = selected_res;
    out_res.norm_m_flags =
        pack_norm_m_flags(unpack_normal(f), out_m, unpack_sample_age(f) + (neighbor_won ? 2u : 0u), unpack_is_sky(f));
    out_res.lo_b_w = lo_b_w;
    out_res.shadow_vis_m = pack2x16float(shd_sum / shd_w_sum, csm);
    out_res.primary_depth = hit_dist;
    j->gi_out[p_idx] = out_res;
  } else {
    j->gi_out[p_idx] = center_res;
  }
  float di_final_w = 0.0f;
  if (selected_di_idx < count && selected_di_p_hat > 1e-5f && di_w_sum > 0.0f) {
    int kept_center_di = (selected_di_idx_seed == center_di.light_idx_seed) && (center_di.w > 0.0f);
    int checkerboard_active = denoiser_taps <= 12u;
    uint32_t checkerboard_slot = (gx & 1u) ^ ((gy & 1u) << 1);
    uint32_t frame_slot = frame_idx & 3u;
    int skip_di_vis_ray = checkerboard_active && kept_center_di && (((checkerboard_slot ^ frame_slot) & 1u) != 0u);
    float vis = 1.0f;
    if (!skip_di_vis_ray) {
      const arc_emissive_source* em = &in->sources[selected_di_idx];
      v3f un_n;
      float eta;
      di_unpack_normal_and_eta(em->normal_packed, &un_n, &eta);
      v3f em_pos = di_emitter_surface_pos(vload(em->pos), un_n, em->radius, eta, selected_di_seed);
      float d_init = vlen(vsub(em_pos, hit_pos));
      float normal_offset = fminf(0.08f, d_init * 0.4f);
      v3f shadow_origin = vadd(hit_pos, vscale(normal, normal_offset));
      v3f to_em = vsub(em_pos, shadow_origin);
      float dist_to_em = vlen(to_em);
      v3f shadow_dir = dist_to_em > 1e-4f ? vdivs(to_em, fmaxf(dist_to_em, 1e-4f)) : V(0.0f, 1.0f, 0.0f);
      float shadow_max_d = fmaxf(0.02f, dist_to_em - fmaxf(em->radius * 0.35f, 0.08f));
      vis = trace_shadow_vis_di(w, u, shadow_origin, shadow_dir, shadow_max_d);
    }
    if (vis < 0.1f) {
      di_w_sum = 0.0f;
      di_final_w = 0.0f;
    } else {
      di_final_w = clampf(di_w_sum / (fmaxf(di_m_sum, 1.0f) * selected_di_p_hat), 0.0f, 2.5f) * vis;
    }
  }
  arc_di_reservoir out_di;
  out_di.light_idx_seed = selected_di_idx_seed;
  out_di.w_sum = di_w_sum;
  out_di.m = fminf(di_m_sum, 12.0f);
  out_di.w = di_final_w;
  j->di_out[p_idx] = out_di;
}

/* ---- pixel loops ------------------------------------------------------------------------------ */
static void* temporal_worker(void* arg) {
  const job* j = (const job*)arg;
  for (uint32_t i = j->first; i < j->last; ++i) temporal_pixel(j, i % j->in->p->width, i / j->in->p->width);
  return NULL;
}
static void* spatial_worker(void* arg) {
  const job* j = (const job*)arg;
  for (uint32_t i = j->first; i < j->last; ++i) spatial_pixel(j, i % j->in->p->width, i / j->in->p->width);
  return NULL;
}
static void run(const orc_world* w, const orc_restir_in* in, arc_gi_reservoir* gi_out, arc_di_reservoir* di_out,
                int nthreads, void* (*fn)(void*)) {
  const uint32_t n = in->p->width * in->p->height;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 64) nthreads = 64;
  job jobs[64];
  pthread_t th[64];
  for (int t = 0; t < nthreads; ++t) {
    jobs[t].w = w;
    jobs[t].in = in;
    jobs[t].gi_out = gi_out;
    jobs[t].di_out = di_out;
    jobs[t].first = (uint32_t)((uint64_t)n * t / nthreads);
    jobs[t].last = (uint32_t)((uint64_t)n * (t + 1) / nthreads);
  }
  for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, fn, &jobs[t]);
  fn(&jobs[0]);
  for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
}

void orc_restir_temporal(const orc_world* w, const orc_restir_in* in, arc_gi_reservoir* gi_out,
                         arc_di_reservoir* di_out, int nthreads) {
  run(w, in, gi_out, di_out, nthreads, temporal_worker);
}
void orc_restir_spatial(const orc_world* w, const orc_restir_in* in, arc_gi_reservoir* gi_out,
                        arc_di_reservoir* di_out, int nthreads) {
  run(w, in, gi_out, di_out, nthreads, spatial_worker);
}
