// ReSTIR GI / DI passes on the device: temporal_main and spatial_main of
// shaders/lighting_restir_gi.wgsl (:62-455, :457-762) with the helpers of shaders/common/types.wgsl
// and shaders/common/di.wgsl.
//
// The reference traces its secondary rays inline, one pixel per invocation. Here a pass is a short
// pipeline of per-pixel kernels around the library's own traversal launches: a kernel writes one
// ray per pixel into a dense buffer (a pixel that needs no ray gets t_max = -1, an immediate miss),
// the persistent-queue traversal traces the whole buffer, and the next kernel reads the results:
//   temporal: k_rt_setup (sun shadow ray, GI ray) -> sun trace, GI trace -> [k_rt_rigid_rays ->
//             rigid-body trace] -> k_rt_bounce (SHARC probe, bounce-sun ray) -> SHARC query,
//             bounce-sun walker -> k_rt_resolve (sun average, GI RIS, DI reservoirs)
//   spatial:  k_rs_rays (neighbour reuse -> GI and DI visibility rays) -> sun / DI walkers ->
//             k_rs_resolve (the same reuse, recomputed, then the visibility decisions)
// Kernels recompute what they need from the G-buffer rather than pass it along; every quantity
// is computed by one shared function, so a recomputation gives the same bits.
//
// Arithmetic: binary32, no contraction, correctly rounded division and square root, operations
// in WGSL source order (vector expressions left to right per component). Where WGSL leaves it
// open: dot = x, y, z left to right; length = sqrt(dot); normalize(v) = v / length(v);
// mix(a, b, t) = a (1 - t) + b t; round() = to nearest, ties to even; sin / cos = the library's
// polynomial (raymath.h); f32 -> f16 packing to nearest even; float -> int conversions truncate.
#include <hip/hip_runtime.h>

#include "raymath.h"
#include "restir.h"
#include "trace_device.h"

namespace arcdev {
namespace {

constexpr float kPi = 3.14159265f;

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 mul(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 mulv(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ V3 divs(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) {
  const float x = a.x * b.x, y = a.y * b.y, z = a.z * b.z;
  return (x + y) + z;
}
__device__ __forceinline__ float len(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) { return divs(a, len(a)); }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  const float x0 = a.y * b.z, x1 = a.z * b.y;
  const float y0 = a.z * b.x, y1 = a.x * b.z;
  const float z0 = a.x * b.y, z1 = a.y * b.x;
  return v3(x0 - x1, y0 - y1, z0 - z1);
}
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
__device__ __forceinline__ float mixf(float a, float b, float t) {
  const float u = a * (1.0f - t);
  const float v = b * t;
  return u + v;
}
__device__ __forceinline__ V3 mix3(V3 a, V3 b, float t) { return v3(mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)); }
__device__ __forceinline__ float fract(float x) { return x - floorf(x); }
// luma (types.wgsl:183-185)
__device__ __forceinline__ float luma(V3 c) { return dot(c, v3(0.2126f, 0.7152f, 0.0722f)); }
__device__ __forceinline__ V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }

// ---- f16 and snorm packing (WGSL pack2x16float / unpack2x16float / unpack2x16snorm) ----------
__device__ __forceinline__ uint32_t f16b(float x) {
  const _Float16 h = (_Float16)x;
  return (uint32_t)__builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float f16v(uint32_t b) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(b & 0xFFFFu)); }
__device__ __forceinline__ uint32_t pack2h(float a, float b) { return f16b(a) | (f16b(b) << 16); }
__device__ __forceinline__ float snorm16v(uint32_t b) {
  const float v = (float)(int16_t)(uint16_t)(b & 0xFFFFu) / 32767.0f;
  return clampf(v, -1.0f, 1.0f);
}

// oct_encode / oct_decode (types.wgsl:108-130)
__device__ __forceinline__ void oct_encode(V3 n, float& px, float& py) {
  const float l1 = (fabsf(n.x) + fabsf(n.y)) + fabsf(n.z);
  const float inv = 1.0f / fmaxf(l1, 1e-6f);
  px = n.x * inv;
  py = n.y * inv;
  if (n.z < 0.0f) {
    const float wx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
    const float wy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
    px = wx;
    py = wy;
  }
}
__device__ __forceinline__ V3 oct_decode(float px, float py) {
  V3 n = v3(px, py, (1.0f - fabsf(px)) - fabsf(py));
  if (n.z < 0.0f) {
    const float sx = px >= 0.0f ? 1.0f : -1.0f, sy = py >= 0.0f ? 1.0f : -1.0f;
    n.x = (1.0f - fabsf(py)) * sx;
    n.y = (1.0f - fabsf(px)) * sy;
  }
  return normalize(n);
}
// unpack_normal_oct (types.wgsl:313-316)
__device__ __forceinline__ V3 gbuf_normal(uint32_t packed) { return oct_decode(snorm16v(packed), snorm16v(packed >> 16)); }

// GiReservoir packing (types.wgsl:132-181)
__device__ __forceinline__ uint32_t pack_norm_m_flags(V3 n, uint32_t m, uint32_t age, bool sky) {
  float ox, oy;
  oct_encode(n, ox, oy);
  const int ix = min(max((int)rintf(ox * 127.0f), -128), 127) + 128;
  const int iy = min(max((int)rintf(oy * 127.0f), -128), 127) + 128;
  return ((uint32_t)ix & 0xFFu) | (((uint32_t)iy & 0xFFu) << 8) | ((min(m, 16u) & 0xFFu) << 16) |
         ((min(age, 127u) & 0x7Fu) << 24) | (sky ? 0x80000000u : 0u);
}
__device__ __forceinline__ V3 unpack_normal(uint32_t f) {
  const float ux = (float)(f & 0xFFu), uy = (float)((f >> 8) & 0xFFu);
  return oct_decode((ux - 128.0f) / 127.0f, (uy - 128.0f) / 127.0f);
}
__device__ __forceinline__ uint32_t unpack_m(uint32_t f) { return (f >> 16) & 0xFFu; }
__device__ __forceinline__ uint32_t unpack_age(uint32_t f) { return (f >> 24) & 0x7Fu; }
__device__ __forceinline__ bool unpack_sky(uint32_t f) { return (f & 0x80000000u) != 0u; }
__device__ __forceinline__ V3 unpack_lo(uint32_t rg, uint32_t bw) { return v3(f16v(rg), f16v(rg >> 16), f16v(bw)); }
__device__ __forceinline__ float unpack_w(uint32_t bw) { return f16v(bw >> 16); }
__device__ __forceinline__ uint2 pack_lo_w(V3 lo, float w) {
  const float r = clampf(lo.x, 0.0f, 65500.0f), g = clampf(lo.y, 0.0f, 65500.0f), b = clampf(lo.z, 0.0f, 65500.0f);
  return make_uint2(pack2h(r, g), pack2h(b, fminf(w, 65500.0f)));
}
__device__ __forceinline__ arc_gi_reservoir empty_gi() {
  arc_gi_reservoir r;
  r.sample_pos[0] = r.sample_pos[1] = r.sample_pos[2] = 0.0f;
  r.norm_m_flags = 0u;
  r.lo_rg = 0u;
  r.lo_b_w = 0u;
  r.primary_depth = 1e5f;
  r.shadow_vis_m = 0u;
  return r;
}
__device__ __forceinline__ arc_di_reservoir empty_di() {
  arc_di_reservoir r;
  r.light_idx_seed = 0xFFFFFFFFu;
  r.w_sum = 0.0f;
  r.m = 0.0f;
  r.w = 0.0f;
  return r;
}

// rand_f32 (types.wgsl:264-267)
__device__ __forceinline__ float rand_f32(uint32_t& seed) {
  seed = pcg_hash(seed);
  return (float)seed / 4294967295.0f;
}

// interleaved_gradient_noise (types.wgsl:483-497)
__device__ __forceinline__ void ign(float px, float py, uint32_t frame, float& u1, float& u2) {
  const float f = (float)frame;
  const uint32_t ux = f2u(px), uy = f2u(py);
  const float a1 = 0.06711056f * px, b1 = 0.00583715f * py;
  const float m1 = 52.9829189f * fract(a1 + b1);
  const float ign1 = fract(m1);
  const float a2 = 0.00583715f * px, b2 = 0.06711056f * py;
  const float m2 = 52.9829189f * fract((a2 - b2) + 0.5f);
  const float ign2 = fract(m2);
  const uint32_t h1 = pcg_hash(ux * 1973u + uy * 9277u + frame * 26699u);
  const uint32_t h2 = pcg_hash(h1 ^ 0x9E3779B9u);
  const float r1 = (float)(h1 & 0xFFFFu) / 65536.0f;
  const float r2 = (float)(h2 & 0xFFFFu) / 65536.0f;
  const float s1 = ign1 + f * 0.754877666f;
  const float s2 = ign2 + f * 0.569840291f;
  u1 = fract(s1 + r1);
  u2 = fract(s2 + r2);
}

// make_basis / sample_cosine_hemisphere (types.wgsl:531-545)
struct Basis {
  V3 t, b;
};
__device__ __forceinline__ Basis make_basis(V3 n) {
  const V3 up = fabsf(n.z) < 0.999f ? v3(0.0f, 0.0f, 1.0f) : v3(1.0f, 0.0f, 0.0f);
  Basis bs;
  bs.t = normalize(cross(up, n));
  bs.b = cross(n, bs.t);
  return bs;
}
__device__ __forceinline__ V3 cosine_sample(float u1, float u2, V3 n, const Basis& bs) {
  const float r = sqrtf(u1);
  const float phi = 6.2831853f * u2;
  const float z = sqrtf(fmaxf(0.0f, 1.0f - u1));
  float s, c;
  arcmath::sincos_small(phi, &s, &c);
  const float rc = r * c, rs = r * s;
  return add(add(mul(bs.t, rc), mul(bs.b, rs)), mul(n, z));
}

// eval_sky_radiance (types.wgsl:500-523); pow(x, 3) and pow(x, 14) as repeated products.
__device__ __forceinline__ V3 sky(V3 d, V3 sun) {
  const float se = sun.y;
  const float day_t = clampf((se + 0.05f) / 0.30f, 0.0f, 1.0f);
  const float twi_t = clampf(1.0f - fabsf(se - 0.02f) / 0.22f, 0.0f, 1.0f);
  const float zen = clampf(d.y, 0.0f, 1.0f);
  const float om = 1.0f - zen;
  const float hor = (om * om) * om;
  const float dz[3] = {0.08f, 0.22f, 0.55f}, dh[3] = {0.38f, 0.50f, 0.68f};
  const float nz[3] = {0.0015f, 0.0025f, 0.0055f}, nh[3] = {0.0035f, 0.0065f, 0.0120f};
  const float tw[3] = {0.65f, 0.25f, 0.10f}, sc[3] = {0.85f, 0.52f, 0.24f};
  const float sc0 = fmaxf(dot(d, sun), 0.0f);
  const float x2 = sc0 * sc0, x4 = x2 * x2, x8 = x4 * x4;
  const float p14 = (x8 * x4) * x2;
  const float tw_w = twi_t * 0.75f;
  const float sun_w = fmaxf(day_t, twi_t);
  float o[3];
  for (int c = 0; c < 3; ++c) {
    const float skz = mixf(nz[c], dz[c], day_t);
    float skh = mixf(nh[c], dh[c], day_t);
    skh = mixf(skh, tw[c], tw_w);
    float s = mixf(skz, skh, hor);
    if (se > -0.12f) {
      const float a = ((sc[c] * p14) * 0.35f) * sun_w;
      s = s + a;
    }
    o[c] = s;
  }
  return v3(o[0], o[1], o[2]);
}

// ---- DI helpers (di.wgsl:42-115) --------------------------------------------------------------
struct EmNormal {
  V3 n;
  float eta;
};
__device__ __forceinline__ EmNormal em_normal(uint32_t np) {
  const float nx = ((float)(np & 1023u) / 1023.0f) * 2.0f - 1.0f;
  const float ny = ((float)((np >> 10) & 1023u) / 1023.0f) * 2.0f - 1.0f;
  const float nz = ((float)((np >> 20) & 1023u) / 1023.0f) * 2.0f - 1.0f;
  const V3 n = v3(nx, ny, nz);
  const float l = len(n);
  EmNormal r;
  r.eta = clampf(l, 0.0f, 1.0f);
  r.n = l > 0.05f ? divs(n, fmaxf(l, 1e-4f)) : v3(0.0f, 1.0f, 0.0f);
  return r;
}
__device__ __forceinline__ V3 em_surface_pos(V3 c, V3 n, float radius, float eta, uint32_t seed) {
  if (radius <= 0.04f) return c;
  const float u1 = (float)(seed & 0xFFu) / 255.0f;
  const float u2 = (float)((seed >> 8) & 0xFFu) / 255.0f;
  const float max_r = fminf(radius * 0.20f, 0.08f);
  if (eta >= 0.35f) {
    const Basis bs = make_basis(n);
    const float r = sqrtf(u1) * max_r;
    const float phi = 6.2831853f * u2;
    float s, co;
    arcmath::sincos_small(phi, &s, &co);
    return add(c, add(mul(bs.t, co * r), mul(bs.b, s * r)));
  }
  const float phi = 6.2831853f * u1;
  const float ct = u2 * 2.0f - 1.0f;
  const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
  float s, co;
  arcmath::sincos_small(phi, &s, &co);
  return add(c, mul(v3(co * st, ct, s * st), max_r));
}
__device__ __forceinline__ V3 di_eval(const arc_emissive_source& em, uint32_t seed, V3 xv, V3 nv) {
  const EmNormal un = em_normal(em.normal_packed);
  const V3 ep = em_surface_pos(ld3(em.pos), un.n, em.radius, un.eta, seed);
  const V3 to_l = sub(ep, xv);
  const float d2 = dot(to_l, to_l);
  if (d2 < 1e-6f || d2 > 1296.0f) return v3(0.0f, 0.0f, 0.0f);
  const float d = sqrtf(d2);
  const V3 l_dir = divs(to_l, d);
  const float cos_v = fmaxf(dot(nv, l_dir), 0.0f);
  if (cos_v <= 0.0f) return v3(0.0f, 0.0f, 0.0f);
  const float dir_factor = (1.0f - un.eta) + un.eta * fmaxf(dot(un.n, neg(l_dir)), 0.0f);
  const float range_fade = clampf(1.0f - d2 / 1296.0f, 0.0f, 1.0f);
  const float geom = ((((cos_v * dir_factor) * 0.31830989f) * range_fade) * range_fade) / fmaxf(d2, 0.20f);
  return mul(ld3(em.intensity), geom);
}

// ---- per-pixel geometry shared by the kernels ---------------------------------------------------
struct Px {
  uint32_t x, y, idx;
  float hit_dist;
  V3 ray_dir, hit_pos, normal;
};
__device__ __forceinline__ V3 pixel_dir(const arc_restir_params& p, uint32_t x, uint32_t y) {
  float d[3];
  arcmath::primary_dir(p.cam, x, y, p.width, p.height, d);  // reconstruct_ray_dir (types.wgsl:334-345)
  return v3(d[0], d[1], d[2]);
}
// false when the pixel has no primary hit (hit_dist >= 1e4)
__device__ __forceinline__ bool load_px(const RestirDev& d, uint32_t i, Px& q) {
  q.x = i % d.p.width;
  q.y = i / d.p.width;
  q.idx = i;
  q.hit_dist = __uint_as_float(d.depth_mat[i].x);
  if (q.hit_dist >= 1e4f) return false;
  q.ray_dir = pixel_dir(d.p, q.x, q.y);
  q.hit_pos = add(ld3(d.p.cam.eye), mul(q.ray_dir, q.hit_dist));
  q.normal = gbuf_normal(d.norm_motion[i].x);
  return true;
}
__device__ __forceinline__ arc_ray make_ray(V3 o, V3 dir, float t_max) {
  arc_ray r;
  r.ox = o.x;
  r.oy = o.y;
  r.oz = o.z;
  r.t_start = 0.0f;
  r.dx = dir.x;
  r.dy = dir.y;
  r.dz = dir.z;
  r.t_max = t_max;
  return r;
}
__device__ __forceinline__ arc_ray no_ray() { return make_ray(v3(0.0f, 0.0f, 0.0f), v3(0.0f, 1.0f, 0.0f), -1.0f); }
__device__ __forceinline__ bool checkerboard_odd(const arc_restir_params& p, uint32_t x, uint32_t y) {
  const uint32_t slot = (x & 1u) ^ ((y & 1u) << 1);
  return ((slot ^ (p.frame_index & 3u)) & 1u) != 0u;
}

// probe_emissive_macro_candidate (lighting_restir_gi.wgsl:19-56)
__device__ __forceinline__ float probe_emissive(const RestirDev& d, V3 pos, V3 dir, float dist) {
  const V3 vp = mul(add(pos, mul(dir, dist)), 16.0f);
  const float wx = (float)d.gx * 256.0f, wy = (float)d.gy * 256.0f, wz = (float)d.gz * 256.0f;
  if (vp.x < 0.0f || vp.y < 0.0f || vp.z < 0.0f || vp.x >= wx || vp.y >= wy || vp.z >= wz) return 0.0f;
  const uint32_t vx = f2u(vp.x), vy = f2u(vp.y), vz = f2u(vp.z);
  const uint32_t ci = (vx >> 8) + (vy >> 8) * d.gx + (vz >> 8) * d.gx * d.gy;
  const uint32_t root_idx = d.roots[ci];
  if (root_idx == ARC_EMPTY_CHUNK_SENTINEL) return 0.0f;
  const uint32_t lx = vx & 255u, ly = vy & 255u, lz = vz & 255u;
  const uint4 root = d.nodes[root_idx];
  const uint32_t b0 = (((lz >> 6) & 3u) << 4) | (((ly >> 6) & 3u) << 2) | ((lx >> 6) & 3u);
  if (!has_child(root.x, root.y, b0)) return 0.0f;
  uint32_t mat;
  if ((root.z & ARC_IS_LEAF_FLAG) != 0u) {
    mat = (root.w >> 24) & 0xFFu;
  } else {
    const uint4 l1 = d.nodes[child_offset(root, b0)];
    const uint32_t b1 = (((lz >> 4) & 3u) << 4) | (((ly >> 4) & 3u) << 2) | ((lx >> 4) & 3u);
    if (!has_child(l1.x, l1.y, b1)) return 0.0f;
    mat = (l1.w >> 24) & 0xFFu;
  }
  const float* m = d.palette + 8u * mat;
  return dot(v3(m[4], m[5], m[6]), v3(0.333f, 0.333f, 0.333f));
}

// ---- temporal pass --------------------------------------------------------------------------
enum : uint32_t {
  S_ACTIVE = 1u,      // primary hit
  S_IN_RANGE = 2u,    // reprojected uv inside the screen: history pixel read
  S_PREV_VALID = 4u,  // history accepted
  S_SKIP = 8u,        // checkerboard: secondary rays skipped
  S_SUN = 16u,        // sun shadow ray traced
};

struct Temporal {
  uint32_t flags, prev_idx;
  float xi0, xi1;
  arc_gi_reservoir prev;  // zero unless S_IN_RANGE (WGSL zero-initialises prev_res)
};

// History validation and checkerboard decision (:75-127).
__device__ __forceinline__ Temporal temporal_prologue(const RestirDev& d, const Px& q) {
  Temporal t;
  t.flags = S_ACTIVE;
  t.prev_idx = 0u;
  t.prev.sample_pos[0] = t.prev.sample_pos[1] = t.prev.sample_pos[2] = 0.0f;
  t.prev.norm_m_flags = t.prev.lo_rg = t.prev.lo_b_w = t.prev.shadow_vis_m = 0u;
  t.prev.primary_depth = 0.0f;
  const uint32_t mot = d.norm_motion[q.idx].y;
  const float uvx = ((float)q.x + 0.5f) / (float)d.p.width, uvy = ((float)q.y + 0.5f) / (float)d.p.height;
  const float pux = uvx + f16v(mot), puy = uvy + f16v(mot >> 16);
  ign((float)q.x, (float)q.y, d.p.frame_index, t.xi0, t.xi1);
  if (pux >= 0.0f && pux < 1.0f && puy >= 0.0f && puy < 1.0f) {
    const uint32_t ppx = min(f2u(pux * (float)d.p.width), d.p.width - 1u);
    const uint32_t ppy = min(f2u(puy * (float)d.p.height), d.p.height - 1u);
    t.prev_idx = ppy * d.p.width + ppx;
    t.prev = d.gi_in[t.prev_idx];
    t.flags |= S_IN_RANGE;
    const uint32_t pm = unpack_m(t.prev.norm_m_flags), page = unpack_age(t.prev.norm_m_flags);
    if (pm > 0u && page <= 12u && t.prev.primary_depth < 1e4f) {
      const float pcd = len(sub(q.hit_pos, ld3(d.p.prev_eye)));
      const float tol = fmaxf(0.12f, q.hit_dist * 0.05f);
      if (fabsf(pcd - t.prev.primary_depth) < tol) t.flags |= S_PREV_VALID;
    }
  }
  if (d.p.denoiser_taps <= 12u && (t.flags & S_PREV_VALID) && checkerboard_odd(d.p, q.x, q.y)) t.flags |= S_SKIP;
  const float ndl = fmaxf(dot(q.normal, ld3(d.p.sun_dir)), 0.0f);
  if (!(t.flags & S_SKIP) && ndl > 0.001f && d.p.sun_intensity > 0.01f) t.flags |= S_SUN;
  return t;
}

struct GiDirs {
  V3 shn, bounce;
  bool guided;
};
// Candidate direction with emissive macro-cell guiding (:193-205).
__device__ __forceinline__ GiDirs gi_dirs(const RestirDev& d, const Px& q, const Temporal& t) {
  GiDirs g;
  g.shn = normalize(add(q.normal, v3(0.0f, 0.22f, 0.0f)));
  const Basis bs = make_basis(g.shn);
  g.bounce = cosine_sample(t.xi0, t.xi1, g.shn, bs);
  const V3 cand = cosine_sample(fract(t.xi0 + 0.5f), fract(t.xi1 + 0.381966f), g.shn, bs);
  const float near = probe_emissive(d, q.hit_pos, cand, 1.25f);
  const float mid = probe_emissive(d, q.hit_pos, cand, 3.0f);
  g.guided = fmaxf(near, mid) > 0.25f;
  if (g.guided) g.bounce = cand;
  return g;
}

// Jittered sun direction (:140-146).
__device__ __forceinline__ V3 jittered_sun(const RestirDev& d, const Temporal& t) {
  const V3 sd = ld3(d.p.sun_dir);
  const float j1 = t.xi0 * 2.0f - 1.0f, j2 = t.xi1 * 2.0f - 1.0f;
  const V3 up = fabsf(sd.y) < 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
  const V3 st = normalize(cross(up, sd));
  const V3 sb = cross(sd, st);
  return normalize(add(sd, mul(add(mul(st, j1), mul(sb, j2)), d.p.shadow_jitter)));
}

__global__ __launch_bounds__(256) void k_rt_setup(RestirDev d) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = d.p.width * d.p.height;
  if (i >= n) return;
  Px q;
  if (!load_px(d, i, q)) {
    d.state[i] = make_uint2(0u, 0u);
    d.rays_a[i] = no_ray();
    d.rays_b[i] = no_ray();
    return;
  }
  const Temporal t = temporal_prologue(d, q);
  d.state[i] = make_uint2(t.flags, t.prev_idx);
  if (t.flags & S_SUN) {
    const V3 js = jittered_sun(d, t);
    d.rays_a[i] = make_ray(add(q.hit_pos, mul(q.normal, 0.08f)), js, d.p.shadow_max_dist);
  } else {
    d.rays_a[i] = no_ray();
  }
  if (t.flags & S_SKIP) {
    d.rays_b[i] = no_ray();
    return;
  }
  const GiDirs g = gi_dirs(d, q, t);
  d.rays_b[i] = make_ray(add(q.hit_pos, mul(q.normal, 0.10f)), g.bounce, d.p.gi_max_dist);
}

// The rigid-body walk of a GI ray stops at its world hit (:210-213).
__global__ __launch_bounds__(256) void k_rt_rigid_rays(RestirDev d) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.p.width * d.p.height) return;
  arc_ray r = d.rays_b[i];
  if (r.t_max >= 0.0f) {
    const arc_trace_result& h = d.full_b[i];
    r.t_max = (h.hit & 1u) ? h.t_voxels * 0.0625f : d.p.gi_max_dist;
  }
  d.rays_c[i] = r;
}

// The GI hit after the rigid-body merge (:214-222).
struct GiHit {
  bool hit;
  float t_voxels;
  V3 pos, normal, lod_rgb;
  uint32_t mat;
};
__device__ __forceinline__ GiHit gi_hit(const RestirDev& d, uint32_t i) {
  const arc_trace_result& h = d.full_b[i];
  GiHit g;
  g.hit = (h.hit & 1u) != 0u;
  g.t_voxels = h.t_voxels;
  g.pos = ld3(h.hit_pos_m);
  g.normal = ld3(h.normal);
  g.lod_rgb = ld3(h.lod_rgb);
  g.mat = h.mat;
  if (d.n_bodies > 0u) {
    const arc_rigid_hit& r = d.rb_c[i];
    if (r.hit) {
      g.hit = true;
      g.t_voxels = r.t_voxels;
      g.pos = ld3(r.hit_pos_m);
      g.normal = ld3(r.normal);
      g.mat = r.mat;
      const float* m = d.palette + 8u * (r.mat & 0xFFu);
      g.lod_rgb = v3(m[0], m[1], m[2]);
    }
  }
  return g;
}

// SHARC probe and bounce-sun shadow ray at the GI hit (:239-257). The bounce-sun ray is traced
// whenever the reference could need it; it is used only if the SHARC lookup misses.
__global__ __launch_bounds__(256) void k_rt_bounce(RestirDev d) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.p.width * d.p.height) return;
  const uint2 st = d.state[i];
  arc_sharc_probe pr;
  pr.pos[0] = pr.pos[1] = pr.pos[2] = 0.0f;
  pr.normal[0] = 0.0f;
  pr.normal[1] = 1.0f;
  pr.normal[2] = 0.0f;
  pr.jitter[0] = pr.jitter[1] = 0.0f;
  arc_ray br = no_ray();
  if ((st.x & S_ACTIVE) && !(st.x & S_SKIP)) {
    const GiHit g = gi_hit(d, i);
    if (g.hit) {
      const uint32_t x = i % d.p.width, y = i / d.p.width;
      float j0, j1;
      ign((float)x + 101.0f, (float)y + 57.0f, d.p.frame_index + 7919u, j0, j1);
      pr.pos[0] = g.pos.x;
      pr.pos[1] = g.pos.y;
      pr.pos[2] = g.pos.z;
      pr.normal[0] = g.normal.x;
      pr.normal[1] = g.normal.y;
      pr.normal[2] = g.normal.z;
      pr.jitter[0] = j0 * 2.0f - 1.0f;
      pr.jitter[1] = j1 * 2.0f - 1.0f;
      const float bndl = fmaxf(dot(g.normal, ld3(d.p.sun_dir)), 0.0f);
      if (bndl > 0.001f && d.p.sun_intensity > 0.01f)
        br = make_ray(add(g.pos, mul(g.normal, 0.10f)), ld3(d.p.sun_dir), 32.0f);
    }
  }
  if (d.has_sharc) d.probes[i] = pr;
  d.rays_d[i] = br;
}

__global__ __launch_bounds__(256) void k_rt_resolve(RestirDev d) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.p.width * d.p.height) return;
  Px q;
  if (!load_px(d, i, q)) {
    d.gi_out[i] = empty_gi();
    d.di_out[i] = empty_di();
    return;
  }
  const Temporal t = temporal_prologue(d, q);
  const bool prev_valid = (t.flags & S_PREV_VALID) != 0u;
  const V3 sd = ld3(d.p.sun_dir);
  // 1. sun shadow and its decoupled temporal average (:136-169)
  float sun_vis = 0.0f, shadow_m = 1.0f, accum = 0.0f;
  const float psx = f16v(t.prev.shadow_vis_m), psm = f16v(t.prev.shadow_vis_m >> 16);
  if (t.flags & S_SKIP) {
    accum = psx;
    shadow_m = fminf(psm, 16.0f);
  } else {
    if (t.flags & S_SUN) {
      if (d.p.shadow_mode == 0u) {
        if (!(d.hits_a[i].flags & ARC_HIT_FLAG_HIT)) {
          const arc_ray r = d.rays_a[i];
          const float ey = r.oy + r.dy * r.t_max;
          if (ey >= 32.0f || r.dy < 0.0f) sun_vis = 1.0f;
        }
      } else {
        sun_vis = d.vis_a[i];
      }
    }
    if (prev_valid) {
      const float pm = fminf(psm, 16.0f);
      shadow_m = pm + 1.0f;
      accum = (psx * pm + sun_vis) / shadow_m;
    } else {
      accum = sun_vis;
      shadow_m = 1.0f;
    }
  }
  if (t.flags & S_SKIP) {  // :171-186
    arc_gi_reservoir o = t.prev;
    o.primary_depth = q.hit_dist;
    o.shadow_vis_m = pack2h(accum, shadow_m);
    const uint32_t f = t.prev.norm_m_flags;
    o.norm_m_flags = pack_norm_m_flags(unpack_normal(f), unpack_m(f), unpack_age(f) + 1u, unpack_sky(f));
    d.gi_out[i] = o;
    d.di_out[i] = prev_valid ? d.di_in[t.prev_idx] : empty_di();
    return;
  }
  // 2. the GI candidate (:188-280)
  const GiDirs g = gi_dirs(d, q, t);
  const GiHit h = gi_hit(d, i);
  V3 cand_pos = add(q.hit_pos, mul(g.bounce, 500.0f));
  V3 cand_norm = neg(g.bounce);
  bool cand_sky = true;
  V3 lo = v3(0.0f, 0.0f, 0.0f);
  if (h.hit) {
    cand_pos = h.pos;
    cand_norm = h.normal;
    cand_sky = false;
    const float* m = d.palette + 8u * (h.mat & 0xFFu);
    const V3 albedo = mix3(v3(m[0], m[1], m[2]), h.lod_rgb, 0.55f);
    const V3 emission = v3(m[4], m[5], m[6]);
    const float gd = h.t_voxels * 0.0625f;
    const float em_atten = 1.0f / ((1.0f + gd * 0.06f) + (gd * gd) * 0.025f);
    const float gw = g.guided ? 0.55f : 1.0f;
    const float dd = d.p.n_emissive > 0u ? 0.25f : 1.0f;
    lo = add(lo, mul(mul(mul(emission, em_atten), gw), dd));
    const float4 sh = d.has_sharc ? d.sharc_out[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (sh.w > 0.5f) {
      lo = add(lo, v3(sh.x, sh.y, sh.z));
    } else {
      const float bndl = fmaxf(dot(h.normal, sd), 0.0f);
      V3 bounce_sun = v3(0.0f, 0.0f, 0.0f);
      if (bndl > 0.001f && d.p.sun_intensity > 0.01f) {
        const float bv = d.vis_d[i];
        if (bv > 0.001f)
          bounce_sun = mul(mul(mul(mul(v3(1.0f, 0.90f, 0.75f), d.p.sun_intensity), bndl), bv), 0.55f);
      }
      lo = add(lo, mulv(albedo, bounce_sun));
    }
    const float ao = clampf(gd / 0.40f, 0.25f, 1.0f);
    lo = mul(lo, ao);
  } else {
    const V3 gi_origin = add(q.hit_pos, mul(q.normal, 0.10f));
    const V3 end_p = add(gi_origin, mul(g.bounce, d.p.gi_max_dist));
    if (g.bounce.y > 0.15f && end_p.y >= 32.0f) {
      const float esc = clampf((g.bounce.y - 0.15f) / 0.35f, 0.0f, 1.0f);
      lo = add(lo, mul(mul(sky(g.bounce, sd), 0.45f), esc));
    }
  }
  const float peak = fmaxf(lo.x, fmaxf(lo.y, lo.z));
  if (peak > 0.001f) lo = divs(lo, 1.0f + peak * 0.16f);
  // 3. temporal RIS (:282-345)
  const float cand_cos = fmaxf(0.0f, dot(g.shn, g.bounce));
  const float cand_p_hat = luma(lo) * (cand_cos / kPi);
  const float cand_w = cand_p_hat > 1e-5f ? luma(lo) : 0.0f;
  float w_sum = cand_w;
  V3 sel_pos = cand_pos, sel_norm = cand_norm, sel_lo = lo;
  float sel_p_hat = cand_p_hat;
  uint32_t sel_age = 0u, m_total = 1u;
  bool sel_sky = cand_sky;
  if (prev_valid) {
    const uint32_t max_m = d.p.target_valid ? 8u : 12u;
    const uint32_t f = t.prev.norm_m_flags;
    const uint32_t prev_m = min(unpack_m(f), max_m);
    const float prev_w = unpack_w(t.prev.lo_b_w);
    const V3 prev_lo = unpack_lo(t.prev.lo_rg, t.prev.lo_b_w);
    const V3 prev_norm = unpack_normal(f);
    const uint32_t prev_age = unpack_age(f);
    const bool prev_sky = unpack_sky(f);
    const V3 to_prev = sub(ld3(t.prev.sample_pos), q.hit_pos);
    const float d_prev = len(to_prev);
    float prev_p_hat = 0.0f;
    if (d_prev > 1e-4f) {
      const V3 dir_prev = divs(to_prev, d_prev);
      const float pc = fmaxf(0.0f, dot(g.shn, dir_prev));
      prev_p_hat = luma(prev_lo) * (pc / kPi);
    }
    const uint32_t eff_m = min(prev_m, max_m - 1u);
    const float prev_weight = prev_p_hat > 1e-5f ? (prev_p_hat * prev_w) * (float)eff_m : 0.0f;
    w_sum = w_sum + prev_weight;
    m_total = eff_m + 1u;
    uint32_t seed = pcg_hash(q.x + pcg_hash(q.y + pcg_hash(d.p.frame_index + 0x9E3779B9u)));
    if (rand_f32(seed) * w_sum < prev_weight) {
      sel_pos = ld3(t.prev.sample_pos);
      sel_norm = prev_norm;
      sel_lo = prev_lo;
      sel_p_hat = prev_p_hat;
      sel_age = prev_age + 1u;
      sel_sky = prev_sky;
    }
    const float blend_t = clampf(prev_weight / fmaxf(w_sum, 1e-5f), 0.0f, 1.0f);
    const V3 mixed = mix3(lo, prev_lo, blend_t);
    const float wl = luma(sel_lo), ml = luma(mixed);
    if (ml > 1e-5f && wl > 1e-5f) sel_lo = mul(mixed, wl / ml);
  }
  float final_w = 0.0f;
  if (sel_p_hat > 1e-5f && w_sum > 0.0f) final_w = clampf(w_sum / ((float)m_total * sel_p_hat), 0.0f, 12.0f);
  const uint2 lw = pack_lo_w(sel_lo, final_w);
  arc_gi_reservoir o;
  o.sample_pos[0] = sel_pos.x;
  o.sample_pos[1] = sel_pos.y;
  o.sample_pos[2] = sel_pos.z;
  o.norm_m_flags = pack_norm_m_flags(sel_norm, m_total, sel_age, sel_sky);
  o.lo_rg = lw.x;
  o.lo_b_w = lw.y;
  o.primary_depth = q.hit_dist;
  o.shadow_vis_m = pack2h(accum, shadow_m);
  d.gi_out[i] = o;
  // 4. DI candidates and temporal reuse (:347-454)
  arc_di_reservoir out_di = empty_di();
  const uint32_t cnt = d.p.n_emissive;
  if (cnt > 0u && d.p.total_flux > 0.0f) {
    arc_di_reservoir cres = empty_di();
    uint32_t rng = pcg_hash(q.x * 1973u + q.y * 9277u + d.p.frame_index * 26699u + 0xA341316Cu);
    arc_di_reservoir prev_di = empty_di();
    uint32_t prev_idx = 0xFFFFu, prev_seed = 0u;
    if (prev_valid) {
      prev_di = d.di_in[t.prev_idx];
      prev_idx = prev_di.light_idx_seed & 0xFFFFu;
      prev_seed = prev_di.light_idx_seed >> 16;
    }
    const float avg_flux = fmaxf(1e-5f, d.p.total_flux / (float)cnt);
    for (uint32_t a = 0; a < 4u; ++a) {
      uint32_t ci = 0u;
      bool alias_pick = a < 2u;
      if (a == 2u) {
        ci = pcg_hash((q.x >> 2) * 1973u + (q.y >> 2) * 9277u + d.p.frame_index * 26699u) % cnt;
      } else if (a == 3u) {
        if (prev_valid && prev_idx < cnt) {
          const int off = (int)(rand_f32(rng) * 33.0f) - 16;
          ci = (uint32_t)min(max((int)prev_idx + off, 0), (int)cnt - 1);
        } else {
          alias_pick = true;
        }
      }
      if (alias_pick) {
        const float u1 = rand_f32(rng);
        const float u2 = rand_f32(rng);
        const uint32_t col = min(f2u(u1 * (float)cnt), cnt - 1u);
        const arc_alias_entry ae = d.alias[col];
        ci = u2 < ae.prob ? col : ae.alias_idx;
      }
      if (ci < cnt) {
        const uint32_t cs = pcg_hash(rng + a * 7919u + 0x517CC1B7u) & 0xFFFFu;
        const arc_emissive_source em = d.sources[ci];
        const float p_hat = luma(di_eval(em, cs, q.hit_pos, q.normal));
        if (p_hat > 1e-5f) {
          const float flux = fmaxf(1e-5f, luma(ld3(em.intensity)));
          const float q_rel = clampf(flux / avg_flux, 0.25f, 4.0f);
          const float cw = p_hat / (a < 2u ? q_rel : 1.0f);
          cres.w_sum = cres.w_sum + cw;
          cres.m = cres.m + 1.0f;
          if (rand_f32(rng) * cres.w_sum < cw) cres.light_idx_seed = ci | (cs << 16);
        }
      }
    }
    const uint32_t wi = cres.light_idx_seed & 0xFFFFu, ws = cres.light_idx_seed >> 16;
    float win_p_hat = 0.0f;
    if (wi < cnt) win_p_hat = luma(di_eval(d.sources[wi], ws, q.hit_pos, q.normal));
    float dws = cres.w_sum, dm = cres.m;
    uint32_t sel = cres.light_idx_seed;
    float sel_ph = win_p_hat;
    if (prev_valid && prev_di.m > 0.0f && prev_di.w > 0.0f && prev_idx < cnt) {
      const float pph = luma(di_eval(d.sources[prev_idx], prev_seed, q.hit_pos, q.normal));
      if (pph > 1e-5f) {
        const float pmc = fminf(prev_di.m, 8.0f);
        const float pw = (pph * prev_di.w) * pmc;
        dws = dws + pw;
        dm = dm + pmc;
        if (rand_f32(rng) * dws < pw) {
          sel = prev_di.light_idx_seed;
          sel_ph = pph;
        }
      }
    }
    float fw = 0.0f;
    if (sel_ph > 1e-5f && dws > 0.0f) fw = dws / (fmaxf(dm, 1.0f) * sel_ph);
    out_di.light_idx_seed = sel;
    out_di.w_sum = dws;
    out_di.m = fminf(dm, 8.0f);
    out_di.w = clampf(fw, 0.0f, 2.5f);
  }
  d.di_out[i] = out_di;
}

// ---- spatial pass ----------------------------------------------------------------------------
struct Spatial {
  bool active, passthrough;  // passthrough: the centre reservoirs are copied
  arc_gi_reservoir center, sel;
  arc_di_reservoir center_di;
  float center_p_hat, center_w;
  uint32_t center_m, m_sum;
  V3 center_lo, accum_lo, won_dir;
  float w_sum, sel_p_hat, shd_sum, shd_w_sum, won_dist, center_shd_m;
  bool neighbor_won;
  float di_w_sum, di_m_sum, sel_di_p_hat;
  uint32_t sel_di_lis, sel_di_idx, sel_di_seed;
};

// Everything of spatial_main up to the visibility rays (:463-645).
__device__ __forceinline__ Spatial spatial_select(const RestirDev& d, const Px& q) {
  Spatial s;
  s.active = true;
  s.passthrough = false;
  s.center = d.gi_in[q.idx];
  s.center_m = unpack_m(s.center.norm_m_flags);
  s.center_di = d.di_in[q.idx];
  const uint32_t cdi = s.center_di.light_idx_seed & 0xFFFFu, cds = s.center_di.light_idx_seed >> 16;
  if (s.center_m == 0u && s.center_di.m == 0.0f) {
    s.passthrough = true;
    return s;
  }
  const V3 shn = normalize(add(q.normal, v3(0.0f, 0.22f, 0.0f)));
  s.center_lo = unpack_lo(s.center.lo_rg, s.center.lo_b_w);
  s.center_w = unpack_w(s.center.lo_b_w);
  const V3 cts = sub(ld3(s.center.sample_pos), q.hit_pos);
  const float cd = len(cts);
  V3 cdir = shn;
  if (cd > 1e-4f) cdir = divs(cts, cd);
  const float ccos = fmaxf(0.0f, dot(shn, cdir));
  s.center_p_hat = luma(s.center_lo) * (ccos / kPi);
  s.w_sum = s.center_p_hat > 1e-5f ? (s.center_p_hat * s.center_w) * (float)s.center_m : 0.0f;
  s.accum_lo = mul(s.center_lo, s.w_sum);
  const float csx = f16v(s.center.shadow_vis_m);
  s.center_shd_m = f16v(s.center.shadow_vis_m >> 16);
  s.shd_sum = csx * 2.0f;
  s.shd_w_sum = 2.0f;
  s.m_sum = s.center_m;
  s.sel = s.center;
  s.sel_p_hat = s.center_p_hat;
  s.neighbor_won = false;
  s.won_dir = cdir;
  s.won_dist = cd;
  const uint32_t cnt = d.p.n_emissive;
  s.di_w_sum = 0.0f;
  s.di_m_sum = 0.0f;
  s.sel_di_lis = s.center_di.light_idx_seed;
  s.sel_di_p_hat = 0.0f;
  s.sel_di_idx = cdi;
  s.sel_di_seed = cds;
  if (cnt > 0u && s.center_di.m > 0.0f && s.center_di.w > 0.0f && cdi < cnt) {
    s.sel_di_p_hat = luma(di_eval(d.sources[cdi], cds, q.hit_pos, q.normal));
    if (s.sel_di_p_hat > 1e-5f) {
      s.di_w_sum = (s.sel_di_p_hat * s.center_di.w) * s.center_di.m;
      s.di_m_sum = s.center_di.m;
    }
  }
  uint32_t seed = pcg_hash(q.x + pcg_hash(q.y + pcg_hash(d.p.frame_index + 0x85EBCA6Bu)));
  const uint32_t taps = d.p.denoiser_taps <= 12u ? 2u : 3u;
  const float angle = rand_f32(seed) * 6.2831853f;
  float sa, ca;
  arcmath::sincos_small(angle, &sa, &ca);
  for (uint32_t k = 0; k < taps; ++k) {
    float bx = 3.0f, by = 1.0f;
    if (k == 1u) {
      bx = -3.5f;
      by = 3.0f;
    } else if (k == 2u) {
      bx = 0.5f;
      by = -5.0f;
    }
    const int rx = (int)rintf(bx * ca - by * sa);
    const int ry = (int)rintf(bx * sa + by * ca);
    const int qx = (int)q.x + rx, qy = (int)q.y + ry;
    if (qx < 0 || qy < 0 || qx >= (int)d.p.width || qy >= (int)d.p.height) continue;
    const uint32_t qi = (uint32_t)qy * d.p.width + (uint32_t)qx;
    const float qd = __uint_as_float(d.depth_mat[qi].x);
    if (qd >= 1e4f) continue;
    if (fabsf(q.hit_dist - qd) / fmaxf(q.hit_dist, 0.001f) > 0.10f) continue;
    const V3 qn = gbuf_normal(d.norm_motion[qi].x);
    if (dot(q.normal, qn) < 0.80f) continue;
    if (s.center_m > 0u) {  // GI neighbour (:552-615)
      const arc_gi_reservoir qr = d.gi_in[qi];
      const float qsx = f16v(qr.shadow_vis_m), qsm = f16v(qr.shadow_vis_m >> 16);
      if (qsm > 0.0f) {
        const float tw = fminf(qsm, 4.0f) * 0.35f;
        s.shd_sum = s.shd_sum + qsx * tw;
        s.shd_w_sum = s.shd_w_sum + tw;
      }
      const uint32_t qm = unpack_m(qr.norm_m_flags);
      if (qm > 0u) {
        const uint32_t qmc = min(qm, 4u);
        s.m_sum += qmc;
        if (unpack_age(qr.norm_m_flags) <= 4u) {
          const V3 q_lo = unpack_lo(qr.lo_rg, qr.lo_b_w);
          const float q_w = unpack_w(qr.lo_b_w);
          const V3 q_ns = unpack_normal(qr.norm_m_flags);
          float jac = 1.0f, p_hat = 0.0f, dist = 0.0f;
          V3 dir = v3(0.0f, 0.0f, 0.0f);
          if (unpack_sky(qr.norm_m_flags)) {
            const V3 to_s = sub(ld3(qr.sample_pos), q.hit_pos);
            dist = len(to_s);
            if (dist >= 1e-4f) {
              dir = divs(to_s, dist);
              p_hat = luma(q_lo) * (fmaxf(0.0f, dot(shn, dir)) / kPi);
            }
          } else {
            const V3 q_pos = add(ld3(d.p.cam.eye), mul(pixel_dir(d.p, (uint32_t)qx, (uint32_t)qy), qd));
            const V3 v_new = sub(q.hit_pos, ld3(qr.sample_pos));
            const V3 v_old = sub(q_pos, ld3(qr.sample_pos));
            const float d_new = len(v_new), d_old = len(v_old);
            if (d_new >= 0.25f && d_old >= 0.25f) {
              const V3 hn = divs(v_new, d_new), ho = divs(v_old, d_old);
              const float cn = dot(q_ns, hn), co = dot(q_ns, ho);
              if (cn > 0.05f && co > 0.05f) {
                const float j_raw = (cn / co) * ((d_old * d_old) / fmaxf(0.0625f, d_new * d_new));
                jac = clampf(j_raw, 0.2f, 2.0f);
                dir = neg(hn);
                dist = d_new;
                p_hat = luma(q_lo) * (fmaxf(0.0f, dot(shn, dir)) / kPi);
              }
            }
          }
          if (p_hat > 1e-5f) {
            const float wq = ((p_hat * jac) * q_w) * (float)qmc;
            s.w_sum = s.w_sum + wq;
            s.accum_lo = add(s.accum_lo, mul(q_lo, wq));
            if (rand_f32(seed) * s.w_sum < wq) {
              s.sel = qr;
              s.sel_p_hat = p_hat;
              s.neighbor_won = true;
              s.won_dir = dir;
              s.won_dist = dist;
            }
          }
        }
      }
    }
    if (cnt > 0u) {  // DI neighbour (:617-642)
      const arc_di_reservoir qdi = d.di_in[qi];
      const uint32_t qix = qdi.light_idx_seed & 0xFFFFu, qsd = qdi.light_idx_seed >> 16;
      const float qdm = fminf(qdi.m, 16.0f);
      if (qdm > 0.0f && qdi.w > 0.0f && qix < cnt) {
        const float p_hat = luma(di_eval(d.sources[qix], qsd, q.hit_pos, q.normal));
        if (p_hat > 1e-5f) {
          const float wq = (p_hat * qdi.w) * qdm;
          s.di_w_sum = s.di_w_sum + wq;
          s.di_m_sum = s.di_m_sum + qdm;
          if (rand_f32(seed) * s.di_w_sum < wq) {
            s.sel_di_lis = qdi.light_idx_seed;
            s.sel_di_p_hat = p_hat;
            s.sel_di_idx = qix;
            s.sel_di_seed = qsd;
          }
        }
      }
    }
  }
  return s;
}

__device__ __forceinline__ bool gi_vis_needed(const Spatial& s) {
  return s.center_m > 0u && s.neighbor_won && !unpack_sky(s.sel.norm_m_flags);
}
__device__ __forceinline__ arc_ray gi_vis_ray(const Px& q, const Spatial& s) {
  return make_ray(add(q.hit_pos, mul(q.normal, 0.08f)), s.won_dir, fmaxf(0.05f, s.won_dist - 0.12f));
}
__device__ __forceinline__ bool di_vis_candidate(const RestirDev& d, const Spatial& s) {
  return s.sel_di_idx < d.p.n_emissive && s.sel_di_p_hat > 1e-5f && s.di_w_sum > 0.0f;
}
__device__ __forceinline__ bool di_vis_skipped(const RestirDev& d, const Px& q, const Spatial& s) {
  const bool kept = s.sel_di_lis == s.center_di.light_idx_seed && s.center_di.w > 0.0f;
  return d.p.denoiser_taps <= 12u && kept && checkerboard_odd(d.p, q.x, q.y);
}
// The DI visibility ray towards the selected emitter sample (:701-713).
__device__ __forceinline__ arc_ray di_vis_ray(const RestirDev& d, const Px& q, const Spatial& s) {
  const arc_emissive_source em = d.sources[s.sel_di_idx];
  const EmNormal un = em_normal(em.normal_packed);
  const V3 ep = em_surface_pos(ld3(em.pos), un.n, em.radius, un.eta, s.sel_di_seed);
  const float d_init = len(sub(ep, q.hit_pos));
  const float off = fminf(0.08f, d_init * 0.4f);
  const V3 so = add(q.hit_pos, mul(q.normal, off));
  const V3 to_em = sub(ep, so);
  const float dist = len(to_em);
  const V3 sdir = dist > 1e-4f ? divs(to_em, fmaxf(dist, 1e-4f)) : v3(0.0f, 1.0f, 0.0f);
  return make_ray(so, sdir, fmaxf(0.02f, dist - fmaxf(em.radius * 0.35f, 0.08f)));
}

__global__ __launch_bounds__(256) void k_rs_rays(RestirDev d) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.p.width * d.p.height) return;
  arc_ray ga = no_ray(), db = no_ray();
  Px q;
  if (load_px(d, i, q)) {
    const Spatial s = spatial_select(d, q);
    if (!s.passthrough) {
      if (gi_vis_needed(s)) ga = gi_vis_ray(q, s);
      if (di_vis_candidate(d, s) && !di_vis_skipped(d, q, s)) db = di_vis_ray(d, q, s);
    }
  }
  d.rays_a[i] = ga;
  d.rays_b[i] = db;
}

__global__ __launch_bounds__(256) void k_rs_resolve(RestirDev d) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.p.width * d.p.height) return;
  Px q;
  if (!load_px(d, i, q)) {
    d.gi_out[i] = empty_gi();
    d.di_out[i] = empty_di();
    return;
  }
  Spatial s = spatial_select(d, q);
  if (s.passthrough) {
    d.gi_out[i] = s.center;
    d.di_out[i] = s.center_di;
    return;
  }
  // finalize GI (:645-687)
  if (s.center_m > 0u) {
    if (gi_vis_needed(s) && d.vis_a[i] < 0.1f) {  // a blocked reconnection keeps the centre sample
      s.sel = s.center;
      s.sel_p_hat = s.center_p_hat;
      s.w_sum = s.center_p_hat > 1e-5f ? (s.center_p_hat * s.center_w) * (float)s.center_m : 0.0f;
      s.accum_lo = mul(s.center_lo, s.w_sum);
      s.m_sum = s.center_m;
    }
    const uint32_t max_m = d.p.target_valid ? 8u : 12u;
    const uint32_t out_m = min(s.m_sum, max_m);
    float final_w = 0.0f;
    if (s.sel_p_hat > 1e-5f && s.w_sum > 0.0f)
      final_w = clampf(s.w_sum / ((float)max(s.m_sum, 1u) * s.sel_p_hat), 0.0f, 12.0f);
    V3 sel_lo = unpack_lo(s.sel.lo_rg, s.sel.lo_b_w);
    const float sl = luma(sel_lo);
    const V3 avg = divs(s.accum_lo, fmaxf(s.w_sum, 1e-5f));
    const float al = luma(avg);
    if (al > 1e-5f && sl > 1e-5f) sel_lo = mul(avg, sl / al);
    const uint2 lw = pack_lo_w(sel_lo, final_w);
    const uint32_t f = s.sel.norm_m_flags;
    arc_gi_reservoir o = s.sel;
    o.norm_m_flags = pack_norm_m_flags(unpack_normal(f), out_m, unpack_age(f) + (s.neighbor_won ? 2u : 0u), unpack_sky(f));
    o.lo_b_w = lw.y;
    o.shadow_vis_m = pack2h(s.shd_sum / s.shd_w_sum, s.center_shd_m);
    o.primary_depth = q.hit_dist;
    d.gi_out[i] = o;
  } else {
    d.gi_out[i] = s.center;
  }
  // finalize DI (:689-761)
  float di_w_sum = s.di_w_sum, fw = 0.0f;
  if (di_vis_candidate(d, s)) {
    const float vis = di_vis_skipped(d, q, s) ? 1.0f : d.vis_b[i];
    if (vis < 0.1f) {
      di_w_sum = 0.0f;
      fw = 0.0f;
    } else {
      fw = clampf(di_w_sum / (fmaxf(s.di_m_sum, 1.0f) * s.sel_di_p_hat), 0.0f, 2.5f) * vis;
    }
  }
  arc_di_reservoir od;
  od.light_idx_seed = s.sel_di_lis;
  od.w_sum = di_w_sum;
  od.m = fminf(s.di_m_sum, 12.0f);
  od.w = fw;
  d.di_out[i] = od;
}

inline dim3 grid_of(const RestirDev& d) { return dim3((d.p.width * d.p.height + 255u) / 256u); }

}  // namespace

#define RESTIR_LAUNCH(name, kern)                                                  \
  hipError_t name(const RestirDev& d, hipStream_t s) {                           \
    hipLaunchKernelGGL(kern, grid_of(d), dim3(256), 0, s, d);                    \
    return hipGetLastError();                                                    \
  }
RESTIR_LAUNCH(launch_rt_setup, k_rt_setup)
RESTIR_LAUNCH(launch_rt_rigid_rays, k_rt_rigid_rays)
RESTIR_LAUNCH(launch_rt_bounce, k_rt_bounce)
RESTIR_LAUNCH(launch_rt_resolve, k_rt_resolve)
RESTIR_LAUNCH(launch_rs_rays, k_rs_rays)
RESTIR_LAUNCH(launch_rs_resolve, k_rs_resolve)
#undef RESTIR_LAUNCH

}  // namespace arcdev
