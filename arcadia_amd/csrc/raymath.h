// Ray generation arithmetic shared by the host generators (rays.cpp) and the device path kernel.
//
// Every function is plain binary32 arithmetic in source order with correctly rounded division
// and square root (host: -ffp-contract=off; device: also -fhip-fp32-correctly-rounded-divide-sqrt),
// so host and device produce the same bits. sin/cos are computed here rather than taken from libm
// or the device math library, whose last bits differ: Cody-Waite reduction by pi/2 in two parts,
// then Taylor polynomials of sin (degree 9) and cos (degree 8) on |r| <= pi/4 (truncation error
// below 2e-9, results within ~1 ulp of the true value).
#pragma once

#include <math.h>
#include <stdint.h>

#include "../../include/arcadia_world.h"

#if defined(__HIPCC__)
#define ARC_HD __host__ __device__ __forceinline__
#else
#define ARC_HD inline
#endif

namespace arcmath {

ARC_HD uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// normalize(v) = v / sqrt(dot(v, v)), dot summed x, y, z left to right.
ARC_HD void normalize3(const float v[3], float out[3]) {
  const float xx = v[0] * v[0];
  const float yy = v[1] * v[1];
  const float zz = v[2] * v[2];
  const float d = (xx + yy) + zz;
  const float len = sqrtf(d);
  out[0] = v[0] / len;
  out[1] = v[1] / len;
  out[2] = v[2] / len;
}

ARC_HD void cross3(const float a[3], const float b[3], float o[3]) {
  const float x0 = a[1] * b[2], x1 = a[2] * b[1];
  const float y0 = a[2] * b[0], y1 = a[0] * b[2];
  const float z0 = a[0] * b[1], z1 = a[1] * b[0];
  o[0] = x0 - x1;
  o[1] = y0 - y1;
  o[2] = z0 - z1;
}

// sin and cos of x for 0 <= x < 8 (the reference's phi = 6.2831853 * u2 with u2 in [0, 1)).
ARC_HD void sincos_small(float x, float* s, float* c) {
  const float kf = floorf(x * 0.636619772f + 0.5f);  // nearest multiple of pi/2
  const int k = (int)kf;
  const float a = kf * 1.5703125f;          // pi/2 high part: 8 significant bits, k * it is exact
  const float b = kf * 4.83826794897e-4f;   // pi/2 - 1.5703125
  const float r = (x - a) - b;
  const float r2 = r * r;
  float ps = 2.75573188e-06f;  // 1/9!
  ps = ps * r2 + -1.98412701e-04f;  // -1/7!
  ps = ps * r2 + 8.33333377e-03f;   // 1/5!
  ps = ps * r2 + -1.66666672e-01f;  // -1/3!
  const float sr = r + (r * r2) * ps;
  float pc = 2.48015876e-05f;  // 1/8!
  pc = pc * r2 + -1.38888892e-03f;  // -1/6!
  pc = pc * r2 + 4.16666679e-02f;   // 1/4!
  pc = pc * r2 + -0.5f;
  const float cr = 1.0f + r2 * pc;
  switch (k & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}

// Pinhole primary ray of pixel (px, py): primary_gbuffer.wgsl:103-115 (tan_half_fov vertical).
ARC_HD void primary_dir(const arc_camera& cam, uint32_t px, uint32_t py, uint32_t width, uint32_t height,
                        float out[3]) {
  const float ux = ((float)px + 0.5f) + cam.jitter[0];
  const float uy = ((float)py + 0.5f) + cam.jitter[1];
  const float uvx = ux / (float)width, uvy = uy / (float)height;
  const float ndx = uvx * 2.0f - 1.0f;
  const float ndy = 1.0f - uvy * 2.0f;
  const float sx = (ndx * cam.aspect) * cam.tan_half_fov;
  const float sy = ndy * cam.tan_half_fov;
  float v[3];
  for (int a = 0; a < 3; ++a) {
    const float rr = cam.right[a] * sx;
    const float uu = cam.up[a] * sy;
    const float s = cam.forward[a] + rr;
    v[a] = s + uu;
  }
  normalize3(v, out);
}

// Diffuse bounce off a hit (lighting_restir_gi.wgsl:195-209; types.wgsl:483-497, 531-545):
// origin = hit + 0.10 m * n, direction = cosine sample around normalize(n + 0.22 y) in the
// make_basis frame. `ovox` = parent origin in voxels (origin_m * 16), `t_vox` the hit t in
// voxels, `face` the compact face code (2*axis + (normal < 0)); (u1, u2) = 24-bit fractions of
// splitmix64(seed ^ splitmix64(index + 0x51ED270B)).
//
// The lobe's frame depends on the face alone (nine codes, 6..8 with no normal), so it is its own
// function: a kernel can tabulate the nine frames once and pass one to bounce_ray_framed, which
// gives the same bits as bounce_ray.
struct BounceFrame {
  float sn[3], t[3], b[3];  // lobe axis normalize(n + 0.22 y) and its make_basis tangents
};

ARC_HD BounceFrame bounce_frame(uint32_t face) {
  float nrm[3] = {0.0f, 0.0f, 0.0f};
  if (face < 6u) nrm[face >> 1] = (face & 1u) ? -1.0f : 1.0f;
  const float lifted[3] = {nrm[0], nrm[1] + 0.22f, nrm[2]};
  BounceFrame f;
  normalize3(lifted, f.sn);
  const bool zup = fabsf(f.sn[2]) < 0.999f;
  const float upv[3] = {zup ? 0.0f : 1.0f, 0.0f, zup ? 1.0f : 0.0f};
  float cr[3];
  cross3(upv, f.sn, cr);
  normalize3(cr, f.t);
  cross3(f.sn, f.t, f.b);
  return f;
}

ARC_HD void bounce_ray_framed(const float ovox[3], const float dir[3], float t_vox, uint32_t face,
                              const BounceFrame& f, uint64_t seed, uint64_t index, float max_dist_m, arc_ray* out) {
  float nrm[3] = {0.0f, 0.0f, 0.0f};
  if (face < 6u) nrm[face >> 1] = (face & 1u) ? -1.0f : 1.0f;
  float org[3];
  for (int a = 0; a < 3; ++a) {
    const float dt = dir[a] * t_vox;
    const float pv = ovox[a] + dt;
    const float hp = pv * 0.0625f;
    const float off = nrm[a] * 0.10f;
    org[a] = hp + off;
  }
  const uint64_t h = splitmix64(seed ^ splitmix64(index + 0x51ED270Bull));
  const float u1 = (float)(uint32_t)(h & 0xFFFFFFu) * (1.0f / 16777216.0f);
  const float u2 = (float)(uint32_t)((h >> 24) & 0xFFFFFFu) * (1.0f / 16777216.0f);
  const float r = sqrtf(u1);
  const float phi = 6.2831853f * u2;
  const float z = sqrtf(fmaxf(0.0f, 1.0f - u1));
  float sphi, cphi;
  sincos_small(phi, &sphi, &cphi);
  const float rc = r * cphi, rs = r * sphi;
  float d[3];
  for (int a = 0; a < 3; ++a) {
    const float p0 = f.t[a] * rc, p1 = f.b[a] * rs, p2 = f.sn[a] * z;
    const float s = p0 + p1;
    d[a] = s + p2;
  }
  out->ox = org[0];
  out->oy = org[1];
  out->oz = org[2];
  out->t_start = 0.0f;
  out->dx = d[0];
  out->dy = d[1];
  out->dz = d[2];
  out->t_max = max_dist_m;
}

ARC_HD void bounce_ray(const float ovox[3], const float dir[3], float t_vox, uint32_t face, uint64_t seed,
                       uint64_t index, float max_dist_m, arc_ray* out) {
  bounce_ray_framed(ovox, dir, t_vox, face, bounce_frame(face), seed, index, max_dist_m, out);
}

// The ray of an empty slot (padding pixel, or a bounce whose parent missed): t_max < 0 makes
// the traversal return a miss with t_voxels = -16 and 0 steps.
ARC_HD arc_ray empty_ray() {
  arc_ray r;
  r.ox = r.oy = r.oz = 0.0f;
  r.t_start = 0.0f;
  r.dx = 0.0f;
  r.dy = 1.0f;
  r.dz = 0.0f;
  r.t_max = -1.0f;
  return r;
}

}  // namespace arcmath
