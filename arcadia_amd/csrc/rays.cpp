// Ray set generation (C ABI in include/arcadia_world.h). Binary32 arithmetic in source order
// (built with -ffp-contract=off), so the device ray generator of the G-buffer pass can repeat
// the primary-ray formula bit for bit.
#include <cmath>
#include <cstring>

#include "world_internal.h"

namespace {

inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// normalize(v) = v / sqrt(dot(v, v)), dot summed x, y, z left to right.
inline void normalize3(const float v[3], float out[3]) {
  float xx = v[0] * v[0];
  float yy = v[1] * v[1];
  float zz = v[2] * v[2];
  float d = xx + yy;
  d = d + zz;
  float len = std::sqrt(d);
  out[0] = v[0] / len;
  out[1] = v[1] / len;
  out[2] = v[2] / len;
}

inline void cross3(const float a[3], const float b[3], float o[3]) {
  float x0 = a[1] * b[2], x1 = a[2] * b[1];
  float y0 = a[2] * b[0], y1 = a[0] * b[2];
  float z0 = a[0] * b[1], z1 = a[1] * b[0];
  o[0] = x0 - x1;
  o[1] = y0 - y1;
  o[2] = z0 - z1;
}

}  // namespace

extern "C" {

int arc_camera_look(const float eye[3], float yaw, float pitch, float fov_y_deg, uint32_t width, uint32_t height,
                    arc_camera* out) {
  if (!eye || !out || width == 0 || height == 0 || !(fov_y_deg > 0.0f && fov_y_deg < 180.0f)) return ARC_ERR_INVALID;
  std::memset(out, 0, sizeof *out);
  std::memcpy(out->eye, eye, sizeof out->eye);
  double cp = std::cos((double)pitch), sp = std::sin((double)pitch);
  double cy = std::cos((double)yaw), sy = std::sin((double)yaw);
  double f[3] = {cp * sy, sp, cp * cy};
  // right = normalize(cross(f, +Y)) = normalize((-f.z, 0, f.x))
  double r[3] = {-f[2], 0.0, f[0]};
  double rl = std::sqrt(r[0] * r[0] + r[2] * r[2]);
  if (rl < 1e-9) return ARC_ERR_INVALID;  // looking straight up/down
  r[0] /= rl;
  r[2] /= rl;
  // up = cross(right, forward)
  double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
  for (int a = 0; a < 3; ++a) {
    out->forward[a] = (float)f[a];
    out->right[a] = (float)r[a];
    out->up[a] = (float)u[a];
  }
  out->tan_half_fov = (float)std::tan(0.5 * (double)fov_y_deg * 3.14159265358979323846 / 180.0);
  out->aspect = (float)width / (float)height;
  return ARC_OK;
}

int arc_gen_primary_rays(const arc_camera* cam, uint32_t width, uint32_t height, float max_dist_m, arc_ray* rays,
                         uint32_t* pixel_of_ray) {
  if (!cam || !rays || width == 0 || height == 0 || width > 65535 || height > 65535) return ARC_ERR_INVALID;
  const uint32_t tiles_x = (width + 7) / 8, tiles_y = (height + 7) / 8;
  const float fw = (float)width, fh = (float)height;
  uint64_t k = 0;
  for (uint32_t ty = 0; ty < tiles_y; ++ty)
    for (uint32_t tx = 0; tx < tiles_x; ++tx)
      for (uint32_t j = 0; j < 64; ++j) {
        uint32_t px = tx * 8 + (j & 7), py = ty * 8 + (j >> 3);
        if (px >= width || py >= height) continue;
        // primary_gbuffer.wgsl:103-115
        float ux = (float)px + 0.5f;
        ux = ux + cam->jitter[0];
        float uy = (float)py + 0.5f;
        uy = uy + cam->jitter[1];
        float uvx = ux / fw, uvy = uy / fh;
        float ndx = uvx * 2.0f - 1.0f;
        float ndy = 1.0f - uvy * 2.0f;
        float sx = ndx * cam->aspect;
        sx = sx * cam->tan_half_fov;
        float sy = ndy * cam->tan_half_fov;
        float v[3];
        for (int a = 0; a < 3; ++a) {
          float rr = cam->right[a] * sx;
          float uu = cam->up[a] * sy;
          float s = cam->forward[a] + rr;
          v[a] = s + uu;
        }
        float d[3];
        normalize3(v, d);
        arc_ray& r = rays[k];
        r.ox = cam->eye[0];
        r.oy = cam->eye[1];
        r.oz = cam->eye[2];
        r.t_start = 0.0f;
        r.dx = d[0];
        r.dy = d[1];
        r.dz = d[2];
        r.t_max = max_dist_m;
        if (pixel_of_ray) pixel_of_ray[k] = py * width + px;
        ++k;
      }
  return ARC_OK;
}

int64_t arc_gen_bounce_rays(const arc_ray* prays, const arc_hit* hits, uint64_t n, uint64_t seed, float max_dist_m,
                            arc_ray* rays, uint32_t* src_index) {
  if ((!prays || !hits || !rays) && n) return ARC_ERR_INVALID;
  int64_t k = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const arc_hit& h = hits[i];
    if (!(h.flags & ARC_HIT_FLAG_HIT)) continue;
    const arc_ray& pr = prays[i];
    const float o[3] = {pr.ox, pr.oy, pr.oz}, dir[3] = {pr.dx, pr.dy, pr.dz};
    float nrm[3] = {0.0f, 0.0f, 0.0f};
    if (h.face < 6) nrm[h.face >> 1] = (h.face & 1) ? -1.0f : 1.0f;
    float hp[3], org[3];
    for (int a = 0; a < 3; ++a) {
      float vo = o[a] * 16.0f;
      float dt = dir[a] * h.t_voxels;
      float pv = vo + dt;
      hp[a] = pv * 0.0625f;
      float off = nrm[a] * 0.10f;
      org[a] = hp[a] + off;
    }
    // lighting_restir_gi.wgsl:195-197; types.wgsl:483-497, 531-545
    const float lifted[3] = {nrm[0], nrm[1] + 0.22f, nrm[2]};
    float sn[3];
    normalize3(lifted, sn);
    const float upv[3] = {std::fabs(sn[2]) < 0.999f ? 0.0f : 1.0f, 0.0f, std::fabs(sn[2]) < 0.999f ? 1.0f : 0.0f};
    float c[3], t[3], b[3];
    cross3(upv, sn, c);
    normalize3(c, t);
    cross3(sn, t, b);
    uint64_t hsh = splitmix64(seed ^ splitmix64(i + 0x51ED270Bull));
    float u1 = (float)(uint32_t)(hsh & 0xFFFFFFu) * (1.0f / 16777216.0f);
    float u2 = (float)(uint32_t)((hsh >> 24) & 0xFFFFFFu) * (1.0f / 16777216.0f);
    float r = std::sqrt(u1);
    float phi = 6.2831853f * u2;
    float z = std::sqrt(std::fmax(0.0f, 1.0f - u1));
    float rc = r * std::cos(phi), rs = r * std::sin(phi);
    float d[3];
    for (int a = 0; a < 3; ++a) {
      float p0 = t[a] * rc, p1 = b[a] * rs, p2 = sn[a] * z;
      float s = p0 + p1;
      d[a] = s + p2;
    }
    arc_ray& out = rays[k];
    out.ox = org[0];
    out.oy = org[1];
    out.oz = org[2];
    out.t_start = 0.0f;
    out.dx = d[0];
    out.dy = d[1];
    out.dz = d[2];
    out.t_max = max_dist_m;
    if (src_index) src_index[k] = (uint32_t)i;
    ++k;
  }
  return k;
}

}  // extern "C"
