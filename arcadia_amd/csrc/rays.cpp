// Ray set generation (C ABI in include/arcadia_world.h). Binary32 arithmetic in source order
// (built with -ffp-contract=off), so the device ray generator of the G-buffer pass can repeat
// the primary-ray formula bit for bit.
#include <cmath>
#include <cstring>

#include "raymath.h"
#include "world_internal.h"

using arcmath::normalize3;

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
        float d[3];
        arcmath::primary_dir(*cam, px, py, width, height, d);
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
    const float ovox[3] = {pr.ox * 16.0f, pr.oy * 16.0f, pr.oz * 16.0f};
    const float dir[3] = {pr.dx, pr.dy, pr.dz};
    arcmath::bounce_ray(ovox, dir, h.t_voxels, h.face, seed, i, max_dist_m, &rays[k]);
    if (src_index) src_index[k] = (uint32_t)i;
    ++k;
  }
  return k;
}

int64_t arc_gen_view_rays(const arc_camera* cam, uint32_t width, uint32_t height, float max_dist_m, arc_ray* rays) {
  if (!cam || !rays || width == 0 || height == 0 || width > 65535 || height > 65535) return ARC_ERR_INVALID;
  const uint32_t tiles_x = (width + 7) / 8, tiles_y = (height + 7) / 8;
  uint64_t k = 0;
  for (uint32_t ty = 0; ty < tiles_y; ++ty)
    for (uint32_t tx = 0; tx < tiles_x; ++tx)
      for (uint32_t j = 0; j < 64; ++j, ++k) {
        const uint32_t px = tx * 8 + (j & 7), py = ty * 8 + (j >> 3);
        if (px >= width || py >= height) {
          rays[k] = arcmath::empty_ray();
          continue;
        }
        float d[3];
        arcmath::primary_dir(*cam, px, py, width, height, d);
        arc_ray& r = rays[k];
        r.ox = cam->eye[0];
        r.oy = cam->eye[1];
        r.oz = cam->eye[2];
        r.t_start = 0.0f;
        r.dx = d[0];
        r.dy = d[1];
        r.dz = d[2];
        r.t_max = max_dist_m;
      }
  return (int64_t)k;
}

int arc_gen_bounce_rays_dense(const arc_ray* prays, const arc_hit* hits, uint64_t n, uint64_t index_base,
                              uint64_t seed, float max_dist_m, arc_ray* rays) {
  if ((!prays || !hits || !rays) && n) return ARC_ERR_INVALID;
  for (uint64_t i = 0; i < n; ++i) {
    const arc_hit& h = hits[i];
    if (!(h.flags & ARC_HIT_FLAG_HIT)) {
      rays[i] = arcmath::empty_ray();
      continue;
    }
    const arc_ray& pr = prays[i];
    const float ovox[3] = {pr.ox * 16.0f, pr.oy * 16.0f, pr.oz * 16.0f};
    const float dir[3] = {pr.dx, pr.dy, pr.dz};
    arcmath::bounce_ray(ovox, dir, h.t_voxels, h.face, seed, index_base + i, max_dist_m, &rays[i]);
  }
  return ARC_OK;
}

}  // extern "C"
