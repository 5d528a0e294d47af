// Path shading of a traced batch: from the hit records of a path launch (primary + bounces,
// k_path) to path radiance, cache training samples and the FP64 accumulation of the judge.
//
// Light transport (no reference counterpart for the judge; assembled from the reference's
// material and sky definitions):
//   * surfaces are Lambertian with the palette albedo (crates/arcadia-core/src/material.rs:
//     99-204, albedo_roughness.rgb) and emit the palette emission (emission_meta.rgb) times
//     `emission_scale`;
//   * a ray that leaves the world (a miss) sees eval_sky_radiance(dir, sun_dir)
//     (shaders/common/types.wgsl:500-523) times `sky_scale`;
//   * bounce directions are the cosine samples of the path launch (arcmath::bounce_ray, the
//     Lambertian pdf cancels the cosine and 1/pi), so a path's radiance is
//     L = Le0 + a0 (Le1 + a1 (Le2 + ...)), truncated after the last bounce of the launch.
// Each slot's rays are regenerated here with the same functions the path launch used, so the
// walk follows exactly the rays that produced the records.
#include <hip/hip_runtime.h>

#include "pathtrace.h"
#include "raymath.h"

namespace arcdev {

__device__ __forceinline__ float mixf(float a, float b, float t) {
  const float u = a * (1.0f - t);
  const float v = b * t;
  return u + v;
}

// eval_sky_radiance (types.wgsl:500-523); pow(x, 3) and pow(x, 14) by repeated products.
__device__ __forceinline__ void sky_radiance(const float d[3], const float sun[3], float out[3]) {
  const float se = sun[1];
  const float day_t = fminf(fmaxf((se + 0.05f) / 0.30f, 0.0f), 1.0f);
  const float twi_t = fminf(fmaxf(1.0f - fabsf(se - 0.02f) / 0.22f, 0.0f), 1.0f);
  const float zen = fminf(fmaxf(d[1], 0.0f), 1.0f);
  const float om = 1.0f - zen;
  const float hor = (om * om) * om;
  const float dz[3] = {0.08f, 0.22f, 0.55f}, dh[3] = {0.38f, 0.50f, 0.68f};
  const float nz[3] = {0.0015f, 0.0025f, 0.0055f}, nh[3] = {0.0035f, 0.0065f, 0.0120f};
  const float tw[3] = {0.65f, 0.25f, 0.10f}, sc[3] = {0.85f, 0.52f, 0.24f};
  const float dp = (d[0] * sun[0] + d[1] * sun[1]) + d[2] * sun[2];
  const float sc0 = fmaxf(dp, 0.0f);
  const float x2 = sc0 * sc0, x4 = x2 * x2, x8 = x4 * x4;
  const float p14 = (x8 * x4) * x2;
  const float tw_w = twi_t * 0.75f;
  const float sun_w = fmaxf(day_t, twi_t);
  for (int c = 0; c < 3; ++c) {
    const float skz = mixf(nz[c], dz[c], day_t);
    float skh = mixf(nh[c], dh[c], day_t);
    skh = mixf(skh, tw[c], tw_w);
    float s = mixf(skz, skh, hor);
    if (se > -0.12f) {
      const float add = ((sc[c] * p14) * 0.35f) * sun_w;
      s = s + add;
    }
    out[c] = s;
  }
}

// One slot: walk the records of depth 0..n_bounces.
__global__ __launch_bounds__(256) void k_shade(ShadeArgs a) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= a.n_items) return;
  // view of the slot and its primary ray (k_path's path_start)
  uint32_t lo = 0u, hi = a.n_views;
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a.view_first[mid] <= slot) lo = mid;
    else hi = mid;
  }
  const ViewDev v = a.views[lo];
  const uint32_t local = slot - a.view_first[lo];
  const uint32_t tile = local >> 6, j = local & 63u;
  const uint32_t px = (tile % v.tiles_x) * 8u + (j & 7u), py = (tile / v.tiles_x) * 8u + (j >> 3);
  arc_path_sample out;
  for (int c = 0; c < 3; ++c) {
    out.pos[c] = out.dir[c] = out.le0[c] = out.a0[c] = out.base[c] = out.target[c] = 0.0f;
  }
  out.flags = 0u;
  out.pad0 = out.pad1 = out.pad2 = out.pad3 = out.pad4 = 0u;
  float L[3] = {0.0f, 0.0f, 0.0f};
  if (px < v.width && py < v.height) {
    out.flags |= ARC_PS_RAY;
    float dir[3];
    arcmath::primary_dir(v.cam, px, py, v.width, v.height, dir);
    float org[3] = {v.cam.eye[0], v.cam.eye[1], v.cam.eye[2]};
    float T[3] = {1.0f, 1.0f, 1.0f};   // throughput from the camera
    float T1[3] = {1.0f, 1.0f, 1.0f};  // throughput from x1
    for (uint32_t d = 0; d <= a.n_bounces; ++d) {
      const uint4 rec = a.hits[(size_t)d * a.n_items + slot];
      const bool hit = ((rec.w >> 16) & ARC_HIT_FLAG_HIT) != 0u;
      if (!hit) {
        float sky[3];
        sky_radiance(dir, a.sun, sky);
        for (int c = 0; c < 3; ++c) {
          const float s = sky[c] * a.sky_scale;
          const float t = T[c] * s;
          L[c] = L[c] + t;
          if (d == 0) out.le0[c] = s;
          if (d == 1) out.base[c] = s;
          if (d >= 2) {
            const float t1 = T1[c] * s;
            out.target[c] = out.target[c] + t1;
          }
        }
        break;
      }
      const float t_vox = __uint_as_float(rec.x);
      const uint32_t mat = (rec.z >> 16) & 0xFFu, face = rec.z >> 24;
      const float* m = a.palette + 8u * mat;
      float le[3], alb[3];
      for (int c = 0; c < 3; ++c) {
        le[c] = m[4 + c] * a.emission_scale;
        alb[c] = m[c];
      }
      for (int c = 0; c < 3; ++c) {
        const float t = T[c] * le[c];
        L[c] = L[c] + t;
      }
      if (d == 0) {
        out.flags |= ARC_PS_X0_HIT;
        for (int c = 0; c < 3; ++c) {
          out.le0[c] = le[c];
          out.a0[c] = alb[c];
        }
      } else if (d == 1) {
        out.flags |= ARC_PS_X1_HIT | (face << 8);
        const float ovox[3] = {org[0] * 16.0f, org[1] * 16.0f, org[2] * 16.0f};
        for (int c = 0; c < 3; ++c) {
          const float dt = dir[c] * t_vox;
          const float pv = ovox[c] + dt;
          out.pos[c] = pv * 0.0625f;
          out.base[c] = le[c];
        }
      } else {
        for (int c = 0; c < 3; ++c) {
          const float t1 = T1[c] * le[c];
          out.target[c] = out.target[c] + t1;
        }
      }
      if (d == a.n_bounces) break;
      if (d >= 1)
        for (int c = 0; c < 3; ++c) T1[c] = T1[c] * alb[c];
      for (int c = 0; c < 3; ++c) T[c] = T[c] * alb[c];
      const float ovox[3] = {org[0] * 16.0f, org[1] * 16.0f, org[2] * 16.0f};
      arc_ray r;
      arcmath::bounce_ray(ovox, dir, t_vox, face, a.seed, (uint64_t)d * a.n_items + slot, a.bounce_tmax, &r);
      org[0] = r.ox;
      org[1] = r.oy;
      org[2] = r.oz;
      dir[0] = r.dx;
      dir[1] = r.dy;
      dir[2] = r.dz;
      if (d == 0)
        for (int c = 0; c < 3; ++c) out.dir[c] = dir[c];
    }
  }
  if (a.samples) a.samples[slot] = out;
  if (a.radiance) {
    a.radiance[3 * (size_t)slot] = L[0];
    a.radiance[3 * (size_t)slot + 1] = L[1];
    a.radiance[3 * (size_t)slot + 2] = L[2];
  }
  if (a.accum) {
    for (int c = 0; c < 3; ++c) a.accum[3 * (size_t)slot + c] += (double)L[c];
  }
}

// est = Le0 + a0 * (base + cache) per slot; cache = cache_rgb (4 floats per slot: rgb, valid) when
// the slot's x1 was hit, ignored otherwise; `mode` 0 = no cache term, 1 = use the cache values.
__global__ __launch_bounds__(256) void k_estimate(const arc_path_sample* __restrict__ s, const float* __restrict__ cache,
                                                  uint32_t n, uint32_t mode, float* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const arc_path_sample p = s[i];
  const bool x1 = (p.flags & ARC_PS_X1_HIT) != 0u;
  for (int c = 0; c < 3; ++c) {
    float inner = p.base[c];
    if (x1 && mode == 1u) {
      const float cv = cache[4 * (size_t)i + c];
      inner = inner + fmaxf(cv, 0.0f);
    }
    const float t = p.a0[c] * inner;
    out[3 * (size_t)i + c] = p.le0[c] + t;
  }
}

hipError_t launch_shade(const ShadeArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_shade, dim3((a.n_items + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_estimate(const arc_path_sample* samples, const float* cache, uint32_t n, uint32_t mode, float* out,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_estimate, dim3((n + 255) / 256), dim3(256), 0, s, samples, cache, n, mode, out);
  return hipGetLastError();
}

}  // namespace arcdev
