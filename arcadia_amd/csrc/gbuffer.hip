// Primary G-buffer pass: device ray generation + G-buffer packing around the traversal.
//
// Restates primary_raymarch (shaders/primary_gbuffer.wgsl:93-312) for the opaque-world path:
// pinhole ray per pixel (:102-115), trace_world_64tree(eye, dir, max_dist, 0, underwater=false)
// (:118-121, done by k_trace in full-record mode), then the 20 B/px G-buffer (:138-143, :145-159,
// :297-311) with the WGSL pack built-ins (types.wgsl:292-332). Water (mat 18) and mirror (mat 28)
// secondary rays, rigid bodies and the underwater camera are not part of this path; their pixels
// take the opaque branch.
//
// Rays are generated in padded 8x8-tile order (ray i -> tile i/64, lane i%64) so a wave traces
// one screen tile; lanes past the image edge get an empty ray (t_max < 0: immediate miss) and
// write nothing.
#include <hip/hip_runtime.h>

#include "gbuffer.h"
#include "trace_device.h"

namespace arcdev {

__device__ __forceinline__ void tile_pixel(uint32_t i, uint32_t w, uint32_t* px, uint32_t* py) {
  const uint32_t tiles_x = (w + 7u) >> 3;
  const uint32_t tile = i >> 6, j = i & 63u;
  *px = (tile % tiles_x) * 8u + (j & 7u);
  *py = (tile / tiles_x) * 8u + (j >> 3);
}

// normalize(v) = v / sqrt(dot(v, v)), dot summed x, y, z left to right (same as the host path).
__device__ __forceinline__ void normalize3(float x, float y, float z, float* o) {
  const float xx = x * x, yy = y * y, zz = z * z;
  const float d = (xx + yy) + zz;
  const float len = sqrtf(d);
  o[0] = x / len;
  o[1] = y / len;
  o[2] = z / len;
}

__global__ __launch_bounds__(256) void k_gbuffer_raygen(GbufArgs g, arc_ray* __restrict__ rays, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t px, py;
  tile_pixel(i, g.width, &px, &py);
  arc_ray r;
  r.ox = g.eye[0];
  r.oy = g.eye[1];
  r.oz = g.eye[2];
  r.t_start = 0.0f;
  if (px >= g.width || py >= g.height) {
    r.dx = 0.0f;
    r.dy = 1.0f;
    r.dz = 0.0f;
    r.t_max = -1.0f;  // outside the image: immediate miss, not written
    rays[i] = r;
    return;
  }
  // primary_gbuffer.wgsl:103-115
  const float ux = ((float)px + 0.5f) + g.jitter[0];
  const float uy = ((float)py + 0.5f) + g.jitter[1];
  const float uvx = ux / (float)g.width, uvy = uy / (float)g.height;
  const float ndx = uvx * 2.0f - 1.0f;
  const float ndy = 1.0f - uvy * 2.0f;
  const float sx = (ndx * g.aspect) * g.tan_half_fov;
  const float sy = ndy * g.tan_half_fov;
  float v[3];
  for (int a = 0; a < 3; ++a) {
    const float rr = g.right[a] * sx;
    const float uu = g.up[a] * sy;
    const float s = g.forward[a] + rr;
    v[a] = s + uu;
  }
  float d[3];
  normalize3(v[0], v[1], v[2], d);
  r.dx = d[0];
  r.dy = d[1];
  r.dz = d[2];
  r.t_max = g.max_dist_m;
  rays[i] = r;
}

// WGSL pack built-ins (floor(0.5 + scale * clamp(e)) and f32 -> f16 round-to-nearest-even).
__device__ __forceinline__ uint32_t snorm16(float e) {
  const float c = fminf(1.0f, fmaxf(-1.0f, e));
  const float s = 32767.0f * c;
  return (uint32_t)(int32_t)floorf(0.5f + s) & 0xFFFFu;
}
__device__ __forceinline__ uint32_t unorm8(float e) {
  const float c = fminf(1.0f, fmaxf(0.0f, e));
  const float s = 255.0f * c;
  return (uint32_t)floorf(0.5f + s);
}
__device__ __forceinline__ uint32_t half_bits(float e) {
  const _Float16 h = (_Float16)e;
  return (uint32_t)__builtin_bit_cast(uint16_t, h);
}

// oct_encode (types.wgsl:112-119) + pack2x16snorm | pack2x16float(motion)
__device__ __forceinline__ uint2 pack_norm_motion(float nx, float ny, float nz, float mx, float my) {
  const float l1 = (fabsf(nx) + fabsf(ny)) + fabsf(nz);
  const float inv = 1.0f / fmaxf(l1, 1e-6f);
  float px = nx * inv, py = ny * inv;
  if (nz < 0.0f) {  // oct_wrap (types.wgsl:108-110)
    const float wx = (1.0f - fabsf(py)) * (px >= 0.0f ? 1.0f : -1.0f);
    const float wy = (1.0f - fabsf(px)) * (py >= 0.0f ? 1.0f : -1.0f);
    px = wx;
    py = wy;
  }
  return make_uint2(snorm16(px) | (snorm16(py) << 16), half_bits(mx) | (half_bits(my) << 16));
}

__global__ __launch_bounds__(256) void k_gbuffer_pack(GbufArgs g, const arc_ray* __restrict__ rays,
                                                      const arc_trace_result* __restrict__ res, uint32_t n,
                                                      uint2* __restrict__ depth_mat, uint2* __restrict__ norm_motion,
                                                      uint32_t* __restrict__ albedo_ao) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t px, py;
  tile_pixel(i, g.width, &px, &py);
  if (px >= g.width || py >= g.height) return;
  const uint32_t p = py * g.width + px;
  const arc_trace_result r = res[i];
  if (!(r.hit & 1u)) {  // :138-143
    const arc_ray ray = rays[i];
    depth_mat[p] = make_uint2(__float_as_uint(1e5f), 0u);
    norm_motion[p] = pack_norm_motion(ray.dx, ray.dy, ray.dz, 0.0f, 0.0f);
    albedo_ao[p] = unorm8(1.0f) << 24;
    return;
  }
  const float hit_dist = r.t_voxels * 0.0625f;
  const uint32_t mat = r.mat & 0xFFu;
  const float* m = g.palette + 8u * mat;
  // :154-162 per-voxel albedo variation, faded with distance
  const float tt = fminf(fmaxf((hit_dist - 18.0f) / (48.0f - 18.0f), 0.0f), 1.0f);
  const float sm = (tt * tt) * (3.0f - 2.0f * tt);
  const float w = 1.0f - sm;
  const float hv = (float)(r.voxel_hash & 0xFFu) / 255.0f;
  const float v_rand = (hv * 2.0f - 1.0f) * w;
  const float k = 1.0f + v_rand * m[7];
  float alb[3];
  for (int a = 0; a < 3; ++a) alb[a] = fminf(fmaxf(m[a] * k, 0.01f), 1.0f);
  // :283-292 motion vector from prev_view_proj (column-major, columns summed left to right)
  const float pu = ((float)px + 0.5f) / (float)g.width, pv = ((float)py + 0.5f) / (float)g.height;
  float clip[4];
  for (int row = 0; row < 4; ++row) {
    const float a0 = g.prev_view_proj[0 * 4 + row] * r.hit_pos_m[0];
    const float a1 = g.prev_view_proj[1 * 4 + row] * r.hit_pos_m[1];
    const float a2 = g.prev_view_proj[2 * 4 + row] * r.hit_pos_m[2];
    const float a3 = g.prev_view_proj[3 * 4 + row] * 1.0f;
    clip[row] = ((a0 + a1) + a2) + a3;
  }
  float mx = -2.0f, my = -2.0f;
  if (clip[3] > 0.001f) {
    const float nxc = clip[0] / clip[3], nyc = clip[1] / clip[3];
    const float ux = nxc * 0.5f + 0.5f;
    const float uy = 0.5f - nyc * 0.5f;
    mx = ux - pu;
    my = uy - pv;
  }
  depth_mat[p] = make_uint2(__float_as_uint(hit_dist), mat & 0xFFFFu);  // flags 0: no rigid-body hit
  norm_motion[p] = pack_norm_motion(r.normal[0], r.normal[1], r.normal[2], mx, my);
  const float ao = fminf(fmaxf(r.contact_ao, 0.0f), 1.0f);
  albedo_ao[p] = unorm8(alb[0]) | (unorm8(alb[1]) << 8) | (unorm8(alb[2]) << 16) | (unorm8(ao) << 24);
}

hipError_t launch_gbuffer_raygen(const GbufArgs& g, arc_ray* rays, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_gbuffer_raygen, dim3((n + 255) / 256), dim3(256), 0, s, g, rays, n);
  return hipGetLastError();
}

hipError_t launch_gbuffer_pack(const GbufArgs& g, const arc_ray* rays, const arc_trace_result* res, uint32_t n,
                               uint32_t* depth_mat, uint32_t* norm_motion, uint32_t* albedo_ao, hipStream_t s) {
  hipLaunchKernelGGL(k_gbuffer_pack, dim3((n + 255) / 256), dim3(256), 0, s, g, rays, res, n,
                     reinterpret_cast<uint2*>(depth_mat), reinterpret_cast<uint2*>(norm_motion), albedo_ao);
  return hipGetLastError();
}

}  // namespace arcdev
