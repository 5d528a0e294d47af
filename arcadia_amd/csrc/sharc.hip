// SHARC voxel-surfel radiance cache on the device: key hashing, 4-way slot claim, fixed-point
// deposits, EMA resolve with eviction, and queries.
//
// Restates shaders/common/sharc.wgsl:36-116 (sharc_make_key, packing), shaders/sharc_update.wgsl:
// 129-169 (sharc_claim_slot, sharc_deposit), shaders/sharc_resolve.wgsl:13-69 (resolve) and
// shaders/common/sharc_query.wgsl:8-29 (sharc_lookup, query_sharc), with the same buffers:
// keys u32[S], accum u32[4S] = {r, g, b fixed-point x1024, count}, resolved 16 B slots
// {key, rg f16x2, b f16 | hist << 16 | age << 24, 0}; S = buckets * 4. The bucket count is a
// power of two chosen at creation (the reference fixes 32768).
//
// Arithmetic fixed where WGSL leaves it open: the cascade level floor(log2(y)) is the binary
// exponent of y (exact, no log2 rounding), distance = sqrt of the left-to-right dot product,
// mix(a, b, t) = a * (1 - t) + b * t, f16 packing rounds to nearest even.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "../../include/arcadia_hip.h"

namespace {

constexpr uint32_t kWays = 4u;
constexpr float kFixedScale = 1024.0f;
constexpr float kMaxRadiance = 32.0f;
constexpr uint32_t kEvictAge = 64u;
constexpr float kHistoryCap = 16.0f;
constexpr uint32_t kKeyValid = 0x80000000u;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t pcg(uint32_t input) {
  const uint32_t state = input * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

// floor(log2(max(d, 4) * 0.25)) clamped to 0..3: y >= 1, so the floor of log2 is y's exponent.
__device__ __forceinline__ uint32_t lod_for_distance(float d) {
  const float y = fmaxf(d, 4.0f) * 0.25f;
  const int e = (int)((__float_as_uint(y) >> 23) & 0xFFu) - 127;
  return (uint32_t)min(max(e, 0), 3);
}

__device__ __forceinline__ uint32_t norm_id(float nx, float ny, float nz) {
  const float ax = fabsf(nx), ay = fabsf(ny), az = fabsf(nz);
  if (ax > ay && ax > az) return nx > 0.0f ? 0u : 1u;
  if (ay > az) return ny > 0.0f ? 2u : 3u;
  return nz > 0.0f ? 4u : 5u;
}

__device__ __forceinline__ void normalize3(float& x, float& y, float& z) {
  const float xx = x * x, yy = y * y, zz = z * z;
  const float len = sqrtf((xx + yy) + zz);
  x = x / len;
  y = y / len;
  z = z / len;
}

struct Key {
  uint32_t bucket, key;
};

// sharc_make_key (sharc.wgsl:71-97)
__device__ __forceinline__ Key make_key(const float p[3], const float n[3], const float cam[3], float jx, float jy,
                                        uint32_t bucket_mask) {
  const float dx = p[0] - cam[0], dy = p[1] - cam[1], dz = p[2] - cam[2];
  const float xx = dx * dx, yy = dy * dy, zz = dz * dz;
  const float dist = sqrtf((xx + yy) + zz);
  const uint32_t lod = lod_for_distance(dist);
  const float s = 0.25f * (float)(1u << lod);
  float a[3];
  for (int k = 0; k < 3; ++k) {
    const float off = n[k] * 0.02f;
    a[k] = p[k] - off;
  }
  if (jx != 0.0f || jy != 0.0f) {  // sharc_tangent_frame (:62-68)
    const bool zup = fabsf(n[2]) < 0.999f;
    const float ux = zup ? 0.0f : 1.0f, uz = zup ? 1.0f : 0.0f;
    // t = normalize(cross(up, n)) with up = (ux, 0, uz), b = cross(n, t)
    const float t0 = 0.0f * n[2], t1 = uz * n[1];
    const float t2 = uz * n[0], t3 = ux * n[2];
    const float t4 = ux * n[1], t5 = 0.0f * n[0];
    float tx = t0 - t1, ty = t2 - t3, tz = t4 - t5;
    normalize3(tx, ty, tz);
    const float b0a = n[1] * tz, b0b = n[2] * ty;
    const float b1a = n[2] * tx, b1b = n[0] * tz;
    const float b2a = n[0] * ty, b2b = n[1] * tx;
    const float bx = b0a - b0b, by = b1a - b1b, bz = b2a - b2b;
    const float sc = 0.42f * s;
    const float t3v[3] = {tx, ty, tz}, b3v[3] = {bx, by, bz};
    for (int k = 0; k < 3; ++k) {
      const float u = t3v[k] * jx, v = b3v[k] * jy;
      const float w = (u + v) * sc;
      a[k] = a[k] + w;
    }
  }
  uint32_t cell[3];
  for (int k = 0; k < 3; ++k) {
    const float q = floorf(a[k] / s);
    cell[k] = (uint32_t)((int)q + 16384);
  }
  const uint32_t nid = norm_id(n[0], n[1], n[2]);
  Key out;
  out.bucket = pcg((cell[0] * 73856093u) ^ (cell[1] * 19349663u) ^ (cell[2] * 83492791u) ^ (nid * 2654435761u) ^
                   (lod * 38069u)) &
               bucket_mask;
  out.key = pcg(out.bucket ^ (cell[0] * 1013904223u) ^ (cell[1] * 1664525u) ^ (cell[2] * 2891336453u) ^ (nid << 20u) ^
                (lod << 24u)) |
            kKeyValid;
  return out;
}

__device__ __forceinline__ uint32_t f16_bits(float x) {
  const _Float16 h = (_Float16)x;
  return (uint32_t)__builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float f16_val(uint32_t b) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)(b & 0xFFFFu));
}

// sharc_claim_slot (sharc_update.wgsl:131-153): an existing surfel first, then the first empty way.
__device__ __forceinline__ uint32_t claim_slot(uint32_t* keys, Key k) {
  const uint32_t base = k.bucket * kWays;
  for (uint32_t i = 0; i < kWays; ++i)
    if (__hip_atomic_load(keys + base + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == k.key) return base + i;
  for (uint32_t i = 0; i < kWays; ++i) {
    const uint32_t old = atomicCAS(keys + base + i, 0u, k.key);
    if (old == 0u || old == k.key) return base + i;
  }
  return kNoSlot;
}

__global__ __launch_bounds__(256) void k_sharc_deposit(const arc_sharc_sample* __restrict__ samples, uint32_t n,
                                                       float cx, float cy, float cz, uint32_t bucket_mask,
                                                       uint32_t* __restrict__ keys, uint32_t* __restrict__ accum) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const arc_sharc_sample sm = samples[i];
  const float cam[3] = {cx, cy, cz};
  const Key k = make_key(sm.pos, sm.normal, cam, 0.0f, 0.0f, bucket_mask);
  const uint32_t slot = claim_slot(keys, k);
  if (slot == kNoSlot) return;
  uint32_t f[3];
  for (int c = 0; c < 3; ++c) {
    const float v = fminf(fmaxf(sm.radiance[c], 0.0f), kMaxRadiance) * kFixedScale;
    f[c] = (uint32_t)v;
  }
  uint32_t* a = accum + 4u * slot;
  atomicAdd(a, f[0]);
  atomicAdd(a + 1, f[1]);
  atomicAdd(a + 2, f[2]);
  atomicAdd(a + 3, 1u);
}

// sharc_resolve.wgsl:13-69, one thread per slot.
__global__ __launch_bounds__(256) void k_sharc_resolve(uint32_t n_slots, uint32_t* __restrict__ keys,
                                                       uint4* __restrict__ accum, uint4* __restrict__ resolved) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= n_slots) return;
  const uint32_t key = keys[slot];
  if (key == 0u) {
    resolved[slot] = make_uint4(0u, 0u, 0u, 0u);
    return;
  }
  const uint4 acc = accum[slot];
  const uint4 prev = resolved[slot];
  const bool prev_valid = prev.x == key;
  float r = 0.0f, g = 0.0f, b = 0.0f;
  uint32_t hist = 0u, age = 0u;
  if (prev_valid) {
    r = f16_val(prev.y);
    g = f16_val(prev.y >> 16);
    b = f16_val(prev.z);
    hist = (prev.z >> 16) & 0xFFu;
    age = (prev.z >> 24) & 0xFFu;
  }
  if (acc.w > 0u) {
    const float den = (float)acc.w * kFixedScale;
    const float inv = 1.0f / den;
    const float sr = (float)acc.x * inv, sg = (float)acc.y * inv, sb = (float)acc.z * inv;
    hist = min(hist + 1u, 255u);
    const float alpha = 1.0f / fminf((float)hist, kHistoryCap);
    const float om = 1.0f - alpha;
    const float r0 = r * om, r1 = sr * alpha;
    const float g0 = g * om, g1 = sg * alpha;
    const float b0 = b * om, b1 = sb * alpha;
    r = r0 + r1;
    g = g0 + g1;
    b = b0 + b1;
    age = 0u;
    accum[slot] = make_uint4(0u, 0u, 0u, 0u);
  } else {
    age = age + 1u;
    if (!prev_valid || age > kEvictAge) {
      keys[slot] = 0u;
      resolved[slot] = make_uint4(0u, 0u, 0u, 0u);
      return;
    }
  }
  const float cr = fminf(fmaxf(r, 0.0f), kMaxRadiance);
  const float cg = fminf(fmaxf(g, 0.0f), kMaxRadiance);
  const float cb = fminf(fmaxf(b, 0.0f), kMaxRadiance);
  const uint32_t rg = f16_bits(cr) | (f16_bits(cg) << 16);
  const uint32_t bm = f16_bits(cb) | ((hist & 0xFFu) << 16) | ((age & 0xFFu) << 24);
  resolved[slot] = make_uint4(key, rg, bm, 0u);
}

__device__ __forceinline__ float4 lookup(const uint4* resolved, Key k) {
  const uint32_t base = k.bucket * kWays;
  for (uint32_t i = 0; i < kWays; ++i) {
    const uint4 s = resolved[base + i];
    if (s.x == k.key) return make_float4(f16_val(s.y), f16_val(s.y >> 16), f16_val(s.z), 1.0f);
  }
  return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// query_sharc (sharc_query.wgsl:20-29)
__global__ __launch_bounds__(256) void k_sharc_query(const arc_sharc_probe* __restrict__ q, uint32_t n, float cx,
                                                     float cy, float cz, uint32_t bucket_mask,
                                                     const uint4* __restrict__ resolved, float4* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const arc_sharc_probe e = q[i];
  const float cam[3] = {cx, cy, cz};
  if (e.jitter[0] != 0.0f || e.jitter[1] != 0.0f) {
    const float4 j = lookup(resolved, make_key(e.pos, e.normal, cam, e.jitter[0], e.jitter[1], bucket_mask));
    if (j.w > 0.5f) {
      out[i] = j;
      return;
    }
  }
  out[i] = lookup(resolved, make_key(e.pos, e.normal, cam, 0.0f, 0.0f, bucket_mask));
}

// Deposits / queries straight from shaded path samples (x1 position, normal from its face code,
// radiance = the path's radiance beyond x1); samples without an x1 hit are skipped.
__device__ __forceinline__ bool path_x1(const arc_path_sample& ps, float n[3]) {
  n[0] = n[1] = n[2] = 0.0f;
  if (!(ps.flags & ARC_PS_X1_HIT)) return false;
  const uint32_t face = (ps.flags >> 8) & 0xFFu;
  if (face < 6u) n[face >> 1] = (face & 1u) ? -1.0f : 1.0f;
  return true;
}

__global__ __launch_bounds__(256) void k_sharc_deposit_paths(const arc_path_sample* __restrict__ s, uint32_t n,
                                                             float cx, float cy, float cz, uint32_t bucket_mask,
                                                             uint32_t* __restrict__ keys, uint32_t* __restrict__ accum) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const arc_path_sample ps = s[i];
  float nrm[3];
  if (!path_x1(ps, nrm)) return;
  const float cam[3] = {cx, cy, cz};
  const Key k = make_key(ps.pos, nrm, cam, 0.0f, 0.0f, bucket_mask);
  const uint32_t slot = claim_slot(keys, k);
  if (slot == kNoSlot) return;
  uint32_t* a = accum + 4u * slot;
  for (int c = 0; c < 3; ++c) {
    const float v = fminf(fmaxf(ps.target[c], 0.0f), kMaxRadiance) * kFixedScale;
    atomicAdd(a + c, (uint32_t)v);
  }
  atomicAdd(a + 3, 1u);
}

__global__ __launch_bounds__(256) void k_sharc_query_paths(const arc_path_sample* __restrict__ s, uint32_t n, float cx,
                                                           float cy, float cz, uint32_t bucket_mask,
                                                           const uint4* __restrict__ resolved, float4* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const arc_path_sample ps = s[i];
  float nrm[3];
  if (!path_x1(ps, nrm)) {
    out[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  const float cam[3] = {cx, cy, cz};
  out[i] = lookup(resolved, make_key(ps.pos, nrm, cam, 0.0f, 0.0f, bucket_mask));
}

// Evicts every surfel whose cell overlaps one of the spheres (x, y, z, r in meters, widened by
// the 2 cm anchor offset): one workgroup per (sphere, cascade level); each thread walks cells of
// the sphere's bounding box at that level and probes the 6 face keys of each. Most probed buckets
// hold none of the keys, so the keys are read first and the atomic runs only on a match.
__global__ __launch_bounds__(256) void k_sharc_invalidate(const float4* __restrict__ spheres, uint32_t bucket_mask,
                                                          uint32_t* __restrict__ keys, uint4* __restrict__ accum,
                                                          uint4* __restrict__ resolved) {
  const float4 sp = spheres[blockIdx.x >> 2];
  const uint32_t lod = blockIdx.x & 3u;
  const float s = 0.25f * (float)(1u << lod);
  const float r = sp.w + 0.05f;
  const int lo[3] = {(int)floorf((sp.x - r) / s), (int)floorf((sp.y - r) / s), (int)floorf((sp.z - r) / s)};
  const int hi[3] = {(int)floorf((sp.x + r) / s), (int)floorf((sp.y + r) / s), (int)floorf((sp.z + r) / s)};
  const uint32_t nx = (uint32_t)(hi[0] - lo[0] + 1), ny = (uint32_t)(hi[1] - lo[1] + 1), nz = (uint32_t)(hi[2] - lo[2] + 1);
  const uint32_t cells = nx * ny * nz;
  for (uint32_t c = threadIdx.x; c < cells; c += blockDim.x) {
    const uint32_t ix = c % nx, iy = (c / nx) % ny, iz = c / (nx * ny);
    const uint32_t cx = (uint32_t)(lo[0] + (int)ix + 16384), cy = (uint32_t)(lo[1] + (int)iy + 16384),
                   cz = (uint32_t)(lo[2] + (int)iz + 16384);
    for (uint32_t nid = 0; nid < 6; ++nid) {
      const uint32_t bucket =
          pcg((cx * 73856093u) ^ (cy * 19349663u) ^ (cz * 83492791u) ^ (nid * 2654435761u) ^ (lod * 38069u)) & bucket_mask;
      const uint32_t key = pcg(bucket ^ (cx * 1013904223u) ^ (cy * 1664525u) ^ (cz * 2891336453u) ^ (nid << 20u) ^
                               (lod << 24u)) |
                           kKeyValid;
      // the bucket's four keys in one load; a compare-and-swap only where one matches (no thread of
      // this launch writes a valid key, so a slot that does not hold `key` now will not later)
      const uint4 held = *reinterpret_cast<const uint4*>(keys + bucket * kWays);
      const uint32_t hk[kWays] = {held.x, held.y, held.z, held.w};
      for (uint32_t w = 0; w < kWays; ++w) {
        const uint32_t slot = bucket * kWays + w;
        if (hk[w] == key && atomicCAS(keys + slot, key, 0u) == key) {
          accum[slot] = make_uint4(0u, 0u, 0u, 0u);
          resolved[slot] = make_uint4(0u, 0u, 0u, 0u);
        }
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

}  // namespace

struct arc_sharc {
  int device = 0;
  uint32_t buckets = 0, slots = 0;
  uint32_t* keys = nullptr;
  uint32_t* accum = nullptr;
  uint4* resolved = nullptr;
  float* inval = nullptr;  // sphere list of arc_sharc_invalidate
  uint32_t inval_cap = 0;
};

extern "C" {

int arc_sharc_create(int device, uint32_t log2_buckets, arc_sharc** out) {
  if (!out || log2_buckets < 2 || log2_buckets > 24) return ARC_ERR_INVALID;
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) return ARC_ERR_HIP;
  arc_sharc* s = new (std::nothrow) arc_sharc();
  if (!s) return ARC_ERR_CAPACITY;
  s->device = device;
  s->buckets = 1u << log2_buckets;
  s->slots = s->buckets * kWays;
  hipError_t e = hipMalloc(&s->keys, (size_t)s->slots * 4);
  if (e == hipSuccess) e = hipMalloc(&s->accum, (size_t)s->slots * 16);
  if (e == hipSuccess) e = hipMalloc(&s->resolved, (size_t)s->slots * 16);
  if (e == hipSuccess) e = hipMemset(s->keys, 0, (size_t)s->slots * 4);
  if (e == hipSuccess) e = hipMemset(s->accum, 0, (size_t)s->slots * 16);
  if (e == hipSuccess) e = hipMemset(s->resolved, 0, (size_t)s->slots * 16);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    (void)hipFree(s->keys);
    (void)hipFree(s->accum);
    (void)hipFree(s->resolved);
    delete s;
    return ARC_ERR_HIP;
  }
  *out = s;
  return ARC_OK;
}

void arc_sharc_destroy(arc_sharc* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  (void)hipDeviceSynchronize();
  (void)hipFree(s->keys);
  (void)hipFree(s->accum);
  (void)hipFree(s->resolved);
  (void)hipFree(s->inval);
  delete s;
}

int arc_sharc_info(const arc_sharc* s, uint32_t* n_slots, uint32_t** keys, uint32_t** accum, void** resolved) {
  if (!s) return ARC_ERR_INVALID;
  if (n_slots) *n_slots = s->slots;
  if (keys) *keys = s->keys;
  if (accum) *accum = s->accum;
  if (resolved) *resolved = s->resolved;
  return ARC_OK;
}

int arc_sharc_read(const arc_sharc* s, uint32_t* keys, uint32_t* accum, uint32_t* resolved) {
  if (!s) return ARC_ERR_INVALID;
  if (hipSetDevice(s->device) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return ARC_ERR_HIP;
  hipError_t e = hipSuccess;
  if (keys) e = hipMemcpy(keys, s->keys, (size_t)s->slots * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && accum) e = hipMemcpy(accum, s->accum, (size_t)s->slots * 16, hipMemcpyDeviceToHost);
  if (e == hipSuccess && resolved) e = hipMemcpy(resolved, s->resolved, (size_t)s->slots * 16, hipMemcpyDeviceToHost);
  return e == hipSuccess ? ARC_OK : ARC_ERR_HIP;
}

int arc_sharc_clear(arc_sharc* s, void* stream) {
  if (!s) return ARC_ERR_INVALID;
  if (hipSetDevice(s->device) != hipSuccess) return ARC_ERR_HIP;
  hipStream_t st = (hipStream_t)stream;
  const uint64_t n[3] = {s->slots, (uint64_t)s->slots * 4, (uint64_t)s->slots * 4};
  uint32_t* p[3] = {s->keys, s->accum, reinterpret_cast<uint32_t*>(s->resolved)};
  for (int k = 0; k < 3; ++k) hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, st, p[k], n[k]);
  return hipGetLastError() == hipSuccess ? ARC_OK : ARC_ERR_HIP;
}

int arc_sharc_deposit(arc_sharc* s, const arc_sharc_sample* d_samples, uint64_t n, const float cam_pos[3],
                      void* stream) {
  if (!s || !cam_pos || (n && !d_samples) || n > 0x7FFFFFFFull) return ARC_ERR_INVALID;
  if (n == 0) return ARC_OK;
  if (hipSetDevice(s->device) != hipSuccess) return ARC_ERR_HIP;
  hipLaunchKernelGGL(k_sharc_deposit, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, d_samples,
                     (uint32_t)n, cam_pos[0], cam_pos[1], cam_pos[2], s->buckets - 1u, s->keys, s->accum);
  return hipGetLastError() == hipSuccess ? ARC_OK : ARC_ERR_HIP;
}

int arc_sharc_resolve(arc_sharc* s, void* stream) {
  if (!s) return ARC_ERR_INVALID;
  if (hipSetDevice(s->device) != hipSuccess) return ARC_ERR_HIP;
  hipLaunchKernelGGL(k_sharc_resolve, dim3((s->slots + 255) / 256), dim3(256), 0, (hipStream_t)stream, s->slots,
                     s->keys, reinterpret_cast<uint4*>(s->accum), s->resolved);
  return hipGetLastError() == hipSuccess ? ARC_OK : ARC_ERR_HIP;
}

int arc_sharc_query(const arc_sharc* s, const arc_sharc_probe* d_queries, uint64_t n, const float cam_pos[3],
                    float* d_out, void* stream) {
  if (!s || !cam_pos || (n && (!d_queries || !d_out)) || n > 0x7FFFFFFFull) return ARC_ERR_INVALID;
  if (n == 0) return ARC_OK;
  if (hipSetDevice(s->device) != hipSuccess) return ARC_ERR_HIP;
  hipLaunchKernelGGL(k_sharc_query, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, d_queries,
                     (uint32_t)n, cam_pos[0], cam_pos[1], cam_pos[2], s->buckets - 1u, s->resolved,
                     reinterpret_cast<float4*>(d_out));
  return hipGetLastError() == hipSuccess ? ARC_OK : ARC_ERR_HIP;
}

int arc_sharc_invalidate(arc_sharc* s, const float* spheres, uint32_t n, void* stream) {
  if (!s || (n && !spheres) || n > (1u << 20)) return ARC_ERR_INVALID;
  if (n == 0) return ARC_OK;
  for (uint32_t i = 0; i < n; ++i)
    if (!(spheres[4 * i + 3] >= 0.0f) || spheres[4 * i + 3] > 64.0f) return ARC_ERR_INVALID;  // <= 64 m
  if (hipSetDevice(s->device) != hipSuccess) return ARC_ERR_HIP;
  if (n > s->inval_cap) {
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return ARC_ERR_HIP;
    (void)hipFree(s->inval);
    s->inval = nullptr;
    s->inval_cap = 0;
    if (hipMalloc(&s->inval, (size_t)n * 16) != hipSuccess) return ARC_ERR_HIP;
    s->inval_cap = n;
  }
  if (hipMemcpyAsync(s->inval, spheres, (size_t)n * 16, hipMemcpyHostToDevice, (hipStream_t)stream) != hipSuccess)
    return ARC_ERR_HIP;
  hipLaunchKernelGGL(k_sharc_invalidate, dim3(4 * n), dim3(256), 0, (hipStream_t)stream,
                     reinterpret_cast<const float4*>(s->inval), s->buckets - 1u, s->keys,
                     reinterpret_cast<uint4*>(s->accum), s->resolved);
  if (hipGetLastError() != hipSuccess) return ARC_ERR_HIP;
  return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? ARC_OK : ARC_ERR_HIP;
}

int arc_sharc_deposit_paths(arc_sharc* s, const arc_path_sample* d_samples, uint64_t n, const float cam_pos[3],
                            void* stream) {
  if (!s || !cam_pos || (n && !d_samples) || n > 0x7FFFFFFFull) return ARC_ERR_INVALID;
  if (n == 0) return ARC_OK;
  if (hipSetDevice(s->device) != hipSuccess) return ARC_ERR_HIP;
  hipLaunchKernelGGL(k_sharc_deposit_paths, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     d_samples, (uint32_t)n, cam_pos[0], cam_pos[1], cam_pos[2], s->buckets - 1u, s->keys, s->accum);
  return hipGetLastError() == hipSuccess ? ARC_OK : ARC_ERR_HIP;
}

int arc_sharc_query_paths(const arc_sharc* s, const arc_path_sample* d_samples, uint64_t n, const float cam_pos[3],
                          float* d_out, void* stream) {
  if (!s || !cam_pos || (n && (!d_samples || !d_out)) || n > 0x7FFFFFFFull) return ARC_ERR_INVALID;
  if (n == 0) return ARC_OK;
  if (hipSetDevice(s->device) != hipSuccess) return ARC_ERR_HIP;
  hipLaunchKernelGGL(k_sharc_query_paths, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     d_samples, (uint32_t)n, cam_pos[0], cam_pos[1], cam_pos[2], s->buckets - 1u, s->resolved,
                     reinterpret_cast<float4*>(d_out));
  return hipGetLastError() == hipSuccess ? ARC_OK : ARC_ERR_HIP;
}

}  // extern "C"
