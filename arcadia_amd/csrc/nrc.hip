// Neural radiance cache: multiresolution hash-grid encoding + fused FP16 MFMA MLP, trained
// online on the GPU (no reference counterpart; Arcadia's only cache is SHARC).
//
// Model: input = 16-level hash-grid encoding of the position (2 features per level, FP16 table
// of 2^T entries per level, trilinear) + surface normal (3) + view direction (3) + a constant 1
// (the first layer's bias), padded to 48; three hidden layers of 64 ReLU units; 3 linear outputs
// (radiance). Weights and table entries are FP16 in the kernels, FP32 master copies are updated
// by Adam; every product accumulates in FP32 on the matrix cores.
//
// MFMA formulation (v_mfma_f32_16x16x16_f16, wave64): a layer computes H^T = W . X^T with the
// weight tile as the A operand (lane l holds W[m = l%16][k = 4(l/16)..+3]) and the activations
// as B (lane l holds X^T[k = 4(l/16)..+3][n = l%16], n = sample). The accumulator D holds
// H^T[m = 4(l/16)+i][n = l%16]: exactly the B layout of the next layer's K tile, so activations
// pass from layer to layer in registers. The encoding is computed by the lanes that consume it:
// the lane group g = l/16 holds input features 4g..4g+3 of K tile 0 (levels 2g, 2g+1) and K tile
// 1 (levels 8+2g, 9+2g), so each lane encodes (and, in the backward pass, scatters the
// gradient of) 4 levels of its sample.
//
// Loss (training): relative L2, mean_c (y_c - t_c)^2 / (y_c^2 + 0.01) with the denominator held
// constant (Instant-NGP / neural radiance caching). Output gradients are not divided by the
// batch size in FP16; the weight gradients are scaled by 1/B in FP32 before Adam.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "../../include/arcadia_hip.h"

namespace {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kLevels = 16;
constexpr int kIn = 48, kHid = 64, kOut = 16;  // padded widths (3 outputs used)
constexpr int kPadIn = kIn + 4, kPadHid = kHid + 4;  // LDS row strides (halves): spread banks
// MLP parameters, in one FP32 / FP16 array: W1 [64][48], W2 [64][64], W3 [64][64], W4 [16][64]
constexpr int kOffW1 = 0, kOffW2 = kOffW1 + kHid * kIn, kOffW3 = kOffW2 + kHid * kHid, kOffW4 = kOffW3 + kHid * kHid;
constexpr int kMlpParams = kOffW4 + kOut * kHid;

struct GridInfo {
  float scale[kLevels];    // cells per unit length of the normalized position
  uint32_t res1[kLevels];  // scale + 1 (vertices per axis)
  uint32_t dense[kLevels]; // 1: direct index (res1^3 <= table), 0: spatial hash
  uint32_t table_mask;     // 2^T - 1
  float bbox_min[3], inv_size;
};

__device__ __forceinline__ floatx4 mfma(half4 a, half4 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// gfx950's double-K form: two 16-deep K tiles in one instruction. Lane group g supplies K slots
// 8g..8g+7 of both operands; giving it the concatenation of its two 16-deep slices (a0 ++ a1 and
// b0 ++ b1, the slices of K tiles t and t+1) pairs every weight with its activation, so the
// product is the sum of the two 16x16x16 products.
__device__ __forceinline__ floatx4 mfma2(half4 a0, half4 a1, half4 b0, half4 b1, floatx4 c) {
  const half8 a = half8{a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
  const half8 b = half8{b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

struct LevelPos {
  uint32_t c[3];
  float f[3];
};

// The 8 corner entries of a level around a position: vertex (x, y, z) of a dense level (res1^3 <=
// table entries) is entry x + y res1 + z res1^2, of a hashed level x ^ y 2654435761 ^ z 805459861,
// both masked to the table. The y and z terms take two values each, so they are formed once per
// level and a corner costs one three-input add or XOR, a select and a mask (a branch per corner on
// the dense flag diverged between the lane groups, which hold different levels).
struct CornerIdx {
  uint32_t x0, y0, y1, z0, z1, mask;
  bool dense;
  __device__ __forceinline__ uint32_t at(int k) const {
    const uint32_t x = x0 + (uint32_t)(k & 1), y = (k & 2) ? y1 : y0, z = (k & 4) ? z1 : z0;
    return (dense ? x + y + z : x ^ y ^ z) & mask;
  }
};

__device__ __forceinline__ LevelPos level_pos(const GridInfo& g, int l, const float p[3]) {
  LevelPos lp;
  for (int a = 0; a < 3; ++a) {
    const float x = p[a] * g.scale[l];
    const float fl = floorf(x);
    lp.c[a] = (uint32_t)fl;
    lp.f[a] = x - fl;
  }
  return lp;
}

__device__ __forceinline__ CornerIdx corner_idx(const GridInfo& g, int l, const LevelPos& lp) {
  CornerIdx c;
  c.dense = g.dense[l] != 0u;
  const uint32_t r = g.res1[l];
  const uint32_t my = c.dense ? r : 2654435761u, mz = c.dense ? r * r : 805459861u;
  c.x0 = lp.c[0];
  c.y0 = lp.c[1] * my;
  c.y1 = (lp.c[1] + 1u) * my;
  c.z0 = lp.c[2] * mz;
  c.z1 = (lp.c[2] + 1u) * mz;
  c.mask = g.table_mask;
  return c;
}

// Trilinear interpolation of the 2 features of level l at normalized position p. Tables hold at
// most 16 x 2^24 entries (arc_nrc_create), so byte offsets into the FP16-pair table (4 B per
// entry) and into its FP32 gradient (8 B per entry) fit 32 bits: the loads and atomics take a
// scalar base and a 32-bit vector offset.
__device__ __forceinline__ void encode_level(const GridInfo& g, const uint32_t* __restrict__ table_h, int l,
                                             const float p[3], float& e0, float& e1) {
  const LevelPos lp = level_pos(g, l, p);
  const CornerIdx ci = corner_idx(g, l, lp);
  const uint32_t lbase = (uint32_t)l * (g.table_mask + 1u);
  e0 = 0.0f;
  e1 = 0.0f;
  for (int k = 0; k < 8; ++k) {
    const uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
    const float w = (dx ? lp.f[0] : 1.0f - lp.f[0]) * (dy ? lp.f[1] : 1.0f - lp.f[1]) * (dz ? lp.f[2] : 1.0f - lp.f[2]);
    const uint32_t v = *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(table_h) +
                                                          ((lbase + ci.at(k)) << 2));
    const float f0 = (float)__builtin_bit_cast(_Float16, (uint16_t)(v & 0xFFFFu));
    const float f1 = (float)__builtin_bit_cast(_Float16, (uint16_t)(v >> 16));
    e0 += w * f0;
    e1 += w * f1;
  }
}

__device__ __forceinline__ void scatter_level(const GridInfo& g, float* __restrict__ grad, int l, const float p[3],
                                              float d0, float d1) {
  if (d0 == 0.0f && d1 == 0.0f) return;
  const LevelPos lp = level_pos(g, l, p);
  const CornerIdx ci = corner_idx(g, l, lp);
  const uint32_t lbase = (uint32_t)l * (g.table_mask + 1u);
  for (int k = 0; k < 8; ++k) {
    const uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
    const float w = (dx ? lp.f[0] : 1.0f - lp.f[0]) * (dy ? lp.f[1] : 1.0f - lp.f[1]) * (dz ? lp.f[2] : 1.0f - lp.f[2]);
    const uint32_t idx = ci.at(k);
    float* t = reinterpret_cast<float*>(reinterpret_cast<char*>(grad) + ((lbase + idx) << 3));
    atomicAdd(t, w * d0);
    atomicAdd(t + 1, w * d1);
  }
}

// Sample features shared by training and inference.
struct SampleIn {
  float p[3];   // normalized position in [0, 1]
  float n[3], d[3];
  bool valid;
};

__device__ __forceinline__ SampleIn load_sample(const arc_path_sample* s, uint32_t i, uint32_t n, const GridInfo& g) {
  SampleIn si;
  si.valid = false;
  for (int a = 0; a < 3; ++a) si.p[a] = si.n[a] = si.d[a] = 0.0f;
  if (i >= n) return si;
  const arc_path_sample ps = s[i];
  si.valid = (ps.flags & ARC_PS_X1_HIT) != 0u;
  const uint32_t face = (ps.flags >> 8) & 0xFFu;
  const float sgn = (face & 1u) ? -1.0f : 1.0f;  // selects, not a dynamic index into the array
  si.n[0] = face < 6u && (face >> 1) == 0u ? sgn : 0.0f;
  si.n[1] = face < 6u && (face >> 1) == 1u ? sgn : 0.0f;
  si.n[2] = face < 6u && (face >> 1) == 2u ? sgn : 0.0f;
  for (int a = 0; a < 3; ++a) {
    const float q = (ps.pos[a] - g.bbox_min[a]) * g.inv_size;
    si.p[a] = fminf(fmaxf(q, 0.0f), 0.99999994f);  // keeps every corner inside the level grid
    si.d[a] = ps.dir[a];
  }
  return si;
}

// B operands of the input layer for this lane's sample: K tiles 0, 1 (encoding), 2 (n, d, 1).
__device__ __forceinline__ void input_tiles(const GridInfo& g, const uint32_t* table_h, const SampleIn& si,
                                            uint32_t grp, half4 xb[3]) {
  for (int kt = 0; kt < 2; ++kt) {
    float e[4];
    encode_level(g, table_h, kt * 8 + 2 * grp, si.p, e[0], e[1]);
    encode_level(g, table_h, kt * 8 + 2 * grp + 1, si.p, e[2], e[3]);
    xb[kt] = half4{(_Float16)e[0], (_Float16)e[1], (_Float16)e[2], (_Float16)e[3]};
  }
  float x[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (grp == 0) {
    x[0] = si.n[0];
    x[1] = si.n[1];
    x[2] = si.n[2];
    x[3] = si.d[0];
  } else if (grp == 1) {
    x[0] = si.d[1];
    x[1] = si.d[2];
    x[2] = 1.0f;
  }
  xb[2] = half4{(_Float16)x[0], (_Float16)x[1], (_Float16)x[2], (_Float16)x[3]};
}

// A operand: rows 16*mt + l%16, columns 16*kt + 4*grp .. +3 of a row-major LDS matrix.
__device__ __forceinline__ half4 lds_a(const _Float16* w, int stride, int mt, int kt, uint32_t lane) {
  const _Float16* p = w + (16 * mt + (lane & 15)) * stride + 16 * kt + 4 * (lane >> 4);
  return *reinterpret_cast<const half4*>(p);
}

__device__ __forceinline__ half4 relu_h(floatx4 v) {
  return half4{(_Float16)fmaxf(v[0], 0.0f), (_Float16)fmaxf(v[1], 0.0f), (_Float16)fmaxf(v[2], 0.0f),
               (_Float16)fmaxf(v[3], 0.0f)};
}

// Stage W (row-major [rows][cols], FP16 in global) into LDS with row stride `stride`, and
// optionally its transpose [cols][rows] with stride `tstride`.
__device__ __forceinline__ void stage(const _Float16* __restrict__ src, int rows, int cols, _Float16* dst, int stride,
                                     _Float16* dstT, int tstride) {
  for (int i = threadIdx.x; i < rows * cols; i += blockDim.x) {
    const int r = i / cols, c = i % cols;
    const _Float16 v = src[i];
    dst[r * stride + c] = v;
    if (dstT) dstT[c * tstride + r] = v;
  }
}

struct alignas(16) MlpLds {
  _Float16 w1[kHid * kPadIn], w2[kHid * kPadHid], w3[kHid * kPadHid], w4[kOut * kPadHid];
};
struct alignas(16) MlpLdsT {
  _Float16 w1t[kIn * kPadHid], w2t[kHid * kPadHid], w3t[kHid * kPadHid], w4t[kHid * (kOut + 4)];
};

// Forward of one 16-sample tile; returns the output accumulator (rows 4*grp+i, sample lane%16)
// and, when `act` is given, the post-ReLU activations of the hidden layers (B layout).
__device__ __forceinline__ floatx4 forward_tile(const MlpLds& L, const half4 xb[3], uint32_t lane, half4 a1[4],
                                                half4 a2[4], half4 a3[4]) {
  const floatx4 z = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int mt = 0; mt < 4; ++mt) {
    // the third 16-deep input tile is paired with a zero tile: an accumulator chained from a
    // 32-deep into a 16-deep MFMA (either order) gave wrong sums on MI355X, so every product here
    // is the 32-deep form
    const half4 z4 = {0, 0, 0, 0};
    floatx4 acc = mfma2(lds_a(L.w1, kPadIn, mt, 0, lane), lds_a(L.w1, kPadIn, mt, 1, lane), xb[0], xb[1], z);
    acc = mfma2(lds_a(L.w1, kPadIn, mt, 2, lane), z4, xb[2], z4, acc);
    a1[mt] = relu_h(acc);
  }
  for (int mt = 0; mt < 4; ++mt) {
    floatx4 acc = z;
    for (int kt = 0; kt < 4; kt += 2)
      acc = mfma2(lds_a(L.w2, kPadHid, mt, kt, lane), lds_a(L.w2, kPadHid, mt, kt + 1, lane), a1[kt], a1[kt + 1], acc);
    a2[mt] = relu_h(acc);
  }
  for (int mt = 0; mt < 4; ++mt) {
    floatx4 acc = z;
    for (int kt = 0; kt < 4; kt += 2)
      acc = mfma2(lds_a(L.w3, kPadHid, mt, kt, lane), lds_a(L.w3, kPadHid, mt, kt + 1, lane), a2[kt], a2[kt + 1], acc);
    a3[mt] = relu_h(acc);
  }
  floatx4 y = z;
  for (int kt = 0; kt < 4; kt += 2)
    y = mfma2(lds_a(L.w4, kPadHid, 0, kt, lane), lds_a(L.w4, kPadHid, 0, kt + 1, lane), a3[kt], a3[kt + 1], y);
  return y;
}

// Spatial sort key of a query: Morton order of its normalized position on a 256^3 lattice (the
// order inference visits the queries in, so the lanes of a tile gather neighbouring table
// entries); queries without an x1 hit sort last.
constexpr int kSortBits = 24;

__device__ __forceinline__ uint32_t spread3_8(uint32_t v) {  // bit i -> bit 3i, 8 bits
  v = (v | (v << 8)) & 0x0F00Fu;
  v = (v | (v << 4)) & 0x0C30C3u;
  v = (v | (v << 2)) & 0x249249u;
  return v;
}

__global__ __launch_bounds__(256) void k_nrc_keys(GridInfo g, const arc_path_sample* __restrict__ s, uint32_t n,
                                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ idx) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SampleIn si = load_sample(s, i, n, g);
  uint32_t k = (1u << kSortBits) - 1u;
  if (si.valid) {
    const uint32_t x = (uint32_t)(si.p[0] * 256.0f), y = (uint32_t)(si.p[1] * 256.0f), z = (uint32_t)(si.p[2] * 256.0f);
    k = spread3_8(x) | (spread3_8(y) << 1) | (spread3_8(z) << 2);
  }
  keys[i] = k;
  idx[i] = i;
}

// perm (may be null): tile lane j handles query perm[16 * tile + j]; each query's output depends
// on that query alone, so the visiting order changes no result.
__global__ __launch_bounds__(256) void k_nrc_infer(GridInfo g0, const uint32_t* __restrict__ table_h,
                                                   const _Float16* __restrict__ mlp_h,
                                                   const arc_path_sample* __restrict__ s, uint32_t n,
                                                   const uint32_t* __restrict__ perm, float4* __restrict__ out) {
  __shared__ MlpLds L;
  // The level arrays are indexed per lane (each lane group encodes its own levels): read from this
  // LDS copy; indexed in the kernel argument they were copied to a per-lane stack first.
  __shared__ GridInfo sg;
  if (threadIdx.x == 0) sg = g0;
  stage(mlp_h + kOffW1, kHid, kIn, L.w1, kPadIn, nullptr, 0);
  stage(mlp_h + kOffW2, kHid, kHid, L.w2, kPadHid, nullptr, 0);
  stage(mlp_h + kOffW3, kHid, kHid, L.w3, kPadHid, nullptr, 0);
  stage(mlp_h + kOffW4, kOut, kHid, L.w4, kPadHid, nullptr, 0);
  __syncthreads();
  const GridInfo& g = sg;
  const uint32_t lane = threadIdx.x & 63u, grp = lane >> 4;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t tile = wave; tile * 16u < n; tile += nwaves) {
    const uint32_t j = tile * 16u + (lane & 15u);
    const uint32_t i = perm && j < n ? perm[j] : j;
    const SampleIn si = load_sample(s, i, n, g);
    half4 xb[3], a1[4], a2[4], a3[4];
    input_tiles(g, table_h, si, grp, xb);
    const floatx4 y = forward_tile(L, xb, lane, a1, a2, a3);
    if (grp == 0 && j < n) out[i] = make_float4(y[0], y[1], y[2], si.valid ? 1.0f : 0.0f);
  }
}

// Forward + backward of one batch; writes the activations and deltas ([feature][sample], FP16)
// for the weight-gradient kernel and accumulates the loss. The encoding gradient either goes to
// `enc_grad` ([level][sample] pairs, with the samples' grid positions in `enc_pos`) for
// k_nrc_scatter, or, when enc_grad is null, straight into the table gradient with global atomics.
// Two waves per SIMD (256 registers, a 12 B spill): 0.263 instead of 0.281 ms per training step
// of 66,048 samples on MI355X.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_nrc_train_fb(
    GridInfo g0, const uint32_t* __restrict__ table_h, const _Float16* __restrict__ mlp_h,
    const arc_path_sample* __restrict__ s, uint32_t n, uint32_t ld, _Float16* __restrict__ act,
    float* __restrict__ table_grad, float2* __restrict__ enc_grad, float4* __restrict__ enc_pos,
    float* __restrict__ loss) {
  __shared__ MlpLds L;
  __shared__ MlpLdsT T;
  __shared__ GridInfo sg;  // per-lane level lookups from LDS (as in k_nrc_infer)
  if (threadIdx.x == 0) sg = g0;
  stage(mlp_h + kOffW1, kHid, kIn, L.w1, kPadIn, T.w1t, kPadHid);
  stage(mlp_h + kOffW2, kHid, kHid, L.w2, kPadHid, T.w2t, kPadHid);
  stage(mlp_h + kOffW3, kHid, kHid, L.w3, kPadHid, T.w3t, kPadHid);
  stage(mlp_h + kOffW4, kOut, kHid, L.w4, kPadHid, T.w4t, kOut + 4);
  __syncthreads();
  const GridInfo& g = sg;
  const uint32_t lane = threadIdx.x & 63u, grp = lane >> 4;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  // activation / delta planes in `act`, each [rows][ld] halves
  _Float16* X = act;                 // 48 rows
  _Float16* A1 = X + (size_t)kIn * ld;   // 64
  _Float16* A2 = A1 + (size_t)kHid * ld;
  _Float16* A3 = A2 + (size_t)kHid * ld;
  _Float16* D1 = A3 + (size_t)kHid * ld;
  _Float16* D2 = D1 + (size_t)kHid * ld;
  _Float16* D3 = D2 + (size_t)kHid * ld;
  _Float16* D4 = D3 + (size_t)kHid * ld;  // 16 rows
  float lsum = 0.0f;
  const floatx4 z = {0.0f, 0.0f, 0.0f, 0.0f};
  for (uint32_t tile = wave; tile * 16u < n; tile += nwaves) {
    const uint32_t col = tile * 16u + (lane & 15u);
    const uint32_t i = col;
    const SampleIn si = load_sample(s, i, n, g);
    if (enc_pos && grp == 0u && i < n) enc_pos[i] = make_float4(si.p[0], si.p[1], si.p[2], si.valid ? 1.0f : 0.0f);
    half4 xb[3], a1[4], a2[4], a3[4];
    input_tiles(g, table_h, si, grp, xb);
    const floatx4 y = forward_tile(L, xb, lane, a1, a2, a3);
    // output gradient: rows 0..2 live in lane group 0
    float tgt[3] = {0.0f, 0.0f, 0.0f};
    if (grp == 0 && si.valid) {
      const arc_path_sample ps = s[i];
      tgt[0] = ps.target[0];
      tgt[1] = ps.target[1];
      tgt[2] = ps.target[2];
    }
    float dy[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (grp == 0 && si.valid) {
      for (int c = 0; c < 3; ++c) {
        const float e = y[c] - tgt[c];
        const float den = y[c] * y[c] + 0.01f;
        lsum += e * e / den * (1.0f / 3.0f);
        dy[c] = 2.0f * e / den * (1.0f / 3.0f);
      }
    }
    const half4 d4 = half4{(_Float16)dy[0], (_Float16)dy[1], (_Float16)dy[2], (_Float16)dy[3]};
    // backward through W4: dA3^T (64 x s) = W4^T (64 x 16) . D4 (16 x s)
    half4 d3[4], d2[4], d1[4];
    for (int mt = 0; mt < 4; ++mt) {
      floatx4 acc = mfma(lds_a(T.w4t, kOut + 4, mt, 0, lane), d4, z);
      for (int k = 0; k < 4; ++k) acc[k] = (float)a3[mt][k] > 0.0f ? acc[k] : 0.0f;
      d3[mt] = half4{(_Float16)acc[0], (_Float16)acc[1], (_Float16)acc[2], (_Float16)acc[3]};
    }
    for (int mt = 0; mt < 4; ++mt) {
      floatx4 acc = z;
      for (int kt = 0; kt < 4; ++kt) acc = mfma(lds_a(T.w3t, kPadHid, mt, kt, lane), d3[kt], acc);
      for (int k = 0; k < 4; ++k) acc[k] = (float)a2[mt][k] > 0.0f ? acc[k] : 0.0f;
      d2[mt] = half4{(_Float16)acc[0], (_Float16)acc[1], (_Float16)acc[2], (_Float16)acc[3]};
    }
    for (int mt = 0; mt < 4; ++mt) {
      floatx4 acc = z;
      for (int kt = 0; kt < 4; ++kt) acc = mfma(lds_a(T.w2t, kPadHid, mt, kt, lane), d2[kt], acc);
      for (int k = 0; k < 4; ++k) acc[k] = (float)a1[mt][k] > 0.0f ? acc[k] : 0.0f;
      d1[mt] = half4{(_Float16)acc[0], (_Float16)acc[1], (_Float16)acc[2], (_Float16)acc[3]};
    }
    // encoding gradient: dX^T rows 0..31 (K tiles 0, 1 of the input) = W1^T . D1
    for (int mt = 0; mt < 2; ++mt) {
      floatx4 acc = z;
      for (int kt = 0; kt < 4; ++kt) acc = mfma(lds_a(T.w1t, kPadHid, mt, kt, lane), d1[kt], acc);
      if (enc_grad) {
        if (i < n) {
          const bool v = si.valid;
          enc_grad[(size_t)(mt * 8 + 2 * grp) * ld + i] = make_float2(v ? acc[0] : 0.0f, v ? acc[1] : 0.0f);
          enc_grad[(size_t)(mt * 8 + 2 * grp + 1) * ld + i] = make_float2(v ? acc[2] : 0.0f, v ? acc[3] : 0.0f);
        }
      } else if (si.valid) {
        scatter_level(g, table_grad, mt * 8 + 2 * grp, si.p, acc[0], acc[1]);
        scatter_level(g, table_grad, mt * 8 + 2 * grp + 1, si.p, acc[2], acc[3]);
      }
    }
    // store activations and deltas for the weight gradients (rows 16*t + 4*grp + k)
    if (col < ld) {
      for (int k = 0; k < 4; ++k) {
        const size_t r = 4 * grp + k;
        for (int t = 0; t < 3; ++t) X[(16 * t + r) * ld + col] = xb[t][k];
        for (int t = 0; t < 4; ++t) {
          A1[(16 * t + r) * ld + col] = a1[t][k];
          A2[(16 * t + r) * ld + col] = a2[t][k];
          A3[(16 * t + r) * ld + col] = a3[t][k];
          D1[(16 * t + r) * ld + col] = d1[t][k];
          D2[(16 * t + r) * ld + col] = d2[t][k];
          D3[(16 * t + r) * ld + col] = d3[t][k];
        }
        D4[r * ld + col] = d4[k];
      }
    }
  }
  // loss: wave sum then one atomic per wave
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
  if (lane == 0 && lsum != 0.0f) atomicAdd(loss, lsum);
}

// Table-gradient scatter through LDS, for tables of at most kLdsScatterEntries entries per level.
// Random FP32 atomics into device memory ran at about 21 G/s on MI355X, with the same rate from
// 1024 to 16384 workgroups and with one table copy per XCD (measured); the batch's 16 levels x
// 8 corners x 2 features per sample (16.9 M at 2^16 paths) took 0.70 of a 0.88 ms training step
// that way. Here workgroup (level l, chunk c) adds the gradients of its samples' 8 corners at
// level l into an LDS copy of that level's table with LDS atomics, then stores the copy to its
// partial slice, and k_nrc_reduce adds the slices into the table gradient: no global atomics.
constexpr uint32_t kLdsScatterEntries = 1u << 14;  // 128 KB of FP32 pairs per workgroup (of 1024 lanes: 16 waves per CU)

__global__ __launch_bounds__(1024) void k_nrc_scatter(GridInfo g0, const float2* __restrict__ enc_grad,
                                                     const float4* __restrict__ enc_pos, uint32_t n, uint32_t ld,
                                                     uint32_t chunk, float* __restrict__ partial) {
  __shared__ float tab[2 * kLdsScatterEntries];
  __shared__ GridInfo sg;  // level lookups from LDS (a dynamically indexed kernel argument goes to the stack)
  if (threadIdx.x == 0) sg = g0;
  const uint32_t l = blockIdx.x % (uint32_t)kLevels, c = blockIdx.x / (uint32_t)kLevels;
  const uint32_t words = 2u * (g0.table_mask + 1u);  // <= 2 * kLdsScatterEntries (checked by the host)
  for (uint32_t j = threadIdx.x; j < words; j += blockDim.x) tab[j] = 0.0f;
  __syncthreads();
  const GridInfo& g = sg;
  const uint32_t i0 = c * chunk, i1 = i0 + chunk < n ? i0 + chunk : n;
  for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const float4 q = enc_pos[i];
    const float2 d = enc_grad[(size_t)l * ld + i];
    if (q.w == 0.0f || (d.x == 0.0f && d.y == 0.0f)) continue;
    const float p[3] = {q.x, q.y, q.z};
    const LevelPos lp = level_pos(g, (int)l, p);
    const CornerIdx ci = corner_idx(g, (int)l, lp);
    for (int k = 0; k < 8; ++k) {
      const uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
      const float w = (dx ? lp.f[0] : 1.0f - lp.f[0]) * (dy ? lp.f[1] : 1.0f - lp.f[1]) * (dz ? lp.f[2] : 1.0f - lp.f[2]);
      const uint32_t e = ci.at(k);
      atomicAdd(&tab[2u * e], w * d.x);
      atomicAdd(&tab[2u * e + 1u], w * d.y);
    }
  }
  __syncthreads();
  float* out = partial + ((size_t)c * kLevels + l) * words;
  for (uint32_t j = threadIdx.x; j < words; j += blockDim.x) out[j] = tab[j];
}

// grad[j] += sum over the nparts partial slices (each `words` floats, the gradient's layout).
__global__ __launch_bounds__(256) void k_nrc_reduce(const float4* __restrict__ partial, uint32_t nparts, uint64_t words4,
                                                    float4* __restrict__ grad) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < words4; j += (uint64_t)gridDim.x * blockDim.x) {
    float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t c = 0; c < nparts; ++c) {
      const float4 v = partial[(size_t)c * words4 + j];
      a.x += v.x;
      a.y += v.y;
      a.z += v.z;
      a.w += v.w;
    }
    if (a.x != 0.0f || a.y != 0.0f || a.z != 0.0f || a.w != 0.0f) {
      float4 gv = grad[j];
      gv.x += a.x;
      gv.y += a.y;
      gv.z += a.z;
      gv.w += a.w;
      grad[j] = gv;
    }
  }
}

// Weight gradient of one layer over a chunk of samples: dW (rows x cols) += D (rows x s) . A^T
// (s x cols), both planes [feature][sample] FP16; one wave per chunk, all 16x16 tiles.
template <int ROWS, int COLS>
__global__ __launch_bounds__(64) void k_nrc_dw(const _Float16* __restrict__ D, const _Float16* __restrict__ A,
                                               uint32_t ld, uint32_t n, uint32_t chunk, float* __restrict__ dw) {
  constexpr int MT = ROWS / 16, NT = COLS / 16;
  const uint32_t lane = threadIdx.x, grp = lane >> 4;
  const uint32_t s0 = blockIdx.x * chunk, s1 = min(n, s0 + chunk);
  floatx4 acc[MT][NT];
  for (int m = 0; m < MT; ++m)
    for (int c = 0; c < NT; ++c) acc[m][c] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  for (uint32_t k0 = s0; k0 < s1; k0 += 16) {
    const uint32_t ks = k0 + 4 * grp;  // this lane's 4 samples
    half4 a[MT], b[NT];
    for (int m = 0; m < MT; ++m) a[m] = *reinterpret_cast<const half4*>(D + (size_t)(16 * m + (lane & 15)) * ld + ks);
    for (int c = 0; c < NT; ++c) b[c] = *reinterpret_cast<const half4*>(A + (size_t)(16 * c + (lane & 15)) * ld + ks);
    for (int m = 0; m < MT; ++m)
      for (int c = 0; c < NT; ++c) acc[m][c] = mfma(a[m], b[c], acc[m][c]);
  }
  for (int m = 0; m < MT; ++m)
    for (int c = 0; c < NT; ++c)
      for (int k = 0; k < 4; ++k) atomicAdd(dw + (16 * m + 4 * grp + k) * COLS + 16 * c + (lane & 15), acc[m][c][k]);
}

// Adam on FP32 masters with an FP16 (or packed FP16x2) copy for the kernels; zero gradients are
// skipped (entries no sample touched keep their moments), gradients are cleared.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                              float* __restrict__ grad, _Float16* __restrict__ ph, uint64_t n,
                                              float gscale, float lr, float b1, float b2, float eps, float bc1,
                                              float bc2) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const float gr = grad[i];
    if (gr == 0.0f) continue;
    const float gv = gr * gscale;
    const float mi = b1 * m[i] + (1.0f - b1) * gv;
    const float vi = b2 * v[i] + (1.0f - b2) * gv * gv;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1, vh = vi / bc2;
    const float pi = p[i] - lr * mh / (sqrtf(vh) + eps);
    p[i] = pi;
    ph[i] = (_Float16)pi;
    grad[i] = 0.0f;
  }
}

__global__ __launch_bounds__(256) void k_to_half(const float* __restrict__ p, _Float16* __restrict__ ph, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    ph[i] = (_Float16)p[i];
}

__global__ __launch_bounds__(256) void k_zero_f32(float* __restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

uint64_t splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

float uniform01(uint64_t& st) {
  st = splitmix(st);
  return (float)(st >> 40) * (1.0f / 16777216.0f);
}

}  // namespace

struct arc_nrc {
  int device = 0;
  arc_nrc_params prm{};
  GridInfo g{};
  uint64_t table_entries = 0;  // per level
  uint64_t grid_params = 0;    // levels * entries * 2
  float *grid_p = nullptr, *grid_m = nullptr, *grid_v = nullptr, *grid_g = nullptr;
  _Float16* grid_h = nullptr;
  float *mlp_p = nullptr, *mlp_m = nullptr, *mlp_v = nullptr, *mlp_g = nullptr;
  _Float16* mlp_h = nullptr;
  _Float16* act = nullptr;  // training planes
  uint32_t act_ld = 0;
  // LDS table-gradient scatter (tables of <= kLdsScatterEntries entries per level): encoding
  // gradients [level][act_ld] and grid positions [act_ld] of a batch, and the workgroups' partial
  // gradient slices (scatter_parts x grid_params floats)
  float2* enc_grad = nullptr;
  float4* enc_pos = nullptr;
  float* scatter_part = nullptr;
  uint32_t scatter_parts = 0;
  float* loss = nullptr;
  uint64_t step = 0;
  uint32_t n_cu = 0;
  // query sort workspace: keys in/out, indices in/out (sort_cap entries each) + radix-sort scratch
  uint32_t* sort_buf = nullptr;
  uint64_t sort_cap = 0;
  void* sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
};

namespace {

int nrc_alloc(arc_nrc* r) {
  const uint64_t gp = r->grid_params;
  hipError_t e = hipSuccess;
  float** f32[] = {&r->grid_p, &r->grid_m, &r->grid_v, &r->grid_g};
  for (float** p : f32)
    if (e == hipSuccess) e = hipMalloc(p, gp * 4);
  if (e == hipSuccess) e = hipMalloc(&r->grid_h, gp * 2);
  float** m32[] = {&r->mlp_p, &r->mlp_m, &r->mlp_v, &r->mlp_g};
  for (float** p : m32)
    if (e == hipSuccess) e = hipMalloc(p, kMlpParams * 4);
  if (e == hipSuccess) e = hipMalloc(&r->mlp_h, kMlpParams * 2);
  if (e == hipSuccess) e = hipMalloc(&r->loss, 4);
  return e == hipSuccess ? ARC_OK : ARC_ERR_HIP;
}

void nrc_free(arc_nrc* r) {
  void* ps[] = {r->grid_p, r->grid_m, r->grid_v,  r->grid_g,   r->grid_h,   r->mlp_p,    r->mlp_m,
                r->mlp_v,  r->mlp_g, r->mlp_h,   r->act,      r->loss,     r->sort_buf, r->sort_tmp,
                r->enc_grad, r->enc_pos, r->scatter_part};
  for (void* p : ps) (void)hipFree(p);
}

// Sorts the queries of an inference launch by their Morton key (k_nrc_keys) into the object's
// workspace (grown on demand) and returns the visiting order: the lanes of a tile then gather
// neighbouring table entries (measured on MI355X at 1080p: 0.84 -> 0.52 ms of inference for
// 0.1 ms of sorting). Calls on one cache object share the workspace: issue them on one stream.
int sort_queries(arc_nrc* r, const arc_path_sample* d_samples, uint64_t n, hipStream_t s, const uint32_t** perm) {
  if (n > r->sort_cap) {
    size_t tmp = 0;
    if (rocprim::radix_sort_pairs(nullptr, tmp, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                  (uint32_t*)nullptr, (size_t)n, 0, kSortBits, s) != hipSuccess)
      return ARC_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return ARC_ERR_HIP;
    (void)hipFree(r->sort_buf);
    (void)hipFree(r->sort_tmp);
    r->sort_buf = nullptr;
    r->sort_tmp = nullptr;
    r->sort_cap = 0;
    if (hipMalloc(&r->sort_buf, n * 4 * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&r->sort_tmp, tmp > 0 ? tmp : 1) != hipSuccess)
      return ARC_ERR_HIP;
    r->sort_cap = n;
    r->sort_tmp_bytes = tmp;
  }
  uint32_t* keys_in = r->sort_buf;
  uint32_t* keys_out = keys_in + r->sort_cap;
  uint32_t* idx_in = keys_out + r->sort_cap;
  uint32_t* idx_out = idx_in + r->sort_cap;
  hipLaunchKernelGGL(k_nrc_keys, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, r->g, d_samples, (uint32_t)n,
                     keys_in, idx_in);
  size_t tmp = r->sort_tmp_bytes;
  if (rocprim::radix_sort_pairs(r->sort_tmp, tmp, keys_in, keys_out, idx_in, idx_out, (size_t)n, 0, kSortBits, s) !=
      hipSuccess)
    return ARC_ERR_HIP;
  *perm = idx_out;
  return hipGetLastError() == hipSuccess ? ARC_OK : ARC_ERR_HIP;
}

int launch_zero(float* p, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_f32, dim3(1024), dim3(256), 0, s, p, n);
  return hipGetLastError() == hipSuccess ? ARC_OK : ARC_ERR_HIP;
}

}  // namespace

extern "C" {

int arc_nrc_create(int device, const arc_nrc_params* prm, arc_nrc** out) {
  if (!prm || !out) return ARC_ERR_INVALID;
  *out = nullptr;
  if (prm->log2_table < 10 || prm->log2_table > 24 || !(prm->bbox_size > 0.0f) || !(prm->base_res >= 1.0f) ||
      !(prm->max_res >= prm->base_res))
    return ARC_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return ARC_ERR_HIP;
  arc_nrc* r = new (std::nothrow) arc_nrc();
  if (!r) return ARC_ERR_CAPACITY;
  r->device = device;
  r->prm = *prm;
  r->table_entries = 1ull << prm->log2_table;
  r->grid_params = (uint64_t)kLevels * r->table_entries * 2;
  GridInfo& g = r->g;
  const double b = std::exp((std::log((double)prm->max_res) - std::log((double)prm->base_res)) / (kLevels - 1));
  for (int l = 0; l < kLevels; ++l) {
    const double res = std::floor((double)prm->base_res * std::pow(b, l));
    g.scale[l] = (float)res;
    g.res1[l] = (uint32_t)res + 1u;
    const double cells = std::pow((double)g.res1[l], 3.0);
    g.dense[l] = cells <= (double)r->table_entries ? 1u : 0u;
  }
  g.table_mask = (uint32_t)(r->table_entries - 1);
  for (int a = 0; a < 3; ++a) g.bbox_min[a] = prm->bbox_min[a];
  g.inv_size = 1.0f / prm->bbox_size;
  int dev_cu = 0;
  (void)hipDeviceGetAttribute(&dev_cu, hipDeviceAttributeMultiprocessorCount, device);
  r->n_cu = dev_cu > 0 ? (uint32_t)dev_cu : 256u;
  int rc = nrc_alloc(r);
  if (rc == ARC_OK) rc = arc_nrc_reset(r, prm->seed);
  if (rc != ARC_OK) {
    nrc_free(r);
    delete r;
    return rc;
  }
  *out = r;
  return ARC_OK;
}

void arc_nrc_destroy(arc_nrc* r) {
  if (!r) return;
  (void)hipSetDevice(r->device);
  (void)hipDeviceSynchronize();
  nrc_free(r);
  delete r;
}

int arc_nrc_reset(arc_nrc* r, uint64_t seed) {
  if (!r) return ARC_ERR_INVALID;
  if (hipSetDevice(r->device) != hipSuccess) return ARC_ERR_HIP;
  // table: U(-1e-4, 1e-4); MLP: uniform He-style init on the used fan-in, padding columns zero
  std::vector<float> grid(r->grid_params), mlp(kMlpParams, 0.0f);
  uint64_t st = splitmix(seed ^ 0xA5A5A5A5ull);
  for (auto& x : grid) x = (uniform01(st) * 2.0f - 1.0f) * 1e-4f;
  auto init = [&](int off, int rows, int cols, int used_rows, int used_cols) {
    const float lim = std::sqrt(6.0f / (float)used_cols);
    for (int i = 0; i < rows; ++i)
      for (int j = 0; j < cols; ++j)
        mlp[off + i * cols + j] = (i < used_rows && j < used_cols) ? (uniform01(st) * 2.0f - 1.0f) * lim : 0.0f;
  };
  init(kOffW1, kHid, kIn, kHid, 39);  // 32 encoding + 3 normal + 3 dir + 1 constant
  init(kOffW2, kHid, kHid, kHid, kHid);
  init(kOffW3, kHid, kHid, kHid, kHid);
  init(kOffW4, kOut, kHid, 3, kHid);
  hipError_t e = hipMemcpy(r->grid_p, grid.data(), r->grid_params * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(r->mlp_p, mlp.data(), kMlpParams * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) return ARC_ERR_HIP;
  int rc = ARC_OK;
  float* zs[] = {r->grid_m, r->grid_v, r->grid_g};
  for (float* z : zs)
    if (rc == ARC_OK) rc = launch_zero(z, r->grid_params, nullptr);
  float* mz[] = {r->mlp_m, r->mlp_v, r->mlp_g};
  for (float* z : mz)
    if (rc == ARC_OK) rc = launch_zero(z, kMlpParams, nullptr);
  if (rc != ARC_OK) return rc;
  hipLaunchKernelGGL(k_to_half, dim3(1024), dim3(256), 0, nullptr, r->grid_p, r->grid_h, r->grid_params);
  hipLaunchKernelGGL(k_to_half, dim3(16), dim3(256), 0, nullptr, r->mlp_p, r->mlp_h, (uint64_t)kMlpParams);
  if (hipDeviceSynchronize() != hipSuccess) return ARC_ERR_HIP;
  r->step = 0;
  return ARC_OK;
}

int arc_nrc_infer(arc_nrc* r, const arc_path_sample* d_samples, uint64_t n, float* d_out, void* stream) {
  if (!r || (n && (!d_samples || !d_out)) || n > 0x7FFFFFFFull) return ARC_ERR_INVALID;
  if (n == 0) return ARC_OK;
  if (hipSetDevice(r->device) != hipSuccess) return ARC_ERR_HIP;
  hipStream_t s = (hipStream_t)stream;
  const uint32_t* perm = nullptr;
  if (sort_queries(r, d_samples, n, s, &perm) != ARC_OK) return ARC_ERR_HIP;
  const uint64_t tiles = (n + 15) / 16;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((tiles + 3) / 4, (uint64_t)r->n_cu * 8);
  hipLaunchKernelGGL(k_nrc_infer, dim3(blocks), dim3(256), 0, s, r->g, reinterpret_cast<const uint32_t*>(r->grid_h),
                     r->mlp_h, d_samples, (uint32_t)n, perm, reinterpret_cast<float4*>(d_out));
  return hipGetLastError() == hipSuccess ? ARC_OK : ARC_ERR_HIP;
}

// Gradients of the summed loss over the batch into the FP32 gradient buffers (no update).
int arc_nrc_backward(arc_nrc* r, const arc_path_sample* d_samples, uint64_t n, void* stream) {
  if (!r || (n && !d_samples) || n > 0x7FFFFFFFull) return ARC_ERR_INVALID;
  if (n == 0) return ARC_OK;
  hipStream_t s = (hipStream_t)stream;
  if (hipSetDevice(r->device) != hipSuccess) return ARC_ERR_HIP;
  const uint32_t ld = (uint32_t)((n + 15) / 16 * 16);
  const bool lds_scatter = r->table_entries <= kLdsScatterEntries;
  if (ld > r->act_ld) {
    if (hipStreamSynchronize(s) != hipSuccess) return ARC_ERR_HIP;
    void* old[] = {r->act, r->enc_grad, r->enc_pos};
    for (void* p : old) (void)hipFree(p);
    r->act = nullptr;
    r->enc_grad = nullptr;
    r->enc_pos = nullptr;
    r->act_ld = 0;
    const size_t rows = kIn + 6 * kHid + kOut;
    if (hipMalloc(&r->act, rows * ld * 2) != hipSuccess) return ARC_ERR_HIP;
    if (lds_scatter && (hipMalloc(&r->enc_grad, (size_t)kLevels * ld * sizeof(float2)) != hipSuccess ||
                        hipMalloc(&r->enc_pos, (size_t)ld * sizeof(float4)) != hipSuccess))
      return ARC_ERR_HIP;
    r->act_ld = ld;
  }
  if (lds_scatter && !r->scatter_part) {
    const uint32_t parts = std::max<uint32_t>(1u, r->n_cu / (uint32_t)kLevels);  // one workgroup per CU
    if (hipMalloc(&r->scatter_part, (size_t)parts * r->grid_params * 4) != hipSuccess) return ARC_ERR_HIP;
    r->scatter_parts = parts;
  }
  const uint64_t tiles = ld / 16;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((tiles + 3) / 4, (uint64_t)r->n_cu * 4);
  // Training keeps the batch order: in Morton order the lanes of a tile hit the same coarse-level
  // entries, and with the global-atomic scatter those atomics serialize (measured slower at 16K
  // and 66K samples).
  hipLaunchKernelGGL(k_nrc_train_fb, dim3(blocks), dim3(256), 0, s, r->g, reinterpret_cast<const uint32_t*>(r->grid_h),
                     r->mlp_h, d_samples, (uint32_t)n, r->act_ld, r->act, r->grid_g, lds_scatter ? r->enc_grad : nullptr,
                     lds_scatter ? r->enc_pos : nullptr, r->loss);
  if (lds_scatter) {
    // chunks of at least 256 samples, at most one per partial slice
    const uint32_t parts = (uint32_t)std::min<uint64_t>(r->scatter_parts, std::max<uint64_t>(1, (n + 255) / 256));
    const uint32_t chunk = (uint32_t)((n + parts - 1) / parts);
    hipLaunchKernelGGL(k_nrc_scatter, dim3(parts * (uint32_t)kLevels), dim3(1024), 0, s, r->g, r->enc_grad, r->enc_pos,
                       (uint32_t)n, r->act_ld, chunk, r->scatter_part);
    hipLaunchKernelGGL(k_nrc_reduce, dim3(2048), dim3(256), 0, s, reinterpret_cast<const float4*>(r->scatter_part),
                       parts, r->grid_params / 4, reinterpret_cast<float4*>(r->grid_g));
  }
  const uint32_t ldv = r->act_ld;
  _Float16* X = r->act;
  _Float16* A1 = X + (size_t)kIn * ldv;
  _Float16* A2 = A1 + (size_t)kHid * ldv;
  _Float16* A3 = A2 + (size_t)kHid * ldv;
  _Float16* D1 = A3 + (size_t)kHid * ldv;
  _Float16* D2 = D1 + (size_t)kHid * ldv;
  _Float16* D3 = D2 + (size_t)kHid * ldv;
  _Float16* D4 = D3 + (size_t)kHid * ldv;
  const uint32_t chunk = 512;
  const uint32_t nb = (ld + chunk - 1) / chunk;
  hipLaunchKernelGGL((k_nrc_dw<kHid, kIn>), dim3(nb), dim3(64), 0, s, D1, X, ldv, ld, chunk, r->mlp_g + kOffW1);
  hipLaunchKernelGGL((k_nrc_dw<kHid, kHid>), dim3(nb), dim3(64), 0, s, D2, A1, ldv, ld, chunk, r->mlp_g + kOffW2);
  hipLaunchKernelGGL((k_nrc_dw<kHid, kHid>), dim3(nb), dim3(64), 0, s, D3, A2, ldv, ld, chunk, r->mlp_g + kOffW3);
  hipLaunchKernelGGL((k_nrc_dw<kOut, kHid>), dim3(nb), dim3(64), 0, s, D4, A3, ldv, ld, chunk, r->mlp_g + kOffW4);
  return hipGetLastError() == hipSuccess ? ARC_OK : ARC_ERR_HIP;
}

int arc_nrc_train(arc_nrc* r, const arc_path_sample* d_samples, uint64_t n, void* stream) {
  if (!r) return ARC_ERR_INVALID;
  int rc = arc_nrc_backward(r, d_samples, n, stream);
  if (rc != ARC_OK || n == 0) return rc;
  hipStream_t s = (hipStream_t)stream;
  r->step++;
  const float b1 = r->prm.beta1, b2 = r->prm.beta2, eps = r->prm.eps;
  const float bc1 = 1.0f - std::pow(b1, (float)r->step), bc2 = 1.0f - std::pow(b2, (float)r->step);
  const float gs = 1.0f / (float)n;
  hipLaunchKernelGGL(k_adam, dim3(2048), dim3(256), 0, s, r->grid_p, r->grid_m, r->grid_v, r->grid_g, r->grid_h,
                     r->grid_params, gs, r->prm.lr_grid, b1, b2, eps, bc1, bc2);
  hipLaunchKernelGGL(k_adam, dim3(48), dim3(256), 0, s, r->mlp_p, r->mlp_m, r->mlp_v, r->mlp_g, r->mlp_h,
                     (uint64_t)kMlpParams, gs, r->prm.lr_mlp, b1, b2, eps, bc1, bc2);
  return hipGetLastError() == hipSuccess ? ARC_OK : ARC_ERR_HIP;
}

int arc_nrc_loss(arc_nrc* r, float* loss_sum, int reset) {
  if (!r || !loss_sum) return ARC_ERR_INVALID;
  if (hipSetDevice(r->device) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return ARC_ERR_HIP;
  if (hipMemcpy(loss_sum, r->loss, 4, hipMemcpyDeviceToHost) != hipSuccess) return ARC_ERR_HIP;
  if (reset && hipMemset(r->loss, 0, 4) != hipSuccess) return ARC_ERR_HIP;
  return ARC_OK;
}

int arc_nrc_info(const arc_nrc* r, uint64_t* grid_params, uint32_t* mlp_params, float scale[16], uint32_t dense[16]) {
  if (!r) return ARC_ERR_INVALID;
  if (grid_params) *grid_params = r->grid_params;
  if (mlp_params) *mlp_params = (uint32_t)kMlpParams;
  for (int l = 0; l < kLevels; ++l) {
    if (scale) scale[l] = r->g.scale[l];
    if (dense) dense[l] = r->g.dense[l];
  }
  return ARC_OK;
}

// Host copies of the FP32 masters and gradients (any may be NULL), or load masters (set != 0).
int arc_nrc_params_io(arc_nrc* r, float* grid, float* mlp, float* grid_grad, float* mlp_grad, int set) {
  if (!r) return ARC_ERR_INVALID;
  if (hipSetDevice(r->device) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return ARC_ERR_HIP;
  hipError_t e = hipSuccess;
  if (set) {
    if (grid) e = hipMemcpy(r->grid_p, grid, r->grid_params * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && mlp) e = hipMemcpy(r->mlp_p, mlp, kMlpParams * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) return ARC_ERR_HIP;
    hipLaunchKernelGGL(k_to_half, dim3(1024), dim3(256), 0, nullptr, r->grid_p, r->grid_h, r->grid_params);
    hipLaunchKernelGGL(k_to_half, dim3(16), dim3(256), 0, nullptr, r->mlp_p, r->mlp_h, (uint64_t)kMlpParams);
    return hipDeviceSynchronize() == hipSuccess ? ARC_OK : ARC_ERR_HIP;
  }
  if (grid) e = hipMemcpy(grid, r->grid_p, r->grid_params * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && mlp) e = hipMemcpy(mlp, r->mlp_p, kMlpParams * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && grid_grad) e = hipMemcpy(grid_grad, r->grid_g, r->grid_params * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && mlp_grad) e = hipMemcpy(mlp_grad, r->mlp_g, kMlpParams * 4, hipMemcpyDeviceToHost);
  return e == hipSuccess ? ARC_OK : ARC_ERR_HIP;
}

int arc_nrc_zero_grads(arc_nrc* r) {
  if (!r) return ARC_ERR_INVALID;
  if (hipSetDevice(r->device) != hipSuccess) return ARC_ERR_HIP;
  int rc = launch_zero(r->grid_g, r->grid_params, nullptr);
  if (rc == ARC_OK) rc = launch_zero(r->mlp_g, kMlpParams, nullptr);
  return rc;
}

}  // extern "C"
