/*
 * arcadia_hip.h — C ABI of the MI355X (gfx950) 64-tree DDA voxel ray traversal.
 *
 * The library keeps Arcadia's world memory layout byte for byte (Node64, LeafBrick64,
 * flat chunk_roots grid) and replaces the WGSL ray cast `trace_world_64tree_from`
 * (shaders/common/traversal64.wgsl:226-556) with hand-written HIP kernels. Every entry
 * point returns ARC_OK (0) or a negative ARC_ERR_* code; no C++ exception crosses
 * the ABI. arc_last_error() gives the text of the last failure on a context.
 *
 * Ownership: the caller owns host arrays; the context owns the device pools.
 * A context is not thread-safe: use one per host thread or lock around it.
 * Streams: each context keeps one self-resetting work queue per kernel family (traversal, path
 * batch, shadow walkers) in device memory, so two launches of one family on one context must
 * not run at the same time on different streams (order them, or use a context per stream).
 */
#ifndef ARCADIA_HIP_H
#define ARCADIA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------- */
#define ARC_OK 0
#define ARC_ERR_INVALID (-1)   /* bad argument (null pointer, size out of range, bad enum) */
#define ARC_ERR_HIP (-2)       /* a HIP runtime call failed */
#define ARC_ERR_NO_WORLD (-3)  /* trace before arc_world_upload */
#define ARC_ERR_CAPACITY (-4)  /* patch or batch exceeds the device pool capacity */
#define ARC_ERR_IO (-5)        /* a snapshot file could not be read or written */
#define ARC_ERR_CORRUPT (-6)   /* a snapshot file fails its format, checksum or tree checks */

/* ---- layout constants (crates/arcadia-core/src/node64.rs:9-16, world_map.rs:16) ----- */
#define ARC_IS_LEAF_FLAG 0x80000000u
#define ARC_IS_UNIFORM_FLAG 0x40000000u
#define ARC_PTR_INDEX_MASK 0x3FFFFFFFu
#define ARC_EMPTY_CHUNK_SENTINEL 0xFFFFFFFFu
#define ARC_CHUNK_DIM 256u
#define ARC_MAT_AIR 0u
#define ARC_MAT_WATER 18u

/* Node64 (crates/arcadia-core/src/node64.rs:25-36, shaders/common/types.wgsl:65-70), 16 B. */
typedef struct arc_node64 {
  uint32_t mask_low;    /* child/voxel occupancy bits 0..31 */
  uint32_t mask_high;   /* child/voxel occupancy bits 32..63 */
  uint32_t child_ptr;   /* bit31 leaf, bit30 uniform, bits 0..29 node/brick base index */
  uint32_t lod_payload; /* RGB8 in bits 0..23, material id in bits 24..31 */
} arc_node64;

/* LeafBrick64 (crates/arcadia-core/src/node64.rs:116-129), 64 B: one u8 material per voxel. */
typedef struct arc_brick64 {
  uint32_t packed_materials[16];
} arc_brick64;

/* One ray. ARC_SEM_WORLD takes meters (origin_m, dir, start_t_m, max_dist_m: the arguments
 * of trace_world_64tree_from); ARC_SEM_CHUNK takes chunk-local voxel units (origin, dir,
 * max_dist of raycast_chunk_tree, crates/arcadia-core/src/raycast_cpu.rs:28-33; t_start is
 * ignored). 32 B. */
typedef struct arc_ray {
  float ox, oy, oz, t_start;
  float dx, dy, dz, t_max;
} arc_ray;

/* Compact hit record, 16 B.
 *  t_voxels: hit distance in voxel units (miss: max_dist*16 for ARC_SEM_WORLD, 0 for CHUNK)
 *  vx,vy,vz: hit voxel in global voxel coordinates (chunk-local for ARC_SEM_CHUNK)
 *  mat:      material id (0 on a miss)
 *  face:     normal code: 2*axis + (component < 0), 6 + axis when the chosen axis has a zero
 *            normal component (WGSL sign(0) == 0), ARC_FACE_NONE on a miss
 *  steps:    outer DDA iterations that passed the loop guards (a first-step hit gives 1)
 *  flags:    ARC_HIT_* bits */
typedef struct arc_hit {
  float t_voxels;
  uint16_t vx, vy, vz;
  uint8_t mat;
  uint8_t face;
  uint16_t steps;
  uint16_t flags;
} arc_hit;

#define ARC_FACE_NONE 255u
#define ARC_HIT_FLAG_HIT 1u
#define ARC_HIT_FLAG_LOD 2u    /* hit resolved on a coarse cell (shift > 0) */

/* Full per-ray result: the WGSL TraceResult (shaders/common/traversal64.wgsl:4-17) plus the
 * step count and hit voxel, 80 B.
 *  hit: bit 0 = hit; bits 8..15 = axis of the normal; bits 16..23 = shift of the hit cell
 *       (0 = single voxel, 2/4/6 = collapsed or LOD cell)
 *  hit_pos_m: meters for ARC_SEM_WORLD (voxel units for ARC_SEM_CHUNK) */
typedef struct arc_trace_result {
  float t_voxels;
  float hit_pos_m[3];
  float normal[3];
  float lod_rgb[3];
  float contact_ao;
  uint32_t mat;
  uint32_t voxel_hash;
  uint32_t steps;
  uint32_t hit;
  uint32_t vx, vy, vz;
  uint32_t reserved[2];
} arc_trace_result;

/* Traversal semantics. */
#define ARC_SEM_WORLD 0u  /* trace_world_64tree_from (traversal64.wgsl:226-556), the GPU path */
#define ARC_SEM_CHUNK 1u  /* raycast_chunk_tree (raycast_cpu.rs:28-218) on chunk 0 of the grid */

/* Per-launch parameters: the traversal uniforms of CameraUniforms
 * (shaders/common/types.wgsl:19-63; filled at crates/arcadia-render/src/renderer.rs:614-640). */
typedef struct arc_trace_params {
  uint32_t semantics;      /* ARC_SEM_* */
  uint32_t lod_bias;       /* 0 primary, 1 secondary (water/mirror), >=2 GI */
  uint32_t ignore_water;   /* skip water (mat 18) cells */
  uint32_t max_steps[3];   /* ray_steps.xyz: cap for lod_bias 0, 1, >=2 */
  float lod_dist_m[3];     /* lod_settings.xyz (meters) */
  uint32_t lod_step[3];    /* quality_flags.x, quality_flags.y, (quality_flags.w >> 16) & 0xFF */
  float ao_max_dist_bits;  /* point_light_count.w reinterpreted as f32 (contact AO range) */
  uint32_t reserved[3];
} arc_trace_params;

/* A ray set inside one batched launch. Rays and hits are device pointers. */
typedef struct arc_ray_set {
  const arc_ray* rays;
  arc_hit* hits;
  uint64_t n;
  arc_trace_params params;
} arc_ray_set;

/* Incremental pool update (crates/arcadia-world/src/world_map.rs:20-27). */
typedef struct arc_gpu_patch {
  uint32_t chunk_flat_idx;
  uint32_t new_root_node_idx;
  uint32_t node_offset;
  uint32_t n_nodes;
  const arc_node64* nodes;
  uint32_t brick_offset;
  uint32_t n_bricks;
  const arc_brick64* bricks;
} arc_gpu_patch;

/* Pinhole camera (FrameParams eye/forward/right/up, crates/arcadia-render/src/renderer.rs:19-38):
 * tan of half the vertical fov, aspect = width / height. */
typedef struct arc_camera {
  float eye[3];
  float forward[3];
  float right[3];
  float up[3];
  float tan_half_fov;
  float aspect;
  float jitter[2];
} arc_camera;

typedef struct arc_ctx arc_ctx;
typedef struct arc_batch arc_batch; /* batched views, see arc_batch_create */

/* ---- context --------------------------------------------------------------------------- */
/* Replaces the wgpu device/queue owned by Renderer (crates/arcadia-render/src/context.rs:25-146). */
int arc_create(int device, arc_ctx** out);
void arc_destroy(arc_ctx* ctx);
const char* arc_last_error(const arc_ctx* ctx);
/* Number of compute units of the context's device (sizes persistent launches). */
int arc_device_info(const arc_ctx* ctx, uint32_t* n_cu, uint32_t* lds_bytes_per_block);

/* ---- world pools ----------------------------------------------------------------------- */
/* Upload chunk_roots / node_pool / brick_pool (the @group(0) bindings 1..3 created by
 * create_group0_resources, crates/arcadia-render/src/gpu_pools.rs:124-261). Host pointers.
 * node_capacity/brick_capacity >= n_nodes/n_bricks reserve headroom for later patches
 * (0 = exact size). */
int arc_world_upload(arc_ctx* ctx, const uint32_t grid[3], const uint32_t* chunk_roots,
                     const arc_node64* nodes, uint64_t n_nodes, uint64_t node_capacity,
                     const arc_brick64* bricks, uint64_t n_bricks, uint64_t brick_capacity);
/* Apply patches in order (Renderer::apply_world_patches, crates/arcadia-render/src/renderer.rs:147-315):
 * node/brick slices are written at their offsets, then chunk_roots[chunk_flat_idx] is set.
 * Slices beyond the reserved capacity grow the pools (device copy), as the reference does. */
int arc_world_patch(arc_ctx* ctx, const arc_gpu_patch* patches, uint64_t n);
/* The same patch application on host pools (roots[n_chunks], nodes[node_cap], bricks[brick_cap]),
 * with the same validation and the same last-writer-wins staging plan as arc_world_patch; no
 * reference counterpart (the reference's host mirror is WorldMap::gpu_nodes). ARC_ERR_CAPACITY if
 * a patch writes past the given capacities. */
int arc_world_patch_host(const arc_gpu_patch* patches, uint64_t n, uint32_t n_chunks, uint32_t* roots,
                         arc_node64* nodes, uint64_t node_cap, arc_brick64* bricks, uint64_t brick_cap);
/* World snapshots (no reference counterpart: the reference keeps worlds in memory only,
 * ARCHITECTURE.md:243). An open snapshot file; see arcadia_world.h for writing and opening. */
typedef struct arc_snapshot arc_snapshot;
/* Streaming pool builder: uploads a snapshot's chunks straight to their final pool offsets (the
 * layout of WorldMap::rebuild_all_gpu_pools, world_map.rs:158-205). Batches of up to 128 MiB are
 * staged through two pinned buffers (kept by the context for later uploads), filled by `nthreads`
 * host threads while the previous batch is copied; the node pointers are then rebased on the
 * device. Every chunk's checksum and tree rules are checked as it is staged (ARC_ERR_CORRUPT: the
 * context is then left without a world). The resulting pools equal arc_world_host_pack +
 * arc_world_upload of the same world. Blocking. */
int arc_snapshot_upload(arc_ctx* ctx, const arc_snapshot* snap, uint64_t node_capacity, uint64_t brick_capacity,
                        int nthreads);
/* Device addresses of the pools (for callers that build their own launches). */
int arc_world_device_pools(const arc_ctx* ctx, const uint32_t** roots, const arc_node64** nodes,
                           const arc_brick64** bricks);
/* Read back the device pools (chunk_roots whole, node / brick index ranges) into host arrays
 * (no reference counterpart: the reference never reads its pools back). Blocking. */
int arc_world_read_pools(arc_ctx* ctx, uint32_t* roots, arc_node64* nodes, uint64_t node_first, uint64_t n_nodes,
                         arc_brick64* bricks, uint64_t brick_first, uint64_t n_bricks);

/* ---- traversal ------------------------------------------------------------------------- */
/* Host-pointer convenience: copies rays in, traces, copies hits out, synchronises. */
int arc_trace(arc_ctx* ctx, const arc_trace_params* params, const arc_ray* rays, uint64_t n,
              arc_hit* hits_out);
/* Device-pointer variant, asynchronous on `stream` (a hipStream_t; NULL = the null stream). */
int arc_trace_device(arc_ctx* ctx, const arc_trace_params* params, const arc_ray* d_rays,
                     uint64_t n, arc_hit* d_hits, void* stream);
/* Allocation-free trace. No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It runs arc_trace_device's kernels with the default launch shape on world pools the
 * caller owns (the layout arc_world_upload takes: grid dims 1..256, chunk_roots, node and brick
 * pools, already checked by the caller, for instance through arc_world_upload once), on the
 * current HIP device, and needs no context. Its only scratch is the ray queue's claim counters, a
 * static 32 KiB device array per GPU; nothing is allocated or freed per call. Limit: n <= 2^31-1
 * rays (ARC_ERR_CAPACITY beyond). Calls that overlap on one GPU (different streams) share that
 * array and race; calls in one stream order are safe. Errors: ARC_ERR_INVALID (null pointer,
 * bad grid or params), ARC_ERR_HIP (a launch or device query failed). */
int arc_trace_pools_noalloc(const uint32_t grid[3], const uint32_t* d_roots, const arc_node64* d_nodes,
                            const arc_brick64* d_bricks, const arc_trace_params* params, const arc_ray* d_rays,
                            uint64_t n, arc_hit* d_hits, void* stream);
/* Full TraceResult records (hit position, normal, lod rgb, voxel hash, contact AO). */
int arc_trace_full_device(arc_ctx* ctx, const arc_trace_params* params, const arc_ray* d_rays,
                          uint64_t n, arc_trace_result* d_out, void* stream);
/* A ray set and one cosine-weighted diffuse bounce off each of its hits, in ONE persistent launch.
 * Replaces the primary trace followed by the GI pass's bounce trace (the reference dispatches
 * primary_raymarch, then traces the bounce of lighting_restir_gi.wgsl:195-209 in its own pass).
 * The bounce of ray i is arc_gen_bounce_rays_dense's ray for index i (bounce_ray with the hit's
 * t_voxels and face, `seed`, hash index i, t_max = bounce_max_dist_m), generated on the device
 * once ray i has finished. d_hits receives 2n compact records: [0, n) the rays' records, equal
 * to arc_trace_device with `primary`; n + i the record of ray i's bounce traced with `bounce`,
 * or the empty-ray record (t_voxels -16, no hit, 0 steps) where ray i missed. Both classes use
 * world semantics. The context keeps one 8 B word per ray (grown on the first larger call).
 * Launches on one context must not overlap (one queue). */
int arc_trace_bounce_device(arc_ctx* ctx, const arc_trace_params* primary, const arc_trace_params* bounce,
                            const arc_ray* d_rays, uint64_t n, uint64_t seed, float bounce_max_dist_m,
                            arc_hit* d_hits, void* stream);

/* ---- shadow any-hit walkers ------------------------------------------------------------- */
/* trace_shadow_vis (shaders/common/traversal64.wgsl:571-773): sun / point-light visibility,
 * 1.0 unoccluded, 0.0 occluded. trace_shadow_vis_di (:781-989): visibility of a sampled
 * emitter, emissive voxels within 16 voxels of the target distance count as unoccluded. Both
 * read origin_m = (ox, oy, oz), dir = (dx, dy, dz), max_dist_m = t_max of each arc_ray; the
 * reference walkers start at the world entry, so t_start is ignored. */
#define ARC_SHADOW_SUN 0u /* trace_shadow_vis */
#define ARC_SHADOW_DI 1u  /* trace_shadow_vis_di */
typedef struct arc_shadow_params {
  uint32_t mode;      /* ARC_SHADOW_SUN or ARC_SHADOW_DI */
  uint32_t max_steps; /* u_cam.ray_steps.w = quality shadow_max_steps (72 at high quality);
                         ARC_SHADOW_DI clamps it to 32..48 (:803) */
} arc_shadow_params;
/* How a walker ended (the reference's return paths). */
#define ARC_SHADOW_MISS_WORLD 0u  /* ray misses the world box or starts past max_dist: 1.0 */
#define ARC_SHADOW_CLEAR 1u       /* reached max_dist or left the world: 1.0 */
#define ARC_SHADOW_BLOCKED 2u     /* solid cell at the walker's LOD level: 0.0 */
#define ARC_SHADOW_EMITTER 3u     /* DI: emissive cell within 16 voxels of max_dist: 1.0 */
#define ARC_SHADOW_BUDGET_DARK 4u /* step budget spent: SUN below 32 m height / DI short of the target: 0.0 */
#define ARC_SHADOW_BUDGET_LIT 5u  /* step budget spent: SUN at or above 32 m / DI within 2 voxels: 1.0 */
/* Per-ray record of a walker (no reference counterpart: the reference returns only the
 * visibility). t_voxels = the walker's t when it returned (max_dist when it left the world),
 * voxel = its current voxel then, steps = outer iterations begun. 16 B. */
typedef struct arc_shadow_rec {
  float t_voxels;
  uint32_t steps;
  uint16_t vx, vy, vz;
  uint8_t outcome; /* ARC_SHADOW_* */
  uint8_t reserved;
} arc_shadow_rec;
/* Device pointers, asynchronous on `stream`: d_vis[i] = visibility of ray i (0.0 / 1.0);
 * d_rec may be NULL. The walkers launch with their own shape, launch class
 * ARC_LAUNCH_CLASS_SHADOW (default: 6 workgroups of 256 per CU, refill 16, windows of 64 rays;
 * arc_autotune_shadow tunes it), whatever arc_set_launch_config / arc_autotune set for traces. */
int arc_trace_shadow_device(arc_ctx* ctx, const arc_shadow_params* params, const arc_ray* d_rays, uint64_t n,
                            float* d_vis, arc_shadow_rec* d_rec, void* stream);

/* ---- rigid voxel bodies and point queries ------------------------------------------------- */
/* RigidBodyInstance (shaders/common/types.wgsl:219-226; GpuRigidBodyInstance,
 * crates/arcadia-physics/src/rigid_voxel/types.rs:20-35 and transform.rs:80-112), 224 B. The
 * matrices are column-major 4x4 (glam to_cols_array_2d): m[4 * col + row]. The body's 64-tree
 * lives in the node/brick pools of the context like the chunks' trees. */
typedef struct arc_rigid_body {
  float world_to_local[16];        /* world metres -> body-local tree metres */
  float local_to_world[16];
  float prev_local_to_world[16];   /* previous frame's, for the motion of the hit point */
  float local_bounds_and_root[4];  /* local box size in metres (x, y, z); w = bits of the root node index */
  float world_sphere[4];           /* bounding sphere: centre (m), radius (m) */
} arc_rigid_body;
/* RigidTraceResult (traversal64.wgsl:994-1003) plus the index of the body hit. 64 B. */
typedef struct arc_rigid_hit {
  float t_voxels;          /* hit body's local voxel units; max_dist * 16 without a hit */
  float hit_pos_m[3];      /* world metres; origin + dir * max_dist without a hit */
  float prev_hit_pos_m[3]; /* the hit point moved with prev_local_to_world */
  float normal[3];         /* world space, normalized; (0, 1, 0) without a hit */
  float contact_ao;        /* eval_leaf_bitwise_ao at voxel-level hits, else 1 */
  uint32_t mat;            /* material, ARC_MAT_AIR without a hit */
  uint32_t voxel_hash;     /* pcg_hash of the hit body's local voxel */
  uint32_t hit;            /* 1 when a body was hit */
  uint32_t body;           /* index of the body hit, 0xFFFFFFFF without (no reference counterpart) */
  uint32_t reserved;
} arc_rigid_hit;
#define ARC_MAX_RIGID_BODIES 32u /* trace_rigid_bodies reads min(count, 32) instances (:1219) */
/* trace_rigid_bodies (traversal64.wgsl:1204-1244) over the first min(n_bodies, 32) of `bodies`
 * (a host array, copied to the device with the launch): for each ray (origin_m, dir, t_max =
 * max_dist_m; t_start unused) the nearest hit over the bodies, each traced by
 * trace_single_rigid_body (:1005-1202). Device rays and output, asynchronous on `stream`.
 * Node and brick indices are checked against the pool sizes; a body whose tree leaves the pools
 * traces as empty. */
int arc_trace_rigid_device(arc_ctx* ctx, const arc_rigid_body* bodies, uint32_t n_bodies, const arc_ray* d_rays,
                           uint64_t n, arc_rigid_hit* d_out, void* stream);
/* query_voxel_world_meters (traversal64.wgsl:182-220): the material at each point (x, y, z
 * metres, 12 B per point), ARC_MAT_AIR outside the world or in empty space. Device pointers,
 * asynchronous on `stream`. */
int arc_query_voxels_device(arc_ctx* ctx, const float* d_points_m, uint64_t n, uint32_t* d_mat, void* stream);

/* ---- primary G-buffer pass -------------------------------------------------------------- */
/* Camera and frame inputs of primary_raymarch (the CameraUniforms fields it reads,
 * shaders/common/types.wgsl:19-63), 144 B. prev_view_proj is column-major. */
typedef struct arc_gbuffer_params {
  float eye[3], forward[3], right[3], up[3];
  float tan_half_fov;  /* tan of half the vertical fov */
  float aspect;        /* width / height */
  float jitter[2];     /* sub-pixel jitter in [-0.5, 0.5] */
  float prev_view_proj[16];
  uint32_t width, height;
  float max_dist_m;    /* 250 m in the reference */
  uint32_t reserved;
} arc_gbuffer_params;

/* The opaque-world path of record_primary_raymarch_pass / primary_raymarch
 * (crates/arcadia-render/src/pipeline.rs:150-166, shaders/primary_gbuffer.wgsl:93-312): one
 * pinhole ray per pixel traced with `params` (lod_bias 0), written as the 20 B/px G-buffer:
 * depth_mat 2 u32/px (depth f32 bits, mat | flags << 16), norm_motion 2 u32/px
 * (pack2x16snorm(oct(normal)), pack2x16float(motion)), albedo_ao 1 u32/px (pack4x8unorm).
 * Device pointers, asynchronous on `stream`. Water/mirror secondary rays and rigid bodies are not
 * traced (their pixels take the opaque branch). */
int arc_primary_gbuffer(arc_ctx* ctx, const arc_trace_params* params, const arc_gbuffer_params* g,
                        uint32_t* d_depth_mat, uint32_t* d_norm_motion, uint32_t* d_albedo_ao,
                        void* stream);

/* ---- path shading (radiance cache workloads) -------------------------------------------- */
/* Light transport of the radiance-cache workloads (no reference counterpart; assembled from the
 * reference's materials and sky): Lambertian palette albedo, palette emission x emission_scale,
 * escaped rays see eval_sky_radiance (types.wgsl:500-523) x sky_scale. */
typedef struct arc_shade_params {
  float sun_dir[3];      /* unit vector towards the sun */
  float emission_scale;
  float sky_scale;
  uint32_t reserved[3];
} arc_shade_params;
/* One slot of a shaded path launch. x0 = primary hit, x1 = hit of the first bounce. The path's
 * radiance is le0 + a0 * (base + target) (x1 hit), le0 + a0 * base (bounce escaped: base = sky),
 * le0 (primary miss: le0 = sky, a0 = 0). A radiance cache at x1 replaces `target`, the rest of the
 * path beyond x1 (its emission excluded, as SHARC stores it, sharc_update.wgsl:63-64). */
#define ARC_PS_RAY 1u     /* the slot has a primary ray */
#define ARC_PS_X0_HIT 2u
#define ARC_PS_X1_HIT 4u  /* bits 8..15: face code of x1 */
typedef struct arc_path_sample {
  float pos[3];    /* x1, meters */
  uint32_t flags;
  float dir[3];    /* direction x0 -> x1 */
  uint32_t pad0;
  float le0[3];
  uint32_t pad1;
  float a0[3];
  uint32_t pad2;
  float base[3];   /* Le1, or the sky seen by the first bounce */
  uint32_t pad3;
  float target[3]; /* a1 (Le2 + a2 (...)): the path's radiance beyond x1 */
  uint32_t pad4;
} arc_path_sample;
/* Shade the records of arc_batch_trace (the batch's current seed): per slot the path sample,
 * its radiance (3 floats) and d_accum += radiance (3 doubles). Any output may be NULL. */
int arc_batch_shade(arc_batch* b, const arc_hit* d_hits, const arc_shade_params* params, arc_path_sample* d_samples,
                    float* d_radiance, double* d_accum, void* stream);
/* Per slot: le0 + a0 * (base + max(cache, 0)) with cache = d_cache[4 i .. 4 i + 2] where x1 was
 * hit (mode 1), or le0 + a0 * base (mode 0, no cache). 3 floats per slot into d_out. */
int arc_estimate(const arc_path_sample* d_samples, const float* d_cache, uint64_t n, uint32_t mode, float* d_out,
                 void* stream);

/* ---- neural radiance cache ---------------------------------------------------------------- */
/* Learned radiance cache (no reference counterpart): 16-level hash-grid encoding of the
 * position (2 FP16 features per level, 2^log2_table entries per level) + normal + view
 * direction -> MLP 3 x 64 ReLU -> RGB, FP16 weights with FP32 accumulation on the matrix cores,
 * trained online with a relative L2 loss (mean over RGB of (y - t)^2 / (y^2 + 0.01), denominator held
 * constant) and Adam. It predicts `target` of arc_path_sample
 * (the radiance beyond x1) from x1's position, normal and the incoming direction. */
typedef struct arc_nrc arc_nrc;
typedef struct arc_nrc_params {
  float bbox_min[3];   /* cube enclosing the scene, meters */
  float bbox_size;
  uint32_t log2_table; /* entries per level (10..24) */
  float base_res;      /* coarsest level resolution (cells per bbox edge) */
  float max_res;       /* finest level resolution */
  float lr_grid, lr_mlp;
  float beta1, beta2, eps;
  uint64_t seed;       /* initialization */
} arc_nrc_params;
int arc_nrc_create(int device, const arc_nrc_params* params, arc_nrc** out);
void arc_nrc_destroy(arc_nrc* r);
/* Re-initialize weights and table (seeded), clear moments and gradients. */
int arc_nrc_reset(arc_nrc* r, uint64_t seed);
/* Prediction per sample: d_out[4 i ..] = (r, g, b, 1 if x1 was hit else 0). */
int arc_nrc_infer(arc_nrc* r, const arc_path_sample* d_samples, uint64_t n, float* d_out, void* stream);
/* One training step on n samples (samples without an x1 hit carry no loss): gradients of the
 * summed loss, then Adam with the gradients scaled by 1/n. */
int arc_nrc_train(arc_nrc* r, const arc_path_sample* d_samples, uint64_t n, void* stream);
/* Gradients only (accumulated into the FP32 gradient buffers; no update). */
int arc_nrc_backward(arc_nrc* r, const arc_path_sample* d_samples, uint64_t n, void* stream);
int arc_nrc_zero_grads(arc_nrc* r);
/* Summed loss since the last reset (blocking). */
int arc_nrc_loss(arc_nrc* r, float* loss_sum, int reset);
/* Parameter counts and the level resolutions / dense-indexing flags. */
int arc_nrc_info(const arc_nrc* r, uint64_t* grid_params, uint32_t* mlp_params, float scale[16], uint32_t dense[16]);
/* FP32 masters (table [16][2^T][2], MLP W1[64][48] W2[64][64] W3[64][64] W4[16][64]) and gradients
 * to host arrays (set = 0; any may be NULL), or load masters from them (set = 1). Blocking. */
int arc_nrc_params_io(arc_nrc* r, float* grid, float* mlp, float* grid_grad, float* mlp_grad, int set);

/* ---- SHARC radiance cache --------------------------------------------------------------- */
/* The voxel-surfel hash cache of shaders/common/sharc.wgsl (2^log2_buckets 4-way buckets; the
 * reference uses 2^15 = 131072 slots, 4.5 MiB; 2^20 buckets = 2^22 slots = 144 MiB). Buffers are
 * the reference's: keys u32[S], accum u32[4S] ({r, g, b} x1024 fixed point + count), resolved
 * 16 B slots {key, f16 r|g, f16 b | hist_frames << 16 | stale_age << 24, 0}. */
typedef struct arc_sharc arc_sharc;
typedef struct arc_sharc_sample { /* one deposit: surfel position / normal and its radiance */
  float pos[3];
  float normal[3];
  float radiance[3];
  uint32_t reserved;
} arc_sharc_sample;
typedef struct arc_sharc_probe { /* one lookup; jitter (0, 0) = unjittered */
  float pos[3];
  float normal[3];
  float jitter[2];
} arc_sharc_probe;
int arc_sharc_create(int device, uint32_t log2_buckets, arc_sharc** out);
void arc_sharc_destroy(arc_sharc* s);
/* Slot count and the device buffers (keys, accum, resolved). */
int arc_sharc_info(const arc_sharc* s, uint32_t* n_slots, uint32_t** keys, uint32_t** accum, void** resolved);
/* Copy the buffers to host arrays (any may be NULL; blocking). */
int arc_sharc_read(const arc_sharc* s, uint32_t* keys, uint32_t* accum, uint32_t* resolved);
/* Zero all three buffers with a fill kernel (capturable; no memset node). */
int arc_sharc_clear(arc_sharc* s, void* stream);
/* sharc_make_key + sharc_deposit (sharc.wgsl:71-97, sharc_update.wgsl:129-169) per sample,
 * unjittered key: claim the surfel's way (existing key first, then the first empty way by CAS),
 * atomically add u32(clamp(radiance, 0, 32) * 1024) and 1. Which way a new key takes depends on
 * the order of racing claims; the contents per key do not. */
int arc_sharc_deposit(arc_sharc* s, const arc_sharc_sample* d_samples, uint64_t n, const float cam_pos[3],
                      void* stream);
/* sharc_resolve.wgsl: fold each slot's deposits into its EMA (alpha = 1 / min(hist, 16)), age
 * untouched surfels and evict them after 64 idle resolves, clear the accumulators. */
int arc_sharc_resolve(arc_sharc* s, void* stream);
/* query_sharc (sharc_query.wgsl:20-29): d_out[4 * i ..] = (r, g, b, 1) on a hit, zeros on a miss. */
int arc_sharc_query(const arc_sharc* s, const arc_sharc_probe* d_queries, uint64_t n, const float cam_pos[3],
                    float* d_out, void* stream);

/* Evict every surfel (all cascade levels and faces) whose cell overlaps one of n spheres
 * {x, y, z, radius} in meters (radius <= 64 m), e.g. the carve spheres of an edit, so no stale
 * radiance is served where the geometry changed. Blocking (the sphere list is copied). */
int arc_sharc_invalidate(arc_sharc* s, const float* spheres, uint32_t n, void* stream);
/* The same deposit / query fed by shaded path samples: the surfel of x1 (normal from its face)
 * receives `target` (the radiance beyond x1); samples without an x1 hit are skipped (query: 0). */
int arc_sharc_deposit_paths(arc_sharc* s, const arc_path_sample* d_samples, uint64_t n, const float cam_pos[3],
                            void* stream);
int arc_sharc_query_paths(const arc_sharc* s, const arc_path_sample* d_samples, uint64_t n, const float cam_pos[3],
                          float* d_out, void* stream);

/* Kernel launch shape of the persistent ray-queue kernel; 0 fields = autotuned defaults. */
typedef struct arc_launch_config {
  uint32_t blocks_per_cu;   /* resident workgroups per compute unit */
  uint32_t block_threads;   /* 64..1024, multiple of 64 */
  uint32_t refill_lanes;    /* a wave refills when at least this many lanes are idle (1..64) */
  uint32_t batch_rays;      /* rays claimed per queue fetch (per wave) */
} arc_launch_config;
/* Set the launch shape of every ray class (lod_bias 0, 1, >=2) / read the class-0 shape. */
int arc_set_launch_config(arc_ctx* ctx, const arc_launch_config* cfg);
int arc_get_launch_config(const arc_ctx* ctx, arc_launch_config* cfg);
/* Launch shape of one ray class (lod_class 0, 1, 2 = the lod_bias classes of the params), or of
 * arc_trace_bounce_device's launch (ARC_LAUNCH_CLASS_BOUNCE). */
#define ARC_LAUNCH_CLASS_BOUNCE 3u
/* arc_trace_shadow_device's launch (both walker modes). */
#define ARC_LAUNCH_CLASS_SHADOW 4u
int arc_set_launch_config_class(arc_ctx* ctx, uint32_t lod_class, const arc_launch_config* cfg);
int arc_get_launch_config_class(const arc_ctx* ctx, uint32_t lod_class, arc_launch_config* cfg);
/* Autotuner (no reference counterpart; the reference dispatches fixed 8x8 workgroups): times
 * candidate launch shapes of the persistent kernel on this ray set with HIP events on
 * `stream` (every shape with `reps` launches, then the four fastest again with max(reps, 8) * 4
 * launches, at most 64), keeps the fastest for the ray class of `params` and returns it (and
 * its time in ms, the median of the second pass). d_hits receives the last traced results.
 * Blocking. */
int arc_autotune(arc_ctx* ctx, const arc_trace_params* params, const arc_ray* d_rays, uint64_t n,
                 arc_hit* d_hits, void* stream, uint32_t reps, arc_launch_config* best,
                 float* best_ms);
/* The same search for arc_trace_bounce_device's launch on this ray set (d_hits: 2n records). */
int arc_autotune_bounce(arc_ctx* ctx, const arc_trace_params* primary, const arc_trace_params* bounce,
                        const arc_ray* d_rays, uint64_t n, uint64_t seed, float bounce_max_dist_m,
                        arc_hit* d_hits, void* stream, uint32_t reps, arc_launch_config* best,
                        float* best_ms);
/* The same search for arc_trace_shadow_device's launch on this ray set (class
 * ARC_LAUNCH_CLASS_SHADOW; d_vis / d_rec receive the last launch's results). */
int arc_autotune_shadow(arc_ctx* ctx, const arc_shadow_params* sp, const arc_ray* d_rays, uint64_t n, float* d_vis,
                        arc_shadow_rec* d_rec, void* stream, uint32_t reps, arc_launch_config* best,
                        float* best_ms);

/* ---- batched views (one persistent launch) ----------------------------------------------- */
/* A batch of camera views of uneven sizes traced with their diffuse bounces in ONE persistent
 * launch (no reference counterpart: the reference records one primary_raymarch dispatch per view
 * and one GI dispatch per bounce, renderer.rs:667 / lighting_restir_gi.wgsl:195-209). Each pixel
 * slot's lane generates the primary ray (primary_gbuffer.wgsl:103-115), and after every hit the
 * next cosine bounce in-lane, bit-identical to arc_gen_view_rays / arc_gen_bounce_rays_dense
 * (bounce d of slot i hashed with index (d-1) * slots + i). Views are laid out one after another
 * in padded 8x8-tile order; hits[d * slots + i] is the compact record of depth d (0 = primary);
 * slots with no ray at that depth (padding pixels, bounces of misses) hold the empty-ray record
 * (t_voxels -16, no hit, 0 steps). */
typedef struct arc_view {
  arc_camera cam;      /* pinhole camera */
  uint32_t width, height;
  float max_dist_m;    /* primary ray length (250 m in the reference) */
  uint32_t reserved;
} arc_view;
/* n_bounces 0..4. */
int arc_batch_create(arc_ctx* ctx, const arc_view* views, uint32_t n_views, const arc_trace_params* primary,
                     const arc_trace_params* bounce, uint32_t n_bounces, float bounce_max_dist_m, uint64_t seed,
                     arc_batch** out);
void arc_batch_destroy(arc_batch* b);
/* Seed of the bounce hash for the following launches (a captured graph keeps its own). */
int arc_batch_set_seed(arc_batch* b, uint64_t seed);
/* Pixel slots (padded) of the batch; the hit buffer holds (1 + n_bounces) * slots records. */
uint64_t arc_batch_slots(const arc_batch* b);
/* First slot of view v (v = n_views gives the slot count). */
uint64_t arc_batch_view_first(const arc_batch* b, uint32_t v);
/* One launch on `stream` (capturable). */
int arc_batch_trace(arc_batch* b, arc_hit* d_hits, void* stream);
/* Capture one launch into a hipGraph bound to d_hits; arc_batch_replay launches it. Upload or pool
 * growth moves the pools: replay then fails until the batch is captured again. */
int arc_batch_capture(arc_batch* b, arc_hit* d_hits);
int arc_batch_replay(arc_batch* b, void* stream);
/* Launch shape of the path kernel, and its autotuner (arc_autotune's search, plus 1024-lane
 * workgroups at the occupancy limit). */
int arc_batch_set_launch_config(arc_batch* b, const arc_launch_config* cfg);
int arc_batch_get_launch_config(const arc_batch* b, arc_launch_config* cfg);
int arc_batch_autotune(arc_batch* b, arc_hit* d_hits, void* stream, uint32_t reps, arc_launch_config* best,
                       float* best_ms);

/* ---- ReSTIR GI / DI passes (shaders/lighting_restir_gi.wgsl) ------------------------------- */
/* GiReservoir (types.wgsl:90-106; crates/arcadia-render/src/gbuffer.rs), 32 B:
 *   norm_m_flags = oct snorm8x2 secondary normal | M << 16 | sample_age << 24 | is_sky << 31,
 *   lo_rg / lo_b_w = f16 pairs (Lo.r, Lo.g) / (Lo.b, W), shadow_vis_m = f16 pair (sun vis, count). */
typedef struct arc_gi_reservoir {
  float sample_pos[3];
  uint32_t norm_m_flags;
  uint32_t lo_rg;
  uint32_t lo_b_w;
  float primary_depth;
  uint32_t shadow_vis_m;
} arc_gi_reservoir;
/* DiReservoir (shaders/common/di.wgsl:29-35): light index | seed << 16, w_sum, M, W. 16 B. */
typedef struct arc_di_reservoir {
  uint32_t light_idx_seed;
  float w_sum;
  float m;
  float w;
} arc_di_reservoir;
/* EmissiveSource (di.wgsl:7-12; crates/arcadia-world/src/emissive_registry.rs), 32 B, and its
 * alias-table entry (di.wgsl:15-18), 8 B: produced by the caller's emissive registry. */
typedef struct arc_emissive_source {
  float pos[3];
  float radius;
  float intensity[3];
  uint32_t normal_packed; /* 10-bit x | y << 10 | z << 20, each (n + 1) / 2 * 1023 */
} arc_emissive_source;
typedef struct arc_alias_entry {
  float prob;
  uint32_t alias_idx;
} arc_alias_entry;
/* The CameraUniforms / EmissiveParams fields the passes read (types.wgsl:19-63, di.wgsl:21-26). */
typedef struct arc_restir_params {
  arc_camera cam;            /* eye, forward, right, up, tan_half_fov, aspect, jitter_brush_dt.xy */
  float prev_eye[3];         /* prev_cam_pos.xyz */
  uint32_t frame_index;      /* bitcast(cam_up_frame.w) */
  float sun_dir[3];          /* sun_dir_intensity.xyz (towards the sun, unit) */
  float sun_intensity;       /* sun_dir_intensity.w */
  uint32_t width, height;    /* world_grid_dims.w */
  float gi_max_dist;         /* ray_distances.z (m) */
  float shadow_max_dist;     /* ray_distances.w (m) */
  float shadow_jitter;       /* lod_settings.w */
  uint32_t shadow_mode;      /* quality_flags.z: 0 = 64-tree trace, 1 = any-hit sun walker */
  uint32_t shadow_max_steps; /* ray_steps.w: step budget of the any-hit walkers */
  uint32_t denoiser_taps;    /* quality_flags.w & 0xFF (<= 12: checkerboard half-rate mode) */
  uint32_t target_valid;     /* target_voxel_info.w > 0.5 */
  uint32_t n_emissive;       /* EmissiveParams.count (<= 16384) */
  float total_flux;          /* EmissiveParams.total_flux */
  uint32_t reserved;
} arc_restir_params;
/* Pass 2A, temporal_main (lighting_restir_gi.wgsl:62-455): sun shadow with its temporal average,
 * one GI candidate per pixel (cosine bounce, emissive macro-cell guiding, world + rigid-body
 * hit, SHARC or bounce-sun radiance, sky), RIS merge with the reprojected history, and the DI
 * candidates + temporal reuse. G-buffer: d_depth_mat / d_norm_motion as arc_primary_gbuffer writes
 * them (2 u32 per pixel each, row-major). Reservoirs: width * height each; *_in is the previous
 * frame's spatial output. sun_trace / gi_trace: traversal parameters of the lod_bias 1 sun trace
 * (shadow_mode 0) and the lod_bias 2 GI trace (ignore_water set by the pass). sharc may be NULL
 * (every lookup misses); d_sources / d_alias may be NULL when n_emissive is 0; bodies are host
 * instances (n_bodies <= 32 used). Asynchronous on `stream`. */
int arc_restir_temporal(arc_ctx* ctx, const arc_restir_params* p, const arc_trace_params* sun_trace,
                        const arc_trace_params* gi_trace, const uint32_t* d_depth_mat, const uint32_t* d_norm_motion,
                        const arc_gi_reservoir* d_gi_in, arc_gi_reservoir* d_gi_out,
                        const arc_di_reservoir* d_di_in, arc_di_reservoir* d_di_out,
                        const arc_emissive_source* d_sources, const arc_alias_entry* d_alias,
                        const arc_sharc* sharc, const arc_rigid_body* bodies, uint32_t n_bodies, void* stream);
/* Pass 2B, spatial_main (:457-762): 2-3 rotated neighbour taps with depth / normal rejection,
 * Jacobian-weighted GI reconnection with a visibility ray for a winning neighbour, DI neighbour
 * reuse and the DI visibility ray. *_in = the temporal pass's output. Asynchronous. */
int arc_restir_spatial(arc_ctx* ctx, const arc_restir_params* p, const uint32_t* d_depth_mat,
                       const uint32_t* d_norm_motion, const arc_gi_reservoir* d_gi_in, arc_gi_reservoir* d_gi_out,
                       const arc_di_reservoir* d_di_in, arc_di_reservoir* d_di_out,
                       const arc_emissive_source* d_sources, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* ARCADIA_HIP_H */
