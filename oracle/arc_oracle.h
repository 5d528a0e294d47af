/*
 * CPU oracle for the 64-tree DDA traversal — TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference traversal, used as the checker of the HIP kernels
 * (tests/, the smoke check and the CPU-baseline leg of bench.py). No product path links,
 * loads or calls this library.
 *
 * Restated functions (each cites the reference lines it follows in arc_oracle.c):
 *   orc_trace_world   trace_world_64tree_from   shaders/common/traversal64.wgsl:226-556
 *   orc_trace_chunk   raycast_chunk_tree        crates/arcadia-core/src/raycast_cpu.rs:28-218
 *   orc_trace_meters  WorldMap::raycast_meters  crates/arcadia-world/src/world_map.rs:464-545
 *   orc_query_voxel   query_voxel_world_meters  shaders/common/traversal64.wgsl:182-220
 *   orc_trace_shadow  trace_shadow_vis / _di    shaders/common/traversal64.wgsl:571-773 / 781-989
 *   orc_trace_rigid   trace_rigid_bodies        shaders/common/traversal64.wgsl:1005-1244
 *   orc_query_point   query_voxel_world_meters  shaders/common/traversal64.wgsl:182-220
 *   orc_restir_*      temporal_main / spatial_main  shaders/lighting_restir_gi.wgsl:62-762 (restir_oracle.c)
 * plus an independent checker that shares no code with the above:
 *   orc_brute_dense   Amanatides-Woo unit-voxel walk in double precision over a dense grid.
 *
 * Pinning (what checks each function): the four reference KATs of raycast_cpu.rs:334-446;
 * the dense brute-force walker (hit voxel/face away from ties, t within tolerance) on
 * seeded worlds; closed forms for axis-aligned rays; agreement between orc_trace_world and
 * orc_trace_chunk where the reference semantics coincide (one chunk, no LOD, no tiny
 * direction components); and a second, pure-Python restatement (oracle/py_trace.py) that
 * must match step counts bit for bit on small worlds; hand-derived step counts, LOD stops,
 * lod_rgb and contact-AO closed forms on constructed worlds (tests/test_step_kat.py).
 * orc_trace_meters is pinned by the KATs, the brute-force walker and the rule of glam
 * 0.29.3's Vec3::normalize_or_zero (rcp = 1 / length; self * rcp when rcp is finite and
 * > 0, else zero), checked on zero, NaN, infinite and overflowing directions and on
 * power-of-two scalings (tests/test_kat.py); glam's compiled code cannot run here.
 */
#ifndef ARC_ORACLE_H
#define ARC_ORACLE_H

#include <stdint.h>
#include "../include/arcadia_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_world {
  uint32_t grid[3];
  const uint32_t* roots;
  const arc_node64* nodes;
  uint64_t n_nodes;
  const arc_brick64* bricks;
  uint64_t n_bricks;
} orc_world;

/* Per-ray memory-traffic counters: loads the algorithm performs from each pool. */
typedef struct orc_counters {
  uint32_t node_loads;   /* node_pool[...] reads (root + descents; register-cached ancestors: 0) */
  uint32_t brick_loads;  /* brick word reads */
  uint32_t root_loads;   /* chunk_roots reads on a chunk-cache miss */
  uint32_t steps;
  uint32_t level_loads[4]; /* node_loads split by tree level of the loaded node (0 = chunk root) */
  uint32_t restarts[4];    /* outer iterations starting at level 0 (root) .. 3 (cached leaf) */
  uint32_t empty_chunk_steps;
} orc_counters;

/* All trace functions run over rays [0, n) on `nthreads` threads (<=0: one). */
void orc_trace_world(const orc_world* w, const arc_trace_params* p, const arc_ray* rays,
                     uint64_t n, arc_trace_result* out, orc_counters* counters, int nthreads);
void orc_trace_chunk(const orc_world* w, const arc_ray* rays, uint64_t n, arc_trace_result* out,
                     orc_counters* counters, int nthreads);
void orc_trace_meters(const orc_world* w, const arc_ray* rays, uint64_t n, arc_trace_result* out,
                      int nthreads);
/* Shadow any-hit walkers (mode ARC_SHADOW_SUN / ARC_SHADOW_DI): visibility and the per-ray
 * record the kernel writes. */
void orc_trace_shadow(const orc_world* w, const arc_shadow_params* sp, const arc_ray* rays, uint64_t n, float* vis,
                      arc_shadow_rec* rec, int nthreads);
/* Rigid voxel bodies (first min(n_bodies, 32) of `bodies`), single-threaded; point query in
 * world metres. */
void orc_trace_rigid(const orc_world* w, const arc_rigid_body* bodies, uint32_t n_bodies, const arc_ray* rays,
                     uint64_t n, arc_rigid_hit* out);
uint32_t orc_query_point(const orc_world* w, const float p[3]);
/* One ray of orc_trace_world / orc_trace_shadow / orc_trace_rigid. */
void orc_trace_world_ray(const orc_world* w, const arc_trace_params* p, const arc_ray* ray, arc_trace_result* res);
float orc_shadow_ray(const orc_world* w, uint32_t mode, uint32_t ray_steps_w, const arc_ray* ray);
void orc_rigid_ray(const orc_world* w, const arc_rigid_body* bodies, uint32_t n_bodies, const arc_ray* ray,
                   arc_rigid_hit* out);

/* ReSTIR GI / DI passes, temporal_main and spatial_main of shaders/lighting_restir_gi.wgsl
 * (:62-455, :457-762), one pixel at a time with the traversals inline, as the shader is written
 * (restir_oracle.c). The material palette (256 x {albedo rgb, roughness, emission rgb, variance})
 * and the SHARC table are inputs; sharc_resolved NULL = no SHARC (every lookup misses). */
typedef struct orc_restir_in {
  const arc_restir_params* p;
  const arc_trace_params* sun_trace;
  const arc_trace_params* gi_trace;
  const uint32_t* depth_mat;
  const uint32_t* norm_motion;
  const arc_gi_reservoir* gi_in;
  const arc_di_reservoir* di_in;
  const arc_emissive_source* sources;
  const arc_alias_entry* alias;
  const float* palette;
  uint32_t sharc_log2_buckets;
  const uint32_t* sharc_resolved;
  const arc_rigid_body* bodies;
  uint32_t n_bodies;
} orc_restir_in;
void orc_restir_temporal(const orc_world* w, const orc_restir_in* in, arc_gi_reservoir* gi_out,
                         arc_di_reservoir* di_out, int nthreads);
void orc_restir_spatial(const orc_world* w, const orc_restir_in* in, arc_gi_reservoir* gi_out,
                        arc_di_reservoir* di_out, int nthreads);
/* Compact records as the kernels write them. */
void orc_compact(const arc_trace_result* full, uint64_t n, int world_units, arc_hit* out);

/* primary_raymarch's opaque path (shaders/primary_gbuffer.wgsl:93-312): rays in padded 8x8-tile
 * order, traced with orc_trace_world, packed into depth_mat (2 u32/px), norm_motion (2 u32/px)
 * and albedo_ao (1 u32/px). `palette` = 256 x {albedo rgb, roughness, emission rgb, variance}. */
void orc_primary_gbuffer(const orc_world* w, const arc_trace_params* p, const arc_gbuffer_params* g,
                         const float* palette, uint32_t* depth_mat, uint32_t* norm_motion,
                         uint32_t* albedo_ao, int nthreads);
/* f32 -> f16 bits, round to nearest even (the conversion of pack2x16float). */
uint16_t orc_f32_to_f16(float x);

/* SHARC cache (shaders/common/sharc.wgsl, sharc_update.wgsl:129-169, sharc_resolve.wgsl,
 * sharc_query.wgsl): keys u32[S], accum u32[4S], resolved u32[4S], S = 4 << log2_buckets.
 * Deposits claim slots sequentially in input order. */
float orc_f16_to_f32(uint16_t h);
void orc_sharc_make_key(const float pos[3], const float n[3], const float cam[3], float jx, float jy,
                        uint32_t bucket_mask, uint32_t* bucket, uint32_t* key);
void orc_sharc_deposit(uint32_t log2_buckets, uint32_t* keys, uint32_t* accum, const arc_sharc_sample* s, uint64_t n,
                       const float cam[3]);
void orc_sharc_resolve(uint32_t n_slots, uint32_t* keys, uint32_t* accum, uint32_t* resolved);
void orc_sharc_query(uint32_t log2_buckets, const uint32_t* resolved, const arc_sharc_probe* q, uint64_t n,
                     const float cam[3], float* out);

uint32_t orc_query_voxel(const orc_world* w, uint32_t x, uint32_t y, uint32_t z);
/* Dense material grid (x fastest) of the whole world, size (gx*256)*(gy*256)*(gz*256). */
void orc_decode_dense(const orc_world* w, uint8_t* dense, int nthreads);
/* Brute-force unit-voxel walk. Output per ray: hit (0/1), voxel, entry axis/sign code in
 * face (2*axis + (dir>0)), t in voxel units (double), and tie = 1 when the entry plane is
 * ambiguous to within 1e-9 relative (face then unchecked). */
typedef struct orc_brute_hit {
  double t;
  int32_t vx, vy, vz;
  int32_t hit;
  int32_t face;
  int32_t tie;
  int32_t mat;
  int32_t pad;
} orc_brute_hit;
void orc_brute_dense(const uint8_t* dense, const uint32_t dims[3], const arc_ray* rays, uint64_t n,
                     int world_units, double max_t_scale, orc_brute_hit* out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
