/*
 * arcadia_world.h — host-side world building for the 64-tree traversal (C ABI).
 *
 * Produces the exact pools the traversal reads, with the reference's construction rules:
 *   - per-chunk 4-level 64-tree: ChunkTree64::from_generator -> build_recursive ->
 *     pack_build_tree (crates/arcadia-core/src/chunk_tree/builder.rs:86-378)
 *   - global pools: WorldMap::rebuild_all_gpu_pools (crates/arcadia-world/src/world_map.rs:158-205)
 *     with Morton chunk order, slot-0 dummy node/brick and pointer relocation (:689-735)
 *   - the 256-entry material palette (crates/arcadia-core/src/material.rs:99-204)
 * plus the seeded synthetic worlds and ray sets of the benchmark workloads.
 * Return codes are the ARC_* codes of arcadia_hip.h.
 */
#ifndef ARCADIA_WORLD_H
#define ARCADIA_WORLD_H

#include <stddef.h>
#include <stdint.h>

#include "arcadia_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct arc_world_host arc_world_host;

/* A world of grid[0] x grid[1] x grid[2] chunks of 256^3 voxels, all empty. */
int arc_world_host_create(const uint32_t grid[3], arc_world_host** out);
void arc_world_host_destroy(arc_world_host* w);

/* Replace chunk `chunk_flat_idx` (x + y*gx + z*gx*gy) by the tree built from a dense 256^3
 * u8 material array (index (z*256 + y)*256 + x), with build_recursive's collapse rules. */
int arc_world_host_set_chunk_dense(arc_world_host* w, uint32_t chunk_flat_idx, const uint8_t* dense);
/* ChunkTree64::from_sparse_voxels (builder.rs:422-484): (lx,ly,lz,mat) quadruples, u32 each. */
int arc_world_host_set_chunk_sparse(arc_world_host* w, uint32_t chunk_flat_idx,
                                    const uint32_t* voxels_xyzm, uint64_t n);
/* Local (unrelocated) tree of one chunk: node/brick counts and copies. */
int arc_world_host_chunk_tree(const arc_world_host* w, uint32_t chunk_flat_idx, arc_node64* nodes,
                              uint64_t* n_nodes, arc_brick64* bricks, uint64_t* n_bricks);

/* Pack all chunks into the global pools (rebuild_all_gpu_pools). */
int arc_world_host_pack(arc_world_host* w);
/* Borrowed views of the packed pools (valid until the next pack/destroy). */
int arc_world_host_pools(const arc_world_host* w, uint32_t grid[3], const uint32_t** roots,
                         const arc_node64** nodes, uint64_t* n_nodes, const arc_brick64** bricks,
                         uint64_t* n_bricks);
/* The palette entry of `mat` as 8 floats (albedo rgb, roughness, emission rgb, variance). */
int arc_palette_material(uint32_t mat, float out8[8]);
/* Material::pack_lod_payload (material.rs:70-75). */
uint32_t arc_palette_lod_payload(uint32_t mat);

/* ---- seeded synthetic worlds --------------------------------------------------------- */
#define ARC_GEN_HEIGHTMAP 0u /* 2D value-noise fBm heightmap, random mats per voxel        */
#define ARC_GEN_CAVES 1u     /* 3D fBm terrain with caves + random axis-aligned cuboids      */
#define ARC_GEN_SPARSE 2u    /* low heightmap terrain for large sparse worlds                */

typedef struct arc_gen_params {
  uint32_t kind;
  uint32_t octaves;        /* fBm octaves */
  uint64_t seed;
  float height_min;        /* heightmap range, voxels */
  float height_max;
  float wavelength;        /* base noise wavelength, voxels */
  float density_threshold; /* ARC_GEN_CAVES: solid when density > threshold */
  uint32_t n_cuboids;      /* ARC_GEN_CAVES: random cuboids */
  uint32_t cuboid_min;     /* cuboid edge range (voxels, inclusive) */
  uint32_t cuboid_max;
  uint32_t mat_min;        /* random material range (inclusive) */
  uint32_t mat_max;
} arc_gen_params;

/* Fill `w` (all chunks) from `params`; chunks are built on `nthreads` threads. */
int arc_world_generate(arc_world_host* w, const arc_gen_params* params, int nthreads);
/* Material of global voxel (x,y,z) the generator defines (before tree building). */
uint32_t arc_gen_voxel(const arc_world_host* w, const arc_gen_params* params, uint32_t x, uint32_t y,
                       uint32_t z);

/* ---- rays ------------------------------------------------------------------------------ */
/* arc_camera (declared in arcadia_hip.h) is the pinhole camera of the reference convention
 * (types.wgsl:334-345, primary_gbuffer.wgsl:102-115): forward/right/up unit vectors, tan of HALF
 * the VERTICAL fov, aspect = width/height. */

/* Camera looking along yaw/pitch (radians): forward = (cos p sin y, sin p, cos p cos y),
 * right = normalize(cross(forward, +Y)), up = cross(right, forward). */
int arc_camera_look(const float eye[3], float yaw, float pitch, float fov_y_deg, uint32_t width,
                    uint32_t height, arc_camera* out);
/* Primary rays of a width x height image in 8x8-tile order (a wave64 traces one tile).
 * pixel_of_ray[i] = py*width + px (may be NULL). Ray fields: origin = eye (meters), dir,
 * t_start = 0, t_max = max_dist_m. Direction: normalize(f + r*(ndc.x*aspect*tan) + u*(ndc.y*tan))
 * with normalize(v) = v / sqrt(dot(v,v)). */
int arc_gen_primary_rays(const arc_camera* cam, uint32_t width, uint32_t height, float max_dist_m,
                         arc_ray* rays, uint32_t* pixel_of_ray);
/* One cosine-weighted bounce ray per hit primary ray (lighting_restir_gi.wgsl:195-209):
 * hit position = (origin*16 + dir*t_voxels) * 0.0625 m (the traversal's hit_pos_meters),
 * normal from the hit's face code, origin hit + n*0.10 m, direction cosine-distributed around
 * normalize(n + 0.22 y) with (u1,u2) from a seeded hash of (seed, ray index). Returns the
 * number of rays written; src_index[k] = primary ray index of bounce ray k (may be NULL). */
int64_t arc_gen_bounce_rays(const arc_ray* primary_rays, const arc_hit* primary_hits, uint64_t n,
                            uint64_t seed, float max_dist_m, arc_ray* rays, uint32_t* src_index);

/* Primary rays of a whole view in padded 8x8-tile order: ceil(w/8)*ceil(h/8)*64 rays, the
 * out-of-image slots holding the empty ray (t_max = -1). Returns the count or a negative code. */
int64_t arc_gen_view_rays(const arc_camera* cam, uint32_t width, uint32_t height, float max_dist_m,
                          arc_ray* rays);
/* One bounce ray per parent (dense): parent i with a hit gets the diffuse bounce of
 * arc_gen_bounce_rays hashed with index index_base + i, a miss gets the empty ray. */
int arc_gen_bounce_rays_dense(const arc_ray* parent_rays, const arc_hit* parent_hits, uint64_t n,
                              uint64_t index_base, uint64_t seed, float max_dist_m, arc_ray* rays);

/* ---- in-place edits and GPU patches ----------------------------------------------------- */
/* ChunkEditReport (crates/arcadia-core/src/chunk_tree/edit/set.rs:17-80): dirty index ranges of a
 * chunk's local node/brick arrays; min = UINT64_MAX and max = 0 when nothing is dirty. */
typedef struct arc_edit_report {
  uint64_t min_dirty_node, max_dirty_node;
  uint64_t min_dirty_brick, max_dirty_brick;
  uint32_t added_voxels;          /* air -> solid transitions */
  uint32_t had_structural_growth; /* nodes or bricks were appended */
  uint32_t reserved[2];
} arc_edit_report;

/* Chunk-local edits (no pool sync). Coordinates are chunk-local voxels in 0..256; entries outside
 * are skipped, as the reference does.
 *   arc_chunk_set_voxels   ChunkTree64::set_voxels_local_tracked (set.rs:84-138); xyzm = 4 u32
 *                          per voxel (x, y, z, material).
 *   arc_chunk_clear_voxels ChunkTree64::clear_voxels_local_tracked (clear.rs:24-53); xyz = 3 u32.
 *   arc_chunk_carve_sphere ChunkTree64::carve_sphere_tracked (carve.rs:39-177); center and radius
 *                          in chunk-local voxels; scorch_rim as the reference (mats 23/24 rim).
 *   arc_chunk_compact      ChunkTree64::compact (set.rs:156-209).
 *   arc_chunk_info         node/brick counts, solid voxel count, pool allocation
 *                          {node_base, node_cap, brick_base, brick_cap} (world_map.rs chunk_allocs). */
int arc_chunk_set_voxels(arc_world_host* w, uint32_t chunk_flat_idx, const uint32_t* xyzm, uint64_t n,
                         arc_edit_report* report);
int arc_chunk_clear_voxels(arc_world_host* w, uint32_t chunk_flat_idx, const uint32_t* xyz, uint64_t n,
                           uint32_t* removed, arc_edit_report* report);
int arc_chunk_carve_sphere(arc_world_host* w, uint32_t chunk_flat_idx, const float center[3], float radius,
                           int scorch_rim, uint32_t* destroyed, arc_edit_report* report);
int arc_chunk_compact(arc_world_host* w, uint32_t chunk_flat_idx);
int arc_chunk_info(const arc_world_host* w, uint32_t chunk_flat_idx, uint64_t* n_nodes, uint64_t* n_bricks,
                   uint64_t* solid_voxels, uint32_t alloc[4]);
/* Material at global voxel coordinates (WorldMap::get_voxel, world_map.rs:387-395); 0 outside. */
uint32_t arc_world_host_voxel(const arc_world_host* w, int32_t x, int32_t y, int32_t z);

/* Pool sync after edits (the world must be packed): WorldMap::sync_dirty_chunk
 * (world_map.rs:208-315) and sync_dirty_chunk_partial (:319-383). They update the host pools
 * returned by arc_world_host_pools (whose arrays may move) and queue GPU patches. */
int arc_world_host_sync_chunk(arc_world_host* w, uint32_t chunk_flat_idx);
int arc_world_host_sync_chunk_partial(arc_world_host* w, uint32_t chunk_flat_idx, const arc_edit_report* report);

/* World-level edits in global coordinates, each followed by the partial sync of every touched
 * chunk when the world is packed:
 *   arc_world_host_set_voxels    WorldMap::set_voxels_global (world_map.rs:586-621); xyzm = 4 i32.
 *   arc_world_host_clear_voxels  WorldMap::clear_voxels_global (:549-583); chunks in ascending index.
 *   arc_world_host_carve_crater  the tree part of WorldMap::carve_crater
 *                                (crates/arcadia-world/src/destruction.rs:47-155): center and radius in
 *                                meters; no debris particles, no island detection. */
int arc_world_host_set_voxels(arc_world_host* w, const int32_t* xyzm, uint64_t n, uint32_t* added);
int arc_world_host_clear_voxels(arc_world_host* w, const int32_t* xyz, uint64_t n, uint32_t* removed);
int arc_world_host_carve_crater(arc_world_host* w, const float center_m[3], float radius_m, int scorch_rim,
                                uint32_t* destroyed);

/* Queued GPU patches (WorldMap::pending_patches), in order, as views into the world valid until its
 * next edit, sync, pack or arc_world_host_clear_patches. Feed them to arc_world_patch. */
int arc_world_host_pending_patches(arc_world_host* w, const arc_gpu_patch** patches, uint64_t* n);
int arc_world_host_clear_patches(arc_world_host* w);

/* ---- world snapshots (new: the reference has no on-disk format, ARCHITECTURE.md:243) ---------
 * A snapshot file holds every non-empty chunk's local 64-tree (nodes[0] = root, chunk-local
 * pointers, then its bricks) in the Morton chunk order of rebuild_all_gpu_pools
 * (world_map.rs:158-205), each payload page-aligned and covered by a CRC-32 (zlib's), behind a
 * 64-byte header and a 32-byte-per-chunk index that carry their own CRC-32s.
 *   arc_snapshot_write      writes the world's chunk trees (to path.tmp, then renamed over path).
 *   arc_snapshot_open       maps the file and checks header, index order, offsets and checksums
 *                           (ARC_ERR_IO: unreadable; ARC_ERR_CORRUPT: fails a check).
 *   arc_snapshot_verify     checks every chunk payload (checksum and tree rules) on nthreads threads.
 *   arc_snapshot_load_host  a host world with the stored chunk trees (checked as they are read);
 *                           arc_world_host_pack then gives the original world's pools.
 *   arc_snapshot_upload     (arcadia_hip.h) streams the chunks straight into the device pools.
 * arc_snapshot_last_error gives the message of the calling thread's last failing snapshot call. */
int arc_snapshot_write(const arc_world_host* w, const char* path);
int arc_snapshot_open(const char* path, arc_snapshot** out);
void arc_snapshot_close(arc_snapshot* snap);
int arc_snapshot_info(const arc_snapshot* snap, uint32_t grid[3], uint32_t* n_chunks_stored, uint64_t* total_nodes,
                      uint64_t* total_bricks);
int arc_snapshot_verify(const arc_snapshot* snap, int nthreads);
int arc_snapshot_load_host(const arc_snapshot* snap, int nthreads, arc_world_host** out);
const char* arc_snapshot_last_error(void);
/* CRC-32 as the snapshot format uses it (zlib.crc32(data, crc)). */
uint32_t arc_crc32(const void* data, uint64_t n, uint32_t crc);

#ifdef __cplusplus
}
#endif
#endif /* ARCADIA_WORLD_H */
