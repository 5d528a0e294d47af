// Host-side launch arguments of the traversal kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/arcadia_world.h"

namespace arcdev {

// Queue of one persistent launch: kQueueCounters claim counters kQueueStride words apart, each on
// its own cache lines, and a finished-workgroup counter beside counter 0. Counter c hands out
// windows c, c + K, c + 2K, ... of the item array (K = kQueueCounters), so the waves homed on the
// K counters still sweep the array together while their claims go to K addresses: one atomic
// address serialises its claims (about 12 ns each on MI355X), which bounded the launches of short
// rays. All words are zero between launches.
constexpr uint32_t kQueueCounters = 8u;
constexpr uint32_t kQueueStride = 1024u;  // 4 KiB
constexpr uint32_t kQueueDone = 1u;
constexpr uint32_t kQueueRegionWords = kQueueCounters * kQueueStride;
static_assert((kQueueCounters & (kQueueCounters - 1u)) == 0u, "kQueueCounters is a power of two");

struct LaunchArgs {
  uint32_t grid[3];
  const uint32_t* roots;
  const arc_node64* nodes;
  const arc_brick64* bricks;
  arc_trace_params params;
  const arc_ray* rays;
  void* out;
  uint64_t n;
  bool full;
  bool off32;       // node and brick pools each within 4 GiB (see pool_node)
  uint32_t* queue;  // kQueueRegionWords u32 words, zero between launches
  uint32_t blocks;
  uint32_t block_threads;
  uint32_t refill;
  uint32_t batch;
};

hipError_t launch_trace(const LaunchArgs& a, hipStream_t stream);

// The static queue scratch of the allocation-free trace on the current device and that device's
// compute-unit count (the scratch's device address is looked up once per device).
hipError_t noalloc_scratch(uint32_t** queue, uint32_t* n_cu);

// One view of a path batch (device copy).
struct ViewDev {
  arc_camera cam;
  uint32_t width, height, tiles_x;
  float max_dist_m;
};

struct PathLaunchArgs {
  uint32_t grid[3];
  const uint32_t* roots;
  const arc_node64* nodes;
  const arc_brick64* bricks;
  arc_trace_params primary, bounce;
  const ViewDev* views;
  const uint32_t* view_first;  // n_views + 1 entries
  uint32_t n_views, n_items, n_bounces;
  float bounce_max_dist_m;
  uint64_t seed;
  arc_hit* out;      // (1 + n_bounces) * n_items records
  bool off32;        // node and brick pools each within 4 GiB (see pool_node)
  uint32_t* queue;   // kQueueRegionWords u32 words, zero between launches
  uint32_t blocks, block_threads, refill, batch;
};
hipError_t launch_path(const PathLaunchArgs& a, hipStream_t stream);

// One launch of a ray set and one diffuse bounce off each of its hits (k_trace_bounce).
struct BounceLaunchArgs {
  uint32_t grid[3];
  const uint32_t* roots;
  const arc_node64* nodes;
  const arc_brick64* bricks;
  arc_trace_params primary, bounce;
  const arc_ray* rays;
  uint64_t n;
  uint64_t seed;
  float bounce_max_dist_m;
  arc_hit* out;      // 2n records: primaries, then the bounce of primary i at n + i
  uint64_t* msg;     // n words, zero between launches
  bool off32;        // node and brick pools each within 4 GiB (see pool_node)
  uint32_t* queue;   // 2 * kQueueRegionWords u32 words, zero between launches
  uint32_t blocks, block_threads, refill, batch;
};
hipError_t launch_trace_bounce(const BounceLaunchArgs& a, hipStream_t stream);
hipError_t bounce_occupancy(uint32_t block_threads, uint32_t* blocks_per_cu);
// Chunk roots staged in LDS (root index + root node, 20 B per chunk): kLdsChunks covers the
// 4x4x4-chunk worlds in 5 KB per workgroup of 256 lanes; kLdsChunksLarge covers a 16x16x16-chunk
// world in 80 KB, held by the path kernel's large shape, one workgroup of kPathLargeBlock lanes
// per CU.
constexpr uint32_t kLdsChunks = 256u;
constexpr uint32_t kLdsChunksLarge = 4096u;
constexpr uint32_t kPathLargeBlock = 1024u;

// One launch of a shadow any-hit walker over a ray set.
struct ShadowLaunchArgs {
  uint32_t grid[3];
  const uint32_t* roots;
  const arc_node64* nodes;
  const arc_brick64* bricks;
  uint32_t mode, max_steps;  // arc_shadow_params
  const arc_ray* rays;
  float* vis;
  arc_shadow_rec* rec;  // may be null
  uint64_t n;
  uint32_t* queue;  // kQueueRegionWords u32 words, zero between launches
  uint32_t blocks, block_threads, refill, batch;
};
hipError_t launch_shadow(const ShadowLaunchArgs& a, hipStream_t stream);

// One launch of the rigid-body walker: `bodies` is a device copy of at most 32 instances.
struct RigidLaunchArgs {
  const arc_node64* nodes;
  const arc_brick64* bricks;
  uint64_t n_nodes, n_bricks;
  const arc_rigid_body* bodies;
  uint32_t n_bodies;
  const arc_ray* rays;
  uint64_t n;
  arc_rigid_hit* out;
};
hipError_t launch_rigid(const RigidLaunchArgs& a, hipStream_t stream);

// One launch of the point query.
struct QueryLaunchArgs {
  uint32_t grid[3];
  const uint32_t* roots;
  const arc_node64* nodes;
  const arc_brick64* bricks;
  uint64_t n_nodes, n_bricks;
  const float* points;
  uint64_t n;
  uint32_t* mat;
};
hipError_t launch_query(const QueryLaunchArgs& a, hipStream_t stream);
hipError_t shadow_occupancy(uint32_t block_threads, uint32_t* blocks_per_cu);
hipError_t path_occupancy(uint32_t block_threads, uint32_t* blocks_per_cu);
// Resident workgroups per CU of the traversal kernel at `block_threads`.
hipError_t kernel_occupancy(uint32_t semantics, bool full, uint32_t block_threads, uint32_t* blocks_per_cu);

}  // namespace arcdev
