// Host-side launch arguments of the traversal kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/arcadia_hip.h"

namespace arcdev {

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
  uint32_t* queue;  // two u32 counters, zero between launches
  uint32_t blocks;
  uint32_t block_threads;
  uint32_t refill;
  uint32_t batch;
};

hipError_t launch_trace(const LaunchArgs& a, hipStream_t stream);
// Resident workgroups per CU of the traversal kernel at `block_threads`.
hipError_t kernel_occupancy(uint32_t semantics, bool full, uint32_t block_threads, uint32_t* blocks_per_cu);

}  // namespace arcdev
