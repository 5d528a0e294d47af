// Host/device arguments of the primary G-buffer pass.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/arcadia_hip.h"

namespace arcdev {

struct GbufArgs {
  float eye[3], forward[3], right[3], up[3];
  float tan_half_fov, aspect, jitter[2];
  float prev_view_proj[16];
  uint32_t width, height;
  float max_dist_m;
  const float* palette;  // device copy of the 256 x 8-float material palette
};

hipError_t launch_gbuffer_raygen(const GbufArgs& g, arc_ray* rays, uint32_t n, hipStream_t s);
hipError_t launch_gbuffer_pack(const GbufArgs& g, const arc_ray* rays, const arc_trace_result* res, uint32_t n,
                               uint32_t* depth_mat, uint32_t* norm_motion, uint32_t* albedo_ao, hipStream_t s);

}  // namespace arcdev
