// Host/device arguments of the path shading kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "trace_launch.h"

namespace arcdev {

struct ShadeArgs {
  const ViewDev* views;
  const uint32_t* view_first;
  uint32_t n_views, n_items, n_bounces;
  float bounce_tmax;
  uint64_t seed;
  const uint4* hits;       // (1 + n_bounces) * n_items compact records of the path launch
  const float* palette;    // 256 x 8 floats
  float sun[3];
  float emission_scale, sky_scale;
  arc_path_sample* samples;  // per slot, or null
  float* radiance;           // 3 floats per slot, or null
  double* accum;             // 3 doubles per slot, += radiance, or null
};

hipError_t launch_shade(const ShadeArgs& a, hipStream_t s);
hipError_t launch_estimate(const arc_path_sample* samples, const float* cache, uint32_t n, uint32_t mode, float* out,
                           hipStream_t s);

}  // namespace arcdev
