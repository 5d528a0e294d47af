// Device side of the ReSTIR GI / DI passes (arc_restir_temporal / arc_restir_spatial).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/arcadia_hip.h"

namespace arcdev {

// Everything a pass kernel reads or writes. The traversal between the kernels is done by the
// library's trace entry points on the ray buffers `rays_*`, whose results come back in the
// matching result buffers.
struct RestirDev {
  arc_restir_params p;
  uint32_t gx, gy, gz;
  const uint32_t* roots;
  const uint4* nodes;
  const float* palette;  // 256 x {albedo rgb, roughness, emission rgb, variance}
  const uint2* depth_mat;
  const uint2* norm_motion;
  const arc_gi_reservoir* gi_in;
  arc_gi_reservoir* gi_out;
  const arc_di_reservoir* di_in;
  arc_di_reservoir* di_out;
  const arc_emissive_source* sources;
  const arc_alias_entry* alias;
  uint32_t has_sharc;
  uint32_t n_bodies;
  // per-pixel scratch
  uint2* state;                   // temporal: flags, history pixel
  arc_ray* rays_a;                // temporal: sun shadow rays; spatial: GI reconnection visibility rays
  arc_ray* rays_b;                // temporal: GI rays; spatial: DI visibility rays
  arc_ray* rays_c;                // temporal: GI rays cut at the world hit, for the rigid bodies
  arc_ray* rays_d;                // temporal: sun shadow rays from the GI hits
  const arc_hit* hits_a;          // temporal, shadow_mode 0: 64-tree sun shadow trace
  const float* vis_a;             // temporal, shadow_mode 1: sun walker; spatial: GI visibility
  const float* vis_b;             // spatial: DI visibility
  const arc_trace_result* full_b; // temporal: GI hits
  const arc_rigid_hit* rb_c;      // temporal: rigid-body GI hits
  arc_sharc_probe* probes;        // temporal: SHARC lookups at the GI hits
  const float4* sharc_out;
  const float* vis_d;             // temporal: bounce sun visibility
};

hipError_t launch_rt_setup(const RestirDev& d, hipStream_t s);
hipError_t launch_rt_rigid_rays(const RestirDev& d, hipStream_t s);
hipError_t launch_rt_bounce(const RestirDev& d, hipStream_t s);
hipError_t launch_rt_resolve(const RestirDev& d, hipStream_t s);
hipError_t launch_rs_rays(const RestirDev& d, hipStream_t s);
hipError_t launch_rs_resolve(const RestirDev& d, hipStream_t s);

}  // namespace arcdev
