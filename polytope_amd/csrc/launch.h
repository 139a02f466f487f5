// Host-side declarations of the kernel launchers (defined in the .hip files).
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include "ptx_internal.h"

namespace ptx {

int launch_closest_hit(const DevScene &sc, const float *d_rays, uint32_t n, ptx_hit *d_hits,
                       ptx_ray_counters *d_counters, ptx_hit_frame *d_frames, uint32_t *d_work, uint4 *d_spill,
                       uint32_t grid, uint32_t lds_depth, uint32_t refill_min, hipStream_t stream);
int closest_hit_blocks_per_cu(uint32_t lds_depth);

}  // namespace ptx
