// Host-buffer ray queries (ptx_intersect_closest): per-device staging kept
// between calls.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "ptx_internal.h"

struct ptx_scene;

namespace ptx {

// The device copy of a host query's rays, hits and counters, and the pinned
// host slots its copies go through.  Grown on demand, freed when the device is
// closed; guarded by the device's `mu`.
struct HostQueryStaging {
   static constexpr size_t kSlotBytes = size_t(2) << 20;  // one pinned chunk
   static constexpr uint32_t kMaxWorkers = 4;             // host threads of one call
   int ordinal = 0;
   char *d_buf = nullptr;
   size_t d_cap = 0;
   char *pinned = nullptr;                  // kMaxWorkers x 2 slots of kSlotBytes
   hipEvent_t events[2 * kMaxWorkers] = {};  // one per slot: its last copy
   bool ready = false;                      // pinned slots and events exist
   ~HostQueryStaging();
};

// Closest hit of n host rays into host hits (and counters when not NULL).
// The caller holds the device's `mu` and has made the device current.
int intersect_host(ptx_scene *s, const float *rays, uint32_t n, ptx_hit *hits, ptx_ray_counters *counters);

}  // namespace ptx
