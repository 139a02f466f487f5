// Copies between the caller's host buffers and HBM through pinned slots kept
// per device: host-buffer ray queries (ptx_intersect_closest) and render film
// readback.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "ptx_internal.h"

struct ptx_scene;
struct ptx_device;

namespace ptx {

// The device copy of a host query's rays, hits and counters, and the pinned
// host slots that host queries and film readback copy through.  Grown on demand, freed when the device is
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

// A host buffer and its device copy.
struct Span {
   char *host;
   char *dev;
   size_t bytes;
};
// A piece of a span that one copy moves (at most kSlotBytes).
struct Piece {
   char *host;
   char *dev;
   size_t len;
};

// Closest hit of n host rays into host hits (and counters when not NULL).
// The caller holds the device's `mu` and has made the device current.
int intersect_host(ptx_scene *s, const float *rays, uint32_t n, ptx_hit *hits, ptx_ray_counters *counters);

// Copies every span's device bytes to its host buffer (pageable memory is
// fine) once the work queued on `stream` before the call has finished, and
// waits for it.  The caller holds the device's `mu` and has made the device
// current.
int download_host(ptx_device *dev, hipStream_t stream, const std::vector<Span> &spans);

}  // namespace ptx
