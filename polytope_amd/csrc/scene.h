// Host-side scene image and the device scene object behind ptx_scene.
#pragma once

#include <atomic>
#include <memory>
#include <mutex>
#include <vector>

#include "launch.h"
#include "ptx_internal.h"

// On failure: clears the calling thread's HIP error state (so that the caller's
// own later HIP calls do not report it again) and returns PTX_ERR_HIP.
#define PTX_HIP_TRY(expr)                                                                           \
   do {                                                                                             \
      hipError_t e_ = (expr);                                                                       \
      if (e_ != hipSuccess) {                                                                       \
         (void)hipGetLastError();                                                                   \
         return ::ptx::set_error(PTX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
      }                                                                                             \
   } while (0)

struct ptx_device;
struct ptx_scene;
namespace ptx {
struct RenderEngine;
struct HostQueryStaging;
}

namespace ptx {

struct HostImage {
   std::vector<Node64> nodes;
   std::vector<float> tris;
   std::vector<float> normals;
   std::vector<MeshRec> meshes;
   float root_box[6] = {0, 0, 0, 0, 0, 0};
   uint32_t root_ref = 0, root_count = 0;
   uint32_t max_stack = 1;
   uint32_t num_interior = 0, num_leaves = 0;
   uint32_t num_tris = 0;
};

int layout_scene(const std::vector<WorldMesh> &meshes, const HostBvh &bvh, HostImage &img);

double now_s();
const char *last_error_cstr();
// The device's tuning, copied under its lock: every launch works from such a
// snapshot, so ptx_set_tuning on another thread never changes a call in flight.
ptx_tuning tuning_of(const ptx_device *dev);
void set_tuning(ptx_device *dev, const ptx_tuning &t);
uint32_t lds_depth_of(const ptx_tuning &t);
// Lanes of a wave that gather at leaves before the wave runs a leaf step:
// measured best 1 for ray queries (leaf and interior steps are balanced) and
// 16 for the render kernels (whose leaf and shading steps share the wave).
constexpr uint32_t kDefaultLeafLanes = 1;
constexpr uint32_t kDefaultRenderLeafLanes = 16;
// Idle lanes a ray-query wave waits for before it takes new rays: 2, 4 and 8
// measured 2,377-2,383 Mrays/s against 2,342 for 1 and 2,349 for 16 (2^21
// incoherent rays on 1.3 M faces).
constexpr uint32_t kDefaultRefillLanes = 4;
inline uint32_t refill_lanes_of(const ptx_tuning &t) {
   if (!t.refill_lanes) return kDefaultRefillLanes;
   return t.refill_lanes < 64u ? t.refill_lanes : 64u;
}
uint32_t leaf_lanes_of(const ptx_tuning &t, bool render = false);
uint32_t closest_hit_grid(const ptx_device *dev, const ptx_tuning &t, uint32_t lds, bool extra, bool windows);
// The scene's window records for the closest-hit kernel, built on first use
// when the tuning asks for them (ray_windows = ON); null when it does not, or
// when they could not be built (the kernel then steps over the Node64 records,
// with the same results).  Allocates on first use.
const Node64 *scene_windows(ptx_scene *s, const ptx_tuning &t);
// The window records if already built (never allocates), else null.
const Node64 *built_windows(ptx_scene *s, const ptx_tuning &t);
size_t closest_hit_scratch_bytes(const ptx_scene *s, uint32_t grid, uint32_t lds);
int run_closest_hit(ptx_scene *s, const float *d_rays, uint32_t n, ptx_hit *d_hits, ptx_ray_counters *d_counters,
                    ptx_hit_frame *d_frames, hipStream_t stream, const ptx_tuning &t);
void device_retain(ptx_device *dev);
void device_release(ptx_device *dev);
// Scenes are reference-counted: ptx_scene_create hands out one reference,
// every server job holds one until it leaves the server, and the device
// buffers are freed when the last goes.
void scene_retain(ptx_scene *s);
void scene_release(ptx_scene *s);
// PTX_ERR_INVALID with a message when dev is NULL or closed, else PTX_OK.
int device_usable(const ptx_device *dev);
// Stream-ordered allocation from the device's own pool, whose freed blocks stay
// reserved for the library's next allocations: the default pool hands them back
// to the driver at every synchronisation, and the next call maps them again
// (7.5 ms per batch of the 16-job render workload, whose films and primary
// records take 1 GB).  The default pool when the device has none.  Free with
// hipFreeAsync.
hipError_t pool_malloc(const ptx_device *dev, void **p, size_t bytes, hipStream_t stream);
// Hands the pool's unused blocks back to the driver (before an allocation
// outside the pool is retried).
void pool_trim(const ptx_device *dev);

}  // namespace ptx

// A device handle lives until ptx_device_close and the destruction of every
// scene and server made on it, whichever comes last: each holds a reference.
struct ptx_device {
   std::atomic<int> refs{1};
   std::atomic<bool> closed{false};  // ptx_device_close was called: no new work
   int ordinal = 0;
   int num_cus = 0;
   std::string arch;
   hipStream_t stream = nullptr;     // library stream (default priority)
   void *noalloc_scratch = nullptr;  // this device's static scratch (ptx_intersect_closest_device_noalloc)
   hipMemPool_t pool = nullptr;      // the library's stream-ordered allocations (see pool_malloc)
   mutable std::mutex tuning_mu;     // guards tuning (see tuning_of)
   ptx_tuning tuning{};
   std::mutex mu;                            // serialises use of the library stream
   std::unique_ptr<ptx::RenderEngine> engine;  // path queues and graphs, created on first render
   std::unique_ptr<ptx::HostQueryStaging> host_query;  // host-buffer query staging, created on first use
   ~ptx_device();
};

struct ptx_scene {
   std::atomic<int> refs{1};
   ptx_device *dev = nullptr;
   ptx::HostBvh bvh;                 // reference-layout BVH (kept for ptx_scene_copy_bvh)
   ptx::DevScene d{};                // device pointers + root
   std::vector<void *> allocations;
   ptx_scene_info info{};
   uint32_t num_meshes = 0;
   bool has_glossy = false;          // a mesh that is not a light has PTX_BRDF_GLOSSY
   bool has_texture = false;         // a Lambert mesh that is not a light has a texture
   uint32_t root_axis = 0;           // split axis of the root when it is interior
   std::mutex windows_mu;            // guards windows and windows_failed
   ptx::Node64 *windows = nullptr;   // closest-hit window records (scene_windows), owned
   bool windows_failed = false;      // their allocation failed once: not retried
   ~ptx_scene();                     // frees every device buffer in `allocations`
};
