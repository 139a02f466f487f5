// Closest-hit ray queries for gfx950 (ptx_intersect_closest*): the
// persistent traversal of traverse.h over an array of rays, one scene.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "launch.h"
#include "traverse.h"

#pragma clang fp contract(off)

namespace ptx {
namespace dev {

// kExtra: the caller may want the per-ray counters or frames; without it the
// kernel writes hits only, and the counter and face-test bookkeeping of the
// traversal is compiled out.
template <bool kExtra>
__device__ __forceinline__ void write_result(const Trav &s, const DevScene &sc, uint32_t id, ptx_hit *hits,
                                             ptx_ray_counters *counters, ptx_hit_frame *frames) {
   ptx_hit h;
   if (s.hit != 0xFFFFFFFFu) {
      h.mesh = sc.pairs[2 * (size_t)s.hit];
      h.face = sc.pairs[2 * (size_t)s.hit + 1];
      h.t = s.min_t; h.u = s.hu; h.v = s.hv; h.w = s.hw;
   } else {
      h.mesh = 0xFFFFFFFFu; h.face = 0xFFFFFFFFu;
      h.t = kInf; h.u = h.v = h.w = 0.f;
   }
   hits[id] = h;
   if (!kExtra) return;
   if (counters) counters[id] = ptx_ray_counters{s.bb_hits, s.bb_misses, s.tris};
   if (frames) {
      ptx_hit_frame fr;
      if (s.hit != 0xFFFFFFFFu) {
         Frame f;
         const float *cn = sc.normals ? sc.normals + 9 * (size_t)s.hit : nullptr;
         if (cn && !(sc.meshes[h.mesh].flags & kRecNormals)) cn = nullptr;
         hit_frame(sc.tris + 9 * (size_t)s.hit, cn, s.r, s.hu, s.hv, s.hw, s.after, f);
         fr = ptx_hit_frame{{f.lx, f.ly, f.lz}, {f.nx, f.ny, f.nz}, {f.t1x, f.t1y, f.t1z},
                            {f.t2x, f.t2y, f.t2z}, {f.ex, f.ey, f.ez}};
      } else {
         fr = ptx_hit_frame{};
      }
      frames[id] = fr;
   }
}

// kWin: step over the scene's window records (win, root_cnt = the root's
// split axis when it is interior) instead of its Node64 records.
template <bool kExtra, bool kWin>
struct RayQuery {
   struct Ctx {};
   static constexpr bool kStaged = true;
   static constexpr bool kWindows = kWin;
   DevScene sc;
   const float *rays;
   ptx_hit *hits;
   ptx_ray_counters *counters;
   ptx_hit_frame *frames;
   const Node64 *win;
   uint32_t root_cnt;
   __device__ __forceinline__ void load(uint32_t id, Fetched &f) const {
      const float *rp = rays + 6 * (size_t)id;
#pragma unroll
      for (int k = 0; k < 6; k++) f.q[k] = rp[k];
   }
   __device__ __forceinline__ int start(uint32_t, const Fetched &f, Ctx &, Trav &s) const {
      make_ray(s.r, f.q[0], f.q[1], f.q[2], f.q[3], f.q[4], f.q[5]);
      return 0;
   }
   __device__ __forceinline__ void load2(Ctx &, Fetched &) const {}
   __device__ __forceinline__ void start2(Ctx &, const Fetched &, Trav &) const {}
   __device__ __forceinline__ void begin_root(Ctx &, Trav &s) const {
      begin(s, sc.root_box, sc.root_ref, sc.root_count);
   }
   __device__ __forceinline__ uint32_t root_count() const { return root_cnt; }
   __device__ __forceinline__ const Node64 *nodes(const Ctx &) const { return sc.nodes; }
   __device__ __forceinline__ const Node64 *windows() const { return win; }
   __device__ __forceinline__ const float *tris(const Ctx &) const { return sc.tris; }
   __device__ __forceinline__ bool debug_pixel(const Ctx &) const { return false; }
   __device__ __forceinline__ void finish(uint32_t id, const Ctx &, const Trav &s) const {
      write_result<kExtra>(s, sc, id, hits, counters, frames);
   }
};

// The hits-only node-record form is held to at most 96 VGPRs, five waves per
// SIMD (with a 12- or 8-entry LDS stack five blocks fit a CU's LDS too): four
// blocks per CU ran 2,154 against 2,307-2,326 Mrays/s with five on 2^21
// incoherent rays.  The counters/frames and window forms need more registers
// and keep four waves rather than spill.
template <int kLds, bool kExtra, bool kWin>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kExtra || kWin ? 4 : 5, 8)))
k_closest_hit(DevScene sc, const float *__restrict__ rays, uint32_t n, ptx_hit *__restrict__ hits,
              ptx_ray_counters *counters, ptx_hit_frame *frames, uint32_t *work, uint4 *spill, uint32_t refill_min,
              uint32_t leaf_min, const Node64 *win, uint32_t root_cnt) {
   const RayQuery<kExtra, kWin> q{sc, rays, hits, counters, frames, win, root_cnt};
   persistent_trace<kLds>(q, n, work, spill, refill_min, leaf_min, sc.max_stack);
}

// One thread per window: window w of interior node i is win[4 i + 2 w] (the
// node's record) and win[4 i + 2 w + 1] (the record of the child visited first
// for direction sign w on the node's axis: child 1 when w = 1, child 0 when
// w = 0; zeros when that child is a leaf).
__global__ void k_build_windows(const Node64 *__restrict__ nodes, uint32_t n_int, Node64 *__restrict__ win) {
   const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= 2 * (size_t)n_int) return;
   const uint32_t i = (uint32_t)(t >> 1), w = (uint32_t)(t & 1);
   const Node64 a = nodes[i];
   const uint32_t child = a.ref[w];
   Node64 b{};
   if (!(child & kLeafBit)) b = nodes[child];
   win[2 * t] = a;
   win[2 * t + 1] = b;
}

__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
   const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) p[i] = 0u;
}

}  // namespace dev

// Scratch of ptx_intersect_closest_device_noalloc: the traversal's work
// counters (first 256 bytes) and its stack spill, one copy per device, held
// until the process ends.  Host code never passes `scratch` itself to a
// kernel (on the host it is only a stand-in for the symbol): it uses the
// address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) uint4 scratch[kNoAllocScratchBytes / sizeof(uint4)];

namespace {
constexpr int kMaxOrdinals = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxOrdinals] = {};
}  // namespace

int noalloc_scratch_address(int ordinal, void **ptr) {
   *ptr = nullptr;
   if (ordinal < 0 || ordinal >= kMaxOrdinals) return -1;
   std::lock_guard<std::mutex> lock(g_scratch_mu);
   if (!g_scratch_ptr[ordinal]) {
      void *p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
         (void)hipGetLastError();
         return -1;
      }
      g_scratch_ptr[ordinal] = p;
   }
   *ptr = g_scratch_ptr[ordinal];
   return 0;
}

// Host launcher.  work: >= 8 u32 of device memory; spill sized by the caller
// for (max_stack - lds) * grid * 256 entries.
namespace {

template <int kLds, bool kWin>
void launch_form(const DevScene &sc, const float *d_rays, uint32_t n, ptx_hit *d_hits, ptx_ray_counters *d_counters,
                 ptx_hit_frame *d_frames, uint32_t *d_work, uint4 *d_spill, uint32_t grid, uint32_t refill_min,
                 uint32_t leaf_min, const Node64 *win, uint32_t root_cnt, hipStream_t stream) {
   if (d_counters || d_frames)
      hipLaunchKernelGGL((dev::k_closest_hit<kLds, true, kWin>), dim3(grid), dim3(256), 0, stream, sc, d_rays, n,
                         d_hits, d_counters, d_frames, d_work, d_spill, refill_min, leaf_min, win, root_cnt);
   else
      hipLaunchKernelGGL((dev::k_closest_hit<kLds, false, kWin>), dim3(grid), dim3(256), 0, stream, sc, d_rays, n,
                         d_hits, d_counters, d_frames, d_work, d_spill, refill_min, leaf_min, win, root_cnt);
}

template <int kLds>
void launch_lds(const DevScene &sc, const float *d_rays, uint32_t n, ptx_hit *d_hits, ptx_ray_counters *d_counters,
                ptx_hit_frame *d_frames, uint32_t *d_work, uint4 *d_spill, uint32_t grid, uint32_t refill_min,
                uint32_t leaf_min, const Node64 *win, uint32_t root_axis, hipStream_t stream) {
   if (win) {
      // the root's count slot: its face count when it is a leaf, else its split axis
      const uint32_t root_cnt = (sc.root_ref & kLeafBit) ? sc.root_count : root_axis;
      launch_form<kLds, true>(sc, d_rays, n, d_hits, d_counters, d_frames, d_work, d_spill, grid, refill_min,
                              leaf_min, win, root_cnt, stream);
   } else {
      launch_form<kLds, false>(sc, d_rays, n, d_hits, d_counters, d_frames, d_work, d_spill, grid, refill_min,
                               leaf_min, nullptr, sc.root_count, stream);
   }
}

template <int kLds, bool kWin>
int blocks_form(bool extra) {
   int b = 0;
   const hipError_t e =
         extra ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, dev::k_closest_hit<kLds, true, kWin>, 256, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, dev::k_closest_hit<kLds, false, kWin>, 256, 0);
   return e == hipSuccess && b > 0 ? b : 1;
}

template <int kLds>
int blocks_lds(bool extra, bool windows) {
   return windows ? blocks_form<kLds, true>(extra) : blocks_form<kLds, false>(extra);
}

}  // namespace

int launch_closest_hit(const DevScene &sc, const float *d_rays, uint32_t n, ptx_hit *d_hits,
                       ptx_ray_counters *d_counters, ptx_hit_frame *d_frames, uint32_t *d_work, uint4 *d_spill,
                       uint32_t grid, uint32_t lds_depth, uint32_t refill_min, uint32_t leaf_min,
                       const Node64 *windows, uint32_t root_axis, hipStream_t stream) {
   hipLaunchKernelGGL(dev::k_zero_u32, dim3(1), dim3(64), 0, stream, d_work, (uint32_t)dev::kWorkWords);
   if (n == 0) return 0;
   switch (lds_depth) {
      case 8:
         launch_lds<8>(sc, d_rays, n, d_hits, d_counters, d_frames, d_work, d_spill, grid, refill_min, leaf_min,
                       windows, root_axis, stream);
         break;
      case 16:
         launch_lds<16>(sc, d_rays, n, d_hits, d_counters, d_frames, d_work, d_spill, grid, refill_min, leaf_min,
                        windows, root_axis, stream);
         break;
      case 24:
         launch_lds<24>(sc, d_rays, n, d_hits, d_counters, d_frames, d_work, d_spill, grid, refill_min, leaf_min,
                        windows, root_axis, stream);
         break;
      default:
         launch_lds<12>(sc, d_rays, n, d_hits, d_counters, d_frames, d_work, d_spill, grid, refill_min, leaf_min,
                        windows, root_axis, stream);
         break;
   }
   return hipGetLastError() == hipSuccess ? 0 : -1;
}

int closest_hit_blocks_per_cu(uint32_t lds_depth, bool extra, bool windows) {
   switch (lds_depth) {
      case 8: return blocks_lds<8>(extra, windows);
      case 16: return blocks_lds<16>(extra, windows);
      case 24: return blocks_lds<24>(extra, windows);
      default: return blocks_lds<12>(extra, windows);
   }
}

int launch_build_windows(const Node64 *nodes, uint32_t n_int, Node64 *win, hipStream_t stream) {
   if (n_int == 0) return 0;
   const size_t threads = 2 * (size_t)n_int;
   hipLaunchKernelGGL(dev::k_build_windows, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, stream, nodes,
                      n_int, win);
   return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace ptx
