// Closest-hit kernels for gfx950.
//
// One persistent kernel, one lane per ray in flight.  Waves fetch rays in
// batches from eight per-XCD segments of the ray list (blockIdx % 8 is the
// XCD the dispatcher places a block on), so coherent rays stay in one L2, and
// refill idle lanes whenever enough have finished (ballot + mbcnt compaction:
// one atomic per wave per refill).  The deferred-child stack lives in LDS
// (`kLds` entries per lane); deeper entries spill to a global array sized from
// the tree depth.
#include <hip/hip_runtime.h>

#include "trace_device.h"

#pragma clang fp contract(off)

namespace ptx {
namespace dev {

struct StackView {
   uint2 *lds;         // [kLds][blockDim]
   uint2 *spill;       // [max_stack - kLds][stride]
   uint32_t tid, gtid, stride, lds_depth, max_stack;
   __device__ __forceinline__ void put(uint32_t k, uint2 e) const {
      if (k < lds_depth) lds[k * blockDim.x + tid] = e;
      else spill[(size_t)(k - lds_depth) * stride + gtid] = e;
   }
   __device__ __forceinline__ uint2 get(uint32_t k) const {
      if (k >= max_stack) return make_uint2(0xFFFFFFFFu, 0x7F800000u);   // unreachable: depth bounds sp
      if (k < lds_depth) return lds[k * blockDim.x + tid];
      return spill[(size_t)(k - lds_depth) * stride + gtid];
   }
};

// Pops deferred children until one passes its pop-time test; returns false
// when the stack is empty (traversal done).
__device__ __forceinline__ bool pop_next(Trav &s, const StackView &st) {
   while (s.sp > 0) {
      s.sp--;
      const uint2 e = st.get(s.sp);
      const float near = __uint_as_float(e.y);
      if (near < 0.f || near <= s.min_t) {
         s.bb_hits++;
         s.cur = e.x;
         return true;
      }
      s.bb_misses++;
   }
   return false;
}

__device__ __forceinline__ void begin(Trav &s, const DevScene &sc) {
   s.min_t = kInf;
   s.hit = 0xFFFFFFFFu;
   s.after = 0;
   s.sp = 0;
   s.bb_hits = s.bb_misses = s.tris = 0;
   s.hu = s.hv = s.hw = 0.f;
   bool inside;
   float near, far;
   slab(sc.root_box[0], sc.root_box[1], sc.root_box[2], sc.root_box[3], sc.root_box[4], sc.root_box[5],
        s.r, inside, near, far);
   if (inside || (near <= far && near <= s.min_t)) {
      s.bb_hits++;
      s.cur = sc.root_ref;
   } else {
      s.bb_misses++;
      s.cur = 0xFFFFFFFFu;
   }
}

// One step: visit the interior node or leaf in s.cur.  Returns false when the
// traversal has finished.
__device__ __forceinline__ bool step(Trav &s, const DevScene &sc, const StackView &st, bool debug_pixel) {
   if (s.cur == 0xFFFFFFFFu) return false;
   if (s.cur & kLeafBit) {
      const LeafRec lf = sc.leaves[s.cur & ~kLeafBit];
      for (uint32_t i = 0; i < lf.count; i++) {
         const uint32_t t = lf.first + i;
         if (test_triangle(sc.tris + 9 * (size_t)t, s.r, s.min_t, s.hu, s.hv, s.hw, debug_pixel)) {
            s.hit = t;
            s.after = 0;
         }
         s.after++;
      }
      s.tris += lf.count;
      if (!pop_next(s, st)) { s.cur = 0xFFFFFFFFu; return false; }
      return true;
   }
   const float4 *rec = reinterpret_cast<const float4 *>(sc.nodes + s.cur);
   const float4 qa = rec[0], qb = rec[1], qc = rec[2];
   const uint4 qm = reinterpret_cast<const uint4 *>(rec)[3];
   bool in0, in1;
   float n0, f0, n1, f1;
   slab(qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, s.r, in0, n0, f0);
   slab(qb.z, qb.w, qc.x, qc.y, qc.z, qc.w, s.r, in1, n1, f1);
   const bool h0 = in0 || (n0 <= f0 && n0 <= s.min_t);
   const bool h1 = in1 || (n1 <= f1 && n1 <= s.min_t);
   // bvh.cpp:510-517: with neg_dir[axis] the high child (child 1) is popped first.
   const bool high_first = (s.r.neg >> qm.z) & 1u;
   const bool fh = high_first ? h1 : h0, sh = high_first ? h0 : h1;
   const uint32_t fref = high_first ? qm.y : qm.x, sref = high_first ? qm.x : qm.y;
   const float snear = high_first ? (in0 ? -1.f : n0) : (in1 ? -1.f : n1);
   if (fh) {
      s.bb_hits++;
      if (sh) {
         if (s.sp < st.max_stack) st.put(s.sp, make_uint2(sref, __float_as_uint(snear)));
         s.sp++;
      } else {
         s.bb_misses++;
      }
      s.cur = fref;
      return true;
   }
   s.bb_misses++;
   if (sh) {
      s.bb_hits++;
      s.cur = sref;
      return true;
   }
   s.bb_misses++;
   if (!pop_next(s, st)) { s.cur = 0xFFFFFFFFu; return false; }
   return true;
}

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
   if (counters) counters[id] = ptx_ray_counters{s.bb_hits, s.bb_misses, s.tris};
   if (frames) {
      ptx_hit_frame fr;
      if (s.hit != 0xFFFFFFFFu) {
         Frame f;
         const float *cn = sc.normals ? sc.normals + 9 * (size_t)s.hit : nullptr;
         if (cn && !sc.meshes[h.mesh].has_normals) cn = nullptr;
         hit_frame(sc.tris + 9 * (size_t)s.hit, cn, s.r, s.hu, s.hv, s.hw, s.after, f);
         fr = ptx_hit_frame{{f.lx, f.ly, f.lz}, {f.nx, f.ny, f.nz}, {f.t1x, f.t1y, f.t1z},
                            {f.t2x, f.t2y, f.t2z}, {f.ex, f.ey, f.ez}};
      } else {
         fr = ptx_hit_frame{};
      }
      frames[id] = fr;
   }
}

constexpr int kSegments = 8;

template <int kLds>
__global__ void __launch_bounds__(256) k_closest_hit(DevScene sc, const float *__restrict__ rays, uint32_t n,
                                                     ptx_hit *__restrict__ hits, ptx_ray_counters *counters,
                                                     ptx_hit_frame *frames, uint32_t *work, uint2 *spill,
                                                     uint32_t refill_min) {
   __shared__ uint2 lds[kLds * 256];
   const uint32_t lane = __lane_id();
   const StackView st{lds, spill, threadIdx.x, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x,
                      (uint32_t)kLds, sc.max_stack};
   const uint32_t xcd = blockIdx.x % kSegments;
   Trav s;
   bool active = false;
   bool exhausted = false;
   uint32_t id = 0;
   while (true) {
      const uint64_t idle = __ballot(!active);
      const uint32_t n_idle = __popcll(idle);
      if (!exhausted && n_idle >= (idle == __ballot(1) ? 1u : refill_min)) {
         uint32_t base = 0, got = 0;
         if (lane == 0) {
            for (int a = 0; a < kSegments && got == 0; a++) {
               const uint32_t seg = (xcd + a) % kSegments;
               const uint32_t lo = (uint32_t)((uint64_t)n * seg / kSegments);
               const uint32_t hi = (uint32_t)((uint64_t)n * (seg + 1) / kSegments);
               if (__hip_atomic_load(&work[seg], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= hi - lo) continue;
               const uint32_t b = atomicAdd(&work[seg], n_idle);
               if (b < hi - lo) { base = lo + b; got = min(n_idle, hi - lo - b); }
            }
         }
         base = __shfl(base, 0);
         got = __shfl(got, 0);
         if (got == 0) exhausted = true;
         if (!active) {
            const uint32_t slot = __popcll(idle & ((1ull << lane) - 1ull));
            if (slot < got) {
               id = base + slot;
               const float *rp = rays + 6 * (size_t)id;
               make_ray(s.r, rp[0], rp[1], rp[2], rp[3], rp[4], rp[5]);
               begin(s, sc);
               active = true;
            }
         }
      }
      if (!__any(active)) break;
      if (active) {
         // run until this lane finishes or enough lanes of the wave are idle
         while (true) {
            if (!step(s, sc, st, false)) {
               write_result(s, sc, id, hits, counters, frames);
               active = false;
            }
            // lanes outside this loop are idle; count the ones still active in it
            const uint32_t idle_now = 64u - (uint32_t)__popcll(__ballot(active));
            if (!active || (!exhausted && idle_now >= refill_min)) break;
         }
      }
   }
}

__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
   const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) p[i] = 0u;
}

}  // namespace dev

// Host launcher.  work: >= 8 u32 of device memory; spill sized by the caller
// for (max_stack - lds) * grid * 256 entries.
int launch_closest_hit(const DevScene &sc, const float *d_rays, uint32_t n, ptx_hit *d_hits,
                       ptx_ray_counters *d_counters, ptx_hit_frame *d_frames, uint32_t *d_work, uint2 *d_spill,
                       uint32_t grid, uint32_t lds_depth, uint32_t refill_min, hipStream_t stream) {
   hipLaunchKernelGGL(dev::k_zero_u32, dim3(1), dim3(64), 0, stream, d_work, (uint32_t)dev::kSegments);
   if (n == 0) return 0;
   switch (lds_depth) {
      case 8:
         hipLaunchKernelGGL(dev::k_closest_hit<8>, dim3(grid), dim3(256), 0, stream, sc, d_rays, n, d_hits,
                            d_counters, d_frames, d_work, d_spill, refill_min);
         break;
      case 16:
         hipLaunchKernelGGL(dev::k_closest_hit<16>, dim3(grid), dim3(256), 0, stream, sc, d_rays, n, d_hits,
                            d_counters, d_frames, d_work, d_spill, refill_min);
         break;
      case 24:
         hipLaunchKernelGGL(dev::k_closest_hit<24>, dim3(grid), dim3(256), 0, stream, sc, d_rays, n, d_hits,
                            d_counters, d_frames, d_work, d_spill, refill_min);
         break;
      default:
         hipLaunchKernelGGL(dev::k_closest_hit<12>, dim3(grid), dim3(256), 0, stream, sc, d_rays, n, d_hits,
                            d_counters, d_frames, d_work, d_spill, refill_min);
         break;
   }
   return hipGetLastError() == hipSuccess ? 0 : -1;
}

int closest_hit_blocks_per_cu(uint32_t lds_depth) {
   int blocks = 0;
   hipError_t e;
   switch (lds_depth) {
      case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, dev::k_closest_hit<8>, 256, 0); break;
      case 16: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, dev::k_closest_hit<16>, 256, 0); break;
      case 24: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, dev::k_closest_hit<24>, 256, 0); break;
      default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, dev::k_closest_hit<12>, 256, 0); break;
   }
   return e == hipSuccess && blocks > 0 ? blocks : 1;
}

}  // namespace ptx
