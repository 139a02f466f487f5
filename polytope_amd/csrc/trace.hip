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

// Deferred-child stack of one lane: entries [0, kLds) in LDS, deeper ones in a
// global spill array.  The LDS part is three arrays (reference, slab near
// distance, leaf face count) indexed [entry][thread] so a wave's accesses hit
// consecutive banks; spill entries go through non-temporal global accesses
// (and so never share an instruction with the LDS path).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int kLds>
struct Stack {
   uint32_t *ref;      // __shared__ [kLds * 256]
   float *near;        // __shared__ [kLds * 256]
   uint16_t *cnt;      // __shared__ [kLds * 256]
   u32x4 *spill;       // [max_stack - kLds][stride]
   uint32_t tid, gtid, stride, max_stack;
   __device__ __forceinline__ void put(uint32_t k, uint32_t r, float nr, uint32_t c) const {
      if (k < (uint32_t)kLds) {
         ref[k * 256 + tid] = r;
         near[k * 256 + tid] = nr;
         cnt[k * 256 + tid] = (uint16_t)c;
      } else if (k < max_stack) {
         const u32x4 e = {r, __float_as_uint(nr), c, 0u};
         __builtin_nontemporal_store(e, &spill[(size_t)(k - kLds) * stride + gtid]);
      }
   }
   __device__ __forceinline__ void get(uint32_t k, uint32_t &r, float &nr, uint32_t &c) const {
      if (k < (uint32_t)kLds) {
         r = ref[k * 256 + tid];
         nr = near[k * 256 + tid];
         c = cnt[k * 256 + tid];
      } else if (k < max_stack) {
         const u32x4 e = __builtin_nontemporal_load(&spill[(size_t)(k - kLds) * stride + gtid]);
         r = e.x; nr = __uint_as_float(e.y); c = e.z;
      } else {                     // unreachable: the tree depth bounds the stack
         r = 0xFFFFFFFFu; nr = kInf; c = 0;
      }
   }
};

// Pops deferred children until one passes its pop-time test; returns false
// when the stack is empty (traversal done).
template <int kLds>
__device__ __forceinline__ bool pop_next(Trav &s, const Stack<kLds> &st) {
   while (s.sp > 0) {
      s.sp--;
      uint32_t r, c;
      float nr;
      st.get(s.sp, r, nr, c);
      if (nr < 0.f || nr <= s.min_t) {
         s.bb_hits++;
         s.cur = r;
         s.cnt = c;
         return true;
      }
      s.bb_misses++;
   }
   s.cur = 0xFFFFFFFFu;
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
      s.cnt = sc.root_count;
   } else {
      s.bb_misses++;
      s.cur = 0xFFFFFFFFu;
      s.cnt = 0;
   }
}

// One step: visit the interior node in s.cur, or test up to two faces of the
// leaf in s.cur.  Loads for either kind are issued before any arithmetic so a
// wave whose lanes are split between nodes and leaves waits once per step.
// Returns false when the traversal has finished.
template <int kLds>
__device__ __forceinline__ bool step(Trav &s, const DevScene &sc, const Stack<kLds> &st, bool debug_pixel) {
   if (s.cur == 0xFFFFFFFFu) return false;
   const bool leaf = (s.cur & kLeafBit) != 0;
   float4 qa, qb, qc;
   uint4 qm;
   float ta[9], tb[9];
   const uint32_t first = s.cur & ~kLeafBit;
   const bool two = leaf && s.cnt > 1;
   if (!leaf) {
      const float4 *rec = reinterpret_cast<const float4 *>(sc.nodes + s.cur);
      qa = rec[0]; qb = rec[1]; qc = rec[2];
      qm = reinterpret_cast<const uint4 *>(rec)[3];
   } else {
      const float *p = sc.tris + 9 * (size_t)first;
#pragma unroll
      for (int k = 0; k < 9; k++) ta[k] = p[k];
      if (two) {
#pragma unroll
         for (int k = 0; k < 9; k++) tb[k] = p[9 + k];
      }
   }
   if (leaf) {
      if (test_triangle(ta, s.r, s.min_t, s.hu, s.hv, s.hw, debug_pixel)) {
         s.hit = first;
         s.after = 0;
      }
      s.after++;
      if (two) {
         if (test_triangle(tb, s.r, s.min_t, s.hu, s.hv, s.hw, debug_pixel)) {
            s.hit = first + 1;
            s.after = 0;
         }
         s.after++;
      }
      const uint32_t done = two ? 2u : 1u;
      s.tris += done;
      if (s.cnt > done) {            // more faces in this leaf: continue it next step
         s.cur += done;
         s.cnt -= done;
         return true;
      }
      return pop_next(s, st);
   }
   bool in0, in1;
   float n0, f0, n1, f1;
   slab(qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, s.r, in0, n0, f0);
   slab(qb.z, qb.w, qc.x, qc.y, qc.z, qc.w, s.r, in1, n1, f1);
   const bool h0 = in0 || (n0 <= f0 && n0 <= s.min_t);
   const bool h1 = in1 || (n1 <= f1 && n1 <= s.min_t);
   // bvh.cpp:510-517: with neg_dir[axis] the high child (child 1) is popped first.
   const bool high_first = (s.r.neg >> qm.w) & 1u;
   const bool fh = high_first ? h1 : h0, sh = high_first ? h0 : h1;
   const uint32_t c0 = qm.z & 0xFFFFu, c1 = qm.z >> 16;
   const uint32_t fref = high_first ? qm.y : qm.x, sref = high_first ? qm.x : qm.y;
   const uint32_t fcnt = high_first ? c1 : c0, scnt = high_first ? c0 : c1;
   const float snear = high_first ? (in0 ? -1.f : n0) : (in1 ? -1.f : n1);
   if (fh) {
      s.bb_hits++;
      if (sh) {
         st.put(s.sp, sref, snear, scnt);
         s.sp++;
      } else {
         s.bb_misses++;
      }
      s.cur = fref;
      s.cnt = fcnt;
      return true;
   }
   s.bb_misses++;
   if (sh) {
      s.bb_hits++;
      s.cur = sref;
      s.cnt = scnt;
      return true;
   }
   s.bb_misses++;
   return pop_next(s, st);
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
                                                     ptx_hit_frame *frames, uint32_t *work, uint4 *spill,
                                                     uint32_t refill_min) {
   __shared__ uint32_t lds_ref[kLds * 256];
   __shared__ float lds_near[kLds * 256];
   __shared__ uint16_t lds_cnt[kLds * 256];
   const uint32_t lane = __lane_id();
   const Stack<kLds> st{lds_ref, lds_near, lds_cnt, reinterpret_cast<u32x4 *>(spill), threadIdx.x, blockIdx.x * blockDim.x + threadIdx.x,
                        gridDim.x * blockDim.x, sc.max_stack};
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
            if (!step<kLds>(s, sc, st, false)) {
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
                       ptx_ray_counters *d_counters, ptx_hit_frame *d_frames, uint32_t *d_work, uint4 *d_spill,
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
