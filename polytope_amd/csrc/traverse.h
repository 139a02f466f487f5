// Persistent closest-hit traversal shared by the ray-query kernel and the
// render path's extend stage.
//
// One lane per ray in flight.  Each loop iteration is one step for every lane
// of the wave: lanes that finished take new work ids from a per-wave pool,
// then every lane issues the loads of its step (a new ray, an interior record
// or up to two leaf faces) before any lane computes, so the wave waits on
// memory once per step.  The deferred-child stack lives in LDS (kLds entries
// per lane), deeper entries spill to a global array sized from the tree depth.
#pragma once

#include <type_traits>

#include "trace_device.h"

#pragma clang fp contract(off)

namespace ptx {
namespace dev {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

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
         spill[(size_t)(k - kLds) * stride + gtid] = e;
      }
   }
   __device__ __forceinline__ void get(uint32_t k, uint32_t &r, float &nr, uint32_t &c) const {
      if (k < (uint32_t)kLds) {
         r = ref[k * 256 + tid];
         nr = near[k * 256 + tid];
         c = cnt[k * 256 + tid];
      } else if (k < max_stack) {
         const u32x4 e = spill[(size_t)(k - kLds) * stride + gtid];
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
      if (nr <= s.min_t) {
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

// Starts a traversal of s.r from the root.  tmax < inf bounds it (shadow rays:
// boxes and faces beyond tmax are culled as a closer hit would cull them).
__device__ __forceinline__ void begin(Trav &s, const float *root_box, uint32_t root_ref, uint32_t root_count,
                                      float tmax = kInf) {
   s.min_t = tmax;
   s.hit = 0xFFFFFFFFu;
   s.after = 0;
   s.sp = 0;
   s.bb_hits = s.bb_misses = s.tris = 0;
   s.hu = s.hv = s.hw = 0.f;
   float near, far;
   slab(root_box[0], root_box[1], root_box[2], root_box[3], root_box[4], root_box[5], s.r, near, far);
   if (near <= far && near <= s.min_t) {
      s.bb_hits++;
      s.cur = root_ref;
      s.cnt = root_count;
   } else {
      s.bb_misses++;
      s.cur = 0xFFFFFFFFu;
      s.cnt = 0;
   }
}

// Memory operands of one traversal step, in one block of registers whatever
// their kind: an interior record (q[0..11] child boxes, q[12..15] refs, counts,
// axis), up to two leaf faces (q[0..8], q[9..17]) or a new ray (q[0..5]).
struct Fetched {
   float q[18];
};

// 16- and 8-byte loads at 4-byte alignment (faces are 36 bytes apart).
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));
typedef float f32x2u __attribute__((ext_vector_type(2), aligned(4)));

// Pointers read from memory (scene tables, LDS) are generic; loads through
// them would be flat loads, which also count against the LDS counter and so
// make every later LDS access wait for them.  Scene arrays live in device
// memory, so they are read through global-address-space pointers.
#define PTX_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const PTX_GLOBAL T *as_global(const T *p) {
   return (const PTX_GLOBAL T *)p;
}

// One load shape for both kinds of node: 64 bytes from a per-lane address (the
// interior record, or the leaf's first face and the start of its second), plus
// 8 bytes for a leaf's second face.  The operands land in the same registers
// whatever the kind, so no lane waits on its loads before the compute phase.
// A one-face leaf reads 28 bytes past its face: the face array carries 128
// bytes of padding.
__device__ __forceinline__ void fetch(const Trav &s, const Node64 *nodes, const float *tris, Fetched &f) {
   const bool leaf = (s.cur & kLeafBit) != 0;
   const PTX_GLOBAL float *p = as_global(leaf ? tris + 9 * (size_t)(s.cur & ~kLeafBit)
                                              : reinterpret_cast<const float *>(nodes + s.cur));
   const f32x4u a = *reinterpret_cast<const PTX_GLOBAL f32x4u *>(p);
   const f32x4u b = *reinterpret_cast<const PTX_GLOBAL f32x4u *>(p + 4);
   const f32x4u c = *reinterpret_cast<const PTX_GLOBAL f32x4u *>(p + 8);
   const f32x4u d = *reinterpret_cast<const PTX_GLOBAL f32x4u *>(p + 12);
   f.q[0] = a.x; f.q[1] = a.y; f.q[2] = a.z; f.q[3] = a.w;
   f.q[4] = b.x; f.q[5] = b.y; f.q[6] = b.z; f.q[7] = b.w;
   f.q[8] = c.x; f.q[9] = c.y; f.q[10] = c.z; f.q[11] = c.w;
   f.q[12] = d.x; f.q[13] = d.y; f.q[14] = d.z; f.q[15] = d.w;
   if (leaf && s.cnt > 1) {
      const f32x2u e = *reinterpret_cast<const PTX_GLOBAL f32x2u *>(p + 16);
      f.q[16] = e.x; f.q[17] = e.y;
   }
}

// Operands of a step over window records: a 128-byte window (the node's record
// and its first child's) or 128 bytes from a leaf's first face (three faces
// and part of a fourth).
struct FetchedWin {
   float q[32];
};

// Window records: four Node64 per interior node i, at win[4 i + 2 w]: the
// node's record, then the record of the child the traversal visits first when
// the ray's direction sign on the node's split axis is w (the high child when
// w = 1, bvh.cpp:510-517), or zeros when that child is a leaf.  The count slot
// of an interior child holds the child's split axis (in the Node64 records
// too), so a lane that arrives at an interior node (from its parent, the stack
// or the root) knows which window to load before it loads anything, and a
// traversal can move between the two record forms at any step.  Each window is
// one 128-byte aligned block, one L2 line.  Every lane issues the same eight
// 16-byte loads from its own address whatever its node's kind, so no lane
// waits on its loads before the compute phase; the face array's 128 bytes of
// padding cover the 92 bytes a one-face leaf at its end reads past its face.
__device__ __forceinline__ void fetch_window(const Trav &s, const Node64 *win, const float *tris, FetchedWin &f) {
   const bool leaf = (s.cur & kLeafBit) != 0;
   const PTX_GLOBAL float *p =
         as_global(leaf ? tris + 9 * (size_t)(s.cur & ~kLeafBit)
                        : reinterpret_cast<const float *>(win + 4 * (size_t)s.cur + 2 * ((s.r.nk >> s.cnt) & 1u)));
#pragma unroll
   for (int k = 0; k < 8; k++) {
      const f32x4u v = *reinterpret_cast<const PTX_GLOBAL f32x4u *>(p + 4 * k);
      f.q[4 * k] = v.x; f.q[4 * k + 1] = v.y; f.q[4 * k + 2] = v.z; f.q[4 * k + 3] = v.w;
   }
}

enum : int { kStepDone = 0, kStepFirst = 1, kStepOther = 2 };

// The two children of the interior node s.cur from its record q[0..15]: the
// child visited first is entered when its box is hit, the other is deferred
// with its entry distance (pop-time test) or entered when only it is hit,
// else the stack is popped.  Returns kStepFirst when the lane entered the
// first child, kStepDone when the traversal has finished, else kStepOther.
template <int kLds>
__device__ __forceinline__ int interior_step(Trav &s, const float *q, const Stack<kLds> &st) {
   float n0, f0, n1, f1;
   slab(q[0], q[1], q[2], q[3], q[4], q[5], s.r, n0, f0);
   slab(q[6], q[7], q[8], q[9], q[10], q[11], s.r, n1, f1);
   const uint32_t mref0 = __float_as_uint(q[12]), mref1 = __float_as_uint(q[13]);
   const uint32_t mcounts = __float_as_uint(q[14]), maxis = __float_as_uint(q[15]);
   const bool h0 = n0 <= f0 && n0 <= s.min_t;
   const bool h1 = n1 <= f1 && n1 <= s.min_t;
   // bvh.cpp:510-517: with neg_dir[axis] the high child (child 1) is popped first.
   const bool high_first = (s.r.nk >> maxis) & 1u;
   const bool fh = high_first ? h1 : h0, sh = high_first ? h0 : h1;
   const uint32_t c0 = mcounts & 0xFFFFu, c1 = mcounts >> 16;
   const uint32_t fref = high_first ? mref1 : mref0, sref = high_first ? mref0 : mref1;
   const uint32_t fcnt = high_first ? c1 : c0, scnt = high_first ? c0 : c1;
   const float snear = high_first ? n0 : n1;
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
      return kStepFirst;
   }
   s.bb_misses++;
   if (sh) {
      s.bb_hits++;
      s.cur = sref;
      s.cnt = scnt;
      return kStepOther;
   }
   s.bb_misses++;
   return pop_next(s, st) ? kStepOther : kStepDone;
}

// Up to kMax faces of the leaf s.cur (face k at q[9 k .. 9 k + 8]), in order.
// Returns false when the traversal has finished.
template <int kLds, int kMax>
__device__ __forceinline__ bool leaf_step(Trav &s, const float *q, const Stack<kLds> &st, bool debug_pixel) {
   const uint32_t first = s.cur & ~kLeafBit;
   const uint32_t n = s.cnt < (uint32_t)kMax ? s.cnt : (uint32_t)kMax;
#pragma unroll
   for (int k = 0; k < kMax; k++) {
      if (k == 0 || (uint32_t)k < n) {
         if (test_triangle(q + 9 * k, s.r, s.min_t, s.hu, s.hv, s.hw, debug_pixel)) {
            s.hit = first + k;
            s.after = 0;
         }
         s.after++;
      }
   }
   s.tris += n;
   if (s.cnt > n) {               // more faces in this leaf: continue it next step
      s.cur += n;
      s.cnt -= n;
      return true;
   }
   return pop_next(s, st);
}

// One step over a window: the node's two children, then, when the lane
// entered the first child and it is interior, that child's two children from
// the window's second record; at a leaf, up to three faces.  The same tests in
// the same order as the advance() steps they replace, so hits, counters and
// stack contents are the same.
template <int kLds>
__device__ __forceinline__ bool advance_window(Trav &s, const FetchedWin &f, const Stack<kLds> &st,
                                               bool debug_pixel) {
   if (s.cur & kLeafBit) return leaf_step<kLds, 3>(s, f.q, st, debug_pixel);
   int r = interior_step(s, f.q, st);
   if (r == kStepFirst && !(s.cur & kLeafBit)) r = interior_step(s, f.q + 16, st);
   return r != kStepDone;
}

// One step on fetched operands: test up to two faces of the leaf s.cur, or the
// two children of the interior node s.cur.  Returns false when the traversal
// has finished.
template <int kLds>
__device__ __forceinline__ bool advance(Trav &s, const Fetched &f, const Stack<kLds> &st, bool debug_pixel) {
   if (s.cur & kLeafBit) return leaf_step<kLds, 2>(s, f.q, st, debug_pixel);
   return interior_step(s, f.q, st) != kStepDone;
}

constexpr uint32_t kBatch = 64;        // ids handed to a wave at a time
constexpr uint32_t kWorkWords = 8;     // u32 words of work counters a launch clears

// Per-wave source of work ids, in batches of 64 consecutive ids.  The first
// three quarters of the batches are dealt out without atomics: wave w of W
// takes batches w, w + W, w + 2W, ..., so every wave samples the whole id
// range and a coherent batch (neighbouring pixels) still lands on one wave.
// The last quarter is handed out by one atomic counter (work[0]), which evens
// out what the static share leaves uneven; lane 0 requests the next dynamic
// batch while the current one is handed out (and the first one while the
// wave takes its last static batch), so the atomic's latency overlaps
// traversal steps.  When the counter passes the end, the wave is done: no
// wave ever waits on more than one atomic at the end of a launch.
struct RaySource {
   uint32_t n;                    // ids
   uint32_t batches, static_batches;
   uint32_t waves;                // waves in the grid
   uint32_t next;                 // this wave's next static batch
   uint32_t pool_base, pool_left;
   uint32_t prefetch;             // pending dynamic request's result (lane 0)
   bool done;
   __device__ __forceinline__ void init(uint32_t n_ids, uint32_t *work, uint32_t lane) {
      n = n_ids;
      batches = (n + kBatch - 1) / kBatch;
      static_batches = batches - batches / 4;
      waves = gridDim.x * (blockDim.x / 64u);
      next = (blockIdx.x * blockDim.x + threadIdx.x) / 64u;
      pool_base = pool_left = 0;
      prefetch = 0;
      done = n == 0;
      if (!done && next >= static_batches) request(work, lane);
   }
   __device__ __forceinline__ void request(uint32_t *work, uint32_t lane) {
      if (lane == 0) prefetch = atomicAdd(&work[0], 1u);
   }
   // Makes pool_left > 0 unless every batch is taken (then done = true).
   __device__ __forceinline__ void refill(uint32_t *work, uint32_t lane) {
      if (pool_left != 0 || done) return;
      uint32_t b;
      if (next < static_batches) {
         b = next;
         next += waves;
         if (next >= static_batches) request(work, lane);     // the first dynamic batch, ahead of time
      } else {
         b = static_batches + __builtin_amdgcn_readlane(prefetch, 0);
         if (b >= batches) {
            done = true;
            return;
         }
         request(work, lane);
      }
      pool_base = b * kBatch;
      pool_left = min(kBatch, n - pool_base);
   }
};

enum : uint32_t { kIdle = 0, kPending = 1, kStarting = 2, kActive = 3, kHeld = 4 };

// Reads v from lane `from` of the wave (every lane active).
__device__ __forceinline__ uint32_t lane_read(uint32_t from, uint32_t v) {
   return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(from << 2), (int)v);
}
__device__ __forceinline__ float lane_read(uint32_t from, float v) {
   return __uint_as_float(lane_read(from, __float_as_uint(v)));
}

// Policies whose rays start from their id alone (no per-lane context, no
// second load) set kStaged: their ray set-up is then done wave-wide.
template <class P, class = void>
struct staged_policy {
   static constexpr bool value = false;
};
template <class P>
struct staged_policy<P, decltype(void(P::kStaged))> {
   static constexpr bool value = P::kStaged;
};

// Policies that set kWindows step over window records (fetch_window,
// advance_window, from windows()) once the ray queue is drained, and over
// node records before; they start at the root with s.cnt = the root's split
// axis when the root is interior (root_count()).
template <class P, class = void>
struct window_policy {
   static constexpr bool value = false;
};
template <class P>
struct window_policy<P, decltype(void(P::kWindows))> {
   static constexpr bool value = P::kWindows;
};

// Ray set-up (reciprocal direction, Wald shear, root box test: six correctly
// rounded divisions and a slab test) done for a whole batch of ids at once:
// lane j of the wave loads and prepares ray base + j, and lanes that need a
// ray later copy a prepared one from its lane.  Set up per refilling lane it
// would run for the whole wave at almost every step, since some lane starts a
// ray at almost every step.
struct Staged {
   float ox, oy, oz, dx, dy, dz, ix, iy, iz, Sx, Sy, Sz;
   uint32_t nk;
   uint32_t cur;                      // root reference, or kNone when the root box is missed
   uint32_t base, left, off;          // wave-uniform: batch base id, rays not yet handed out, next one
};

// Runs the persistent loop.  Policy supplies:
//   Ctx                                   per-lane work context
//   load(id, f)                           issue the loads of work item id into f.q
//   start(id, f, ctx, s) -> int           set s.r and ctx from the loaded item; return
//                                         0 = ready (uniform scene roots), 1 = needs a
//                                         second load (load2), 2 = done without tracing
//   load2(ctx, f) / start2(ctx, f, s)     optional second load (per-lane scene roots)
//   begin_root(ctx, s)                    root box test for uniform-scene policies
//   nodes(ctx), tris(ctx)                 the lane's scene arrays
//   debug_pixel(ctx)                      reference debug-pixel quirk flag
//   finish(id, ctx, s)                    write the result
// A leaf step (up to two Wald tests) costs several interior steps and runs for
// the whole wave when any lane takes it, so lanes whose next node is a leaf
// hold until at least leaf_min of them wait or no other lane can step; every
// lane still performs its own steps in its own order.
template <int kLds, class Policy>
__device__ __forceinline__ void persistent_trace(const Policy &P, uint32_t n, uint32_t *work, uint4 *spill,
                                                 uint32_t refill_min, uint32_t leaf_min, uint32_t max_stack) {
   __shared__ uint32_t lds_ref[kLds * 256];
   __shared__ float lds_near[kLds * 256];
   __shared__ uint16_t lds_cnt[kLds * 256];
   const uint32_t lane = __lane_id();
   const Stack<kLds> st{lds_ref, lds_near, lds_cnt, reinterpret_cast<u32x4 *>(spill), threadIdx.x,
                        blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, max_stack};
   RaySource src;
   src.init(n, work, lane);
   Trav s;
   typename Policy::Ctx ctx;
   uint32_t mode = kIdle;
   uint32_t id = 0;
   refill_min = min(refill_min, 64u);     // a wave that is all idle always refills
   constexpr bool kStaged = staged_policy<Policy>::value;
   constexpr bool kWin = window_policy<Policy>::value;
   static_assert(!kStaged || std::is_empty<typename Policy::Ctx>::value, "staged policies carry no context");
   static_assert(!kWin || kStaged, "window records are stepped by staged policies only");
   Staged g;
   g.ox = g.oy = g.oz = g.dx = g.dy = g.dz = g.ix = g.iy = g.iz = g.Sx = g.Sy = g.Sz = 0.f;
   g.nk = 0;
   g.cur = 0xFFFFFFFFu;
   g.left = 0;
   g.base = g.off = 0;
   while (true) {
      uint64_t idle = __ballot(mode == kIdle);
      if constexpr (kStaged) {
         if (idle && !(src.done && g.left == 0) && (uint32_t)__popcll(idle) >= refill_min) {
            while (idle) {
               if (g.left == 0) {
                  src.refill(work, lane);
                  if (src.done) break;
                  g.base = src.pool_base;
                  g.left = src.pool_left;
                  g.off = 0;
                  src.pool_left = 0;
                  if (lane < g.left) {
                     Fetched f;
                     Trav t;
                     typename Policy::Ctx c;
                     P.load(g.base + lane, f);
                     P.start(g.base + lane, f, c, t);
                     P.begin_root(c, t);
                     g.ox = t.r.ox; g.oy = t.r.oy; g.oz = t.r.oz; g.dx = t.r.dx; g.dy = t.r.dy; g.dz = t.r.dz;
                     g.ix = t.r.ix; g.iy = t.r.iy; g.iz = t.r.iz; g.Sx = t.r.Sx; g.Sy = t.r.Sy; g.Sz = t.r.Sz;
                     g.nk = t.r.nk;
                     g.cur = t.cur;
                  }
               }
               const uint32_t take = min((uint32_t)__popcll(idle), g.left);
               const uint32_t rank = __popcll(idle & ((1ull << lane) - 1ull));
               const bool mine = ((idle >> lane) & 1ull) && rank < take;
               const uint32_t from = g.off + min(rank, take - 1u);
               Ray r;
               r.ox = lane_read(from, g.ox); r.oy = lane_read(from, g.oy); r.oz = lane_read(from, g.oz);
               r.dx = lane_read(from, g.dx); r.dy = lane_read(from, g.dy); r.dz = lane_read(from, g.dz);
               r.ix = lane_read(from, g.ix); r.iy = lane_read(from, g.iy); r.iz = lane_read(from, g.iz);
               r.Sx = lane_read(from, g.Sx); r.Sy = lane_read(from, g.Sy); r.Sz = lane_read(from, g.Sz);
               const uint32_t nk = lane_read(from, g.nk), cur = lane_read(from, g.cur);
               r.nk = nk;
               if (mine) {
                  // the state begin() leaves, with the root test already done
                  id = g.base + from;
                  s.r = r;
                  s.min_t = kInf;
                  s.hit = 0xFFFFFFFFu;
                  s.after = 0;
                  s.sp = 0;
                  s.tris = 0;
                  s.hu = s.hv = s.hw = 0.f;
                  s.cur = cur;
                  s.cnt = cur != 0xFFFFFFFFu ? P.root_count() : 0u;
                  s.bb_hits = cur != 0xFFFFFFFFu ? 1u : 0u;
                  s.bb_misses = 1u - s.bb_hits;
                  mode = kActive;
                  if (cur == 0xFFFFFFFFu) {
                     P.finish(id, ctx, s);
                     mode = kIdle;
                  }
               }
               idle &= ~__ballot(mine);
               g.off += take;
               g.left -= take;
            }
         }
      } else {
         if (idle && !src.done && (uint32_t)__popcll(idle) >= refill_min) {
            while (idle) {
               src.refill(work, lane);
               if (src.done) break;
               const uint32_t take = min((uint32_t)__popcll(idle), src.pool_left);
               const uint32_t rank = __popcll(idle & ((1ull << lane) - 1ull));
               const bool mine = ((idle >> lane) & 1ull) && rank < take;
               if (mine) {
                  id = src.pool_base + rank;
                  mode = kPending;
               }
               idle &= ~__ballot(mine);
               src.pool_base += take;
               src.pool_left -= take;
            }
         }
      }
      if (!__any(mode != kIdle)) {
         // staged lanes whose ray missed the root finished at once: more may be left
         if (kStaged && !(src.done && g.left == 0)) continue;
         break;
      }
      const bool at_leaf = mode == kActive && (s.cur & kLeafBit) != 0;
      const bool hold = at_leaf && (uint32_t)__popcll(__ballot(at_leaf)) < leaf_min &&
                        __any(mode != kIdle && !at_leaf);
      const uint32_t m = hold ? kHeld : mode;
      if constexpr (kWin) {
         // Once the ray queue is drained the wave only finishes the rays it
         // holds, the longest ones of the launch, and each step is a wait on
         // one dependent load: from then on it steps over window records, two
         // levels per load.  While rays remain, node records cost less per
         // step.
         if (src.done && g.left == 0) {
            FetchedWin f;
            if (m == kActive) fetch_window(s, P.windows(), P.tris(ctx), f);
            if (m == kActive && !advance_window<kLds>(s, f, st, P.debug_pixel(ctx))) {
               P.finish(id, ctx, s);
               mode = kIdle;
            }
            continue;
         }
      }
      Fetched f;
      if (!kStaged && m == kPending) {
         P.load(id, f);
      } else if (!kStaged && m == kStarting) {
         P.load2(ctx, f);
      } else if (m == kActive) {
         fetch(s, P.nodes(ctx), P.tris(ctx), f);
      }
      if (!kStaged && m == kPending) {
         const int r = P.start(id, f, ctx, s);
         if (r == 2) {
            P.finish(id, ctx, s);
            mode = kIdle;
         } else if (r == 1) {
            mode = kStarting;
         } else {
            P.begin_root(ctx, s);
            mode = kActive;
            if (s.cur == 0xFFFFFFFFu) {
               P.finish(id, ctx, s);
               mode = kIdle;
            }
         }
      } else if (!kStaged && m == kStarting) {
         P.start2(ctx, f, s);
         mode = kActive;
         if (s.cur == 0xFFFFFFFFu) {
            P.finish(id, ctx, s);
            mode = kIdle;
         }
      } else if (m == kActive) {
         if (!advance<kLds>(s, f, st, P.debug_pixel(ctx))) {
            P.finish(id, ctx, s);
            mode = kIdle;
         }
      }
   }
}

}  // namespace dev
}  // namespace ptx
