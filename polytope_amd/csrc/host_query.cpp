// Host-buffer ray queries (ptx_intersect_closest).  The rays go to HBM and the
// hits come back through pinned host slots.  Up to four host threads each
// take a contiguous share of the caller's arrays and copy it chunk by chunk
// into their two slots while the DMA engine moves the chunk copied before, so
// the caller's pageable memory is read and written at host-memory speed and
// the link runs at the pinned DMA rate (about 57 GB/s each way on MI355X,
// against 28 GB/s host-to-device and 16 GB/s device-to-host for copies from
// pageable memory).  Chunks are 2 MiB: each copy call costs about 18 us that
// the DMA does not hide, so 1 MiB and 512 KiB chunks were slower.  The kernel
// is the device query's (run_closest_hit), launched once all rays are
// resident.
#include "host_query.h"

#include <sched.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "scene.h"

namespace ptx {

HostQueryStaging::~HostQueryStaging() {
   (void)hipSetDevice(ordinal);
   for (hipEvent_t &e : events)
      if (e) (void)hipEventDestroy(e);
   if (pinned) (void)hipHostFree(pinned);
   if (d_buf) (void)hipFree(d_buf);
}

namespace {

constexpr size_t kSlot = HostQueryStaging::kSlotBytes;

// Host threads for one call: half the cores this process may run on, at most
// kMaxWorkers (four threads copy pageable to pinned memory at 60-75 GB/s,
// past what the link takes; eight gave no more).
uint32_t host_workers() {
   cpu_set_t cs;
   CPU_ZERO(&cs);
   unsigned c = sched_getaffinity(0, sizeof cs, &cs) == 0 ? unsigned(CPU_COUNT(&cs)) : std::thread::hardware_concurrency();
   return std::clamp<uint32_t>(c / 2, 1u, HostQueryStaging::kMaxWorkers);
}

// Worker w's share of the spans: bytes [total * w / workers, total * (w + 1) /
// workers) of their concatenation, cut at the spans' ends and every kSlot
// bytes.
std::vector<Piece> share(const std::vector<Span> &spans, uint32_t w, uint32_t workers) {
   size_t total = 0;
   for (const Span &a : spans) total += a.bytes;
   const size_t lo = total / workers * w + total % workers * w / workers;
   const size_t hi = total / workers * (w + 1) + total % workers * (w + 1) / workers;
   std::vector<Piece> out;
   size_t base = 0;
   for (const Span &a : spans) {
      const size_t b0 = std::max(lo, base), b1 = std::min(hi, base + a.bytes);
      for (size_t off = b0; off < b1; off += kSlot)
         out.push_back(Piece{a.host + (off - base), a.dev + (off - base), std::min(kSlot, b1 - off)});
      base += a.bytes;
   }
   return out;
}

// State one call's workers share.  Phase 0: host spans to the device; the
// calling thread runs `between` (the launch) once every worker has queued
// its copies; phase 1: device spans back to the host; -1: a step failed, the
// rest is skipped.
struct HostCall {
   HostQueryStaging *st;
   hipStream_t stream;
   int ordinal;
   uint32_t workers = 1;
   std::vector<Span> in, out;
   std::mutex mu;
   std::condition_variable cv;
   uint32_t arrived = 0;
   int phase = 0;
   std::string error;

   void fail(const std::string &m) {
      std::lock_guard<std::mutex> lock(mu);
      if (error.empty()) error = m;
   }
   bool ok(hipError_t e, const char *what) {
      if (e == hipSuccess) return true;
      (void)hipGetLastError();
      fail(std::string(what) + ": " + hipGetErrorString(e));
      return false;
   }
   char *slot(uint32_t w, uint32_t k) const { return st->pinned + (2 * size_t(w) + (k & 1)) * kSlot; }
   hipEvent_t event(uint32_t w, uint32_t k) const { return st->events[2 * w + (k & 1)]; }

   // Worker w's share going up: copy a piece into a slot, then queue its DMA;
   // a slot is refilled once the DMA that read it has finished.
   bool send(uint32_t w) {
      const std::vector<Piece> ps = share(in, w, workers);
      for (uint32_t k = 0; k < ps.size(); k++) {
         if (k >= 2 && !ok(hipEventSynchronize(event(w, k)), "hipEventSynchronize")) return false;
         std::memcpy(slot(w, k), ps[k].host, ps[k].len);
         if (!ok(hipMemcpyAsync(ps[k].dev, slot(w, k), ps[k].len, hipMemcpyHostToDevice, stream), "hipMemcpyAsync"))
            return false;
         if (!ok(hipEventRecord(event(w, k), stream), "hipEventRecord")) return false;
      }
      return true;
   }

   // Worker w's share coming back: the DMA of piece k + 1 is queued before
   // piece k is copied out of its slot.
   bool receive(uint32_t w) {
      const std::vector<Piece> ps = share(out, w, workers);
      auto queue = [&](uint32_t k) {
         return ok(hipMemcpyAsync(slot(w, k), ps[k].dev, ps[k].len, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync") &&
                ok(hipEventRecord(event(w, k), stream), "hipEventRecord");
      };
      if (ps.empty()) return true;
      if (!queue(0)) return false;
      for (uint32_t k = 0; k < ps.size(); k++) {
         if (k + 1 < ps.size() && !queue(k + 1)) return false;
         if (!ok(hipEventSynchronize(event(w, k)), "hipEventSynchronize")) return false;
         std::memcpy(ps[k].host, slot(w, k), ps[k].len);
      }
      return true;
   }

   void worker(uint32_t w) {
      bool good = ok(hipSetDevice(ordinal), "hipSetDevice") && send(w);
      std::unique_lock<std::mutex> lock(mu);
      if (!good && error.empty()) error = "host copy failed";
      arrived++;
      cv.notify_all();
      cv.wait(lock, [&] { return phase != 0; });
      if (phase < 0) return;
      lock.unlock();
      receive(w);
   }

   // The whole call on the calling thread and up to `workers` threads.
   template <class F>
   int run(F between) {
      size_t in_bytes = 0, out_bytes = 0;
      for (const Span &a : in) in_bytes += a.bytes;
      for (const Span &a : out) out_bytes += a.bytes;
      const size_t pieces = (std::max(in_bytes, out_bytes) + kSlot - 1) / kSlot;
      workers = uint32_t(std::min<size_t>(host_workers(), std::max<size_t>(1, pieces)));
      int rc = PTX_OK;
      if (workers == 1) {
         if (send(0)) {
            rc = between();
            if (rc == PTX_OK) receive(0);
         }
      } else {
         std::vector<std::thread> ts;
         ts.reserve(workers);
         try {
            for (uint32_t w = 0; w < workers; w++) ts.emplace_back([this, w] { worker(w); });
         } catch (const std::exception &) {
            fail("could not start the host copy threads");
         }
         std::unique_lock<std::mutex> lock(mu);
         // threads that did not start never arrive: count them as arrived
         arrived += workers - uint32_t(ts.size());
         cv.wait(lock, [&] { return arrived == workers; });
         if (error.empty()) {
            lock.unlock();
            rc = between();
            lock.lock();
         }
         phase = (rc == PTX_OK && error.empty()) ? 1 : -1;
         cv.notify_all();
         lock.unlock();
         for (std::thread &th : ts) th.join();
      }
      // every queued copy has finished before the slots serve another call
      const hipError_t se = hipStreamSynchronize(stream);
      if (rc != PTX_OK) return rc;
      if (!error.empty()) return set_error(PTX_ERR_HIP, error);
      if (se != hipSuccess) {
         (void)hipGetLastError();
         return set_error(PTX_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(se));
      }
      return PTX_OK;
   }
};

int ensure_staging(ptx_device *dev, size_t need) {
   if (!dev->host_query) {
      dev->host_query = std::make_unique<HostQueryStaging>();
      dev->host_query->ordinal = dev->ordinal;
   }
   HostQueryStaging &st = *dev->host_query;
   if (!st.ready) {
      // what an earlier failed call made is kept and not made again
      if (!st.pinned)
         PTX_HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&st.pinned), 2 * HostQueryStaging::kMaxWorkers * kSlot,
                                   hipHostMallocDefault));
      for (hipEvent_t &e : st.events)
         if (!e) PTX_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      st.ready = true;
   }
   if (need > st.d_cap) {
      if (st.d_buf) PTX_HIP_TRY(hipFree(st.d_buf));
      st.d_buf = nullptr;
      st.d_cap = 0;
      PTX_HIP_TRY(hipMalloc(reinterpret_cast<void **>(&st.d_buf), need));
      st.d_cap = need;
   }
   return PTX_OK;
}

size_t round_up(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

int intersect_host(ptx_scene *s, const float *rays, uint32_t n, ptx_hit *hits, ptx_ray_counters *counters) {
   ptx_device *dev = s->dev;
   const size_t b_rays = round_up(size_t(n) * 24), b_hits = round_up(size_t(n) * sizeof(ptx_hit));
   const size_t b_cnt = counters ? round_up(size_t(n) * sizeof(ptx_ray_counters)) : 0;
   if (int rc = ensure_staging(dev, b_rays + b_hits + b_cnt)) return rc;
   HostQueryStaging &st = *dev->host_query;
   HostCall call;
   call.st = &st;
   call.stream = dev->stream;
   call.ordinal = dev->ordinal;
   char *d_hits = st.d_buf + b_rays, *d_cnt = counters ? d_hits + b_hits : nullptr;
   call.in.push_back(Span{reinterpret_cast<char *>(const_cast<float *>(rays)), st.d_buf, size_t(n) * 24});
   call.out.push_back(Span{reinterpret_cast<char *>(hits), d_hits, size_t(n) * sizeof(ptx_hit)});
   if (counters)
      call.out.push_back(Span{reinterpret_cast<char *>(counters), d_cnt, size_t(n) * sizeof(ptx_ray_counters)});
   const ptx_tuning t = tuning_of(dev);
   return call.run([&] {
      return run_closest_hit(s, reinterpret_cast<const float *>(st.d_buf), n, reinterpret_cast<ptx_hit *>(d_hits),
                             reinterpret_cast<ptx_ray_counters *>(d_cnt), nullptr, call.stream, t);
   });
}

int download_host(ptx_device *dev, hipStream_t stream, const std::vector<Span> &spans) {
   if (int rc = ensure_staging(dev, 0)) return rc;
   HostCall call;
   call.st = dev->host_query.get();
   call.stream = stream;
   call.ordinal = dev->ordinal;
   call.out = spans;
   return call.run([] { return int(PTX_OK); });
}

}  // namespace ptx
