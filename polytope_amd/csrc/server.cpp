// Continuous-batching render service (ptx_server_*).
//
// A service thread owns a RenderEngine on its own stream.  Jobs are queued by
// ptx_server_submit and admitted in arrival order when a job-table slot is
// free and the job's device memory (film + primary-hit records) fits the HBM
// budget next to the engine's path pool.  Admitted jobs trace their primary
// hits, then take part in rounds: every round gives the jobs in flight one
// chunk of samples each (their share of the path capacity), least recently
// served first, runs the bounces, and adds each chunk to its job's film in
// sample order.  Jobs that arrive while a round runs join the next round.
//
// The host stays one round ahead of the device: round N+1 is scheduled and
// enqueued while round N runs (two chunk tables, each reused only after the
// event of the round that read it), and the thread waits on the older round's
// event, never on the whole stream.  A job whose last chunk has been enqueued
// has its film finalised and copied into pinned host memory behind that
// round; it is done when that copy's event completes, and its device buffers
// go back in stream order.  A job's film depends only on its own samples and
// their order, so it is the same alone and in any mix.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "launch.h"
#include "render_host.h"

namespace ptx {
namespace {

constexpr uint32_t kMaxServerJobs = 256;
constexpr uint32_t kMaxServerScenes = 64;
constexpr uint32_t kDefaultServerRegenPaths = 1u << 27;   // path slots per round (12 B each), at most
constexpr uint32_t kMinServerRegenPaths = 1u << 16;       // ... and at least, when the budget is small
constexpr int kRoundsAhead = 2;                             // rounds enqueued at once (chunk tables)

struct JobState {
   uint64_t id = 0;
   ptx_job_desc desc{};
   ptx_scene *scene = nullptr;   // holds one scene reference from submit until released
   uint32_t npix = 0;
   int32_t state = PTX_JOB_QUEUED;
   int32_t error = PTX_OK;
   double submit = 0, start = 0, end = 0;
   uint64_t samples_done = 0;    // samples of rounds that have completed
   uint32_t next_sample = 0;     // first sample not yet scheduled
   uint64_t last_round = 0;      // round that last gave the job a chunk (0 = none yet)
   uint32_t inflight = 0;        // enqueued rounds with a chunk of this job that have not completed
   size_t bytes = 0;
   uint32_t slot = 0, scene_slot = 0;
   bool cancel = false;
   bool started = false;         // device buffers allocated (start_job succeeded)
   bool finishing = false;       // film finalised and copy enqueued; done when `done_ev` completes
   bool retired = false;         // left the server: no device work or memory of it remains
   hipEvent_t done_ev = nullptr;
   JobDev jd{};
   DevScene sd{};                // staged copy for the device scene table
   float *film = nullptr;        // pinned host copy of the film (from the server's pool)
   size_t film_bytes = 0;
};

// One enqueued round: its event and the chunk each job got.
struct RoundInFlight {
   hipEvent_t ev = nullptr;
   int table = 0;
   int rc = PTX_OK;
   std::vector<std::pair<std::shared_ptr<JobState>, uint32_t>> parts;   // (job, samples per pixel)
};

}  // namespace
}  // namespace ptx

struct ptx_server {
   ptx_device *dev = nullptr;
   ptx_server_config cfg{};
   hipStream_t stream = nullptr;
   ptx::RenderEngine eng;
   std::thread worker;
   std::mutex mu;
   std::condition_variable cv_work, cv_done;
   bool stop = false;
   uint64_t next_id = 1;
   std::map<uint64_t, std::shared_ptr<ptx::JobState>> jobs;
   std::deque<uint64_t> queue;
   std::vector<std::shared_ptr<ptx::JobState>> running;
   size_t pool_bytes = 0, used_bytes = 0;
   uint32_t paths = 0;
   uint64_t rounds = 0;          // rounds enqueued so far
   std::vector<bool> slot_used;
   std::map<const ptx_scene *, std::pair<uint32_t, uint32_t>> scene_slots;   // scene -> (slot, refs)
   double t0 = 0;
   // pinned host blocks for finished films, reused across jobs (freed at destroy)
   std::vector<std::pair<float *, size_t>> film_pool;
   std::deque<ptx::RoundInFlight> inflight;
   std::vector<hipEvent_t> free_events;
   ptx::Chunk *chunk_host[ptx::kRoundsAhead] = {};   // pinned chunk tables, one per round in flight

   void loop();
   void fail(const std::shared_ptr<ptx::JobState> &j, int code);
   void release(const std::shared_ptr<ptx::JobState> &j);
   void drop_scene(ptx::JobState &j);
   hipEvent_t take_event();
   float *take_film(size_t bytes);
   void give_film(float *p, size_t bytes);
   int complete_oldest_round();
   void retire_done_jobs();
};

using namespace ptx;

void ptx_server::fail(const std::shared_ptr<JobState> &j, int code) {
   if (j->state == PTX_JOB_FAILED) return;
   j->state = PTX_JOB_FAILED;
   j->error = code;
   j->end = now_s() - t0;
}

void ptx_server::drop_scene(JobState &j) {
   if (j.scene) scene_release(j.scene);
   j.scene = nullptr;
}

// Frees a job's device memory (stream-ordered, behind every round that used
// it), its table slots, budget share and scene reference (caller holds mu).
void ptx_server::release(const std::shared_ptr<JobState> &j) {
   if (j->started) eng.free_job(j->jd);
   j->started = false;
   if (j->bytes) {
      used_bytes -= j->bytes;
      j->bytes = 0;
   }
   slot_used[j->slot] = false;
   auto it = scene_slots.find(j->scene);
   if (it != scene_slots.end() && --it->second.second == 0) scene_slots.erase(it);
   if (j->done_ev) free_events.push_back(j->done_ev);
   j->done_ev = nullptr;
   drop_scene(*j);
}

hipEvent_t ptx_server::take_event() {
   if (!free_events.empty()) {
      hipEvent_t e = free_events.back();
      free_events.pop_back();
      return e;
   }
   hipEvent_t e = nullptr;
   if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
   }
   return e;
}

float *ptx_server::take_film(size_t bytes) {
   size_t best = film_pool.size();
   for (size_t i = 0; i < film_pool.size(); i++)
      if (film_pool[i].second >= bytes && (best == film_pool.size() || film_pool[i].second < film_pool[best].second))
         best = i;
   if (best < film_pool.size()) {
      float *p = film_pool[best].first;
      film_pool.erase(film_pool.begin() + best);
      return p;
   }
   void *p = nullptr;
   if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
   }
   return static_cast<float *>(p);
}

void ptx_server::give_film(float *p, size_t bytes) {
   if (p) film_pool.emplace_back(p, bytes);
}

// Waits for the oldest enqueued round and credits its samples.
int ptx_server::complete_oldest_round() {
   RoundInFlight r = std::move(inflight.front());
   inflight.pop_front();
   hipError_t e = hipEventSynchronize(r.ev);
   if (e != hipSuccess) (void)hipGetLastError();
   std::lock_guard<std::mutex> lock(mu);
   free_events.push_back(r.ev);
   for (auto &pj : r.parts) {
      JobState &j = *pj.first;
      j.inflight--;
      j.samples_done += (uint64_t)pj.second * j.npix;
      if (e != hipSuccess || r.rc != PTX_OK) fail(pj.first, r.rc != PTX_OK ? r.rc : PTX_ERR_HIP);
   }
   return e == hipSuccess ? r.rc : PTX_ERR_HIP;
}

// Jobs whose film copy has landed become done; failed and cancelled jobs
// leave once no enqueued round uses them.  A done job's copy event comes after
// all of its device work; a job that ends otherwise may still have work on the
// stream (its primary hits), so the stream is drained before it is released.
void ptx_server::retire_done_jobs() {
   std::vector<std::shared_ptr<JobState>> gone;
   bool drain = false;
   {
      std::lock_guard<std::mutex> lock(mu);
      for (auto &j : running) {
         if (j->inflight) continue;
         if (j->finishing) {
            const hipError_t q = hipEventQuery(j->done_ev);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
               (void)hipGetLastError();
               fail(j, PTX_ERR_HIP);
            } else if (j->state == PTX_JOB_RUNNING && !j->cancel) {
               j->state = PTX_JOB_DONE;
               j->end = now_s() - t0;
            }
         }
         if (j->cancel && j->state == PTX_JOB_RUNNING) {
            j->state = PTX_JOB_CANCELLED;
            j->end = now_s() - t0;
         }
         if (j->state != PTX_JOB_RUNNING) {
            gone.push_back(j);
            drain |= j->state != PTX_JOB_DONE && j->started;
         }
      }
   }
   if (gone.empty()) return;
   if (drain && hipStreamSynchronize(stream) != hipSuccess) (void)hipGetLastError();
   std::lock_guard<std::mutex> lock(mu);
   for (auto &j : gone) {
      if (j->state != PTX_JOB_DONE) {
         give_film(j->film, j->film_bytes);
         j->film = nullptr;
      }
      release(j);
      j->retired = true;
      running.erase(std::remove(running.begin(), running.end(), j), running.end());
   }
   cv_done.notify_all();
}

void ptx_server::loop() {
   (void)hipSetDevice(dev->ordinal);
   std::vector<Chunk> chunks;
   while (true) {
      std::vector<std::shared_ptr<JobState>> admitted;
      {
         std::unique_lock<std::mutex> lock(mu);
         cv_work.wait(lock, [&] { return stop || !queue.empty() || !running.empty() || !inflight.empty(); });
         if (stop) break;
         // admission in arrival order: a free slot, a scene slot and room in the budget
         while (!queue.empty()) {
            auto j = jobs[queue.front()];
            if (j->cancel) {
               queue.pop_front();
               j->state = PTX_JOB_CANCELLED;
               j->end = now_s() - t0;
               drop_scene(*j);
               j->retired = true;
               cv_done.notify_all();
               continue;
            }
            const size_t need = job_device_bytes(j->desc);
            if (cfg.max_jobs_in_flight && running.size() + admitted.size() >= cfg.max_jobs_in_flight) break;
            if (used_bytes + need > cfg.hbm_budget_bytes - pool_bytes) break;
            auto sit = scene_slots.find(j->scene);
            if (sit == scene_slots.end() && scene_slots.size() >= kMaxServerScenes) break;
            uint32_t slot = 0;
            while (slot < kMaxServerJobs && slot_used[slot]) slot++;
            if (slot == kMaxServerJobs) break;
            queue.pop_front();
            slot_used[slot] = true;
            if (sit == scene_slots.end()) {
               std::vector<bool> taken(kMaxServerScenes, false);
               for (auto &kv : scene_slots) taken[kv.second.first] = true;
               uint32_t ss = 0;
               while (taken[ss]) ss++;
               sit = scene_slots.emplace(j->scene, std::make_pair(ss, 0u)).first;
            }
            sit->second.second++;
            j->slot = slot;
            j->scene_slot = sit->second.first;
            j->bytes = need;
            used_bytes += need;
            j->state = PTX_JOB_RUNNING;
            j->start = now_s() - t0;
            admitted.push_back(j);
         }
      }
      // start admitted jobs: tables, film, primary hits (stream-ordered behind
      // the rounds already enqueued, which never read these table entries)
      for (auto &j : admitted) {
         j->jd = job_dev(j->desc, j->scene_slot);
         int rc = eng.start_job(j->jd, j->scene->d, j->npix);
         if (rc == PTX_OK) {
            j->started = true;
            j->sd = j->scene->d;
            if (hipMemcpyAsync(eng.d_scenes + j->scene_slot, &j->sd, sizeof(DevScene), hipMemcpyHostToDevice, stream) !=
                      hipSuccess ||
                hipMemcpyAsync(eng.d_jobs + j->slot, &j->jd, sizeof(JobDev), hipMemcpyHostToDevice, stream) !=
                      hipSuccess) {
               (void)hipGetLastError();
               rc = set_error(PTX_ERR_HIP, "server: job table upload failed");
            }
         }
         std::lock_guard<std::mutex> lock(mu);
         if (rc != PTX_OK) fail(j, rc);
         running.push_back(j);            // a failed job leaves through retire_done_jobs
      }
      // the next round: jobs with samples left, least recently served first
      std::vector<std::shared_ptr<JobState>> cand;
      {
         std::lock_guard<std::mutex> lock(mu);
         for (auto &j : running)
            if (!j->cancel && j->state == PTX_JOB_RUNNING && j->next_sample < j->desc.spp) cand.push_back(j);
      }
      std::stable_sort(cand.begin(), cand.end(),
                       [](const std::shared_ptr<JobState> &a, const std::shared_ptr<JobState> &b) {
                          return a->last_round < b->last_round;
                       });
      if (cand.size() > paths_max_chunks()) cand.resize(paths_max_chunks());
      if (!cand.empty() && (int)inflight.size() >= kRoundsAhead) complete_oldest_round();
      if (!cand.empty()) {
         const int table = (int)(rounds % kRoundsAhead);
         RoundInFlight r;
         r.table = table;
         chunks.clear();
         uint64_t used = 0;
         uint32_t depth = 0;
         bool nee = false, glossy = false, textured = false;
         const uint64_t share = paths / cand.size();
         for (auto &j : cand) {
            const uint32_t left = j->desc.spp - j->next_sample;
            uint64_t k = std::max<uint64_t>(1, share / j->npix);
            k = std::min<uint64_t>(k, left);
            k = std::min<uint64_t>(k, (paths - used) / j->npix);
            if (k == 0) continue;            // the pool is full: first in line next round
            chunks.push_back(Chunk{j->slot, j->next_sample, (uint32_t)k, j->npix, (uint32_t)used});
            r.parts.emplace_back(j, (uint32_t)k);
            used += k * j->npix;
            depth = std::max(depth, j->desc.max_depth);
            nee |= (j->desc.flags & PTX_JOB_NEE) != 0;
            glossy |= j->scene->has_glossy;
            textured |= j->scene->has_texture;
         }
         // the table this round fills was last read by a round that has completed
         std::copy(chunks.begin(), chunks.end(), chunk_host[table]);
         Chunk *d_table = eng.d_chunks + (size_t)table * kMaxRoundChunks;
         hipError_t e = hipMemcpyAsync(d_table, chunk_host[table], chunks.size() * sizeof(Chunk),
                                       hipMemcpyHostToDevice, stream);
         if (e != hipSuccess) {
            (void)hipGetLastError();
            r.rc = set_error(PTX_ERR_HIP, std::string("server: chunk table upload: ") + hipGetErrorString(e));
         } else {
            r.rc = eng.round(d_table, chunks, depth, kMaxServerScenes, nee, glossy, textured);
         }
         rounds++;
         r.ev = take_event();
         if (r.ev && hipEventRecord(r.ev, stream) != hipSuccess) {
            (void)hipGetLastError();
            free_events.push_back(r.ev);
            r.ev = nullptr;
         }
         if (!r.ev && r.rc == PTX_OK) r.rc = set_error(PTX_ERR_HIP, "server: round event");
         {
            std::lock_guard<std::mutex> lock(mu);
            for (auto &pj : r.parts) {
               JobState &j = *pj.first;
               j.next_sample += pj.second;
               j.last_round = rounds;
               j.inflight++;
            }
         }
         // jobs whose last chunk is in this round: finalise and copy the film behind it
         for (auto &pj : r.parts) {
            JobState &j = *pj.first;
            if (j.next_sample < j.desc.spp || r.rc != PTX_OK) continue;
            std::lock_guard<std::mutex> lock(mu);
            j.film_bytes = (size_t)j.npix * 3 * sizeof(float);
            j.film = take_film(j.film_bytes);
            j.done_ev = take_event();
            int rc = j.film && j.done_ev ? PTX_OK : set_error(PTX_ERR_OOM, "server: pinned film or event");
            if (rc == PTX_OK) rc = eng.finish_job(j.jd, j.npix, j.desc.spp, j.film);
            if (rc == PTX_OK && hipEventRecord(j.done_ev, stream) != hipSuccess) {
               (void)hipGetLastError();
               rc = PTX_ERR_HIP;
            }
            if (rc != PTX_OK) fail(pj.first, rc);
            else j.finishing = true;
         }
         if (r.ev) {
            inflight.push_back(std::move(r));
         } else {
            // no event to wait on: finish the round synchronously
            const hipError_t se = hipStreamSynchronize(stream);
            if (se != hipSuccess) (void)hipGetLastError();
            std::lock_guard<std::mutex> lock(mu);
            for (auto &pj : r.parts) {
               pj.first->inflight--;
               pj.first->samples_done += (uint64_t)pj.second * pj.first->npix;
               fail(pj.first, r.rc != PTX_OK ? r.rc : PTX_ERR_HIP);
            }
         }
      } else if (!inflight.empty()) {
         complete_oldest_round();
      } else {
         // nothing to schedule and nothing enqueued: wait for the film copies
         bool waiting = false;
         {
            std::lock_guard<std::mutex> lock(mu);
            for (auto &j : running) waiting |= j->finishing && j->state == PTX_JOB_RUNNING;
         }
         if (waiting) {
            const hipError_t se = hipStreamSynchronize(stream);
            if (se != hipSuccess) (void)hipGetLastError();
         }
      }
      retire_done_jobs();
   }
   // stopping: let the enqueued work end before the buffers go
   (void)hipStreamSynchronize(stream);
   (void)hipGetLastError();
   while (!inflight.empty()) complete_oldest_round();
}

extern "C" {

int ptx_server_create(ptx_device *dev, const ptx_server_config *cfg, ptx_server **out) {
   if (!dev || !cfg || !out) return set_error(PTX_ERR_INVALID, "NULL argument");
   *out = nullptr;
   if (int rc = device_usable(dev)) return rc;
   PTX_HIP_TRY(hipSetDevice(dev->ordinal));
   auto s = std::make_unique<ptx_server>();
   s->dev = dev;
   s->cfg = *cfg;
   PTX_HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
   s->eng.dev = dev;
   s->eng.stream = s->stream;
   s->eng.spare_cap = 0;        // job buffers count against the server's budget only while jobs hold them
   // the server keeps the tuning it was created with: a later ptx_set_tuning
   // changes neither its round form nor the pool its budget was charged for
   s->eng.tun = tuning_of(dev);
   // rounds hold at most paths_max_chunks() jobs, so the default form is always
   // path regeneration and only an explicit wavefront tuning needs queues
   const bool queues = s->eng.tun.render_mode == PTX_RENDER_WAVEFRONT;
   // By default a round holds 2^27 path slots, or as many as a quarter of the
   // budget pays for if that is fewer: measured on a Poisson stream of 64
   // jobs at 64/s, rounds of 2^24, 2^26 and 2^27 slots served 1,609, 1,739 and
   // 1,752 Msamples/s with p99 latencies of 0.33, 0.23 and 0.18 s.
   uint32_t regen_paths = kDefaultServerRegenPaths;
   const uint64_t affordable = cfg->hbm_budget_bytes / 4 / RenderEngine::bytes_for(1, false);
   if (affordable < regen_paths) regen_paths = (uint32_t)std::max<uint64_t>(affordable, kMinServerRegenPaths);
   s->paths = cfg->paths_per_wave ? cfg->paths_per_wave
                                  : (s->eng.tun.paths_per_wave ? s->eng.tun.paths_per_wave
                                     : queues ? RenderEngine::kDefaultPaths : regen_paths);
   // the path pool is sized for scenes up to 64 levels deep; deeper scenes grow the spill
   int rc = s->eng.ensure(s->paths, 64, kMaxServerJobs, kMaxServerScenes, 4 * kMaxServerJobs);
   if (rc == PTX_OK && queues) rc = s->eng.ensure_queues(s->paths);
   if (rc) {
      s->eng.release();
      (void)hipStreamDestroy(s->stream);
      return rc;
   }
   s->pool_bytes = RenderEngine::bytes_for(s->paths, queues) + s->eng.spill_entries * sizeof(uint4);
   if (cfg->hbm_budget_bytes <= s->pool_bytes) {
      s->eng.release();
      (void)hipStreamDestroy(s->stream);
      return set_error(PTX_ERR_BUDGET, "HBM budget is smaller than the server's path pool (" +
                                           std::to_string(s->pool_bytes) + " bytes)");
   }
   for (int t = 0; t < kRoundsAhead; t++) {
      void *p = nullptr;
      if (hipHostMalloc(&p, kMaxRoundChunks * sizeof(Chunk), hipHostMallocDefault) != hipSuccess) {
         (void)hipGetLastError();
         for (int u = 0; u < t; u++) (void)hipHostFree(s->chunk_host[u]);
         s->eng.release();
         (void)hipStreamDestroy(s->stream);
         return set_error(PTX_ERR_OOM, "server: pinned chunk tables");
      }
      s->chunk_host[t] = static_cast<Chunk *>(p);
   }
   s->slot_used.assign(kMaxServerJobs, false);
   s->t0 = now_s();
   device_retain(dev);
   ptx_server *raw = s.release();
   raw->worker = std::thread([raw] { raw->loop(); });
   *out = raw;
   return PTX_OK;
}

void ptx_server_destroy(ptx_server *srv) {
   if (!srv) return;
   {
      std::lock_guard<std::mutex> lock(srv->mu);
      srv->stop = true;
   }
   srv->cv_work.notify_all();
   if (srv->worker.joinable()) srv->worker.join();
   (void)hipSetDevice(srv->dev->ordinal);
   (void)hipStreamSynchronize(srv->stream);
   for (auto &j : srv->running) {
      srv->give_film(j->film, j->film_bytes);
      j->film = nullptr;
      srv->release(j);
   }
   srv->running.clear();
   for (uint64_t id : srv->queue) srv->drop_scene(*srv->jobs[id]);
   srv->queue.clear();
   for (auto &kv : srv->jobs) {
      // done jobs not yet waited for keep their film until now
      srv->give_film(kv.second->film, kv.second->film_bytes);
      kv.second->film = nullptr;
      srv->drop_scene(*kv.second);
   }
   srv->jobs.clear();
   (void)hipStreamSynchronize(srv->stream);
   for (auto &f : srv->film_pool) (void)hipHostFree(f.first);
   for (Chunk *c : srv->chunk_host)
      if (c) (void)hipHostFree(c);
   for (hipEvent_t e : srv->free_events) (void)hipEventDestroy(e);
   srv->eng.release();
   (void)hipStreamDestroy(srv->stream);
   ptx_device *dev = srv->dev;
   delete srv;
   device_release(dev);
}

int ptx_server_submit(ptx_server *srv, ptx_scene *scene, const ptx_job_desc *job, uint64_t *job_id) {
   if (!srv || !scene || !job || !job_id) return set_error(PTX_ERR_INVALID, "NULL argument");
   if (scene->dev != srv->dev) return set_error(PTX_ERR_INVALID, "scene belongs to another device");
   int rc = validate_job(job);
   if (rc) return rc;
   const uint32_t npix = job->camera.width * job->camera.height;
   if (job_device_bytes(*job) > srv->cfg.hbm_budget_bytes - srv->pool_bytes)
      return set_error(PTX_ERR_BUDGET, "job needs more device memory than the HBM budget leaves");
   if (npix > srv->paths) return set_error(PTX_ERR_BUDGET, "job has more pixels than the path pool holds");
   if (scene->d.max_stack > srv->eng.stack_cap)
      return set_error(PTX_ERR_INVALID, "scene BVH is deeper than the server's stack spill supports");
   auto j = std::make_shared<JobState>();
   j->desc = *job;
   scene_retain(scene);          // released when the job leaves the server
   j->scene = scene;
   j->npix = npix;
   {
      std::lock_guard<std::mutex> lock(srv->mu);
      j->id = srv->next_id++;
      j->submit = now_s() - srv->t0;
      srv->jobs[j->id] = j;
      srv->queue.push_back(j->id);
   }
   srv->cv_work.notify_all();
   *job_id = j->id;
   return PTX_OK;
}

static void fill_status(const JobState &j, ptx_job_status *out) {
   out->state = j.state;
   out->error = j.error;
   out->submit_time = j.submit;
   out->start_time = j.start;
   out->end_time = j.end;
   out->samples_done = j.samples_done;
   out->samples_total = (uint64_t)j.npix * j.desc.spp;
}

int ptx_server_status(ptx_server *srv, uint64_t job_id, ptx_job_status *out) {
   if (!srv || !out) return set_error(PTX_ERR_INVALID, "NULL argument");
   std::lock_guard<std::mutex> lock(srv->mu);
   auto it = srv->jobs.find(job_id);
   if (it == srv->jobs.end()) return set_error(PTX_ERR_NOT_FOUND, "unknown job id");
   fill_status(*it->second, out);
   return PTX_OK;
}

int ptx_server_wait(ptx_server *srv, uint64_t job_id, float *film_rgb, ptx_job_status *out) {
   if (!srv) return set_error(PTX_ERR_INVALID, "NULL server");
   std::unique_lock<std::mutex> lock(srv->mu);
   auto it = srv->jobs.find(job_id);
   if (it == srv->jobs.end()) return set_error(PTX_ERR_NOT_FOUND, "unknown job id");
   auto j = it->second;
   srv->cv_done.wait(lock, [&] { return j->retired; });
   if (out) fill_status(*j, out);
   if (j->state == PTX_JOB_DONE && film_rgb && j->film) std::copy(j->film, j->film + (size_t)j->npix * 3, film_rgb);
   srv->give_film(j->film, j->film_bytes);
   j->film = nullptr;
   srv->jobs.erase(job_id);
   if (j->state == PTX_JOB_FAILED) return set_error(j->error, "job failed");
   if (j->state == PTX_JOB_CANCELLED) return set_error(PTX_ERR_CANCELLED, "job was cancelled");
   return PTX_OK;
}

int ptx_server_cancel(ptx_server *srv, uint64_t job_id) {
   if (!srv) return set_error(PTX_ERR_INVALID, "NULL server");
   {
      std::lock_guard<std::mutex> lock(srv->mu);
      auto it = srv->jobs.find(job_id);
      if (it == srv->jobs.end()) return set_error(PTX_ERR_NOT_FOUND, "unknown job id");
      it->second->cancel = true;
   }
   srv->cv_work.notify_all();
   return PTX_OK;
}

}  // extern "C"
