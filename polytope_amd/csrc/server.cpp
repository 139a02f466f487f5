// Continuous-batching render service (ptx_server_*) and the autotuner
// (ptx_autotune).
//
// A service thread owns a RenderEngine on its own stream.  Jobs are queued by
// ptx_server_submit and admitted in arrival order when a job-table slot is
// free and the job's device memory (film + primary-hit records) fits the HBM
// budget next to the engine's path pool.  Admitted jobs trace their primary
// hits, then take part in rounds: every round gives each job in flight one
// chunk of samples (its share of the path capacity), runs the bounces, and
// adds the chunk to the job's film in sample order.  Jobs that arrive while a
// round runs join the next round; finished jobs leave it, their film is copied
// out and their memory returns to the budget.  A job's film depends only on
// its own samples and their order, so it is the same alone and in any mix.
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
constexpr uint32_t kDefaultServerRegenPaths = 1u << 24;   // path slots per round (12 B each)

struct JobState {
   uint64_t id = 0;
   ptx_job_desc desc{};
   ptx_scene *scene = nullptr;
   uint32_t npix = 0;
   int32_t state = PTX_JOB_QUEUED;
   int32_t error = PTX_OK;
   double submit = 0, start = 0, end = 0;
   uint64_t samples_done = 0;
   uint32_t next_sample = 0;
   size_t bytes = 0;
   uint32_t slot = 0, scene_slot = 0;
   bool cancel = false;
   JobDev jd{};
   DevScene sd{};                // staged copy for the device scene table
   std::vector<float> film;      // host copy once done
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
   size_t rotation = 0;         // first running job of the next round when more than fit
   std::vector<bool> slot_used;
   std::map<const ptx_scene *, std::pair<uint32_t, uint32_t>> scene_slots;   // scene -> (slot, refs)
   double t0 = 0;

   void loop();
   void fail(const std::shared_ptr<ptx::JobState> &j, int code);
   void release(const std::shared_ptr<ptx::JobState> &j);
};

using namespace ptx;

void ptx_server::fail(const std::shared_ptr<JobState> &j, int code) {
   j->state = PTX_JOB_FAILED;
   j->error = code;
   j->end = now_s() - t0;
}

// Frees a job's device memory, table slots and budget share (caller holds mu).
void ptx_server::release(const std::shared_ptr<JobState> &j) {
   eng.free_job(j->jd);
   if (j->bytes) {
      used_bytes -= j->bytes;
      j->bytes = 0;
   }
   slot_used[j->slot] = false;
   auto it = scene_slots.find(j->scene);
   if (it != scene_slots.end() && --it->second.second == 0) scene_slots.erase(it);
}

void ptx_server::loop() {
   (void)hipSetDevice(dev->ordinal);
   std::vector<Chunk> chunks;
   while (true) {
      std::vector<std::shared_ptr<JobState>> admitted;
      {
         std::unique_lock<std::mutex> lock(mu);
         cv_work.wait(lock, [&] { return stop || !queue.empty() || !running.empty(); });
         if (stop) break;
         // admission in arrival order: a free slot, a scene slot and room in the budget
         while (!queue.empty()) {
            auto j = jobs[queue.front()];
            if (j->cancel) {
               queue.pop_front();
               j->state = PTX_JOB_CANCELLED;
               j->end = now_s() - t0;
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
      // start admitted jobs: tables, film, primary hits
      for (auto &j : admitted) {
         j->jd = job_dev(j->desc, j->scene_slot);
         int rc = eng.start_job(j->jd, j->scene->d, j->npix);
         if (rc == PTX_OK) {
            j->sd = j->scene->d;
            if (hipMemcpyAsync(eng.d_scenes + j->scene_slot, &j->sd, sizeof(DevScene), hipMemcpyHostToDevice, stream) !=
                      hipSuccess ||
                hipMemcpyAsync(eng.d_jobs + j->slot, &j->jd, sizeof(JobDev), hipMemcpyHostToDevice, stream) !=
                      hipSuccess)
               rc = PTX_ERR_HIP;
         }
         std::lock_guard<std::mutex> lock(mu);
         if (rc != PTX_OK) {
            fail(j, rc);
            release(j);
            cv_done.notify_all();
         } else {
            running.push_back(j);
         }
      }
      // one round over the jobs in flight
      std::vector<std::shared_ptr<JobState>> in_round;
      {
         std::lock_guard<std::mutex> lock(mu);
         for (auto &j : running)
            if (!j->cancel) in_round.push_back(j);
      }
      // at most paths_max_chunks() jobs per round, taken in rotation when more run
      const size_t max_round = paths_max_chunks();
      if (in_round.size() > max_round) {
         std::rotate(in_round.begin(), in_round.begin() + (rotation % in_round.size()), in_round.end());
         in_round.resize(max_round);
         rotation += max_round;
      }
      if (!in_round.empty()) {
         int round_rc = PTX_OK;
         chunks.clear();
         uint64_t used = 0;
         uint32_t depth = 0;
         bool nee = false, glossy = false, textured = false;
         const uint64_t share = paths / in_round.size();
         for (auto &j : in_round) {
            const uint32_t left = j->desc.spp - j->next_sample;
            uint64_t k = std::max<uint64_t>(1, share / j->npix);
            k = std::min<uint64_t>(k, left);
            k = std::min<uint64_t>(k, (paths - used) / j->npix);
            if (k == 0) continue;
            chunks.push_back(Chunk{j->slot, j->next_sample, (uint32_t)k, j->npix, (uint32_t)used});
            used += k * j->npix;
            depth = std::max(depth, j->desc.max_depth);
            nee |= (j->desc.flags & PTX_JOB_NEE) != 0;
            glossy |= j->scene->has_glossy;
            textured |= j->scene->has_texture;
         }
         if (!chunks.empty()) {
            (void)hipMemcpyAsync(eng.d_chunks, chunks.data(), chunks.size() * sizeof(Chunk), hipMemcpyHostToDevice,
                                 stream);
            round_rc = eng.round(eng.d_chunks, chunks, depth, kMaxServerScenes, nee, glossy, textured);
         }
         const hipError_t e = hipStreamSynchronize(stream);
         std::lock_guard<std::mutex> lock(mu);
         for (const Chunk &c : chunks) {
            for (auto &j : in_round)
               if (j->slot == c.job) {
                  j->next_sample += c.k;
                  j->samples_done += (uint64_t)c.k * c.npix;
               }
         }
         if (e != hipSuccess || round_rc != PTX_OK)
            for (auto &j : in_round) fail(j, round_rc != PTX_OK ? round_rc : PTX_ERR_HIP);
      }
      // retire finished and cancelled jobs
      std::vector<std::shared_ptr<JobState>> done;
      {
         std::lock_guard<std::mutex> lock(mu);
         for (auto &j : running)
            if (j->cancel || j->state == PTX_JOB_FAILED || j->next_sample >= j->desc.spp) done.push_back(j);
      }
      for (auto &j : done) {
         if (!j->cancel && j->state == PTX_JOB_RUNNING) {
            j->film.assign((size_t)j->npix * 3, 0.f);
            int rc = eng.finish_job(j->jd, j->npix, j->desc.spp, j->film.data());
            if (rc == PTX_OK && hipStreamSynchronize(stream) != hipSuccess) rc = PTX_ERR_HIP;
            std::lock_guard<std::mutex> lock(mu);
            if (rc != PTX_OK) fail(j, rc);
            else {
               j->state = PTX_JOB_DONE;
               j->end = now_s() - t0;
            }
         }
         std::lock_guard<std::mutex> lock(mu);
         if (j->cancel && j->state == PTX_JOB_RUNNING) {
            j->state = PTX_JOB_CANCELLED;
            j->end = now_s() - t0;
         }
         release(j);
         running.erase(std::remove(running.begin(), running.end(), j), running.end());
         cv_done.notify_all();
      }
   }
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
   // rounds hold at most paths_max_chunks() jobs, so the default form is always
   // path regeneration and only an explicit wavefront tuning needs queues
   const bool queues = dev->tuning.render_mode == PTX_RENDER_WAVEFRONT;
   s->paths = cfg->paths_per_wave ? cfg->paths_per_wave
                                  : (dev->tuning.paths_per_wave ? dev->tuning.paths_per_wave
                                     : queues ? RenderEngine::kDefaultPaths : kDefaultServerRegenPaths);
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
   for (auto &j : srv->running) srv->eng.free_job(j->jd);
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
   srv->cv_done.wait(lock, [&] {
      return j->state == PTX_JOB_DONE || j->state == PTX_JOB_FAILED || j->state == PTX_JOB_CANCELLED;
   });
   if (out) fill_status(*j, out);
   if (j->state == PTX_JOB_DONE && film_rgb) std::copy(j->film.begin(), j->film.end(), film_rgb);
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
