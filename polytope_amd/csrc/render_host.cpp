// Host side of rendering: the plain one-job-per-launch-series path
// (ptx_render), the batched wavefront path (ptx_render_batch) and the
// autotuner (ptx_autotune).
//
// The batched path keeps one set of path queues, hit records, per-path
// results and a stack spill area on the device (the RenderEngine) and reuses
// them across calls.  Work is organised in rounds: a round holds one chunk
// (samples [s0, s0 + k) of every pixel) from each job that still has samples,
// sized so the round fits the path capacity.  A round is: shade every slot at
// its cached primary hit, then one bounce per replay of a captured hipGraph
// (prologue, extend, shade), then add the round's samples to the films in
// sample order.  The host never waits between bounces or rounds.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "render_host.h"

namespace ptx {

namespace {

template <class T>
int dalloc(T **p, size_t n) {
   *p = nullptr;
   if (n == 0) return PTX_OK;
   const hipError_t e = hipMalloc((void **)p, n * sizeof(T));
   if (e != hipSuccess) return set_error(PTX_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
   return PTX_OK;
}

template <class T>
void dfree(T *&p) {
   if (p) (void)hipFree((void *)p);
   p = nullptr;
}

JobDev job_dev(const ptx_job_desc &j, uint32_t scene_index) {
   JobDev J{};
   std::memcpy(J.cam, j.camera.camera_to_world, sizeof(J.cam));
   J.oow = j.camera.one_over_width;
   J.ooh = j.camera.one_over_height;
   J.aspect = j.camera.aspect_ratio;
   J.tanf2 = j.camera.tan_fov_over_2;
   J.width = j.camera.width;
   J.height = j.camera.height;
   J.spp = j.spp;
   J.max_depth = j.max_depth;
   J.seed = j.seed;
   J.scene = scene_index;
   return J;
}

}  // namespace

int validate_job(const ptx_job_desc *j) {
   if (!j) return set_error(PTX_ERR_INVALID, "NULL job");
   if (j->camera.width == 0 || j->camera.height == 0) return set_error(PTX_ERR_INVALID, "job has an empty film");
   if ((uint64_t)j->camera.width * j->camera.height > (1u << 28))
      return set_error(PTX_ERR_INVALID, "job film exceeds 2^28 pixels");
   if (j->spp == 0) return set_error(PTX_ERR_INVALID, "job has zero samples per pixel");
   if (j->max_depth > 1000) return set_error(PTX_ERR_INVALID, "job max_depth above 1000");
   return PTX_OK;
}

size_t job_device_bytes(const ptx_job_desc &j) {
   const size_t npix = (size_t)j.camera.width * j.camera.height;
   return npix * (3 * sizeof(float) + 16 * sizeof(uint32_t));
}

RenderEngine::~RenderEngine() { release(); }

void RenderEngine::release() {
   if (g_ab) (void)hipGraphExecDestroy(g_ab);
   if (g_ba) (void)hipGraphExecDestroy(g_ba);
   g_ab = g_ba = nullptr;
   dfree(qa); dfree(qb); dfree(hits); dfree(res); dfree(words); dfree(spill);
   dfree(d_jobs); dfree(d_scenes); dfree(d_chunks);
   cap = 0;
   spill_entries = 0;
}

size_t RenderEngine::bytes_for(uint32_t paths) {
   return (size_t)paths * (2 * 16 * sizeof(float) + 8 * sizeof(float) + 3 * sizeof(float));
}

int RenderEngine::ensure(uint32_t paths, uint32_t max_stack, uint32_t njobs, uint32_t nscenes, uint32_t nchunks) {
   const uint32_t lds = render_lds_depth();
   if (!ext_grid) {
      ext_grid = (uint32_t)(dev->num_cus * render_blocks_per_cu(0));
      shade_grid = (uint32_t)(dev->num_cus * 8);
      mega_grid = (uint32_t)(dev->num_cus * render_blocks_per_cu(2));
   }
   const uint32_t stride_threads = std::max(ext_grid, mega_grid) * 256;
   const size_t need_spill = max_stack > lds ? (size_t)(max_stack - lds) * stride_threads : 0;
   bool graphs_stale = false;
   int rc;
   if (paths > cap) {
      dfree(qa); dfree(qb); dfree(hits); dfree(res);
      if ((rc = dalloc(&qa, (size_t)paths * 16)) || (rc = dalloc(&qb, (size_t)paths * 16)) ||
          (rc = dalloc(&hits, (size_t)paths * 8)) || (rc = dalloc(&res, (size_t)paths * 3))) {
         release();
         return rc;
      }
      cap = paths;
      graphs_stale = true;
   }
   if (!words) {
      if ((rc = dalloc(&words, 64))) return rc;
      graphs_stale = true;
   }
   if (need_spill > spill_entries) {
      dfree(spill);
      if ((rc = dalloc(&spill, need_spill))) { release(); return rc; }
      spill_entries = need_spill;
      graphs_stale = true;
   }
   if (max_stack > stack_cap) {
      stack_cap = max_stack;
      graphs_stale = true;
   }
   if (njobs > jobs_cap) {
      dfree(d_jobs);
      jobs_cap = std::max(njobs, 64u);
      if ((rc = dalloc(&d_jobs, jobs_cap))) return rc;
      graphs_stale = true;
   }
   if (nscenes > scenes_cap) {
      dfree(d_scenes);
      scenes_cap = std::max(nscenes, 16u);
      if ((rc = dalloc(&d_scenes, scenes_cap))) return rc;
      graphs_stale = true;
   }
   if (nchunks > chunks_cap) {
      dfree(d_chunks);
      chunks_cap = std::max(nchunks, 1024u);
      if ((rc = dalloc(&d_chunks, chunks_cap))) return rc;
   }
   if (graphs_stale || !g_ab) {
      if (g_ab) (void)hipGraphExecDestroy(g_ab);
      if (g_ba) (void)hipGraphExecDestroy(g_ba);
      g_ab = g_ba = nullptr;
      for (int dir = 0; dir < 2; dir++) {
         hipGraph_t g = nullptr;
         PTX_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
         float *qin = dir == 0 ? qa : qb, *qout = dir == 0 ? qb : qa;
         uint32_t *cin = words + (dir == 0 ? kCountA : kCountB), *cout = words + (dir == 0 ? kCountB : kCountA);
         launch_bounce(qin, cin, qout, cout, hits, d_jobs, d_scenes, res, words + kWork, spill, stack_cap,
                       words + kBounce, ext_grid, shade_grid, stream);
         PTX_HIP_TRY(hipStreamEndCapture(stream, &g));
         hipGraphExec_t ex = nullptr;
         const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
         (void)hipGraphDestroy(g);
         if (ie != hipSuccess) return set_error(PTX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
         (dir == 0 ? g_ab : g_ba) = ex;
      }
   }
   return PTX_OK;
}

// Rounds: each holds at most one chunk per job; chunk sizes share the path
// capacity between the jobs that still have samples.
std::vector<std::vector<Chunk>> schedule_rounds(const std::vector<RenderJob> &jobs, uint32_t cap) {
   std::vector<uint32_t> next(jobs.size(), 0);
   std::vector<std::vector<Chunk>> rounds;
   while (true) {
      std::vector<uint32_t> active;
      for (uint32_t j = 0; j < jobs.size(); j++)
         if (next[j] < jobs[j].desc.spp) active.push_back(j);
      if (active.empty()) break;
      std::vector<Chunk> round;
      uint64_t used = 0;
      const uint64_t share = cap / active.size();
      for (uint32_t j : active) {
         const uint32_t npix = jobs[j].npix;
         const uint32_t left = jobs[j].desc.spp - next[j];
         uint64_t k = std::max<uint64_t>(1, share / npix);
         k = std::min<uint64_t>(k, left);
         k = std::min<uint64_t>(k, (cap - used) / npix);
         if (k == 0) continue;
         round.push_back(Chunk{j, next[j], (uint32_t)k, npix, (uint32_t)used});
         used += k * npix;
         next[j] += (uint32_t)k;
      }
      rounds.push_back(std::move(round));
   }
   // chunk bases must be increasing within a round for the device-side search
   for (auto &r : rounds) {
      uint32_t base = 0;
      for (Chunk &c : r) { c.base = base; base += c.k * c.npix; }
   }
   return rounds;
}

int RenderEngine::run_batch(const std::vector<RenderJob> &jobs, const std::vector<const ptx_scene *> &scenes,
                            const std::vector<float *> &films_host) {
   const uint32_t nj = (uint32_t)jobs.size();
   uint32_t max_stack = 1, max_npix = 1;
   uint64_t total_paths = 0;
   for (const auto *s : scenes) max_stack = std::max(max_stack, s->d.max_stack);
   for (const auto &j : jobs) {
      max_npix = std::max(max_npix, j.npix);
      total_paths += (uint64_t)j.npix * j.desc.spp;
   }
   uint32_t want = dev->tuning.paths_per_wave ? dev->tuning.paths_per_wave : kDefaultPaths;
   want = (uint32_t)std::min<uint64_t>(want, std::max<uint64_t>(total_paths, max_npix));
   want = std::max(want, max_npix);
   const auto rounds = schedule_rounds(jobs, std::max(want, cap));
   size_t nchunks = 0;
   for (const auto &r : rounds) nchunks += r.size();
   int rc = ensure(std::max(want, cap), max_stack, nj, (uint32_t)scenes.size(), (uint32_t)nchunks);
   if (rc) return rc;
   hipStream_t st = stream;
   // per-job device buffers
   std::vector<JobDev> jd(nj);
   std::vector<float *> films(nj, nullptr);
   std::vector<uint32_t *> prim(nj, nullptr);
   auto cleanup = [&]() {
      for (auto &f : films) dfree(f);
      for (auto &p : prim) dfree(p);
   };
   for (uint32_t j = 0; j < nj; j++) {
      if ((rc = dalloc(&films[j], (size_t)jobs[j].npix * 3)) || (rc = dalloc(&prim[j], (size_t)jobs[j].npix * 16))) {
         cleanup();
         return rc;
      }
      jd[j] = job_dev(jobs[j].desc, jobs[j].scene_index);
      jd[j].film = films[j];
      jd[j].primary = prim[j];
      launch_zero_floats(films[j], (size_t)jobs[j].npix * 3, st);
   }
   std::vector<DevScene> sd(scenes.size());
   for (size_t i = 0; i < scenes.size(); i++) sd[i] = scenes[i]->d;
   std::vector<Chunk> all;
   all.reserve(nchunks);
   for (const auto &r : rounds) all.insert(all.end(), r.begin(), r.end());
   PTX_HIP_TRY(hipMemcpyAsync(d_jobs, jd.data(), nj * sizeof(JobDev), hipMemcpyHostToDevice, st));
   PTX_HIP_TRY(hipMemcpyAsync(d_scenes, sd.data(), sd.size() * sizeof(DevScene), hipMemcpyHostToDevice, st));
   if (!all.empty())
      PTX_HIP_TRY(hipMemcpyAsync(d_chunks, all.data(), all.size() * sizeof(Chunk), hipMemcpyHostToDevice, st));
   // primary hits, one traversal per pixel and job
   for (uint32_t j = 0; j < nj; j++)
      launch_primary(sd[jobs[j].scene_index], jd[j], words + kWork, spill,
                     std::min<uint32_t>(ext_grid, (jobs[j].npix + 255) / 256), st);
   size_t off = 0;
   for (const auto &r : rounds) {
      uint32_t nslots = 0, depth = 0;
      for (const Chunk &c : r) {
         nslots = std::max(nslots, c.base + c.k * c.npix);
         depth = std::max(depth, jobs[c.job].desc.max_depth);
      }
      launch_round_start(d_chunks + off, (uint32_t)r.size(), nslots, d_jobs, d_scenes, res, qa, words + kCountA,
                         words + kBounce, std::min<uint32_t>(shade_grid, (nslots + 255) / 256), st);
      for (uint32_t b = 1; b <= depth; b++) PTX_HIP_TRY(hipGraphLaunch(b & 1 ? g_ab : g_ba, st));
      launch_accumulate(d_chunks + off, (uint32_t)r.size(), d_jobs, res, shade_grid, st);
      off += r.size();
   }
   for (uint32_t j = 0; j < nj; j++) {
      launch_finalize(films[j], jobs[j].npix, jobs[j].desc.spp, st);
      if (films_host[j])
         PTX_HIP_TRY(hipMemcpyAsync(films_host[j], films[j], (size_t)jobs[j].npix * 3 * sizeof(float),
                                    hipMemcpyDeviceToHost, st));
   }
   const hipError_t se = hipStreamSynchronize(st);
   cleanup();
   if (se != hipSuccess) return set_error(PTX_ERR_HIP, std::string("render batch: ") + hipGetErrorString(se));
   return PTX_OK;
}

int RenderEngine::run_plain(const ptx_scene *scene, const ptx_job_desc &desc, float *film_host) {
   const uint32_t npix = desc.camera.width * desc.camera.height;
   int rc = ensure(cap ? cap : 1u, scene->d.max_stack, 1, 1, 1);
   if (rc) return rc;
   hipStream_t st = stream;
   float *film = nullptr;
   if ((rc = dalloc(&film, (size_t)npix * 3))) return rc;
   JobDev J = job_dev(desc, 0);
   J.film = film;
   launch_zero_floats(film, (size_t)npix * 3, st);
   const uint32_t grid = std::min<uint32_t>(mega_grid, (npix + 255) / 256);
   for (uint32_t s = 0; s < desc.spp; s++) launch_mega(scene->d, J, s, spill, grid, st);
   launch_finalize(film, npix, desc.spp, st);
   hipError_t e = hipMemcpyAsync(film_host, film, (size_t)npix * 3 * sizeof(float), hipMemcpyDeviceToHost, st);
   if (e == hipSuccess) e = hipStreamSynchronize(st);
   dfree(film);
   if (e != hipSuccess) return set_error(PTX_ERR_HIP, std::string("render: ") + hipGetErrorString(e));
   return PTX_OK;
}

RenderEngine *engine_of(ptx_device *dev) {
   if (!dev->engine) {
      dev->engine = std::make_unique<RenderEngine>();
      dev->engine->dev = dev;
      dev->engine->stream = dev->stream;
   }
   return dev->engine.get();
}

}  // namespace ptx

using namespace ptx;

extern "C" {

int ptx_render(ptx_scene *scene, const ptx_job_desc *job, float *film_rgb) {
   if (!scene || !film_rgb) return set_error(PTX_ERR_INVALID, "NULL argument");
   int rc = validate_job(job);
   if (rc) return rc;
   ptx_device *dev = scene->dev;
   PTX_HIP_TRY(hipSetDevice(dev->ordinal));
   std::lock_guard<std::mutex> lock(dev->mu);
   return engine_of(dev)->run_plain(scene, *job, film_rgb);
}

int ptx_render_batch(ptx_scene *const *scenes, const ptx_job_desc *jobs, uint32_t n_jobs, float *const *films) {
   if (!scenes || !jobs || !films || n_jobs == 0) return set_error(PTX_ERR_INVALID, "NULL argument or no jobs");
   ptx_device *dev = scenes[0] ? scenes[0]->dev : nullptr;
   if (!dev) return set_error(PTX_ERR_INVALID, "NULL scene");
   std::vector<const ptx_scene *> uniq;
   std::map<const ptx_scene *, uint32_t> index;
   std::vector<RenderJob> rj(n_jobs);
   std::vector<float *> out(n_jobs);
   for (uint32_t i = 0; i < n_jobs; i++) {
      if (!scenes[i] || scenes[i]->dev != dev) return set_error(PTX_ERR_INVALID, "jobs must use scenes of one device");
      int rc = validate_job(&jobs[i]);
      if (rc) return rc;
      auto it = index.find(scenes[i]);
      if (it == index.end()) {
         it = index.emplace(scenes[i], (uint32_t)uniq.size()).first;
         uniq.push_back(scenes[i]);
      }
      rj[i].desc = jobs[i];
      rj[i].scene_index = it->second;
      rj[i].npix = jobs[i].camera.width * jobs[i].camera.height;
      out[i] = films[i];
   }
   PTX_HIP_TRY(hipSetDevice(dev->ordinal));
   std::lock_guard<std::mutex> lock(dev->mu);
   return engine_of(dev)->run_batch(rj, uniq, out);
}

}  // extern "C"
