// Host side of rendering: the plain one-job-per-launch-series path
// (ptx_render), the batched path (ptx_render_batch) and the autotuner
// (ptx_autotune).
//
// The batched path keeps per-path results, a stack spill area and (for the
// wavefront form) path queues and hit records on the device (the
// RenderEngine) and reuses them across calls.  Work is organised in rounds: a
// round holds one chunk (samples [s0, s0 + k) of every pixel) from each job
// that still has samples, sized so the round fits the path capacity.  A round
// is either one persistent path-regeneration launch (k_paths), or, in the
// wavefront form, shading every slot at its cached primary hit and then one
// bounce per replay of a captured hipGraph (prologue, extend, shade); either
// way the round ends by adding its samples to the films in sample order.  The
// host never waits between bounces or rounds.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "render_host.h"
#include "host_query.h"

namespace ptx {

namespace {

// Device buffer of n T for the engine; when the allocation fails, the
// engine's spare job buffers and the device pool's unused blocks go back to
// the driver and it is tried once more.
template <class T>
int dalloc(RenderEngine *eng, T **p, size_t n) {
   *p = nullptr;
   if (n == 0) return PTX_OK;
   hipError_t e = hipMalloc((void **)p, n * sizeof(T));
   if (e != hipSuccess && eng->dev && (eng->dev->pool || !eng->spares.empty())) {
      (void)hipGetLastError();
      eng->drop_spares();
      (void)hipStreamSynchronize(eng->stream);
      pool_trim(eng->dev);
      e = hipMalloc((void **)p, n * sizeof(T));
   }
   if (e != hipSuccess) {
      (void)hipGetLastError();      // the caller's own later HIP calls must not see this failure
      *p = nullptr;
      return set_error(PTX_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
   }
   return PTX_OK;
}

template <class T>
void dfree(T *&p) {
   if (p) (void)hipFree((void *)p);
   p = nullptr;
}

}  // namespace

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
   J.flags = j.flags;
   return J;
}

int validate_job(const ptx_job_desc *j) {
   if (!j) return set_error(PTX_ERR_INVALID, "NULL job");
   if (j->camera.width == 0 || j->camera.height == 0) return set_error(PTX_ERR_INVALID, "job has an empty film");
   if ((uint64_t)j->camera.width * j->camera.height > (1u << 28))
      return set_error(PTX_ERR_INVALID, "job film exceeds 2^28 pixels");
   if (j->spp == 0) return set_error(PTX_ERR_INVALID, "job has zero samples per pixel");
   if (j->max_depth > 1000) return set_error(PTX_ERR_INVALID, "job max_depth above 1000");
   if (j->flags & ~(uint32_t)PTX_JOB_NEE) return set_error(PTX_ERR_INVALID, "unknown job flags");
   return PTX_OK;
}

size_t job_device_bytes(const ptx_job_desc &j) {
   const size_t npix = (size_t)j.camera.width * j.camera.height;
   return npix * (3 * sizeof(float) + 16 * sizeof(uint32_t));
}

RenderEngine::~RenderEngine() { release(); }

hipError_t RenderEngine::take_buffer(void **p, size_t bytes) {
   size_t best = spares.size();
   for (size_t i = 0; i < spares.size(); i++)
      if (spares[i].bytes >= bytes && spares[i].bytes / 2 <= bytes &&
          (best == spares.size() || spares[i].bytes < spares[best].bytes))
         best = i;
   if (best < spares.size()) {
      *p = spares[best].p;
      lent.push_back(spares[best]);
      spare_bytes -= spares[best].bytes;
      spares.erase(spares.begin() + (ptrdiff_t)best);
      return hipSuccess;
   }
   hipError_t e = pool_malloc(dev, p, bytes, stream);
   if (e != hipSuccess && !spares.empty()) {
      (void)hipGetLastError();
      drop_spares();
      e = pool_malloc(dev, p, bytes, stream);
   }
   if (e == hipSuccess) lent.push_back(Spare{*p, bytes});
   return e;
}

void RenderEngine::give_buffer(void *p) {
   if (!p) return;
   size_t bytes = 0;
   for (size_t i = 0; i < lent.size(); i++)
      if (lent[i].p == p) {
         bytes = lent[i].bytes;
         lent[i] = lent.back();
         lent.pop_back();
         break;
      }
   if (bytes && spare_bytes + bytes <= spare_cap) {
      spares.push_back(Spare{p, bytes});
      spare_bytes += bytes;
   } else {
      (void)hipFreeAsync(p, stream);
   }
}

void RenderEngine::drop_spares() {
   for (const Spare &b : spares) (void)hipFreeAsync(b.p, stream);
   spares.clear();
   spare_bytes = 0;
}

void RenderEngine::release() {
   drop_spares();
   for (hipGraphExec_t *g : {&g_ab, &g_ba, &g_ab_nee, &g_ba_nee}) {
      if (*g) (void)hipGraphExecDestroy(*g);
      *g = nullptr;
   }
   dfree(qa); dfree(qb); dfree(hits); dfree(qs); dfree(direct); dfree(res); dfree(words); dfree(spill);
   dfree(nee_scratch);
   dfree(d_jobs); dfree(d_offsets); dfree(d_scenes); dfree(d_chunks);
   cap = qcap = 0;
   jobs_cap = scenes_cap = chunks_cap = 0;
   spill_entries = 0;
   graphs_stale = true;
}

size_t RenderEngine::bytes_for(uint32_t paths, bool queues) {
   // results; queues: two path queues, hit records, the shadow queue and the next-event sums
   return (size_t)paths * (3 * sizeof(float) + (queues ? (2 * 16 + 8 + 12 + 3) * sizeof(float) : 0));
}

bool RenderEngine::regen_round(uint32_t nchunks, uint32_t nscenes) const {
   return tun.render_mode != PTX_RENDER_WAVEFRONT && nchunks <= paths_max_chunks() &&
          nscenes <= paths_max_scenes();
}

int RenderEngine::ensure(uint32_t paths, uint32_t max_stack, uint32_t njobs, uint32_t nscenes, uint32_t nchunks) {
   const uint32_t lds = render_lds_depth();
   if (!ext_grid) {
      ext_grid = (uint32_t)(dev->num_cus * render_blocks_per_cu(0));
      shade_grid = (uint32_t)(dev->num_cus * 8);
      mega_grid = (uint32_t)(dev->num_cus * render_blocks_per_cu(2));
      mega_nee_grid = (uint32_t)(dev->num_cus * render_blocks_per_cu(3));
      paths_grid = (uint32_t)(dev->num_cus * paths_blocks_per_cu());
      connect_grid = (uint32_t)(dev->num_cus * render_blocks_per_cu(4));
   }
   const uint32_t stride_threads = std::max({ext_grid, mega_grid, mega_nee_grid, paths_grid, connect_grid}) * 256;
   const size_t need_spill = max_stack > lds ? (size_t)(max_stack - lds) * stride_threads : 0;
   int rc;
   if (paths > cap) {
      dfree(res);
      if ((rc = dalloc(this, &res, (size_t)paths * 3))) {
         release();
         return rc;
      }
      cap = paths;
      graphs_stale = true;
   }
   if (!words) {
      if ((rc = dalloc(this, &words, 64))) return rc;
      graphs_stale = true;
   }
   if (need_spill > spill_entries) {
      dfree(spill);
      if ((rc = dalloc(this, &spill, need_spill))) { release(); return rc; }
      spill_entries = need_spill;
      graphs_stale = true;
   }
   if (max_stack > stack_cap) {
      stack_cap = max_stack;
      graphs_stale = true;
   }
   // a table's capacity is recorded only once its allocation succeeded
   if (njobs > jobs_cap) {
      dfree(d_jobs);
      dfree(d_offsets);
      jobs_cap = 0;
      const uint32_t want = std::max(njobs, 64u);
      if ((rc = dalloc(this, &d_jobs, want))) return rc;
      if ((rc = dalloc(this, &d_offsets, want + 1))) return rc;
      jobs_cap = want;
      graphs_stale = true;
   }
   if (nscenes > scenes_cap) {
      dfree(d_scenes);
      scenes_cap = 0;
      const uint32_t want = std::max(nscenes, 16u);
      if ((rc = dalloc(this, &d_scenes, want))) return rc;
      scenes_cap = want;
      graphs_stale = true;
   }
   if (nchunks > chunks_cap) {
      dfree(d_chunks);
      chunks_cap = 0;
      const uint32_t want = std::max(nchunks, 1024u);
      if ((rc = dalloc(this, &d_chunks, want))) return rc;
      chunks_cap = want;
   }
   return PTX_OK;
}

int RenderEngine::ensure_queues(uint32_t need) {
   int rc;
   if (qcap < need) {
      dfree(qa); dfree(qb); dfree(hits); dfree(qs); dfree(direct);
      qcap = 0;
      if ((rc = dalloc(this, &qa, (size_t)need * 16)) || (rc = dalloc(this, &qb, (size_t)need * 16)) ||
          (rc = dalloc(this, &hits, (size_t)need * 8)) || (rc = dalloc(this, &qs, (size_t)need * 3)) ||
          (rc = dalloc(this, &direct, (size_t)need * 3))) {
         dfree(qa); dfree(qb); dfree(hits); dfree(qs); dfree(direct);
         return rc;
      }
      qcap = need;
      graphs_stale = true;
   }
   if (graph_leaf != leaf_lanes_of(tun, true)) {
      graph_leaf = leaf_lanes_of(tun, true);
      graphs_stale = true;
   }
   if (graphs_stale || !g_ab) {
      for (hipGraphExec_t *g : {&g_ab, &g_ba, &g_ab_nee, &g_ba_nee}) {
         if (*g) (void)hipGraphExecDestroy(*g);
         *g = nullptr;
      }
      for (int variant = 0; variant < 4; variant++) {
         const int dir = variant & 1;
         const bool nee = variant >= 2;
         hipGraph_t g = nullptr;
         PTX_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
         float *qin = dir == 0 ? qa : qb, *qout = dir == 0 ? qb : qa;
         uint32_t *cin = words + (dir == 0 ? kCountA : kCountB), *cout = words + (dir == 0 ? kCountB : kCountA);
         launch_bounce(qin, cin, qout, cout, hits, d_jobs, d_scenes, res, words + kWork, spill, stack_cap,
                       leaf_lanes_of(tun, true), words + kBounce, nee ? shadow_args() : ShadowArgs{}, ext_grid,
                       shade_grid, stream);
         PTX_HIP_TRY(hipStreamEndCapture(stream, &g));
         hipGraphExec_t ex = nullptr;
         const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
         (void)hipGraphDestroy(g);
         if (ie != hipSuccess) {
            (void)hipGetLastError();
            return set_error(PTX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
         }
         (nee ? (dir == 0 ? g_ab_nee : g_ba_nee) : (dir == 0 ? g_ab : g_ba)) = ex;
      }
      graphs_stale = false;
   }
   return PTX_OK;
}

// Rounds: each holds at most one chunk per job and at most max_chunks chunks;
// chunk sizes share the path capacity between the jobs that still have samples.
std::vector<std::vector<Chunk>> schedule_rounds(const std::vector<RenderJob> &jobs, uint32_t cap,
                                                uint32_t max_chunks) {
   std::vector<uint32_t> next(jobs.size(), 0);
   std::vector<std::vector<Chunk>> rounds;
   while (true) {
      std::vector<uint32_t> active;
      for (uint32_t j = 0; j < jobs.size() && active.size() < max_chunks; j++)
         if (next[j] < jobs[j].desc.spp) active.push_back(j);
      if (active.empty()) break;
      // Water-filling: the jobs with the least left go first and take what
      // they need up to an equal share of the capacity still free, so the
      // share they leave goes to the larger jobs and the round fills.
      std::stable_sort(active.begin(), active.end(), [&](uint32_t a, uint32_t b) {
         return (uint64_t)(jobs[a].desc.spp - next[a]) * jobs[a].npix <
                (uint64_t)(jobs[b].desc.spp - next[b]) * jobs[b].npix;
      });
      std::vector<Chunk> round;
      uint64_t used = 0;
      for (size_t i = 0; i < active.size(); i++) {
         const uint32_t j = active[i];
         const uint32_t npix = jobs[j].npix;
         const uint32_t left = jobs[j].desc.spp - next[j];
         const uint64_t share = (cap - used) / (active.size() - i);
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

int RenderEngine::start_job(JobDev &J, const DevScene &S, uint32_t npix, bool trace_primary) {
   // stream-ordered: a job's buffers come and go between the rounds that use
   // them without waiting for the device
   J.film = nullptr;
   J.primary = nullptr;
   hipError_t e = take_buffer((void **)&J.film, (size_t)npix * 3 * sizeof(float));
   if (e == hipSuccess) e = take_buffer((void **)&J.primary, (size_t)npix * 16 * sizeof(uint32_t));
   if (e != hipSuccess) {
      (void)hipGetLastError();
      free_job(J);
      return set_error(PTX_ERR_OOM, std::string("job film and primary records: ") + hipGetErrorString(e));
   }
   launch_zero_floats(J.film, (size_t)npix * 3, stream);
   if (trace_primary)
      launch_primary(S, J, words + kWork, spill, leaf_lanes_of(tun, true),
                     std::min<uint32_t>(ext_grid, (npix + 255) / 256), stream);
   return PTX_OK;
}

int RenderEngine::round(const Chunk *d_round, const std::vector<Chunk> &r, uint32_t depth, uint32_t nscenes,
                       bool nee, bool glossy, bool textured) {
   uint32_t nslots = 0;
   for (const Chunk &c : r) nslots = std::max(nslots, c.base + c.k * c.npix);
   if (nee && regen_round((uint32_t)r.size(), nscenes) && !nee_scratch) {
      const int rc = dalloc(this, &nee_scratch, (size_t)paths_grid * 256 * 3);
      if (rc) return rc;
   }
   if (regen_round((uint32_t)r.size(), nscenes)) {
      const uint32_t shade_min = tun.shade_lanes ? tun.shade_lanes : kDefaultShadeLanes;
      launch_paths(d_round, (uint32_t)r.size(), nslots, d_jobs, d_scenes, nscenes, res, words + kWork, spill, stack_cap,
                   shade_min, leaf_lanes_of(tun, true), nee ? nee_scratch : nullptr, glossy, textured,
                   std::min<uint32_t>(paths_grid, (nslots + 255) / 256), stream);
      launch_accumulate(d_round, (uint32_t)r.size(), d_jobs, res, nullptr, shade_grid, stream);
      return PTX_OK;
   }
   const int rc = ensure_queues(nslots);
   if (rc) return rc;
   const ShadowArgs sh = nee ? shadow_args() : ShadowArgs{};
   launch_round_start(d_round, (uint32_t)r.size(), nslots, d_jobs, d_scenes, res, qa, words + kCountA,
                      words + kBounce, sh, spill, stack_cap, std::min<uint32_t>(shade_grid, (nslots + 255) / 256),
                      stream);
   for (uint32_t b = 1; b <= depth; b++) {
      const hipError_t e = hipGraphLaunch(b & 1 ? (nee ? g_ab_nee : g_ab) : (nee ? g_ba_nee : g_ba), stream);
      if (e != hipSuccess) {
         (void)hipGetLastError();
         return set_error(PTX_ERR_HIP, std::string("hipGraphLaunch (bounce ") + std::to_string(b) +
                                             "): " + hipGetErrorString(e));
      }
   }
   launch_accumulate(d_round, (uint32_t)r.size(), d_jobs, res, nee ? direct : nullptr, shade_grid, stream);
   return PTX_OK;
}

int RenderEngine::finish_job(JobDev &J, uint32_t npix, uint32_t spp, float *film_host) {
   launch_finalize(J.film, npix, spp, stream);
   if (film_host)
      PTX_HIP_TRY(hipMemcpyAsync(film_host, J.film, (size_t)npix * 3 * sizeof(float), hipMemcpyDeviceToHost, stream));
   return PTX_OK;
}

void RenderEngine::free_job(JobDev &J) {
   give_buffer(J.film);
   give_buffer(J.primary);
   J.film = nullptr;
   J.primary = nullptr;
}

uint32_t RenderEngine::capacity_for(uint64_t total_paths, uint32_t max_npix, bool regen) const {
   uint32_t want = tun.paths_per_wave ? tun.paths_per_wave : regen ? kDefaultRegenPaths : kDefaultPaths;
   want = (uint32_t)std::min<uint64_t>(want, std::max<uint64_t>(total_paths, max_npix));
   // regeneration keeps using what is allocated; wavefront queues are sized per call
   return std::max(std::max(want, max_npix), regen ? cap : 0u);
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
   const bool regen = regen_round(1, (uint32_t)scenes.size());
   const uint32_t paths = capacity_for(total_paths, max_npix, regen);
   const auto rounds = schedule_rounds(jobs, paths, regen ? paths_max_chunks() : 0xFFFFFFFFu);
   size_t nchunks = 0;
   for (const auto &r : rounds) nchunks += r.size();
   int rc = ensure(paths, max_stack, nj, (uint32_t)scenes.size(), (uint32_t)nchunks);
   if (rc) return rc;
   hipStream_t st = stream;
   std::vector<JobDev> jd(nj);
   auto cleanup = [&]() {
      for (auto &J : jd) free_job(J);
   };
   std::vector<DevScene> sd(scenes.size());
   for (size_t i = 0; i < scenes.size(); i++) sd[i] = scenes[i]->d;
   // every job's primary rays in one launch once the tables are up (one
   // launch per job when the batch has more pixels than 32-bit ids cover)
   std::vector<uint32_t> offsets(nj + 1, 0);
   uint64_t total_pixels = 0;
   for (uint32_t j = 0; j < nj; j++) {
      offsets[j] = (uint32_t)total_pixels;
      total_pixels += jobs[j].npix;
   }
   const bool one_launch = total_pixels < (1ull << 32);
   offsets[nj] = (uint32_t)total_pixels;
   for (uint32_t j = 0; j < nj; j++) {
      jd[j] = job_dev(jobs[j].desc, jobs[j].scene_index);
      if ((rc = start_job(jd[j], sd[jobs[j].scene_index], jobs[j].npix, !one_launch))) {
         (void)hipStreamSynchronize(st);
         cleanup();
         return rc;
      }
   }
   std::vector<Chunk> all;
   all.reserve(nchunks);
   for (const auto &r : rounds) all.insert(all.end(), r.begin(), r.end());
   {
      // the tables go up before any round; a failed copy still frees every job's buffers
      hipError_t e = hipMemcpyAsync(d_jobs, jd.data(), nj * sizeof(JobDev), hipMemcpyHostToDevice, st);
      if (e == hipSuccess)
         e = hipMemcpyAsync(d_scenes, sd.data(), sd.size() * sizeof(DevScene), hipMemcpyHostToDevice, st);
      if (e == hipSuccess && !all.empty())
         e = hipMemcpyAsync(d_chunks, all.data(), all.size() * sizeof(Chunk), hipMemcpyHostToDevice, st);
      if (e == hipSuccess && one_launch)
         e = hipMemcpyAsync(d_offsets, offsets.data(), (nj + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, st);
      if (e != hipSuccess) {
         (void)hipGetLastError();
         (void)hipStreamSynchronize(st);
         cleanup();
         return set_error(PTX_ERR_HIP, std::string("render batch tables: ") + hipGetErrorString(e));
      }
   }
   if (one_launch)
      launch_primaries(d_jobs, d_scenes, d_offsets, nj, words + kWork, spill, leaf_lanes_of(tun, true), max_stack,
                       (uint32_t)std::min<uint64_t>(ext_grid, (total_pixels + 255) / 256), st);
   size_t off = 0;
   for (const auto &r : rounds) {
      uint32_t depth = 0;
      bool nee = false, glossy = false, textured = false;
      for (const Chunk &c : r) {
         depth = std::max(depth, jobs[c.job].desc.max_depth);
         nee |= (jobs[c.job].desc.flags & PTX_JOB_NEE) != 0;
         glossy |= scenes[jobs[c.job].scene_index]->has_glossy;
         textured |= scenes[jobs[c.job].scene_index]->has_texture;
      }
      if ((rc = round(d_chunks + off, r, depth, (uint32_t)scenes.size(), nee, glossy, textured))) {
         (void)hipStreamSynchronize(st);
         cleanup();
         return rc;
      }
      off += r.size();
   }
   // films back through the pinned staging, several host threads at once
   std::vector<Span> films;
   for (uint32_t j = 0; j < nj; j++) {
      if ((rc = finish_job(jd[j], jobs[j].npix, jobs[j].desc.spp, nullptr))) {
         (void)hipStreamSynchronize(st);
         cleanup();
         return rc;
      }
      if (films_host[j])
         films.push_back(Span{reinterpret_cast<char *>(films_host[j]), reinterpret_cast<char *>(jd[j].film),
                              (size_t)jobs[j].npix * 3 * sizeof(float)});
   }
   rc = download_host(dev, st, films);
   cleanup();
   return rc;
}

int RenderEngine::run_plain(const ptx_scene *scene, const ptx_job_desc &desc, float *film_host) {
   const uint32_t npix = desc.camera.width * desc.camera.height;
   int rc = ensure(cap ? cap : 1u, scene->d.max_stack, 1, 1, 1);
   if (rc) return rc;
   hipStream_t st = stream;
   // the default traces each pixel's primary ray once (every sample of a pixel
   // has the same one) and starts every sample from that record
   const bool cached = tun.plain_primary != PTX_PLAIN_RETRACE;
   float *film = nullptr;
   uint32_t *primary = nullptr;
   hipError_t e = take_buffer((void **)&film, (size_t)npix * 3 * sizeof(float));
   if (e == hipSuccess && cached) e = take_buffer((void **)&primary, (size_t)npix * 16 * sizeof(uint32_t));
   if (e != hipSuccess) {
      (void)hipGetLastError();
      give_buffer(film);
      return set_error(PTX_ERR_OOM, std::string("plain film and primary records: ") + hipGetErrorString(e));
   }
   JobDev J = job_dev(desc, 0);
   J.film = film;
   J.primary = primary;
   launch_zero_floats(film, (size_t)npix * 3, st);
   if (cached)
      launch_primary(scene->d, J, words + kWork, spill, leaf_lanes_of(tun, true),
                     std::min<uint32_t>(ext_grid, (npix + 255) / 256), st);
   const uint32_t grid =
         std::min<uint32_t>((desc.flags & PTX_JOB_NEE) ? mega_nee_grid : mega_grid, (npix + 255) / 256);
   for (uint32_t s = 0; s < desc.spp; s++) launch_mega(scene->d, J, s, spill, scene->has_texture, cached, grid, st);
   launch_finalize(film, npix, desc.spp, st);
   rc = download_host(dev, st,
                      {Span{reinterpret_cast<char *>(film_host), reinterpret_cast<char *>(film),
                            (size_t)npix * 3 * sizeof(float)}});
   give_buffer(film);
   give_buffer(primary);
   return rc;
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
   if ((rc = device_usable(dev))) return rc;
   PTX_HIP_TRY(hipSetDevice(dev->ordinal));
   std::lock_guard<std::mutex> lock(dev->mu);
   RenderEngine *eng = engine_of(dev);
   eng->tun = tuning_of(dev);
   return eng->run_plain(scene, *job, film_rgb);
}

int ptx_render_batch(ptx_scene *const *scenes, const ptx_job_desc *jobs, uint32_t n_jobs, float *const *films) {
   if (!scenes || !jobs || !films || n_jobs == 0) return set_error(PTX_ERR_INVALID, "NULL argument or no jobs");
   ptx_device *dev = scenes[0] ? scenes[0]->dev : nullptr;
   if (!dev) return set_error(PTX_ERR_INVALID, "NULL scene");
   if (int rc = device_usable(dev)) return rc;
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
   RenderEngine *eng = engine_of(dev);
   eng->tun = tuning_of(dev);
   const uint32_t max_scenes = paths_max_scenes();
   if (uniq.size() <= max_scenes || eng->tun.render_mode == PTX_RENDER_WAVEFRONT) return eng->run_batch(rj, uniq, out);
   // More scenes than the regeneration kernel's scene table holds: the jobs run
   // in job order in groups of at most that many scenes, each group one batch
   // (a job's film does not depend on the jobs rendered beside it).
   for (uint32_t i = 0; i < n_jobs;) {
      std::vector<const ptx_scene *> group;
      std::map<const ptx_scene *, uint32_t> group_index;
      std::vector<RenderJob> group_jobs;
      std::vector<float *> group_out;
      for (; i < n_jobs; i++) {
         const ptx_scene *sc = uniq[rj[i].scene_index];
         auto it = group_index.find(sc);
         if (it == group_index.end()) {
            if (group.size() == max_scenes) break;
            it = group_index.emplace(sc, (uint32_t)group.size()).first;
            group.push_back(sc);
         }
         group_jobs.push_back(rj[i]);
         group_jobs.back().scene_index = it->second;
         group_out.push_back(out[i]);
      }
      if (int rc = eng->run_batch(group_jobs, group, group_out)) return rc;
   }
   return PTX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- autotuner

namespace ptx {
namespace {

// Camera-to-world looking at `target` from `eye` (left-handed, +z forward).
void look_at(const float eye[3], const float target[3], float m[16]) {
   float d[3] = {target[0] - eye[0], target[1] - eye[1], target[2] - eye[2]};
   float l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
   for (float &x : d) x /= l;
   const float up[3] = {0.f, 1.f, 0.f};
   float r[3] = {up[1] * d[2] - up[2] * d[1], up[2] * d[0] - up[0] * d[2], up[0] * d[1] - up[1] * d[0]};
   l = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
   for (float &x : r) x /= l;
   const float u[3] = {d[1] * r[2] - d[2] * r[1], d[2] * r[0] - d[0] * r[2], d[0] * r[1] - d[1] * r[0]};
   const float mm[16] = {r[0], u[0], d[0], eye[0], r[1], u[1], d[1], eye[1], r[2], u[2], d[2], eye[2], 0, 0, 0, 1};
   std::memcpy(m, mm, sizeof(mm));
}

}  // namespace
}  // namespace ptx

extern "C" int ptx_autotune(ptx_device *dev, ptx_scene *scene, ptx_tuning *chosen) {
   if (!dev || !scene) return set_error(PTX_ERR_INVALID, "NULL argument");
   if (int rc = device_usable(dev)) return rc;
   PTX_HIP_TRY(hipSetDevice(dev->ordinal));
   // Trials run on local tunings (the ray queries take theirs as an argument,
   // the render trials set the engine's); the device's tuning is written once,
   // at the end, so calls on other threads never see a candidate.
   const ptx_tuning saved = tuning_of(dev);
   ptx_tuning best = saved;
   ptx_tuning trial = saved;
   // 1. traversal: LDS stack depth x refill threshold on 2^20 incoherent rays in the root box
   const uint32_t n = 1u << 20;
   std::vector<float> rays((size_t)n * 6);
   uint64_t st = 0x9E3779B97F4A7C15ull;
   auto rnd = [&]() {
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      return (float)((st >> 40) * (1.0 / 16777216.0));
   };
   const float *bb = scene->d.root_box;
   for (uint32_t i = 0; i < n; i++) {
      float *r = &rays[6 * (size_t)i];
      for (int k = 0; k < 3; k++) r[k] = bb[k] + (bb[3 + k] - bb[k]) * rnd();
      float z = 2 * rnd() - 1, ph = 6.2831853f * rnd(), s = std::sqrt(std::max(0.f, 1 - z * z));
      r[3] = s * std::cos(ph); r[4] = s * std::sin(ph); r[5] = z;
   }
   float *d_rays = nullptr;
   ptx_hit *d_hits = nullptr;
   PTX_HIP_TRY(hipMalloc((void **)&d_rays, rays.size() * sizeof(float)));
   if (hipMalloc((void **)&d_hits, (size_t)n * sizeof(ptx_hit)) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(d_rays);
      return set_error(PTX_ERR_OOM, "autotune buffers");
   }
   PTX_HIP_TRY(hipMemcpy(d_rays, rays.data(), rays.size() * sizeof(float), hipMemcpyHostToDevice));
   hipEvent_t e0, e1;
   (void)hipEventCreate(&e0);
   (void)hipEventCreate(&e1);
   float best_ms = 1e30f;
   int rc = PTX_OK;
   {
      std::lock_guard<std::mutex> lock(dev->mu);
      for (uint32_t lds : {8u, 12u, 16u})
         for (uint32_t refill : {1u, 4u, 8u, 32u}) {
            trial.lds_stack = lds;
            trial.refill_lanes = refill;
            float ms_min = 1e30f;
            for (int rep = 0; rep < 3 && rc == PTX_OK; rep++) {
               (void)hipEventRecord(e0, dev->stream);
               rc = run_closest_hit(scene, d_rays, n, d_hits, nullptr, nullptr, dev->stream, trial);
               (void)hipEventRecord(e1, dev->stream);
               (void)hipEventSynchronize(e1);
               float ms = 0;
               (void)hipEventElapsedTime(&ms, e0, e1);
               if (rep) ms_min = std::min(ms_min, ms);
            }
            if (rc == PTX_OK && ms_min < best_ms) {
               best_ms = ms_min;
               best.lds_stack = lds;
               best.refill_lanes = refill;
            }
         }
   }
   (void)hipFree(d_rays);
   (void)hipFree(d_hits);
   trial = best;
   // 2. round form and path slots per round, on a probe render of the scene
   if (rc == PTX_OK) {
      ptx_job_desc job{};
      const float c[3] = {(bb[0] + bb[3]) / 2, (bb[1] + bb[4]) / 2, (bb[2] + bb[5]) / 2};
      const float ext = std::max({bb[3] - bb[0], bb[4] - bb[1], bb[5] - bb[2]});
      const float eye[3] = {c[0] + 0.3f * ext, c[1] + 0.4f * ext, c[2] - 1.5f * ext};
      look_at(eye, c, job.camera.camera_to_world);
      job.camera.fov = 50.f;
      job.camera.width = 256;
      job.camera.height = 256;
      job.camera.one_over_width = 1.f / 256.f;
      job.camera.one_over_height = 1.f / 256.f;
      job.camera.aspect_ratio = 1.f;
      job.camera.tan_fov_over_2 = (float)std::tan(50.0 * 3.14159265358979 / 360.0);
      job.spp = 32;
      job.max_depth = 4;
      job.seed = 1;
      // one film per job: the films come back on several host threads at once
      std::vector<float> film((size_t)4 * 256 * 256 * 3);
      std::vector<ptx::RenderJob> rj(4);
      for (auto &r : rj) { r.desc = job; r.scene_index = 0; r.npix = 256 * 256; }
      std::vector<float *> films(4);
      for (size_t i = 0; i < films.size(); i++) films[i] = film.data() + i * 256 * 256 * 3;
      double best_s = 1e30;
      uint32_t best_paths = best.paths_per_wave, best_mode = best.render_mode;
      std::lock_guard<std::mutex> lock(dev->mu);
      RenderEngine *eng = engine_of(dev);
      for (uint32_t mode : {(uint32_t)PTX_RENDER_REGEN, (uint32_t)PTX_RENDER_WAVEFRONT})
         for (uint32_t paths : {1u << 20, 1u << 21, 1u << 22, 1u << 23}) {
            trial.paths_per_wave = paths;
            trial.render_mode = mode;
            eng->tun = trial;
            double t_min = 1e30;
            for (int rep = 0; rep < 2 && rc == PTX_OK; rep++) {
               const double t0 = now_s();
               rc = eng->run_batch(rj, {scene}, films);
               t_min = std::min(t_min, now_s() - t0);
            }
            if (rc == PTX_OK && t_min < best_s) {
               best_s = t_min;
               best_paths = paths;
               best_mode = mode;
            }
         }
      best.paths_per_wave = best_paths;
      best.render_mode = best_mode;
      // 3. regeneration: lanes a wave gathers before it shades
      if (best_mode == PTX_RENDER_REGEN) {
         trial.paths_per_wave = best_paths;
         trial.render_mode = best_mode;
         uint32_t best_shade = 0;
         for (uint32_t shade : {1u, 8u, 16u, 32u, 48u}) {
            trial.shade_lanes = shade;
            eng->tun = trial;
            double t_min = 1e30;
            for (int rep = 0; rep < 2 && rc == PTX_OK; rep++) {
               const double t0 = now_s();
               rc = eng->run_batch(rj, {scene}, films);
               t_min = std::min(t_min, now_s() - t0);
            }
            if (rc == PTX_OK && t_min < best_s) {
               best_s = t_min;
               best_shade = shade;
            }
         }
         best.shade_lanes = best_shade;
         // 4. lanes a traversal step gathers at leaves (render kernels)
         trial.shade_lanes = best_shade;
         uint32_t best_leaf = 0;
         for (uint32_t leaf : {1u, 8u, 16u, 32u}) {
            trial.leaf_lanes = leaf;
            eng->tun = trial;
            double t_min = 1e30;
            for (int rep = 0; rep < 2 && rc == PTX_OK; rep++) {
               const double t0 = now_s();
               rc = eng->run_batch(rj, {scene}, films);
               t_min = std::min(t_min, now_s() - t0);
            }
            if (rc == PTX_OK && t_min < best_s) {
               best_s = t_min;
               best_leaf = leaf;
            }
         }
         best.leaf_lanes = best_leaf;
      }
   }
   (void)hipEventDestroy(e0);
   (void)hipEventDestroy(e1);
   const ptx_tuning result = rc == PTX_OK ? best : saved;
   set_tuning(dev, result);
   if (chosen) *chosen = result;
   return rc;
}
