// Render engine: device-resident path queues and the batched wavefront driver.
#pragma once

#include <vector>

#include "scene.h"

namespace ptx {

struct RenderJob {
   ptx_job_desc desc{};
   uint32_t scene_index = 0;
   uint32_t npix = 0;
};

int validate_job(const ptx_job_desc *j);
JobDev job_dev(const ptx_job_desc &j, uint32_t scene_index);
// Device bytes a job holds while it runs (film + primary-hit records).
size_t job_device_bytes(const ptx_job_desc &j);

// Rounds of at most max_chunks chunks (one per job) sharing cap path slots.
std::vector<std::vector<Chunk>> schedule_rounds(const std::vector<RenderJob> &jobs, uint32_t cap,
                                                uint32_t max_chunks);

struct RenderEngine {
   static constexpr uint32_t kDefaultPaths = 1u << 22;        // wavefront rounds (232 B per slot)
   static constexpr uint32_t kDefaultRegenPaths = 1u << 30;   // regeneration rounds (12 B per slot)
   static constexpr uint32_t kDefaultShadeLanes = 16;   // lanes a wave gathers before it shades
   // layout of `words`: 8 traversal work words (word 0 counts the dynamic
   // batches handed out), the two queue counts, the bounce index, the
   // shadow-queue count
   static constexpr uint32_t kWork = 0, kCountA = 8, kCountB = 9, kBounce = 10, kCountS = 11;

   ptx_device *dev = nullptr;
   hipStream_t stream = nullptr;
   // The tuning this engine runs with: the device's, copied at the start of a
   // ptx_render / ptx_render_batch call, or once when a server is created.
   ptx_tuning tun{};
   uint32_t cap = 0;                        // path slots (res)
   uint32_t qcap = 0;                       // path slots of the wavefront queues (0 = none)
   float *qa = nullptr, *qb = nullptr;      // queues, 16 words per entry
   float *hits = nullptr;                   // 8 words per path slot
   float4 *qs = nullptr;                    // shadow-connect queue, 3 float4 per entry
   float *direct = nullptr;                 // next-event sums of the wavefront form, 3 floats per slot
   float *res = nullptr;                    // 3 floats per path slot
   uint32_t *words = nullptr;
   uint4 *spill = nullptr;
   float4 *nee_scratch = nullptr;           // per resident lane of the regeneration kernel, 48 bytes
   size_t spill_entries = 0;
   uint32_t stack_cap = 0;
   JobDev *d_jobs = nullptr;
   uint32_t *d_offsets = nullptr;           // jobs_cap + 1 entries: first primary-ray id of each job
   uint32_t jobs_cap = 0;
   DevScene *d_scenes = nullptr;
   uint32_t scenes_cap = 0;
   Chunk *d_chunks = nullptr;
   uint32_t chunks_cap = 0;
   uint32_t ext_grid = 0, shade_grid = 0, mega_grid = 0, mega_nee_grid = 0, paths_grid = 0, connect_grid = 0;
   // Device buffers of finished jobs (films, primary records) kept for later
   // jobs on this engine's stream: freeing them costs about 0.2 ms of host
   // time each (7 ms per batch of the 16-job render workload).  At most
   // spare_cap bytes are kept; the server's engine keeps none (its jobs are
   // admitted against its HBM budget).
   struct Spare {
      void *p;
      size_t bytes;
   };
   std::vector<Spare> spares;
   std::vector<Spare> lent;                 // buffers handed out by take_buffer, with their sizes
   size_t spare_bytes = 0;
   size_t spare_cap = size_t(4) << 30;
   hipGraphExec_t g_ab = nullptr, g_ba = nullptr;        // bounce graphs per queue direction
   hipGraphExec_t g_ab_nee = nullptr, g_ba_nee = nullptr;  // the same with the shadow connection
   bool graphs_stale = true;
   uint32_t graph_leaf = 0;                 // leaf_min the graphs were captured with

   ~RenderEngine();
   void release();
   // A device buffer of at least `bytes`: the smallest spare that fits (and
   // is at most twice as large), else a new one from the device pool.
   hipError_t take_buffer(void **p, size_t bytes);
   // Back from a finished job: kept as a spare up to spare_cap, else freed.
   void give_buffer(void *p);
   void drop_spares();
   // Device bytes of `paths` slots: results, plus queues and hit records when
   // the wavefront form may run.
   static size_t bytes_for(uint32_t paths, bool queues);
   // Whether a round of nchunks chunks over nscenes scenes runs as one
   // path-regeneration launch under the engine's tuning.
   bool regen_round(uint32_t nchunks, uint32_t nscenes) const;
   int ensure(uint32_t paths, uint32_t max_stack, uint32_t njobs, uint32_t nscenes, uint32_t nchunks);
   // Wavefront queues and hit records for at least `need` slots, and the
   // per-bounce graphs.
   int ensure_queues(uint32_t need);
   ShadowArgs shadow_args() const { return ShadowArgs{qs, words + kCountS, direct, connect_grid}; }
   // Allocates the job's film and primary records, zeroes the film and, with
   // trace_primary, traces the primary hits (a batch traces every job's in
   // one launch instead).  J.film / J.primary are set.
   int start_job(JobDev &J, const DevScene &S, uint32_t npix, bool trace_primary = true);
   // One round over the chunks in d_round (a device copy of r); the job and
   // scene tables must hold the round's jobs, the scene table nscenes entries.
   // nee: some job of the round uses next-event estimation;
   // glossy / textured: some scene of the round has glossy / textured meshes.
   int round(const Chunk *d_round, const std::vector<Chunk> &r, uint32_t depth, uint32_t nscenes, bool nee,
             bool glossy, bool textured);
   // Queues the film's finalisation and, when film_host is not NULL (the
   // server's pinned films), its copy there.
   int finish_job(JobDev &J, uint32_t npix, uint32_t spp, float *film_host);
   void free_job(JobDev &J);
   uint32_t capacity_for(uint64_t total_paths, uint32_t max_npix, bool regen) const;
   int run_batch(const std::vector<RenderJob> &jobs, const std::vector<const ptx_scene *> &scenes,
                 const std::vector<float *> &films_host);
   int run_plain(const ptx_scene *scene, const ptx_job_desc &desc, float *film_host);
};

RenderEngine *engine_of(ptx_device *dev);

}  // namespace ptx
