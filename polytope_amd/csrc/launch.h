// Host-side declarations of the kernel launchers (defined in the .hip files).
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include "ptx_internal.h"

namespace ptx {

// windows: the scene's window records (launch_build_windows), or null to step
// over sc.nodes; root_axis: the root's split axis (used with windows).
int launch_closest_hit(const DevScene &sc, const float *d_rays, uint32_t n, ptx_hit *d_hits,
                       ptx_ray_counters *d_counters, ptx_hit_frame *d_frames, uint32_t *d_work, uint4 *d_spill,
                       uint32_t grid, uint32_t lds_depth, uint32_t refill_min, uint32_t leaf_min,
                       const Node64 *windows, uint32_t root_axis, hipStream_t stream);
// Resident blocks per CU of the closest-hit kernel; extra = counters or frames
// requested, windows = the form that steps over window records.
int closest_hit_blocks_per_cu(uint32_t lds_depth, bool extra, bool windows);
// Window records of the closest-hit kernel: win (4 * n_int Node64, 128-byte
// aligned) from the scene's n_int interior records.
int launch_build_windows(const Node64 *nodes, uint32_t n_int, Node64 *win, hipStream_t stream);

// Size of the per-device static scratch of the no-allocation closest-hit entry
// point: 256 bytes of work counters, then the deferred-stack spill.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
// Device address of that scratch on device `ordinal`, which must be current.
// Looked up once per device (hipGetSymbolAddress); 0 on success.
int noalloc_scratch_address(int ordinal, void **ptr);

}  // namespace ptx

namespace ptx {

uint32_t render_lds_depth();
// 0 extend, 1 primary, 2 plain megakernel, 3 plain with next-event estimation, 4 shadow connection
int render_blocks_per_cu(int which);
void launch_primary(const DevScene &S, const JobDev &J, uint32_t *work, uint4 *spill, uint32_t leaf_min, uint32_t grid,
                    hipStream_t st);
// Primary records of every job of a batch in one launch: jobs / scenes are the
// device tables, offsets (njobs + 1 entries, device) the first pixel id of
// each job and the total; max_stack bounds every scene's deferred stack.
void launch_primaries(const JobDev *jobs, const DevScene *scenes, const uint32_t *offsets, uint32_t njobs,
                      uint32_t *work, uint4 *spill, uint32_t leaf_min, uint32_t max_stack, uint32_t grid,
                      hipStream_t st);
// Shadow connection of the wavefront form (rounds with next-event jobs):
// the vertex queue (48 B per entry, one entry per path slot at most), its
// count, the per-slot next-event sums (3 floats) and the connect grid.  q null:
// no next-event job in the round.
struct ShadowArgs {
   float4 *q = nullptr;
   uint32_t *count = nullptr;
   float *direct = nullptr;
   uint32_t grid = 0;
};
void launch_round_start(const Chunk *chunks, uint32_t nchunks, uint32_t nslots, const JobDev *jobs,
                        const DevScene *scenes, float *res, float *q_out, uint32_t *count_out, uint32_t *bounce,
                        const ShadowArgs &sh, uint4 *spill, uint32_t max_stack, uint32_t grid, hipStream_t st);
void launch_bounce(const float *q_in, const uint32_t *count_in, float *q_out, uint32_t *count_out, float *hits,
                   const JobDev *jobs, const DevScene *scenes, float *res, uint32_t *work, uint4 *spill,
                   uint32_t max_stack, uint32_t leaf_min, uint32_t *bounce, const ShadowArgs &sh, uint32_t ext_grid,
                   uint32_t shade_grid, hipStream_t st);
// direct: the per-slot next-event sums of a wavefront round, added to the
// samples of next-event jobs; null for path-regeneration rounds.
void launch_accumulate(const Chunk *chunks, uint32_t nchunks, const JobDev *jobs, const float *res,
                       const float *direct, uint32_t grid, hipStream_t st);
void launch_finalize(float *film, uint32_t npix, uint32_t spp, hipStream_t st);
void launch_zero_floats(float *a, size_t n, hipStream_t st);
int paths_blocks_per_cu();
// Path-regeneration mode: one launch per round; rounds limited to
// paths_max_chunks() chunks and paths_max_scenes() scenes.  nee_scratch (48
// bytes per resident lane, grid * 256 lanes) selects the kernel variant with
// next-event estimation for the jobs that ask for it; null = without.
uint32_t paths_max_chunks();
uint32_t paths_max_scenes();
void launch_paths(const Chunk *chunks, uint32_t nchunks, uint32_t nslots, const JobDev *jobs, const DevScene *scenes,
                  uint32_t nscenes, float *res, uint32_t *work, uint4 *spill, uint32_t max_stack, uint32_t shade_min,
                  uint32_t leaf_min, float4 *nee_scratch, bool glossy, bool textured, uint32_t grid, hipStream_t st);
// textured: the scene has textured meshes; cached: start every path from the
// pixel's primary record in J.primary (written by launch_primary).
void launch_mega(const DevScene &S, const JobDev &J, uint32_t sample, uint4 *spill, bool textured, bool cached,
                 uint32_t grid, hipStream_t st);

}  // namespace ptx
