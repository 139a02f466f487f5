/*
 * ptx.h — C ABI of libptx_hip.so, the MI355X (gfx950) path-tracing hot path
 * for polytope scenes: SAH-BVH closest-hit traversal, Wald watertight
 * ray/triangle intersection and the Lambertian/mirror bounce loop.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes
 * (PTX_OK == 0) with a per-thread message from ptx_last_error().  Nothing in
 * the library calls exit(): a failed call returns an error and leaves the
 * process running.  Scene handles are immutable after creation and may be used
 * from several host threads at once; the caller owns every buffer it passes.
 *
 * Reference interfaces each entry replaces (danielthompson/polytope):
 *   ptx_device_open/close     poly::GPUMemoryManager(w, h) / ~GPUMemoryManager
 *                             (src/cuda/gpu_memory_manager.h:79-119, .cpp:279-286)
 *   ptx_scene_create          GPUMemoryManager::MallocScene(Scene*) (gpu_memory_manager.cpp:19-277),
 *                             fed by bvh::bound + bvh::compact (src/cpu/acceleration/bvh.cpp:36, 695)
 *   ptx_intersect_closest*    Scene::intersect(Ray&, x, y) -> bvh::intersect_compact
 *                             (src/cpu/scenes/Scene.cpp:12-31, src/cpu/acceleration/bvh.cpp:476-523),
 *                             batched over many rays
 *   ptx_render                PathTracerKernel::Trace(spp) + OutputPNG::Output's device->host
 *                             film copy (src/cuda/kernels/path_tracer.cu:899-1018,
 *                             src/cuda/png_output.cpp:12-24), with the CPU path's semantics
 *                             (PathTraceIntegrator::GetSample, BoxFilter), not the CUDA kernel's
 *   ptx_render_batch          no counterpart: many render jobs merged into one persistent launch
 *   ptx_server_*              no counterpart: continuous-batching render service
 */
#ifndef PTX_H
#define PTX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
   PTX_OK = 0,
   PTX_ERR_INVALID = 1,      /* bad argument or scene description            */
   PTX_ERR_HIP = 2,          /* HIP runtime error                             */
   PTX_ERR_OOM = 3,          /* device or host allocation failed              */
   PTX_ERR_BUDGET = 4,       /* job larger than the server's HBM budget       */
   PTX_ERR_NOT_FOUND = 5,    /* unknown job id                                */
   PTX_ERR_CANCELLED = 6,    /* job was cancelled                             */
   PTX_ERR_NO_DEVICE = 7     /* no usable gfx950 device                       */
};

/* PTX_BRDF_GLOSSY: GlossyBRDF, the reference's "plastic" material
 * (src/cpu/shading/brdf/glossy_brdf.h:13-60, pbrt_parser.cpp:791-842). */
enum { PTX_BRDF_LAMBERT = 1, PTX_BRDF_MIRROR = 2, PTX_BRDF_GLOSSY = 3 };

typedef struct ptx_device ptx_device;
typedef struct ptx_scene ptx_scene;
typedef struct ptx_server ptx_server;

/* One mesh instance: poly::mesh_geometry (packed SoA vertices + face index
 * arrays, src/cpu/shapes/mesh.h:17-72) plus poly::Mesh's transform and
 * material (mesh.h:77-113).  Matrices are row-major 4x4, as Matrix4x4. */
typedef struct ptx_mesh_desc {
   uint32_t num_vertices, num_faces;
   const float *x, *y, *z;               /* packed vertex positions             */
   const float *nx, *ny, *nz;            /* optional vertex normals (all NULL)  */
   const uint32_t *fv0, *fv1, *fv2;      /* per-face vertex indices             */
   float object_to_world[16];
   float world_to_object[16];            /* Transform::Inverse, used for normals */
   int32_t brdf;                         /* PTX_BRDF_LAMBERT, _MIRROR or _GLOSSY */
   float refl[3];                        /* Kd (Lambert, glossy) or Kr (mirror)  */
   int32_t is_light;                     /* Mesh::spd != nullptr                 */
   float emission[3];                    /* *Mesh::spd, 0-255 units              */
   float specular[3];                    /* glossy: Ks                           */
   float roughness;                      /* glossy: 0 = mirror-like .. 1        */
   /* Textured Lambert: LambertBRDF(texture) (lambert_brdf.h:31-56).  u, v are
    * mesh_geometry::u_packed / v_packed (mesh.h:56-57), num_uv entries each,
    * indexed by the face vertex indices.  texture = k > 0 takes the reflectance
    * from ptx_scene_desc.textures[k - 1] instead of refl; 0 (a zeroed
    * descriptor) = untextured, u and v ignored. */
   const float *u, *v;
   uint32_t num_uv;
   uint32_t texture;
} ptx_mesh_desc;

/* An image texture, poly::texture (src/cpu/shading/texture.h:12-42) as the
 * parser decodes a "color"/"spectrum" imagemap (pbrt_parser.cpp:730-776):
 * RGBA8 rows top to bottom.  The texel of (u, v) is column (unsigned)(u * W),
 * row (unsigned)((1 - v) * H), reflectance = byte * (1.0f / 255.0f) per RGB
 * channel (spectrum.h:17-18).  Coordinates outside the image clamp to the
 * nearest edge texel, where the reference reads outside its texel array. */
typedef struct ptx_texture_desc {
   uint32_t width, height;               /* 1 .. 32767 each                      */
   const uint8_t *rgba;                  /* width * height * 4 bytes             */
} ptx_texture_desc;

/* poly::compact_bvh_node, byte-compatible (src/cpu/acceleration/bvh.h:24-53). */
typedef struct ptx_bvh_node {
   float p0[3], p1[3];
   uint32_t offset;        /* interior: high_child_offset; leaf: face_index_offset */
   uint16_t num_faces;     /* 0 = interior                                         */
   uint16_t flags;         /* split axis                                           */
} ptx_bvh_node;

/* Point light for next-event estimation (PTX_JOB_NEE).  The reference parses
 * LightSource "point" but never uses it (pbrt_parser.cpp:1348), so scenes
 * rendered for parity carry none; intensity is in the same 0-255 units as
 * emission and the skybox. */
typedef struct ptx_point_light {
   float position[3];
   float intensity[3];
} ptx_point_light;

typedef struct ptx_scene_desc {
   const ptx_mesh_desc *meshes;
   uint32_t num_meshes;
   /* Optional prebuilt reference BVH (compact nodes + leaf (mesh, face) pairs,
    * 2 x u32 each).  NULL: the library builds it on the host with a builder
    * that takes the same decisions as bvh::bound + bvh::compact. */
   const ptx_bvh_node *nodes;
   uint32_t num_nodes;
   const uint32_t *leaf_pairs;
   uint32_t num_pairs;
   int32_t has_skybox;     /* Scene::Skybox != nullptr (ColorSkybox)  */
   float skybox[3];
   uint32_t build_threads; /* 0 = all host cores */
   const ptx_point_light *lights;   /* optional, for jobs with PTX_JOB_NEE */
   uint32_t num_lights;
   const ptx_texture_desc *textures;   /* ptx_mesh_desc.texture = k selects textures[k - 1] */
   uint32_t num_textures;
} ptx_scene_desc;

typedef struct ptx_scene_info {
   uint32_t num_nodes, num_pairs, num_interior, num_leaves, max_depth;
   uint64_t device_bytes;
   double build_seconds;   /* host BVH build + layout, 0 for a prebuilt BVH */
   double upload_seconds;
} ptx_scene_info;

/* PerspectiveCamera (src/cpu/cameras/PerspectiveCamera.cpp:17-49) with the
 * constructor's derived values precomputed by the caller on the host. */
typedef struct ptx_camera_desc {
   float camera_to_world[16];
   float fov;
   uint32_t width, height;
   float one_over_width, one_over_height, aspect_ratio, tan_fov_over_2;
} ptx_camera_desc;

/* One render: CenterSampler, PathTraceIntegrator{MaxDepth}, BoxFilter film.
 * Lambert bounces draw u0, u1 from Philox4x32-10 with key (seed, 0x706F6C79)
 * and counter (y*W + x, sample, bounce, 0): u = (r >> 8) * 2^-24. */
/* ptx_job_desc.flags */
enum {
   /* Next-event estimation (no reference counterpart; the reference's shadow
    * connection is commented out, PathTraceIntegrator.cpp:97-130): at each
    * Lambert vertex, for each point light in order, l = normalize(light - P),
    * cos = n . l; when cos > 0 and a shadow ray (P, l, t < |light - P|) hits
    * nothing, direct += (throughput * Kd) * (I * (cos / (pi * |light - P|^2))).
    * The sample is its terminal radiance plus direct.  Off: the reference's
    * integrator exactly. */
   PTX_JOB_NEE = 1
};

typedef struct ptx_job_desc {
   ptx_camera_desc camera;
   uint32_t spp, max_depth, seed;
   uint32_t flags;         /* PTX_JOB_* */
} ptx_job_desc;

/* Closest hit of one ray.  face == 0xFFFFFFFF: miss (mesh also 0xFFFFFFFF). */
typedef struct ptx_hit {
   uint32_t mesh, face;
   float t, u, v, w;
} ptx_hit;

/* Shading frame the reference leaves in poly::Intersection after traversal
 * (src/cpu/shapes/mesh.cpp:366-461): offset Location, flipped normal,
 * tangents and the error bound. */
typedef struct ptx_hit_frame {
   float location[3], normal[3], tangent1[3], tangent2[3], error[3];
} ptx_hit_frame;

/* The reference's per-ray counters (Intersection::num_bb_hits, num_bb_misses,
 * num_triangle_isects; bvh.cpp:495, 502, 520). */
typedef struct ptx_ray_counters {
   uint32_t bb_hits, bb_misses, tri_isects;
} ptx_ray_counters;

const char *ptx_last_error(void);
const char *ptx_version(void);

/* Opens a gfx950 device.  The handle owns a stream-ordered memory pool that
 * keeps freed blocks for the library's next allocations, and the render
 * engine keeps finished jobs' films and primary records (up to 4 GB) for
 * later jobs; both are released on an allocation failure of the library and
 * when the handle ends. */
int ptx_device_open(int ordinal, ptx_device **out);
/* Ends the handle: no new scene, query, render or server on it.  Scenes and
 * servers made on it stay valid until destroyed (they keep the device state
 * they need); a second close is a no-op. */
void ptx_device_close(ptx_device *dev);

int ptx_scene_create(ptx_device *dev, const ptx_scene_desc *desc, ptx_scene **out);
void ptx_scene_destroy(ptx_scene *scene);
int ptx_scene_get_info(const ptx_scene *scene, ptx_scene_info *out);
/* Copies the reference-layout BVH the scene was built from (num_nodes nodes,
 * 2*num_pairs u32). Either pointer may be NULL. */
int ptx_scene_copy_bvh(const ptx_scene *scene, ptx_bvh_node *nodes, uint32_t *leaf_pairs);

/* Host-only BVH build, no device needed: the decisions of bvh::bound + bvh::compact
 * (src/cpu/acceleration/bvh.cpp:36-395, 652-698) over the meshes' world-space
 * faces.  Query sizes with nodes/leaf_pairs NULL, then call again with buffers
 * of *num_nodes nodes and 2 * *num_pairs u32.  threads 0 = all host cores. */
int ptx_bvh_build(const ptx_mesh_desc *meshes, uint32_t num_meshes, uint32_t threads,
                  ptx_bvh_node *nodes, uint32_t *num_nodes, uint32_t *leaf_pairs, uint32_t *num_pairs);

/* Closest hit for n rays (6 floats each: origin, direction; MinT = +inf),
 * host buffers (pageable memory is fine).  counters may be NULL.  Blocks until
 * the hits are in `hits`.  The copies go through 16 MiB of pinned host memory
 * and device buffers sized to the largest call so far, both kept per device
 * until ptx_device_close; up to four host threads copy while the DMA engine
 * runs.  Calls on one device are serialised. */
int ptx_intersect_closest(ptx_scene *scene, const float *rays, uint32_t n, ptx_hit *hits,
                          ptx_ray_counters *counters);
/* Same, device buffers already resident in HBM, on `stream` (hipStream_t; NULL =
 * the null stream).  counters and frames may be NULL.  Asynchronous. */
int ptx_intersect_closest_device(ptx_scene *scene, const float *d_rays, uint32_t n,
                                 ptx_hit *d_hits, ptx_ray_counters *d_counters,
                                 ptx_hit_frame *d_frames, void *stream);

/* Same as ptx_intersect_closest_device, but allocates nothing: no
 * hipMalloc, hipFree or allocating library call.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  The
 * traversal's work counters and stack spill live in a static device array of
 * 64 MiB per GPU, held from ptx_device_open until the process ends; the
 * launch grid shrinks to what it holds, and a scene whose BVH needs more than
 * one block per CU can get (deferred stacks deeper than the LDS stack by
 * about 64 levels) is refused with PTX_ERR_INVALID.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (on different
 * streams, or from different threads) race; serialise them on one stream.
 * With ray_windows = ON it steps over the scene's window records only when an
 * earlier ptx_intersect_closest* call built them (building allocates).
 * Asynchronous; results equal ptx_intersect_closest_device's bit for bit. */
int ptx_intersect_closest_device_noalloc(ptx_scene *scene, const float *d_rays, uint32_t n,
                                         ptx_hit *d_hits, ptx_ray_counters *d_counters,
                                         ptx_hit_frame *d_frames, void *stream);

/* Plain GPU render of one job: one launch per sample, each pixel's path in one
 * thread with every bounce inline (the launch shape of the reference's CUDA
 * path); each pixel's primary ray is traced once unless ptx_tuning.plain_primary
 * = RETRACE.  film_rgb = W*H*3 floats on the host (BoxFilter::Output values). */
int ptx_render(ptx_scene *scene, const ptx_job_desc *job, float *film_rgb);
/* Batched render: all jobs' samples are merged into rounds of path slots that
 * share the GPU.  A round is one persistent path-regeneration launch (every
 * lane takes the next path slot as soon as its path ends).  The wavefront form
 * (render_mode = WAVEFRONT, a comparison form) keeps the paths in queues and
 * replays one captured hipGraph per bounce.  films[i] = W_i*H_i*3 host floats
 * (distinct, non-overlapping buffers: they are written by several host threads
 * at once, through the pinned staging that ptx_intersect_closest uses). */
int ptx_render_batch(ptx_scene *const *scenes, const ptx_job_desc *jobs, uint32_t n_jobs,
                     float *const *films);

/* Film output (host, no device needed).  film_rgb = W*H*3 averaged floats as
 * ptx_render returns them.  8-bit conversion follows PNGFilm::Output
 * (src/cpu/films/PNGFilm.cpp:20-58): min(v, 255) truncated, alpha 255; a NaN
 * channel becomes 0.  The PNG is 8-bit RGBA; the PFM holds the floats, rows
 * bottom to top, little-endian. */
int ptx_film_to_rgba8(const float *film_rgb, uint32_t width, uint32_t height, uint8_t *rgba);
int ptx_film_write_png(const float *film_rgb, uint32_t width, uint32_t height, const char *path);
int ptx_film_write_pfm(const float *film_rgb, uint32_t width, uint32_t height, const char *path);

/* Round form of ptx_render_batch and the server (ptx_tuning.render_mode). */
enum { PTX_RENDER_AUTO = 0, PTX_RENDER_REGEN = 1, PTX_RENDER_WAVEFRONT = 2 };

/* ptx_tuning.plain_primary */
enum { PTX_PLAIN_DEFAULT = 0, PTX_PLAIN_CACHED = 1, PTX_PLAIN_RETRACE = 2 };

/* ptx_tuning.ray_windows */
enum { PTX_WINDOWS_DEFAULT = 0, PTX_WINDOWS_ON = 1, PTX_WINDOWS_OFF = 2 };

/* Tuning knobs shared by ptx_render, ptx_render_batch and the server; zero = default. */
typedef struct ptx_tuning {
   uint32_t paths_per_wave;  /* path slots per round                                  */
   uint32_t block_size;      /* threads per block of the traversal kernel (256)       */
   uint32_t lds_stack;       /* LDS stack entries per lane: 8, 12, 16 or 24           */
   uint32_t refill_lanes;    /* ray queries: idle lanes that trigger a wave's ray     */
                             /* refill (1..64; 0 = 4)                                */
   uint32_t blocks_per_cu;   /* persistent blocks per CU; 0 = occupancy limit         */
   uint32_t render_mode;     /* PTX_RENDER_*: AUTO and REGEN = path regeneration     */
                             /* (rounds of at most 64 chunks; a batch of more than  */
                             /* 64 scenes runs as consecutive groups of at most 64); */
                             /* WAVEFRONT = path queues, kept as a comparison form   */
   uint32_t shade_lanes;     /* regeneration: lanes of a wave that must wait for a    */
                             /* shading step before the wave shades (1..64; 0 = 16)  */
   uint32_t leaf_lanes;      /* traversal: lanes of a wave that must reach a leaf     */
                             /* before the wave runs a leaf step (1..64; 0 = 1 for   */
                             /* ray queries, 16 for rendering)                       */
   uint32_t plain_primary;   /* ptx_render, PTX_PLAIN_*: CACHED (default) traces each */
                             /* pixel's primary ray once and starts every sample     */
                             /* from it; RETRACE traces it again for every sample,   */
                             /* as the reference's CUDA path does.  Same films.      */
   uint32_t ray_windows;     /* closest-hit queries, PTX_WINDOWS_*: OFF (default)     */
                             /* steps over one node record at a time; ON steps over  */
                             /* window records (a node's record and its first-       */
                             /* visited child's in one 128-byte block, 256 B per     */
                             /* interior node, built by the scene's first query      */
                             /* with ON) once the ray queue is drained.  Same hits   */
                             /* and counters; ON is slower on MI355X as measured.    */
} ptx_tuning;
int ptx_set_tuning(ptx_device *dev, const ptx_tuning *t);
int ptx_get_tuning(ptx_device *dev, ptx_tuning *t);
/* Times candidate tunings on `scene` and keeps the fastest. */
int ptx_autotune(ptx_device *dev, ptx_scene *scene, ptx_tuning *chosen);

/* ---------------------------------------------------------------- serving */
enum { PTX_JOB_QUEUED = 0, PTX_JOB_RUNNING = 1, PTX_JOB_DONE = 2, PTX_JOB_FAILED = 3,
       PTX_JOB_CANCELLED = 4 };

typedef struct ptx_server_config {
   uint64_t hbm_budget_bytes;    /* films + path state of admitted jobs      */
   uint32_t max_jobs_in_flight;  /* 0 = unlimited (budget only)              */
   uint32_t paths_per_wave;      /* path slots per round; 0 = the device     */
                                 /* tuning's paths_per_wave if set, else     */
                                 /* 2^27 or what a quarter of the budget     */
                                 /* pays for (12 B per slot), if fewer       */
} ptx_server_config;

typedef struct ptx_job_status {
   int32_t state;                /* PTX_JOB_*                                */
   int32_t error;                /* PTX_* error for FAILED                   */
   double submit_time, start_time, end_time;  /* seconds, server clock       */
   uint64_t samples_done, samples_total;
} ptx_job_status;

int ptx_server_create(ptx_device *dev, const ptx_server_config *cfg, ptx_server **out);
void ptx_server_destroy(ptx_server *srv);
/* Queues a job.  The job holds a reference to the scene until it leaves the
 * server, so the caller may destroy its scene handle right after submitting.
 * Rejects with PTX_ERR_BUDGET a job that
 * could never fit the budget. */
int ptx_server_submit(ptx_server *srv, ptx_scene *scene, const ptx_job_desc *job, uint64_t *job_id);
int ptx_server_status(ptx_server *srv, uint64_t job_id, ptx_job_status *out);
/* Blocks until the job ends; copies its film (W*H*3 floats) if film_rgb != NULL
 * and releases the job. */
int ptx_server_wait(ptx_server *srv, uint64_t job_id, float *film_rgb, ptx_job_status *out);
int ptx_server_cancel(ptx_server *srv, uint64_t job_id);

#ifdef __cplusplus
}
#endif
#endif /* PTX_H */
