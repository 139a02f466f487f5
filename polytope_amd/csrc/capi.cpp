// extern "C" entry points of libptx_hip.so for devices, scenes and ray
// queries (declared in include/ptx.h); rendering and serving live in
// render_host.cpp and server.cpp.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>

#include "host_query.h"
#include "render_host.h"

namespace ptx {

namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const std::string &msg) {
   g_last_error = msg;
   return code;
}

const char *last_error_cstr() { return g_last_error.c_str(); }

double now_s() {
   return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

namespace {

template <class T>
int upload(ptx_scene *s, const std::vector<T> &v, const T **dst) {
   *dst = nullptr;
   if (v.empty()) return PTX_OK;
   void *p = nullptr;
   PTX_HIP_TRY(hipMalloc(&p, v.size() * sizeof(T)));
   s->allocations.push_back(p);
   PTX_HIP_TRY(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
   s->info.device_bytes += v.size() * sizeof(T);
   *dst = static_cast<const T *>(p);
   return PTX_OK;
}

}  // namespace

static_assert(sizeof(ptx_mesh_desc) == 280, "ptx_mesh_desc layout (Python and oracle mirrors)");
static_assert(sizeof(ptx_texture_desc) == 16, "ptx_texture_desc layout (Python mirror)");

void device_retain(ptx_device *dev) { dev->refs.fetch_add(1); }

void device_release(ptx_device *dev) {
   if (dev->refs.fetch_sub(1) == 1) delete dev;
}

hipError_t pool_malloc(const ptx_device *dev, void **p, size_t bytes, hipStream_t stream) {
   return dev->pool ? hipMallocFromPoolAsync(p, bytes, dev->pool, stream) : hipMallocAsync(p, bytes, stream);
}

void pool_trim(const ptx_device *dev) {
   if (dev->pool && hipMemPoolTrimTo(dev->pool, 0) != hipSuccess) (void)hipGetLastError();
}

int device_usable(const ptx_device *dev) {
   if (!dev) return set_error(PTX_ERR_INVALID, "NULL device");
   if (dev->closed.load()) return set_error(PTX_ERR_INVALID, "the device has been closed");
   return PTX_OK;
}

ptx_tuning tuning_of(const ptx_device *dev) {
   std::lock_guard<std::mutex> lock(dev->tuning_mu);
   return dev->tuning;
}

void set_tuning(ptx_device *dev, const ptx_tuning &t) {
   std::lock_guard<std::mutex> lock(dev->tuning_mu);
   dev->tuning = t;
}

uint32_t lds_depth_of(const ptx_tuning &t) { return t.lds_stack ? t.lds_stack : 12; }

uint32_t leaf_lanes_of(const ptx_tuning &t, bool render) {
   if (t.leaf_lanes) return std::min(64u, t.leaf_lanes);
   return render ? kDefaultRenderLeafLanes : kDefaultLeafLanes;
}

uint32_t closest_hit_grid(const ptx_device *dev, const ptx_tuning &t, uint32_t lds, bool extra, bool windows) {
   int per_cu = closest_hit_blocks_per_cu(lds, extra, windows);
   if (t.blocks_per_cu) per_cu = std::min<int>(per_cu, (int)t.blocks_per_cu);
   return (uint32_t)(dev->num_cus * per_cu);
}

size_t closest_hit_scratch_bytes(const ptx_scene *s, uint32_t grid, uint32_t lds) {
   const size_t spill_entries = s->d.max_stack > lds ? size_t(s->d.max_stack - lds) * grid * 256 : 0;
   return 256 + spill_entries * sizeof(uint4);
}

void scene_retain(ptx_scene *s) { s->refs.fetch_add(1); }

void scene_release(ptx_scene *s) {
   if (s->refs.fetch_sub(1) == 1) {
      ptx_device *dev = s->dev;
      delete s;
      device_release(dev);
   }
}

const Node64 *built_windows(ptx_scene *s, const ptx_tuning &t) {
   if (t.ray_windows != PTX_WINDOWS_ON) return nullptr;
   std::lock_guard<std::mutex> lock(s->windows_mu);
   return s->windows;
}

const Node64 *scene_windows(ptx_scene *s, const ptx_tuning &t) {
   if (t.ray_windows != PTX_WINDOWS_ON || s->info.num_interior == 0) return nullptr;
   std::lock_guard<std::mutex> lock(s->windows_mu);
   if (s->windows || s->windows_failed) return s->windows;
   // built on a stream of its own and waited for, so that launches on any
   // stream see the records
   const size_t bytes = size_t(4) * s->info.num_interior * sizeof(Node64);
   void *w = nullptr;
   hipStream_t st = nullptr;
   bool ok = hipMalloc(&w, bytes) == hipSuccess;
   ok = ok && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
   ok = ok && launch_build_windows(s->d.nodes, s->info.num_interior, static_cast<Node64 *>(w), st) == 0;
   ok = ok && hipStreamSynchronize(st) == hipSuccess;
   if (st) (void)hipStreamDestroy(st);
   if (!ok) {
      (void)hipGetLastError();
      if (w) (void)hipFree(w);
      s->windows_failed = true;
      return nullptr;
   }
   s->windows = static_cast<Node64 *>(w);
   return s->windows;
}

// Closest-hit launch with stack-spill scratch taken from the stream-ordered allocator.
int run_closest_hit(ptx_scene *s, const float *d_rays, uint32_t n, ptx_hit *d_hits, ptx_ray_counters *d_counters,
                    ptx_hit_frame *d_frames, hipStream_t stream, const ptx_tuning &t) {
   ptx_device *dev = s->dev;
   const uint32_t lds = lds_depth_of(t);
   const Node64 *win = n ? scene_windows(s, t) : nullptr;
   uint32_t grid = closest_hit_grid(dev, t, lds, d_counters != nullptr || d_frames != nullptr, win != nullptr);
   const uint32_t needed_blocks = (n + 255) / 256;
   grid = std::max(1u, std::min(grid, needed_blocks));
   void *scratch = nullptr;
   PTX_HIP_TRY(pool_malloc(dev, &scratch, closest_hit_scratch_bytes(s, grid, lds), stream));
   uint32_t *work = static_cast<uint32_t *>(scratch);
   uint4 *spill = reinterpret_cast<uint4 *>(static_cast<char *>(scratch) + 256);
   const uint32_t refill = refill_lanes_of(t);
   const uint32_t leaf = leaf_lanes_of(t);
   const int rc = launch_closest_hit(s->d, d_rays, n, d_hits, d_counters, d_frames, work, spill, grid, lds, refill, leaf,
                                     win, s->root_axis, stream);
   const hipError_t fe = hipFreeAsync(scratch, stream);
   if (rc != 0) return set_error(PTX_ERR_HIP, "closest-hit launch failed");
   if (fe != hipSuccess) return set_error(PTX_ERR_HIP, std::string("hipFreeAsync: ") + hipGetErrorString(fe));
   return PTX_OK;
}

}  // namespace ptx

ptx_device::~ptx_device() {
   // every allocation from the pool has been freed by now (scenes and servers
   // keep the device alive until they are gone)
   if (pool) {
      (void)hipSetDevice(ordinal);
      (void)hipMemPoolDestroy(pool);
   }
}

ptx_scene::~ptx_scene() {
   if ((!allocations.empty() || windows) && dev) (void)hipSetDevice(dev->ordinal);
   for (void *p : allocations) (void)hipFree(p);
   allocations.clear();
   if (windows) (void)hipFree(windows);
   windows = nullptr;
}

using namespace ptx;

extern "C" {

const char *ptx_last_error(void) { return last_error_cstr(); }

const char *ptx_version(void) { return "ptx_hip 0.1 gfx950"; }

int ptx_device_open(int ordinal, ptx_device **out) {
   if (!out) return set_error(PTX_ERR_INVALID, "out is NULL");
   *out = nullptr;
   int count = 0;
   if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
      return set_error(PTX_ERR_NO_DEVICE, "no HIP device");
   if (ordinal < 0 || ordinal >= count) return set_error(PTX_ERR_NO_DEVICE, "device ordinal out of range");
   hipDeviceProp_t prop;
   PTX_HIP_TRY(hipGetDeviceProperties(&prop, ordinal));
   std::string arch = prop.gcnArchName;
   if (arch.rfind("gfx950", 0) != 0)
      return set_error(PTX_ERR_NO_DEVICE, "device " + std::to_string(ordinal) + " is " + arch +
                                              ", this library is built for gfx950 only");
   PTX_HIP_TRY(hipSetDevice(ordinal));
   auto *d = new ptx_device();
   d->ordinal = ordinal;
   d->num_cus = prop.multiProcessorCount;
   d->arch = arch;
   if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      delete d;
      return set_error(PTX_ERR_HIP, "hipStreamCreate failed");
   }
   // set-up of the no-allocation entry point: this device's copy of its static
   // scratch (a failure here only disables that entry point)
   if (noalloc_scratch_address(ordinal, &d->noalloc_scratch) != 0) d->noalloc_scratch = nullptr;
   // the library's own pool keeps freed blocks (a failure here leaves the
   // default pool in use)
   hipMemPoolProps props{};
   props.allocType = hipMemAllocationTypePinned;
   props.handleTypes = hipMemHandleTypeNone;
   props.location.type = hipMemLocationTypeDevice;
   props.location.id = ordinal;
   if (hipMemPoolCreate(&d->pool, &props) == hipSuccess) {
      uint64_t keep = UINT64_MAX;
      if (hipMemPoolSetAttribute(d->pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess) {
         (void)hipMemPoolDestroy(d->pool);
         d->pool = nullptr;
      }
   } else {
      d->pool = nullptr;
   }
   (void)hipGetLastError();
   *out = d;
   return PTX_OK;
}

void ptx_device_close(ptx_device *dev) {
   if (!dev || dev->closed.exchange(true)) return;
   (void)hipSetDevice(dev->ordinal);
   {
      std::lock_guard<std::mutex> lock(dev->mu);
      dev->engine.reset();
      dev->host_query.reset();
      if (dev->stream) (void)hipStreamDestroy(dev->stream);
      dev->stream = nullptr;
   }
   device_release(dev);      // scenes and servers still alive keep the handle until they go
}

int ptx_scene_create(ptx_device *dev, const ptx_scene_desc *desc, ptx_scene **out) {
   if (!dev || !desc || !out) return set_error(PTX_ERR_INVALID, "NULL argument");
   *out = nullptr;
   if (int rc = device_usable(dev)) return rc;
   if (desc->num_meshes == 0 || !desc->meshes) return set_error(PTX_ERR_INVALID, "scene has no meshes");
   if (desc->num_lights && !desc->lights) return set_error(PTX_ERR_INVALID, "num_lights > 0 with NULL lights");
   if (desc->num_lights > 64) return set_error(PTX_ERR_INVALID, "at most 64 point lights");
   if (desc->num_textures && !desc->textures) return set_error(PTX_ERR_INVALID, "num_textures > 0 with NULL textures");
   for (uint32_t t = 0; t < desc->num_textures; t++) {
      const ptx_texture_desc &x = desc->textures[t];
      if (!x.rgba || x.width < 1 || x.height < 1 || x.width > 32767 || x.height > 32767)
         return set_error(PTX_ERR_INVALID, "texture " + std::to_string(t) + " needs 1..32767 x 1..32767 RGBA8 texels");
   }
   for (uint32_t i = 0; i < desc->num_meshes; i++) {
      const ptx_mesh_desc &m = desc->meshes[i];
      if (!m.x || !m.y || !m.z || !m.fv0 || !m.fv1 || !m.fv2)
         return set_error(PTX_ERR_INVALID, "mesh " + std::to_string(i) + " lacks vertex or face arrays");
      if (m.brdf != PTX_BRDF_LAMBERT && m.brdf != PTX_BRDF_MIRROR && m.brdf != PTX_BRDF_GLOSSY)
         return set_error(PTX_ERR_INVALID, "mesh " + std::to_string(i) + " has an unknown brdf");
      if (m.brdf == PTX_BRDF_GLOSSY && !std::isfinite(m.roughness))
         return set_error(PTX_ERR_INVALID, "mesh " + std::to_string(i) + " has a non-finite roughness");
      for (uint32_t f = 0; f < m.num_faces; f++)
         if (m.fv0[f] >= m.num_vertices || m.fv1[f] >= m.num_vertices || m.fv2[f] >= m.num_vertices)
            return set_error(PTX_ERR_INVALID, "mesh " + std::to_string(i) + " face " + std::to_string(f) +
                                                  " indexes past its vertices");
      if (m.texture > desc->num_textures)
         return set_error(PTX_ERR_INVALID, "mesh " + std::to_string(i) + " names a texture the scene lacks");
      if (m.texture > 0 && !m.is_light) {
         if (m.brdf != PTX_BRDF_LAMBERT)
            return set_error(PTX_ERR_INVALID, "mesh " + std::to_string(i) + ": only Lambert meshes take a texture");
         if (!m.u || !m.v)
            return set_error(PTX_ERR_INVALID, "mesh " + std::to_string(i) + " has a texture but no texture coordinates");
         for (uint32_t f = 0; f < m.num_faces; f++)
            if (m.fv0[f] >= m.num_uv || m.fv1[f] >= m.num_uv || m.fv2[f] >= m.num_uv)
               return set_error(PTX_ERR_INVALID, "mesh " + std::to_string(i) + " face " + std::to_string(f) +
                                                     " indexes past its texture coordinates");
      }
   }
   {
      // the texel pool is indexed by u32: each texture that a mesh uses, once
      std::vector<bool> used(desc->num_textures, false);
      uint64_t pool = 0;
      for (uint32_t i = 0; i < desc->num_meshes; i++) {
         const ptx_mesh_desc &m = desc->meshes[i];
         if (m.texture == 0 || m.is_light || used[m.texture - 1]) continue;
         used[m.texture - 1] = true;
         pool += uint64_t(desc->textures[m.texture - 1].width) * desc->textures[m.texture - 1].height;
      }
      if (pool > 0xFFFFFFFFull) return set_error(PTX_ERR_INVALID, "a scene's textures exceed 2^32 texels");
   }
   PTX_HIP_TRY(hipSetDevice(dev->ordinal));
   const double t0 = now_s();
   // a scene that fails part-way frees what it uploaded (~ptx_scene)
   auto s = std::make_unique<ptx_scene>();
   s->dev = dev;
   s->num_meshes = desc->num_meshes;
   std::vector<WorldMesh> meshes(desc->num_meshes);
   for (uint32_t i = 0; i < desc->num_meshes; i++) {
      world_vertices(desc->meshes[i], meshes[i]);
      s->has_glossy |= meshes[i].rec.brdf == PTX_BRDF_GLOSSY && !meshes[i].rec.is_light;
      s->has_texture |= (meshes[i].rec.flags & kRecTextured) != 0;
   }
   double build_s = 0;
   if (desc->nodes && desc->leaf_pairs) {
      s->bvh.nodes.assign(desc->nodes, desc->nodes + desc->num_nodes);
      s->bvh.pairs.assign(desc->leaf_pairs, desc->leaf_pairs + 2 * size_t(desc->num_pairs));
   } else {
      build_reference_bvh(meshes, desc->build_threads, s->bvh);
      build_s = now_s() - t0;
   }
   HostImage img;
   const int lrc = layout_scene(meshes, s->bvh, img);
   if (lrc != PTX_OK) return lrc;
   const double t1 = now_s();
   int rc;
   if ((rc = upload(s.get(), img.nodes, &s->d.nodes)) != PTX_OK) return rc;
   if ((rc = upload(s.get(), img.tris, &s->d.tris)) != PTX_OK) return rc;
   if ((rc = upload(s.get(), s->bvh.pairs, &s->d.pairs)) != PTX_OK) return rc;
   if ((rc = upload(s.get(), img.meshes, &s->d.meshes)) != PTX_OK) return rc;
   if (s->has_texture) {
      // one texture record per leaf-ordered triangle, then the texel pool (each
      // texture that a mesh uses, once), in one allocation
      const size_t nt = img.num_tris;
      std::vector<int64_t> first(desc->num_textures, -1);
      size_t pool = 0;
      for (uint32_t i = 0; i < desc->num_meshes; i++) {
         const int32_t t = meshes[i].texture;
         if (t < 0 || first[t] >= 0) continue;
         first[t] = (int64_t)pool;
         pool += size_t(desc->textures[t].width) * desc->textures[t].height;
      }
      if (pool > 0xFFFFFFFFull) return set_error(PTX_ERR_INVALID, "a scene's textures exceed 2^32 texels");
      std::vector<uint32_t> blob(nt * (sizeof(TexTri) / 4) + pool, 0u);
      TexTri *rec = reinterpret_cast<TexTri *>(blob.data());
      for (size_t k = 0; k < nt; k++) {
         const uint32_t mi = s->bvh.pairs[2 * k], face = s->bvh.pairs[2 * k + 1];
         const WorldMesh &m = meshes[mi];
         if (m.texture < 0) continue;
         const ptx_texture_desc &x = desc->textures[m.texture];
         std::copy(&m.uv[6 * size_t(face)], &m.uv[6 * size_t(face)] + 6, rec[k].uv);
         rec[k].wh = x.width | x.height << 15 | kTexTextured | (m.has_normals ? kTexNormals : 0u);
         rec[k].texel0 = (uint32_t)first[m.texture];
      }
      for (uint32_t t = 0; t < desc->num_textures; t++)
         if (first[t] >= 0)
            std::memcpy(&blob[nt * (sizeof(TexTri) / 4) + first[t]], desc->textures[t].rgba,
                        size_t(desc->textures[t].width) * desc->textures[t].height * 4);
      const uint32_t *dev_blob = nullptr;
      if ((rc = upload(s.get(), blob, &dev_blob)) != PTX_OK) return rc;
      s->d.tex_tris = reinterpret_cast<const TexTri *>(dev_blob);
   }
   if ((rc = upload(s.get(), img.normals, &s->d.normals)) != PTX_OK) return rc;
   if (desc->num_lights) {
      std::vector<float> lights(6 * (size_t)desc->num_lights);
      for (uint32_t i = 0; i < desc->num_lights; i++) {
         std::copy(desc->lights[i].position, desc->lights[i].position + 3, &lights[6 * i]);
         std::copy(desc->lights[i].intensity, desc->lights[i].intensity + 3, &lights[6 * i + 3]);
      }
      if ((rc = upload(s.get(), lights, &s->d.lights)) != PTX_OK) return rc;
   }
   s->d.num_lights = desc->num_lights;
   std::copy(img.root_box, img.root_box + 6, s->d.root_box);
   s->d.root_ref = img.root_ref;
   s->d.root_count = img.root_count;
   s->root_axis = (img.root_ref & kLeafBit) ? 0u : img.nodes[img.root_ref].axis;
   s->d.num_tris = img.num_tris;
   s->d.max_stack = img.max_stack;
   s->d.has_sky = desc->has_skybox;
   std::copy(desc->skybox, desc->skybox + 3, s->d.sky);
   s->info.num_nodes = (uint32_t)s->bvh.nodes.size();
   s->info.num_pairs = (uint32_t)(s->bvh.pairs.size() / 2);
   s->info.num_interior = img.num_interior;
   s->info.num_leaves = img.num_leaves;
   s->info.max_depth = img.max_stack;
   s->info.build_seconds = build_s;
   s->info.upload_seconds = now_s() - t1;
   device_retain(dev);
   *out = s.release();
   return PTX_OK;
}

void ptx_scene_destroy(ptx_scene *scene) {
   if (!scene) return;
   scene_release(scene);       // server jobs still holding the scene keep it until they leave
}

int ptx_scene_get_info(const ptx_scene *scene, ptx_scene_info *out) {
   if (!scene || !out) return set_error(PTX_ERR_INVALID, "NULL argument");
   *out = scene->info;
   return PTX_OK;
}

int ptx_scene_copy_bvh(const ptx_scene *scene, ptx_bvh_node *nodes, uint32_t *leaf_pairs) {
   if (!scene) return set_error(PTX_ERR_INVALID, "NULL scene");
   if (nodes) std::memcpy(nodes, scene->bvh.nodes.data(), scene->bvh.nodes.size() * sizeof(ptx_bvh_node));
   if (leaf_pairs) std::memcpy(leaf_pairs, scene->bvh.pairs.data(), scene->bvh.pairs.size() * sizeof(uint32_t));
   return PTX_OK;
}

int ptx_bvh_build(const ptx_mesh_desc *meshes, uint32_t num_meshes, uint32_t threads, ptx_bvh_node *nodes,
                  uint32_t *num_nodes, uint32_t *leaf_pairs, uint32_t *num_pairs) {
   if (!meshes || !num_meshes || !num_nodes || !num_pairs) return set_error(PTX_ERR_INVALID, "NULL argument");
   std::vector<WorldMesh> wm(num_meshes);
   for (uint32_t i = 0; i < num_meshes; i++) world_vertices(meshes[i], wm[i], false);
   HostBvh bvh;
   build_reference_bvh(wm, threads, bvh);
   const uint32_t nn = (uint32_t)bvh.nodes.size(), np = (uint32_t)(bvh.pairs.size() / 2);
   if (nodes || leaf_pairs) {
      if (*num_nodes < nn || *num_pairs < np) {
         *num_nodes = nn;
         *num_pairs = np;
         return set_error(PTX_ERR_INVALID, "output buffers too small");
      }
      if (nodes) std::memcpy(nodes, bvh.nodes.data(), size_t(nn) * sizeof(ptx_bvh_node));
      if (leaf_pairs) std::memcpy(leaf_pairs, bvh.pairs.data(), bvh.pairs.size() * sizeof(uint32_t));
   }
   *num_nodes = nn;
   *num_pairs = np;
   return PTX_OK;
}

int ptx_intersect_closest_device(ptx_scene *scene, const float *d_rays, uint32_t n, ptx_hit *d_hits,
                                 ptx_ray_counters *d_counters, ptx_hit_frame *d_frames, void *stream) {
   if (!scene || (n && (!d_rays || !d_hits))) return set_error(PTX_ERR_INVALID, "NULL argument");
   if (int rc = device_usable(scene->dev)) return rc;
   PTX_HIP_TRY(hipSetDevice(scene->dev->ordinal));
   return run_closest_hit(scene, d_rays, n, d_hits, d_counters, d_frames, static_cast<hipStream_t>(stream),
                          tuning_of(scene->dev));
}

int ptx_intersect_closest_device_noalloc(ptx_scene *scene, const float *d_rays, uint32_t n, ptx_hit *d_hits,
                                         ptx_ray_counters *d_counters, ptx_hit_frame *d_frames, void *stream) {
   if (!scene || (n && (!d_rays || !d_hits))) return set_error(PTX_ERR_INVALID, "NULL argument");
   ptx_device *dev = scene->dev;
   if (int rc = device_usable(dev)) return rc;
   if (!dev->noalloc_scratch)
      return set_error(PTX_ERR_HIP, "the no-allocation scratch of this device could not be located");
   const ptx_tuning t = tuning_of(dev);
   const uint32_t lds = lds_depth_of(t);
   // window records only when an earlier query built them (building allocates)
   const Node64 *win = built_windows(scene, t);
   // the spill takes (max_stack - lds) entries of 16 bytes per resident lane;
   // the grid shrinks to what the scratch holds, down to one block per CU
   const size_t per_block = scene->d.max_stack > lds ? size_t(scene->d.max_stack - lds) * 256 * sizeof(uint4) : 0;
   uint32_t grid = closest_hit_grid(dev, t, lds, d_counters != nullptr || d_frames != nullptr, win != nullptr);
   if (per_block) {
      const size_t fit = (kNoAllocScratchBytes - 256) / per_block;
      if (fit < (size_t)dev->num_cus)
         return set_error(PTX_ERR_INVALID,
                          "scene BVH too deep for the no-allocation scratch: " + std::to_string(scene->d.max_stack) +
                                " stack levels need " + std::to_string(per_block * dev->num_cus + 256) +
                                " bytes for one block per CU, the scratch holds " +
                                std::to_string(kNoAllocScratchBytes) + "; use ptx_intersect_closest_device");
      grid = (uint32_t)std::min<size_t>(grid, fit);
   }
   grid = std::max(1u, std::min(grid, (n + 255) / 256));
   PTX_HIP_TRY(hipSetDevice(dev->ordinal));
   uint32_t *work = static_cast<uint32_t *>(dev->noalloc_scratch);
   uint4 *spill = reinterpret_cast<uint4 *>(static_cast<char *>(dev->noalloc_scratch) + 256);
   const uint32_t refill = refill_lanes_of(t);
   // the work counters are cleared by a kernel at the head of the launch (no memset)
   if (launch_closest_hit(scene->d, d_rays, n, d_hits, d_counters, d_frames, work, spill, grid, lds, refill,
                          leaf_lanes_of(t), win, scene->root_axis, static_cast<hipStream_t>(stream)) != 0) {
      (void)hipGetLastError();
      return set_error(PTX_ERR_HIP, "closest-hit launch failed");
   }
   return PTX_OK;
}

int ptx_intersect_closest(ptx_scene *scene, const float *rays, uint32_t n, ptx_hit *hits,
                          ptx_ray_counters *counters) {
   if (!scene || (n && (!rays || !hits))) return set_error(PTX_ERR_INVALID, "NULL argument");
   if (int rc = device_usable(scene->dev)) return rc;
   if (n == 0) return PTX_OK;
   ptx_device *dev = scene->dev;
   PTX_HIP_TRY(hipSetDevice(dev->ordinal));
   std::lock_guard<std::mutex> lock(dev->mu);
   return intersect_host(scene, rays, n, hits, counters);
}

int ptx_set_tuning(ptx_device *dev, const ptx_tuning *t) {
   if (!dev || !t) return set_error(PTX_ERR_INVALID, "NULL argument");
   if (int rc = device_usable(dev)) return rc;
   if (t->lds_stack && t->lds_stack != 8 && t->lds_stack != 12 && t->lds_stack != 16 && t->lds_stack != 24)
      return set_error(PTX_ERR_INVALID, "lds_stack must be 0, 8, 12, 16 or 24");
   if (t->render_mode > PTX_RENDER_WAVEFRONT) return set_error(PTX_ERR_INVALID, "render_mode must be 0, 1 or 2");
   if (t->shade_lanes > 64) return set_error(PTX_ERR_INVALID, "shade_lanes must be 0..64");
   if (t->leaf_lanes > 64) return set_error(PTX_ERR_INVALID, "leaf_lanes must be 0..64");
   if (t->plain_primary > PTX_PLAIN_RETRACE) return set_error(PTX_ERR_INVALID, "plain_primary must be 0, 1 or 2");
   if (t->ray_windows > PTX_WINDOWS_OFF) return set_error(PTX_ERR_INVALID, "ray_windows must be 0, 1 or 2");
   set_tuning(dev, *t);
   return PTX_OK;
}

int ptx_get_tuning(ptx_device *dev, ptx_tuning *t) {
   if (!dev || !t) return set_error(PTX_ERR_INVALID, "NULL argument");
   *t = tuning_of(dev);
   if (!t->lds_stack) t->lds_stack = lds_depth_of(*t);
   if (!t->block_size) t->block_size = 256;
   if (!t->refill_lanes) t->refill_lanes = kDefaultRefillLanes;
   // leaf_lanes stays 0 when unset: each kernel family then uses its own default
   if (!t->shade_lanes) t->shade_lanes = RenderEngine::kDefaultShadeLanes;
   return PTX_OK;
}

}  // extern "C"
