// GpuRunner (see GpuRunner.h): translates the runner's scene description into
// the C ABI's descriptors, renders with ptx_render and hands the averaged film
// to the runner's film.
#include "GpuRunner.h"

#include <algorithm>
#include <cstring>
#include <vector>

#include "cpu/acceleration/bvh.h"
#include "cpu/cameras/PerspectiveCamera.h"
#include "cpu/integrators/PathTraceIntegrator.h"
#include "cpu/samplers/samplers.h"
#include "cpu/scenes/Scene.h"
#include "cpu/scenes/skyboxes/ColorSkybox.h"
#include "cpu/shading/brdf/lambert_brdf.h"
#include "cpu/shading/brdf/mirror_brdf.h"
#include "cpu/shading/texture.h"
#include "cpu/shapes/mesh.h"

namespace poly {

namespace {

void copy_matrix(const Transform &t, float out[16]) {
   for (int r = 0; r < 4; r++)
      for (int c = 0; c < 4; c++) out[4 * r + c] = t.Matrix.Matrix[r][c];
}

struct Handles {
   ptx_device *dev = nullptr;
   ptx_scene *scene = nullptr;
   ~Handles() {
      if (scene) ptx_scene_destroy(scene);
      if (dev) ptx_device_close(dev);
   }
};

}  // namespace

int GpuRunner::fail(int code, const std::string &msg) {
   status_ = code;
   error_ = msg;
   return code;
}

void GpuRunner::Run(int threadId) {
   if (threadId != 0) return;
   status_ = PTX_OK;
   error_.clear();

   // the description: sampler, integrator, camera (AbstractRunner.h:49-54)
   if (!dynamic_cast<const CenterSampler *>(Sampler.get())) {
      fail(PTX_ERR_INVALID, "GpuRunner renders with CenterSampler only (the sampler the pbrt parser builds)");
      return;
   }
   const auto *integrator = dynamic_cast<const PathTraceIntegrator *>(Integrator.get());
   if (!integrator) {
      fail(PTX_ERR_INVALID, "GpuRunner renders PathTraceIntegrator only");
      return;
   }
   const auto *camera = dynamic_cast<const PerspectiveCamera *>(Scene->Camera.get());
   if (!camera || !camera->LeftHanded) {
      fail(PTX_ERR_INVALID, "GpuRunner needs the left-handed PerspectiveCamera the pbrt parser builds");
      return;
   }

   // meshes: packed geometry, transforms, material, emission (mesh.h:17-113)
   std::vector<ptx_mesh_desc> meshes(Scene->Shapes.size());
   std::vector<ptx_texture_desc> textures;          // one per distinct poly::texture
   std::vector<const texture *> texture_of;
   for (size_t i = 0; i < Scene->Shapes.size(); i++) {
      const Mesh &m = *Scene->Shapes[i];
      const mesh_geometry &g = *m.mesh_geometry;
      ptx_mesh_desc &d = meshes[i];
      std::memset(&d, 0, sizeof(d));
      d.num_vertices = g.num_vertices_packed;
      d.num_faces = g.num_faces;
      d.x = g.x_packed.data();
      d.y = g.y_packed.data();
      d.z = g.z_packed.data();
      if (g.has_vertex_normals) {
         d.nx = g.nx_packed.data();
         d.ny = g.ny_packed.data();
         d.nz = g.nz_packed.data();
      }
      d.fv0 = g.fv0.data();
      d.fv1 = g.fv1.data();
      d.fv2 = g.fv2.data();
      copy_matrix(*m.object_to_world, d.object_to_world);
      copy_matrix(*m.world_to_object, d.world_to_object);
      const AbstractBRDF *brdf = m.material ? m.material->BRDF.get() : nullptr;
      if (const auto *lam = dynamic_cast<const LambertBRDF *>(brdf)) {
         d.brdf = PTX_BRDF_LAMBERT;
         d.refl[0] = lam->refl.r; d.refl[1] = lam->refl.g; d.refl[2] = lam->refl.b;
         if (lam->texture && !m.spd) {      // LambertBRDF(texture): texels and u/v_packed
            const texture *t = lam->texture.get();
            if (t->data_format != texture::format::RGBA) {
               fail(PTX_ERR_INVALID, "GpuRunner: mesh " + std::to_string(i) + " has a grey texture");
               return;
            }
            if (!g.has_vertex_uvs || g.u_packed.empty()) {
               fail(PTX_ERR_INVALID, "GpuRunner: mesh " + std::to_string(i) + " has a texture but no u, v");
               return;
            }
            size_t k = 0;
            while (k < texture_of.size() && texture_of[k] != t) k++;
            if (k == texture_of.size()) {
               texture_of.push_back(t);
               textures.push_back(ptx_texture_desc{t->width, t->height, t->data.data()});
            }
            d.texture = (uint32_t)k + 1;
            d.u = g.u_packed.data();
            d.v = g.v_packed.data();
            d.num_uv = (uint32_t)std::min(g.u_packed.size(), g.v_packed.size());
         }
      } else if (const auto *mir = dynamic_cast<const MirrorBRDF *>(brdf)) {
         d.brdf = PTX_BRDF_MIRROR;
         d.refl[0] = mir->refl.r; d.refl[1] = mir->refl.g; d.refl[2] = mir->refl.b;
      } else {
         fail(PTX_ERR_INVALID, "GpuRunner: mesh " + std::to_string(i) + " has a BRDF other than Lambert or mirror");
         return;
      }
      if (m.spd) {
         d.is_light = 1;
         d.emission[0] = m.spd->r; d.emission[1] = m.spd->g; d.emission[2] = m.spd->b;
      }
   }

   ptx_scene_desc sd;
   std::memset(&sd, 0, sizeof(sd));
   sd.meshes = meshes.data();
   sd.num_meshes = (uint32_t)meshes.size();
   sd.textures = textures.data();
   sd.num_textures = (uint32_t)textures.size();
   std::vector<uint32_t> pairs;
   if (use_scene_bvh_) {   // the tree bvh::bound + bvh::compact already made for this scene
      const compact_bvh *cb = Scene->bvh_root.compact_root;
      sd.nodes = reinterpret_cast<const ptx_bvh_node *>(cb->nodes);
      sd.num_nodes = cb->num_nodes;
      pairs.reserve(2 * cb->leaf_ordered_indices.size());
      for (const auto &p : cb->leaf_ordered_indices) {
         pairs.push_back(p.first);
         pairs.push_back(p.second);
      }
      sd.leaf_pairs = pairs.data();
      sd.num_pairs = (uint32_t)cb->leaf_ordered_indices.size();
   }
   if (Scene->Skybox) {
      if (!dynamic_cast<const ColorSkybox *>(Scene->Skybox.get())) {
         fail(PTX_ERR_INVALID, "GpuRunner: only ColorSkybox is supported");
         return;
      }
      const SpectralPowerDistribution sky = Scene->Skybox->GetSpd(Vector(0.f, 0.f, 1.f));
      sd.has_skybox = 1;
      sd.skybox[0] = sky.r; sd.skybox[1] = sky.g; sd.skybox[2] = sky.b;
   }

   ptx_job_desc job;
   std::memset(&job, 0, sizeof(job));
   copy_matrix(camera->CameraToWorld, job.camera.camera_to_world);
   job.camera.fov = camera->Settings.FieldOfView;
   job.camera.width = Bounds.x;
   job.camera.height = Bounds.y;
   job.camera.one_over_width = camera->OneOverWidth;
   job.camera.one_over_height = camera->OneOverHeight;
   job.camera.aspect_ratio = camera->AspectRatio;
   job.camera.tan_fov_over_2 = camera->TanFOVOver2;
   job.spp = NumSamples;
   job.max_depth = (uint32_t)integrator->MaxDepth;
   job.seed = seed_;

   Handles h;
   std::vector<float> rgb((size_t)Bounds.x * Bounds.y * 3);
   int rc = ptx_device_open(device_, &h.dev);
   if (rc == PTX_OK) rc = ptx_scene_create(h.dev, &sd, &h.scene);
   if (rc == PTX_OK) rc = ptx_render(h.scene, &job, rgb.data());
   if (rc != PTX_OK) {
      fail(rc, ptx_last_error());
      return;
   }

   // one sample per pixel holding the GPU's average: BoxFilter::Output returns
   // it unchanged (sum of one sample times 1/1), so Output() writes the image
   for (unsigned int y = 0; y < Bounds.y; y++)
      for (unsigned int x = 0; x < Bounds.x; x++) {
         const float *p = &rgb[((size_t)y * Bounds.x + x) * 3];
         Film->AddSample(Point2f((float)x, (float)y), Sample(SpectralPowerDistribution(p[0], p[1], p[2])));
      }
}

}  // namespace poly
