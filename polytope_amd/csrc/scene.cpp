// Scene ingestion on the host and its device image.
//
// ptx_scene_create's counterpart in the reference is GPUMemoryManager::MallocScene
// (src/cuda/gpu_memory_manager.cpp:19-277), which copies per-mesh SoA vertices,
// the compact BVH and the leaf index pairs to the device and makes every face
// test re-transform its vertices.  Here the host transforms each face's
// vertices once, with the reference's ApplyInPlace arithmetic, and lays the
// scene out for the traversal kernel:
//   * 64-byte interior records holding both child boxes, child references and
//     leaf face counts (Node64),
//   * world-space triangles (36 B) in the reference's leaf order, so a leaf's
//     faces are contiguous,
//   * the (mesh, face) id of every leaf-ordered triangle, read once per ray.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <thread>

#include "scene.h"

namespace ptx {

namespace {
inline float sop(float a, float b, float c, float d) {      // src/cpu/constants.h:128-133
   const float cd = c * d;
   const float s = std::fma(a, b, cd);
   const float e = std::fma(c, d, -cd);
   return s + e;
}
}  // namespace

void world_vertices(const ptx_mesh_desc &m, WorldMesh &out, bool shading) {
   out.nf = m.num_faces;
   out.wv.resize(size_t(9) * m.num_faces);
   const float *M = m.object_to_world;
   const float *I = m.world_to_object;
   out.has_normals = m.nx && m.ny && m.nz;
   if (out.has_normals) out.wn.resize(size_t(9) * m.num_faces);
   auto work = [&](size_t lo, size_t hi) {
      for (size_t f = lo; f < hi; f++) {
         const uint32_t ix[3] = {m.fv0[f], m.fv1[f], m.fv2[f]};
         for (int c = 0; c < 3; c++) {
            const float x = m.x[ix[c]], y = m.y[ix[c]], z = m.z[ix[c]];
            // Transform::ApplyInPlace(Point&) (src/cpu/structures/Transform.cpp:69-91)
            const float nx = sop(x, M[0], y, M[1]) + std::fmaf(z, M[2], M[3]);
            const float ny = sop(x, M[4], y, M[5]) + std::fmaf(z, M[6], M[7]);
            const float nz = sop(x, M[8], y, M[9]) + std::fmaf(z, M[10], M[11]);
            const float w = sop(x, M[12], y, M[13]) + std::fmaf(z, M[14], M[15]);
            float *o = &out.wv[9 * f + 3 * c];
            if (w == 1) { o[0] = nx; o[1] = ny; o[2] = nz; }
            else {
               const float d = 1.f / w;
               o[0] = nx * d; o[1] = ny * d; o[2] = nz * d;
            }
            if (out.has_normals) {
               // Transform::ApplyInPlace(Normal&) (Transform.cpp:135-143)
               const float a = m.nx[ix[c]], b = m.ny[ix[c]], cc = m.nz[ix[c]];
               float *q = &out.wn[9 * f + 3 * c];
               q[0] = a * I[0] + b * I[4] + cc * I[8];
               q[1] = a * I[1] + b * I[5] + cc * I[9];
               q[2] = a * I[2] + b * I[6] + cc * I[10];
            }
         }
      }
   };
   const size_t nf = m.num_faces;
   const unsigned nt = nf > 200000 ? std::max(1u, std::thread::hardware_concurrency()) : 1u;
   if (nt == 1) work(0, nf);
   else {
      std::vector<std::thread> ts;
      for (unsigned t = 0; t < nt; t++) ts.emplace_back(work, nf * t / nt, nf * (t + 1) / nt);
      for (auto &t : ts) t.join();
   }
   out.rec.brdf = m.brdf;
   std::copy(m.refl, m.refl + 3, out.rec.refl);
   out.rec.is_light = m.is_light;
   std::copy(m.emission, m.emission + 3, out.rec.emission);
   if (m.brdf == PTX_BRDF_GLOSSY && !m.is_light) {
      // GlossyBRDF::sample's reflectance Ks (1 - r) + Kd r (glossy_brdf.h:52) is
      // constant per mesh; the roughness rides in the unused emission slot
      const float r = m.roughness;
      for (int c = 0; c < 3; c++) out.rec.refl[c] = m.specular[c] * (1 - r) + m.refl[c] * r;
      out.rec.emission[0] = r;
   }
   out.rec.flags = out.has_normals ? kRecNormals : 0u;
   out.texture = -1;
   out.uv.clear();
   if (shading && m.texture > 0 && m.brdf == PTX_BRDF_LAMBERT && !m.is_light && m.u && m.v) {
      // LambertBRDF(texture) (lambert_brdf.h:31-56): the face's corner texture
      // coordinates u_packed / v_packed[fv0, fv1, fv2] (mesh.cpp:342-351)
      out.texture = (int32_t)m.texture - 1;
      out.rec.flags |= kRecTextured;
      out.uv.resize(size_t(6) * m.num_faces);
      for (size_t f = 0; f < m.num_faces; f++) {
         const uint32_t ix[3] = {m.fv0[f], m.fv1[f], m.fv2[f]};
         for (int c = 0; c < 3; c++) {
            out.uv[6 * f + 2 * c] = m.u[ix[c]];
            out.uv[6 * f + 2 * c + 1] = m.v[ix[c]];
         }
      }
   }
}

int layout_scene(const std::vector<WorldMesh> &meshes, const HostBvh &bvh, HostImage &img) {
   const size_t nn = bvh.nodes.size();
   const size_t np = bvh.pairs.size() / 2;
   if (nn == 0) return set_error(PTX_ERR_INVALID, "scene has no faces");
   std::vector<uint32_t> ref(nn);
   uint32_t n_int = 0, n_leaf = 0;
   for (size_t i = 0; i < nn; i++) {
      const ptx_bvh_node &b = bvh.nodes[i];
      if (b.num_faces > 0) {
         if (size_t(b.offset) + b.num_faces > np || b.offset >= kLeafBit)
            return set_error(PTX_ERR_INVALID, "BVH leaf range exceeds the leaf pair array");
         ref[i] = kLeafBit | b.offset;
         n_leaf++;
      } else {
         ref[i] = n_int++;
      }
   }
   img.nodes.assign(n_int, Node64{});
   for (size_t i = 0; i < nn; i++) {
      const ptx_bvh_node &b = bvh.nodes[i];
      if (b.num_faces > 0) continue;
      const size_t lo = i + 1, hi = i + b.offset;
      if (lo >= nn || hi >= nn || b.offset < 2 || b.flags > 2)
         return set_error(PTX_ERR_INVALID, "BVH interior node has an out-of-range child");
      Node64 &r = img.nodes[ref[i]];
      const ptx_bvh_node &c0 = bvh.nodes[lo], &c1 = bvh.nodes[hi];
      for (int k = 0; k < 3; k++) {
         r.box[k] = c0.p0[k]; r.box[3 + k] = c0.p1[k];
         r.box[6 + k] = c1.p0[k]; r.box[9 + k] = c1.p1[k];
      }
      r.ref[0] = ref[lo];
      r.ref[1] = ref[hi];
      // a leaf child's slot holds its face count, an interior child's its split
      // axis (which the closest-hit kernel's window records are selected by)
      const uint32_t k0 = c0.num_faces ? (uint32_t)c0.num_faces : (uint32_t)c0.flags;
      const uint32_t k1 = c1.num_faces ? (uint32_t)c1.num_faces : (uint32_t)c1.flags;
      r.counts = k0 | (k1 << 16);
      r.axis = b.flags;
   }
   // deepest chain of interior nodes bounds the deferred stack
   {
      std::vector<std::pair<uint32_t, uint32_t>> st = {{0u, 0u}};
      uint32_t depth = 0;
      while (!st.empty()) {
         const auto [i, d] = st.back();
         st.pop_back();
         const ptx_bvh_node &b = bvh.nodes[i];
         if (b.num_faces > 0) continue;
         depth = std::max(depth, d + 1);
         st.push_back({i + 1, d + 1});
         st.push_back({i + b.offset, d + 1});
      }
      img.max_stack = depth + 1;
   }
   img.root_ref = ref[0];
   img.root_count = bvh.nodes[0].num_faces;
   for (int k = 0; k < 3; k++) {
      img.root_box[k] = bvh.nodes[0].p0[k];
      img.root_box[3 + k] = bvh.nodes[0].p1[k];
   }
   // 32 floats of padding: a traversal step loads 72 bytes (Node64 form) or 128
   // bytes (window form) from a leaf's first face
   img.tris.assign(9 * np + 32, 0.f);
   img.num_tris = (uint32_t)np;
   bool any_normals = false;
   for (const auto &m : meshes) any_normals |= m.has_normals;
   if (any_normals) img.normals.assign(9 * np, 0.f);
   for (size_t p = 0; p < np; p++) {
      const uint32_t mi = bvh.pairs[2 * p], f = bvh.pairs[2 * p + 1];
      if (mi >= meshes.size() || f >= meshes[mi].nf)
         return set_error(PTX_ERR_INVALID, "leaf pair names a face that does not exist");
      std::copy(&meshes[mi].wv[9 * size_t(f)], &meshes[mi].wv[9 * size_t(f)] + 9, &img.tris[9 * p]);
      if (any_normals && meshes[mi].has_normals)
         std::copy(&meshes[mi].wn[9 * size_t(f)], &meshes[mi].wn[9 * size_t(f)] + 9, &img.normals[9 * p]);
   }
   img.meshes.resize(meshes.size());
   for (size_t i = 0; i < meshes.size(); i++) img.meshes[i] = meshes[i].rec;
   img.num_interior = n_int;
   img.num_leaves = n_leaf;
   return PTX_OK;
}

}  // namespace ptx
