// Internal declarations shared by the host C++ and HIP translation units of
// libptx_hip.so.  Not part of the C ABI (include/ptx.h).
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ptx.h"

namespace ptx {

// ---------------------------------------------------------------- device layout
//
// Interior node record, 64 bytes, one per interior node of the reference BVH in
// DFS order.  Child 0 is the reference's low child (node i+1), child 1 its high
// child (node i + high_child_offset).  A child reference with bit 31 set is a
// leaf and holds the index of its first triangle in the leaf-ordered triangle
// array (the reference's leaf_ordered_indices order); the leaf's face count (the
// reference's u16 num_faces) sits in `counts`.  Otherwise the reference is the
// index of the child's interior record.
//   f[0..5]   child 0 box p0.xyz, p1.xyz
//   f[6..11]  child 1 box p0.xyz, p1.xyz
//   u[12]     child 0 ref, u[13] child 1 ref
//   u[14]     child 0 slot | child 1 slot << 16: a leaf child's face count,
//             an interior child's split axis
//   u[15]     split axis
struct alignas(64) Node64 {
   float box[12];
   uint32_t ref[2];
   uint32_t counts;
   uint32_t axis;
};
static_assert(sizeof(Node64) == 64, "Node64 is 64 bytes");

constexpr uint32_t kLeafBit = 0x80000000u;

// Per-mesh shading record (9 words, read as one block by the shading step).
// Glossy meshes that are not lights keep the lobe's constant reflectance in
// refl and its roughness in emission[0].  A textured Lambert mesh
// (kRecTextured) takes its reflectance from its triangles' texture records.
constexpr uint32_t kRecNormals = 1u, kRecTextured = 2u;
struct MeshRec {
   int32_t brdf;
   float refl[3];
   int32_t is_light;
   float emission[3];
   uint32_t flags;           // kRecNormals | kRecTextured
};
static_assert(sizeof(MeshRec) == 36, "MeshRec layout");

// Texture record of one leaf-ordered triangle (32 bytes), present for every
// triangle of a scene with textured meshes: its corner texture coordinates,
// the texture's size and whether the triangle is textured, and the offset of
// the texture's first texel in the scene's texel pool, which follows the
// num_tris records in the same allocation.
constexpr uint32_t kTexTextured = 1u << 31, kTexNormals = 1u << 30;
struct TexTri {
   float uv[6];              // (u, v) of corners 0, 1, 2 (fv0, fv1, fv2)
   uint32_t wh;              // width | height << 15 | kTexNormals | kTexTextured
   uint32_t texel0;          // first texel of the texture in the pool
};
static_assert(sizeof(TexTri) == 32, "TexTri layout");

// Device view of one scene.  The first 48 bytes are what a traversal needs
// (loaded per lane by the render path's extend stage as three 16-byte loads).
struct DevScene {
   const Node64 *nodes;      // interior records
   const float *tris;        // 9 floats (world v0, v1, v2) per leaf-ordered triangle
   float root_box[6];
   uint32_t root_ref;        // kLeafBit|first triangle, or interior index
   uint32_t root_count;      // face count when the root is a leaf
   const uint32_t *pairs;    // (mesh, face) per leaf-ordered triangle
   const MeshRec *meshes;
   const float *normals;     // world corner normals, 9 floats per leaf-ordered triangle (or null)
   uint32_t num_tris;
   uint32_t max_stack;       // deepest deferred-stack use + 1 (tree depth bound)
   float sky[3];
   uint16_t num_lights;
   uint16_t has_sky;
   const float *lights;      // point lights, 6 floats each (position, intensity), or null
   const TexTri *tex_tris;   // texture records per leaf-ordered triangle, then the texel pool; or null
};
static_assert(sizeof(DevScene) == 112, "DevScene layout");

// Per-job constants of a render, device-resident in the job table.
struct JobDev {
   float cam[16];            // camera-to-world, row-major
   float oow, ooh, aspect, tanf2;
   uint32_t width, height, spp, max_depth;
   uint32_t seed, scene;     // RNG seed; index into the scene table
   uint32_t flags, pad;      // PTX_JOB_*
   float *film;              // W*H*3 running sums (BoxFilter order), then the average
   uint32_t *primary;        // W*H primary-hit records (16 words each)
};

// A contiguous block of path slots: samples [s0, s0 + k) of every pixel of a job.
// Slot of (pixel p, sample s) = base + (s - s0) * npix + p.
struct Chunk {
   uint32_t job, s0, k, npix, base;
};

// Largest number of chunks in a round of the path-regeneration kernel (its
// chunk table lives in LDS).
constexpr uint32_t kMaxRoundChunks = 64;

// ---------------------------------------------------------------- host-side scene

struct HostBvh {
   std::vector<ptx_bvh_node> nodes;       // reference compact layout
   std::vector<uint32_t> pairs;           // 2 per leaf entry (mesh, face)
};

struct WorldMesh {
   uint32_t nf = 0;
   std::vector<float> wv;                 // 9 floats per face (world v0, v1, v2)
   bool has_normals = false;
   std::vector<float> wn;                 // 9 floats per face (world corner normals)
   std::vector<float> uv;                 // textured: 6 floats per face (u, v of corners 0, 1, 2)
   int32_t texture = -1;                  // textured: index into the scene's texture table
   MeshRec rec{};
};

// Transform::ApplyInPlace(Point&) semantics (src/cpu/structures/Transform.cpp:69-91).
// shading: also the mesh record and a textured mesh's corner coordinates (whose
// indices the caller has checked against num_uv).
void world_vertices(const ptx_mesh_desc &m, WorldMesh &out, bool shading = true);

// Decision-identical re-implementation of bvh::bound + bvh::compact
// (src/cpu/acceleration/bvh.cpp:36-395, 652-698), multithreaded.
void build_reference_bvh(const std::vector<WorldMesh> &meshes, unsigned threads, HostBvh &out);

// Error reporting: per-thread message returned by ptx_last_error().
int set_error(int code, const std::string &msg);

}  // namespace ptx
