// ORACLE — CPU restatement of polytope's path-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product in
// polytope_amd/: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
// bench.py may load it.  Nothing in the product links or calls it.
//
// Every function restates the reference algorithm with the same float
// operations in the same order (built with -ffp-contract=off; fused operations
// appear only where the reference calls std::fma/fmaf):
//   world vertices       Transform::ApplyInPlace(Point&)        src/cpu/structures/Transform.cpp:69-91
//   BVH build            bvh::bound                             src/cpu/acceleration/bvh.cpp:36-395
//   BVH flatten          bvh::compact / compact_helper          src/cpu/acceleration/bvh.cpp:652-698
//   traversal            bvh::intersect_compact                 src/cpu/acceleration/bvh.cpp:476-523
//   box test             BoundingBox::Hits                      src/cpu/structures/BoundingBox.cpp:12-85
//   triangle test        Mesh::intersect (test part)            src/cpu/shapes/mesh.cpp:161-355
//   hit frame            Mesh::intersect (post-hit part)        src/cpu/shapes/mesh.cpp:357-462
//   camera               PerspectiveCamera::get_ray_for_pixel   src/cpu/cameras/PerspectiveCamera.cpp:27-49
//   integrator           PathTraceIntegrator::GetSample         src/cpu/integrators/PathTraceIntegrator.cpp:10-135
//   local frame          Intersection::WorldToLocal/LocalToWorld src/cpu/structures/Intersection.cpp:10-23
//   BRDFs                LambertBRDF/AbstractBRDF/MirrorBRDF::sample
//                        src/cpu/shading/brdf/{lambert_brdf.h:48-57,abstract_brdf.h:36-45,mirror_brdf.h:23-35}
//   textures             Mesh::intersect's u/v_tex_lerp (mesh.cpp:342-353), texture::evaluate_rgba
//                        (src/cpu/shading/texture.h:25-31), ReflectanceSpectrum(int,int,int) (spectrum.h:17-18)
//                        GlossyBRDF::sample (the "plastic" material)  src/cpu/shading/brdf/glossy_brdf.h:27-60
//   box filter           BoxFilter::AddSample/Output            src/cpu/filters/BoxFilter.cpp:14-71
//   runner               TileRunner::Run 2x2 tiles              src/cpu/runners/TileRunner.cpp:61-117
//
// Deliberate, documented differences from the reference:
//   * RNG: the reference draws from an unseeded thread-local mt19937
//     (src/cpu/constants.h:86-104), which no run can reproduce.  Both this oracle
//     and the product use Philox4x32-10 (Salmon et al., SC'11) keyed by the job
//     seed, counter (pixel, sample, bounce, 0); u = (r >> 8) * 2^-24.
//   * sin/cos in CosineSampleHemisphere (src/cpu/constants.h:106-115) use a
//     project-owned polynomial (ptx_sincos below) so CPU and GPU agree bitwise.
//   * Texels: where texture::evaluate_rgba's flat index falls outside the
//     image, the reference reads outside its texel array; here the column and
//     row clamp to the image (texel_index below).
//   * Post-hit frame: the reference re-runs the post-hit block after every face
//     test once a hit exists, with the mesh of the face just tested and
//     re-adding the origin offset each time (mesh.cpp:357-462).  Here the frame
//     uses the hit mesh and the offset is applied k times, k = number of face
//     tests from the accepting test to the end of traversal: identical to the
//     reference for single-mesh scenes and for multi-mesh scenes whenever the
//     faces tested after the hit belong to the hit mesh.
//
// Parity pins (tests/): test_oracle.py (reference golden fixtures from the
// compiled reference, the reference's own unit-test cases, closed forms,
// brute force, Philox known-answer vectors), test_glossy.py (the reference's
// GlossyBRDF::sample on its own recorded draws), test_textures.py (the
// reference's texture lookups and films, golden tex_probe_crafted.npz),
// test_pbrt_ingest.py / test_pbrt_reader.py (films and trees from the
// reference's parser).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#pragma STDC FP_CONTRACT OFF

namespace orc {

constexpr float kInf = std::numeric_limits<float>::infinity();
constexpr float kFloatMax = std::numeric_limits<float>::max();
// src/cpu/constants.h:247, 320-333
constexpr float kMachineEpsilon = std::numeric_limits<float>::epsilon() * 0.5f;
constexpr float gamma_n(int n) { return (n * kMachineEpsilon) / (1.f - n * kMachineEpsilon); }
constexpr float kGamma2 = (2 * kMachineEpsilon) / (1 - 2 * kMachineEpsilon);
constexpr float kGamma3 = (3 * kMachineEpsilon) / (1 - 3 * kMachineEpsilon);
constexpr float kGamma5 = (5 * kMachineEpsilon) / (1 - 5 * kMachineEpsilon);
constexpr float kGamma7 = (7 * kMachineEpsilon) / (1 - 7 * kMachineEpsilon);
constexpr float kPI = float(3.14159265358979323846);
constexpr float kTwoPI = 2.0f * kPI;

struct V3 { float x, y, z; };
static inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline float comp(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
static inline float& comp_ref(V3& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
// Vector::Dot(v) = v.x*x + v.y*y + v.z*z (src/cpu/structures/Vectors.cpp:44-50)
static inline float dot(V3 self, V3 v) { return v.x * self.x + v.y * self.y + v.z * self.z; }
// Vector::Cross (src/cpu/structures/Vectors.cpp:101-111)
static inline V3 cross(V3 a, V3 r) {
   return {a.y * r.z - a.z * r.y, a.z * r.x - a.x * r.z, a.x * r.y - a.y * r.x};
}
// Vector::Normalize / Normal::Normalize (src/cpu/structures/Vectors.cpp:11-16, 94-99)
static inline V3 normalize(V3 v) {
   const float ld = 1.0f / std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
   return {v.x * ld, v.y * ld, v.z * ld};
}

// src/cpu/constants.h:121-133
static inline float difference_of_products(float a, float b, float c, float d) {
   float cd = c * d;
   float dop = std::fma(a, b, -cd);
   float error = std::fma(-c, d, cd);
   return dop + error;
}
static inline float sum_of_products(float a, float b, float c, float d) {
   float cd = c * d;
   float sop = std::fma(a, b, cd);
   float error = std::fma(c, d, -cd);
   return sop + error;
}

// Transform::ApplyInPlace(Point&) (src/cpu/structures/Transform.cpp:69-91); m row-major.
static inline V3 apply_point_inplace(const float* m, V3 p) {
   const float nx = sum_of_products(p.x, m[0], p.y, m[1]) + std::fmaf(p.z, m[2], m[3]);
   const float ny = sum_of_products(p.x, m[4], p.y, m[5]) + std::fmaf(p.z, m[6], m[7]);
   const float nz = sum_of_products(p.x, m[8], p.y, m[9]) + std::fmaf(p.z, m[10], m[11]);
   const float w = sum_of_products(p.x, m[12], p.y, m[13]) + std::fmaf(p.z, m[14], m[15]);
   if (w == 1) return {nx, ny, nz};
   const float div = 1.f / w;
   return {nx * div, ny * div, nz * div};
}
// Transform::Apply(const Point&) (Transform.cpp:93-111): plain mul/add.
static inline V3 apply_point(const float* m, V3 p) {
   const float nx = p.x * m[0] + p.y * m[1] + p.z * m[2] + m[3];
   const float ny = p.x * m[4] + p.y * m[5] + p.z * m[6] + m[7];
   const float nz = p.x * m[8] + p.y * m[9] + p.z * m[10] + m[11];
   const float w = p.x * m[12] + p.y * m[13] + p.z * m[14] + m[15];
   if (w == 1) return {nx, ny, nz};
   const float div = 1.f / w;
   return {nx * div, ny * div, nz * div};
}
// Transform::Apply(const Vector&) (Transform.cpp:123-133)
static inline V3 apply_vector(const float* m, V3 v) {
   return {v.x * m[0] + v.y * m[1] + v.z * m[2],
           v.x * m[4] + v.y * m[5] + v.z * m[6],
           v.x * m[8] + v.y * m[9] + v.z * m[10]};
}
// Transform::ApplyInPlace(Normal&) (Transform.cpp:135-143): uses the inverse, transposed.
static inline V3 apply_normal(const float* inv, V3 n) {
   return {n.x * inv[0] + n.y * inv[4] + n.z * inv[8],
           n.x * inv[1] + n.y * inv[5] + n.z * inv[9],
           n.x * inv[2] + n.y * inv[6] + n.z * inv[10]};
}

struct Box { V3 p0, p1; };
// BoundingBox::UnionInPlace(const BoundingBox&) (src/cpu/structures/BoundingBox.cpp:136-144)
static inline void union_in_place(Box& a, const Box& b) {
   a.p0.x = a.p0.x < b.p0.x ? a.p0.x : b.p0.x;
   a.p0.y = a.p0.y < b.p0.y ? a.p0.y : b.p0.y;
   a.p0.z = a.p0.z < b.p0.z ? a.p0.z : b.p0.z;
   a.p1.x = a.p1.x > b.p1.x ? a.p1.x : b.p1.x;
   a.p1.y = a.p1.y > b.p1.y ? a.p1.y : b.p1.y;
   a.p1.z = a.p1.z > b.p1.z ? a.p1.z : b.p1.z;
}
// BoundingBox::surface_area (BoundingBox.cpp:146-152)
static inline float surface_area(const Box& b) {
   const float x = b.p1.x - b.p0.x;
   const float y = b.p1.y - b.p0.y;
   const float z = b.p1.z - b.p0.z;
   return (2 * x * y) + (2 * y * z) + (2 * z * x);
}

// BoundingBox::Hits (BoundingBox.cpp:12-85), counters removed.
static inline bool box_hits(const Box& b, V3 o, V3 inv, float min_t) {
   const bool inside = o.x > b.p0.x && o.x < b.p1.x && o.y > b.p0.y && o.y < b.p1.y &&
                       o.z > b.p0.z && o.z < b.p1.z;
   if (inside) return true;
   float max_far = min_t;
   float min_near = 0;
   const float gm = 1 + 2 * gamma_n(3);
   for (int a = 0; a < 3; a++) {
      float t_near = (comp(b.p0, a) - comp(o, a)) * comp(inv, a);
      float t_far = (comp(b.p1, a) - comp(o, a)) * comp(inv, a);
      const float swap = t_near;
      t_near = t_near > t_far ? t_far : t_near;
      t_far = swap > t_far ? swap : t_far;
      t_far *= gm;
      min_near = (t_near > min_near) ? t_near : min_near;
      max_far = (t_far < max_far) ? t_far : max_far;
      if (a < 2 && min_near > max_far) return false;
   }
   return min_near <= max_far;
}

// ----------------------------------------------------------------------------- scene

struct MeshIn {            // mirrors polytope_amd's ptx_mesh_desc (C layout, see include/ptx.h)
   uint32_t num_vertices, num_faces;
   const float *x, *y, *z;
   const float *nx, *ny, *nz;        // optional vertex normals (all null if absent)
   const uint32_t *fv0, *fv1, *fv2;
   float object_to_world[16];
   float world_to_object[16];        // Transform::Inverse of object_to_world (used for normals)
   int32_t brdf;                     // 1 lambert, 2 mirror
   float refl[3];
   int32_t is_light;
   float emission[3];
   float specular[3];                // glossy: Ks (refl holds Kd)
   float roughness;                  // glossy: 0 specular .. 1 diffuse
   const float *u, *v;               // mesh_geometry::u_packed / v_packed (num_uv each), or null
   uint32_t num_uv;
   uint32_t texture;                 // textured Lambert: 1 + index into the scene's textures; 0: none
};

static_assert(sizeof(MeshIn) == 280, "MeshIn mirrors ptx_mesh_desc");

struct TexIn {                       // mirrors ptx_texture_desc
   uint32_t width, height;
   const uint8_t *rgba;              // rows top to bottom
};

struct Texture {                     // poly::texture (src/cpu/shading/texture.h:12-42), RGBA data
   uint32_t w = 0, h = 0;
   std::vector<uint8_t> rgba;
};

struct Mesh {
   uint32_t nf = 0;
   std::vector<V3> wv;                 // world vertices, 3 per face (v0,v1,v2 of face f at 3f..3f+2)
   bool has_normals = false;
   std::vector<V3> vn;                 // world-transformed vertex normals per face corner
   int brdf = 1;
   float refl[3] = {1, 1, 1};        // glossy: the mixed reflectance Ks (1 - r) + Kd r
   float rough = 0;                  // glossy roughness
   bool light = false;
   float L[3] = {0, 0, 0};
   int tex = -1;                     // textured Lambert: texture index
   std::vector<float> uv;            // textured: u, v of corners fv0, fv1, fv2 per face (6 floats)
};

struct CompactNode {      // compact_bvh_node (src/cpu/acceleration/bvh.h:24-53), 32 bytes
   Box bb;
   uint32_t offset;       // high_child_offset (interior) / face_index_offset (leaf)
   uint16_t num_faces;    // 0 = interior
   uint16_t flags;        // split axis
};
static_assert(sizeof(CompactNode) == 32, "compact node is 32 bytes");

using Pair = std::pair<unsigned int, unsigned int>;

struct TriInfo { Box bb; V3 c; };

struct BuildNode {
   Box bb;
   int axis = 0;
   int32_t low = -1, high = -1;      // child indices into the node pool
   size_t lo = 0, hi = 0;            // leaf index range into the shared index array
};

struct PointLight {
   V3 p;
   float I[3];
};

struct Scene {
   std::vector<Mesh> meshes;
   std::vector<CompactNode> nodes;
   std::vector<Pair> pairs;
   bool has_sky = false;
   float sky[3] = {0, 0, 0};
   std::vector<PointLight> lights;     // used only by next-event estimation
   std::vector<Texture> textures;
};

// bvh::bound (src/cpu/acceleration/bvh.cpp:36-395).  The reference copies each
// node's index vector; here every node owns a contiguous range of one shared
// index array, and a node's children are formed by a stable partition of that
// range in bucket order — the same sequence the reference builds by
// concatenating its bucket vectors (bvh.cpp:352-360), so every float reduction
// runs over the same elements in the same order.
static void bound(Scene& s, std::vector<BuildNode>& pool, std::vector<Pair>& idx) {
   std::vector<std::vector<TriInfo>> info(s.meshes.size());
   V3 rmin = {kInf, kInf, kInf}, rmax = {-kInf, -kInf, -kInf};
   size_t total = 0;
   for (auto& m : s.meshes) total += m.nf;
   idx.clear();
   idx.reserve(total);
   for (uint32_t mi = 0; mi < s.meshes.size(); mi++) {
      const Mesh& m = s.meshes[mi];
      info[mi].resize(m.nf);
      for (uint32_t f = 0; f < m.nf; f++) {
         const V3 v0 = m.wv[3 * f], v1 = m.wv[3 * f + 1], v2 = m.wv[3 * f + 2];
         V3 mn, mx;
         mn.x = v0.x < v1.x ? v0.x : v1.x; mn.x = mn.x < v2.x ? mn.x : v2.x;
         mn.y = v0.y < v1.y ? v0.y : v1.y; mn.y = mn.y < v2.y ? mn.y : v2.y;
         mn.z = v0.z < v1.z ? v0.z : v1.z; mn.z = mn.z < v2.z ? mn.z : v2.z;
         mx.x = v0.x > v1.x ? v0.x : v1.x; mx.x = mx.x > v2.x ? mx.x : v2.x;
         mx.y = v0.y > v1.y ? v0.y : v1.y; mx.y = mx.y > v2.y ? mx.y : v2.y;
         mx.z = v0.z > v1.z ? v0.z : v1.z; mx.z = mx.z > v2.z ? mx.z : v2.z;
         rmin.x = rmin.x < mn.x ? rmin.x : mn.x;
         rmin.y = rmin.y < mn.y ? rmin.y : mn.y;
         rmin.z = rmin.z < mn.z ? rmin.z : mn.z;
         rmax.x = rmax.x > mx.x ? rmax.x : mx.x;
         rmax.y = rmax.y > mx.y ? rmax.y : mx.y;
         rmax.z = rmax.z > mx.z ? rmax.z : mx.z;
         info[mi][f] = {{mn, mx}, {(mn.x + mx.x) * .5f, (mn.y + mx.y) * .5f, (mn.z + mx.z) * .5f}};
         idx.emplace_back(mi, f);
      }
   }
   pool.clear();
   pool.push_back(BuildNode{});
   pool[0].bb = {rmin, rmax};
   pool[0].lo = 0; pool[0].hi = idx.size();

   std::vector<Pair> tmp;
   std::vector<uint32_t> bucket_of;
   std::vector<size_t> work = {0};
   while (!work.empty()) {
      const size_t ni = work.back();
      work.pop_back();
      const size_t lo = pool[ni].lo, hi = pool[ni].hi, n = hi - lo;
      if (n == 1) continue;                                   // stop case 1
      const Box nbb = pool[ni].bb;
      const float node_sa = surface_area(nbb);
      if (node_sa == 0.f) continue;                           // stop case 2
      const float inv_sa = 1.f / node_sa;
      float cmin[3] = {kInf, kInf, kInf}, cmax[3] = {-kInf, -kInf, -kInf};
      for (size_t i = lo; i < hi; i++) {
         const V3 c = info[idx[i].first][idx[i].second].c;
         cmin[0] = cmin[0] < c.x ? cmin[0] : c.x;
         cmin[1] = cmin[1] < c.y ? cmin[1] : c.y;
         cmin[2] = cmin[2] < c.z ? cmin[2] : c.z;
         cmax[0] = cmax[0] > c.x ? cmax[0] : c.x;
         cmax[1] = cmax[1] > c.y ? cmax[1] : c.y;
         cmax[2] = cmax[2] > c.z ? cmax[2] : c.z;
      }
      const float ex = cmax[0] - cmin[0], ey = cmax[1] - cmin[1], ez = cmax[2] - cmin[2];
      if (ex == ey && ey == ez && ez == 0.f) continue;        // stop case 3
      int axis;
      if (ex >= ey && ex >= ez) axis = 0;
      else if (ey >= ex && ey >= ez) axis = 1;
      else axis = 2;
      const float extent = axis == 0 ? ex : (axis == 1 ? ey : ez);
      const float amin = cmin[axis];
      const unsigned nb = 12;
      const float bucket_width = extent / (float)nb;
      bucket_of.resize(n);
      size_t counts[12] = {0};
      for (size_t i = lo; i < hi; i++) {
         const V3 c = info[idx[i].first][idx[i].second].c;
         const float raw = (comp(c, axis) - amin) / (float)bucket_width;
         unsigned b = (unsigned)raw;
         if (b == nb) b--;
         bucket_of[i - lo] = b;
         counts[b]++;
      }
      // stable partition of the range into bucket order
      size_t start[12];
      size_t acc = 0;
      for (unsigned b = 0; b < nb; b++) { start[b] = acc; acc += counts[b]; }
      tmp.resize(n);
      {
         size_t cur[12];
         std::memcpy(cur, start, sizeof(cur));
         for (size_t i = lo; i < hi; i++) tmp[cur[bucket_of[i - lo]]++] = idx[i];
      }
      float fwd[12], bwd[12];
      Box low_bbs[12], high_bbs[12];
      Box low_bb = info[tmp[0].first][tmp[0].second].bb;      // bucket 0 is never empty
      float num_low = 0.f;
      for (unsigned b = 0; b < nb; b++) {
         for (size_t k = start[b]; k < start[b] + counts[b]; k++)
            union_in_place(low_bb, info[tmp[k].first][tmp[k].second].bb);
         low_bbs[b] = low_bb;
         num_low += counts[b];
         const float cost = surface_area(low_bb) * inv_sa * num_low * 1.f;
         fwd[b] = cost == 0 ? kInf : cost;
      }
      const Pair& h0 = tmp[start[nb - 1]];                   // bucket 11 is never empty
      Box high_bb = info[h0.first][h0.second].bb;
      float num_high = 0.f;
      for (int b = (int)nb - 1; b >= 0; b--) {
         for (size_t k = start[b]; k < start[b] + counts[b]; k++)
            union_in_place(high_bb, info[tmp[k].first][tmp[k].second].bb);
         high_bbs[b] = high_bb;
         num_high += counts[b];
         const float cost = surface_area(high_bb) * inv_sa * num_high * 1.f;
         bwd[nb - 1 - b] = cost == 0 ? kInf : cost;
      }
      float total_cost[11];
      for (unsigned b = 0; b < nb - 1; b++)
         total_cost[b] = 1.f + fwd[b] + bwd[nb - 1 - 1 - b];
      const float leaf_cost = n * 1.f;
      bool make_leaf = true;
      for (unsigned b = 0; b < nb - 1; b++)
         if (leaf_cost > total_cost[b]) make_leaf = false;
      if (make_leaf && n <= 4) continue;                      // stop case 5
      unsigned best = 0;
      float best_cost = kFloatMax;
      for (unsigned b = 0; b < nb - 1; b++)
         if (total_cost[b] < best_cost) { best_cost = total_cost[b]; best = b; }
      size_t n_low = 0;
      for (unsigned b = 0; b <= best; b++) n_low += counts[b];
      pool[ni].axis = axis;
      if (n_low == n || n_low == 0) continue;                 // no progress
      std::copy(tmp.begin(), tmp.begin() + n, idx.begin() + lo);
      BuildNode hc, lc;
      hc.bb = high_bbs[best + 1]; hc.lo = lo + n_low; hc.hi = hi;
      lc.bb = low_bbs[best];      lc.lo = lo;         lc.hi = lo + n_low;
      const int32_t hi_i = (int32_t)pool.size();
      pool.push_back(hc);
      const int32_t lo_i = (int32_t)pool.size();
      pool.push_back(lc);
      pool[ni].high = hi_i;
      pool[ni].low = lo_i;
      work.push_back((size_t)hi_i);
      work.push_back((size_t)lo_i);
   }
}

// bvh::compact / compact_helper (bvh.cpp:652-698): DFS pre-order, low child at
// i+1, high child at i + high_child_offset; leaf pairs sorted with the
// reference's comparator (bvh.cpp:666-671) using the same std::sort.
static uint32_t compact_helper(Scene& s, std::vector<BuildNode>& pool, std::vector<Pair>& idx,
                               int32_t ni, uint32_t index) {
   const BuildNode& b = pool[ni];
   CompactNode& cn = s.nodes[index];
   cn.bb = b.bb;
   cn.flags = (uint16_t)(short)b.axis;
   uint32_t n_child = 0;
   if (b.low < 0) {
      cn.offset = (uint32_t)s.pairs.size();
      std::vector<Pair> leaf(idx.begin() + b.lo, idx.begin() + b.hi);
      std::sort(leaf.begin(), leaf.end(), [](const Pair& lhs, const Pair& rhs) {
         if (lhs.first == rhs.first) return lhs.second < rhs.second;
         return lhs.first < lhs.second;
      });
      for (const auto& p : leaf) s.pairs.push_back(p);
      cn.num_faces = (uint16_t)leaf.size();
   } else {
      n_child += 1u + compact_helper(s, pool, idx, b.low, index + 1);
      s.nodes[index].offset = n_child + 1;
      s.nodes[index].num_faces = 0;
      n_child += 1u + compact_helper(s, pool, idx, b.high, index + 1 + n_child);
   }
   return n_child;
}

static void build_bvh(Scene& s) {
   std::vector<BuildNode> pool;
   std::vector<Pair> idx;
   bound(s, pool, idx);
   s.nodes.assign(pool.size(), CompactNode{});
   s.pairs.clear();
   s.pairs.reserve(idx.size());
   compact_helper(s, pool, idx, 0, 0);
}

// ----------------------------------------------------------------------------- traversal

struct Ray { V3 o, d; float min_t; };

struct Isect {
   bool hits = false;
   uint32_t mesh = 0, face = 0;
   float u = 0, v = 0, w = 0;
   V3 loc = {0, 0, 0}, err = {0, 0, 0};
   V3 n = {0, 0, 0}, t1 = {0, 0, 0}, t2 = {0, 0, 0};
   uint32_t bb_hits = 0, bb_misses = 0, tri_isects = 0;
   uint32_t tests_after_hit = 0;     // face tests from the accepting test to the end (>= 1 on hit)
   int px = -1, py = -1;
};

static inline float max_component(V3 v) {   // src/cpu/shapes/mesh.cpp:149-155
   if (v.x >= v.y && v.x >= v.z) return v.x;
   else if (v.y >= v.x && v.y >= v.z) return v.y;
   return v.z;
}
static inline V3 vabs(V3 v) { return {std::fabs(v.x), std::fabs(v.y), std::fabs(v.z)}; }

// Mesh::intersect test part for one face (src/cpu/shapes/mesh.cpp:169-355).
static inline void test_face(const Mesh& m, uint32_t mesh_index, uint32_t f, Ray& ray, Isect& is) {
   const V3 v0 = m.wv[3 * f], v1 = m.wv[3 * f + 1], v2 = m.wv[3 * f + 2];
   if (is.px == 128 && is.py == 128 && std::isnan(ray.d.x)) return;
   const float dxa = std::fabs(ray.d.x), dya = std::fabs(ray.d.y), dza = std::fabs(ray.d.z);
   int kz = 0;
   if (dya > dxa && dya > dza) kz = 1;
   else if (dza > dxa && dza > dya) kz = 2;
   int kx = kz + 1; if (kx == 3) kx = 0;
   int ky = kx + 1; if (ky == 3) ky = 0;
   V3 p0 = sub(v0, ray.o), p1 = sub(v1, ray.o), p2 = sub(v2, ray.o);
   const V3 d = {comp(ray.d, kx), comp(ray.d, ky), comp(ray.d, kz)};
   p0 = {comp(p0, kx), comp(p0, ky), comp(p0, kz)};
   p1 = {comp(p1, kx), comp(p1, ky), comp(p1, kz)};
   p2 = {comp(p2, kx), comp(p2, ky), comp(p2, kz)};
   const float Sx = -d.x / d.z, Sy = -d.y / d.z, Sz = 1.f / d.z;
   p0.x = std::fma(Sx, p0.z, p0.x); p0.y = std::fma(Sy, p0.z, p0.y);
   p1.x = std::fma(Sx, p1.z, p1.x); p1.y = std::fma(Sy, p1.z, p1.y);
   p2.x = std::fma(Sx, p2.z, p2.x); p2.y = std::fma(Sy, p2.z, p2.y);
   float U = difference_of_products(p1.x, p2.y, p1.y, p2.x);
   float V = difference_of_products(p2.x, p0.y, p2.y, p0.x);
   float W = difference_of_products(p0.x, p1.y, p0.y, p1.x);
   if (U == 0.0f || V == 0.0f || W == 0.0f) {
      const double CxBy = (double)p2.x * (double)p1.y;
      const double CyBx = (double)p2.y * (double)p1.x;
      U = (float)(CyBx - CxBy);
      const double AxCy = (double)p0.x * (double)p2.y;
      const double AyCx = (double)p0.y * (double)p2.x;
      V = (float)(AyCx - AxCy);
      const double BxAy = (double)p1.x * (double)p0.y;
      const double ByAx = (double)p1.y * (double)p0.x;
      W = (float)(ByAx - BxAy);
   }
   if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return;
   const float det = U + V + W;
   if (det == 0.0f) return;
   p0.z *= Sz; p1.z *= Sz; p2.z *= Sz;
   const float T = std::fma(U, p0.z, sum_of_products(V, p1.z, W, p2.z));
   if (det < 0.f && (T >= 0 || T < ray.min_t * det)) return;
   if (det > 0.f && (T <= 0 || T > ray.min_t * det)) return;
   const float rcp = 1.0f / det;
   const float u = U * rcp, v = V * rcp, w = W * rcp;
   const float t = T * rcp;
   if (std::isnan(t) && is.px == 128 && is.py == 128) return;
   const float maxZt = max_component(vabs({p0.z, p1.z, p2.z}));
   const float deltaZ = kGamma3 * maxZt;
   const float maxXt = max_component(vabs({p0.x, p1.x, p2.x}));
   const float maxYt = max_component(vabs({p0.y, p1.y, p2.y}));
   const float deltaX = kGamma5 * (maxXt + maxZt);
   const float deltaY = kGamma5 * (maxYt + maxZt);
   const float deltaE = 2 * (kGamma2 * maxXt * maxYt + deltaY * maxXt + deltaX * maxYt);
   const float maxE = max_component(vabs({U, V, W}));
   const float deltaT = 3 * (kGamma3 * maxE * maxZt + deltaE * maxZt + deltaZ * maxE) * std::fabs(rcp);
   if (t <= deltaT) return;
   const float xe = std::fabs(u * v0.x) + std::fabs(v * v1.x) + std::fabs(w * v2.x);
   const float ye = std::fabs(u * v0.y) + std::fabs(v * v1.y) + std::fabs(w * v2.y);
   const float ze = std::fabs(u * v0.z) + std::fabs(v * v1.z) + std::fabs(w * v2.z);
   is.err = {xe * kGamma7, ye * kGamma7, ze * kGamma7};
   ray.min_t = t;
   is.hits = true;
   is.face = f;
   is.mesh = mesh_index;
   is.loc = {std::fmaf(v0.x, u, sum_of_products(v1.x, v, v2.x, w)),
             std::fmaf(v0.y, u, sum_of_products(v1.y, v, v2.y, w)),
             std::fmaf(v0.z, u, sum_of_products(v1.z, v, v2.z, w))};
   is.u = u; is.v = v; is.w = w;
   is.tests_after_hit = 0;
}

// Mesh::intersect post-hit part (src/cpu/shapes/mesh.cpp:366-461), run
// is.tests_after_hit times on the hit mesh (see header note).
static void hit_frame(const Scene& s, const Ray& ray, Isect& is) {
   const Mesh& m = s.meshes[is.mesh];
   const uint32_t f = is.face;
   const V3 v0 = m.wv[3 * f], v1 = m.wv[3 * f + 1], v2 = m.wv[3 * f + 2];
   const V3 e0 = sub(v1, v0), e1 = sub(v2, v1), e2 = sub(v0, v2);
   V3 n;
   if (m.has_normals) {
      const V3 n0 = m.vn[3 * f], n1 = m.vn[3 * f + 1], n2 = m.vn[3 * f + 2];
      n = {n0.x * is.u + n1.x * is.v + n2.x * is.w,
           n0.y * is.u + n1.y * is.v + n2.y * is.w,
           n0.z * is.u + n1.z * is.v + n2.z * is.w};
   } else {
      n = cross(e0, e1);
   }
   const float rdn = n.x * ray.d.x + n.y * ray.d.y + n.z * ray.d.z;
   const float flip = rdn > 0 ? -1 : 1;
   n = {n.x * flip, n.y * flip, n.z * flip};
   n = normalize(n);
   const V3 an = vabs(n);
   const float dd = dot(is.err, an);
   V3 off = {n.x * dd, n.y * dd, n.z * dd};
   if ((n.x * ray.d.x + n.y * ray.d.y + n.z * ray.d.z) > 0) off = {-off.x, -off.y, -off.z};
   V3 p = is.loc;
   for (uint32_t k = 0; k < is.tests_after_hit; k++) {
      p = {p.x + off.x, p.y + off.y, p.z + off.z};
      for (int i = 0; i < 3; i++) {
         const float oi = comp(off, i);
         if (oi > 0) comp_ref(p, i) = std::nextafter(comp(p, i), kInf);
         else if (oi < 0) comp_ref(p, i) = std::nextafter(comp(p, i), -kInf);
      }
   }
   is.loc = p;
   is.n = n;
   const float ed0 = std::fabs(dot(e0, e1)), ed1 = std::fabs(dot(e1, e2)), ed2 = std::fabs(dot(e2, e0));
   V3 t1;
   if (ed0 > ed1 && ed0 > ed2) t1 = e0;
   else if (ed1 > ed0 && ed1 > ed2) t1 = e1;
   else t1 = e2;
   V3 t2 = cross(t1, n);
   is.t1 = normalize(t1);
   is.t2 = normalize(t2);
}

// bvh::intersect_compact (src/cpu/acceleration/bvh.cpp:476-523)
static void intersect(const Scene& s, Ray& ray, Isect& is) {
   const V3 inv = {1.f / ray.d.x, 1.f / ray.d.y, 1.f / ray.d.z};
   const bool neg[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
   thread_local std::vector<uint32_t> stack;
   stack.clear();
   stack.push_back(0);
   while (!stack.empty()) {
      const uint32_t ni = stack.back();
      stack.pop_back();
      const CompactNode& node = s.nodes[ni];
      if (box_hits(node.bb, ray.o, inv, ray.min_t)) {
         is.bb_hits++;
         if (node.num_faces > 0) {
            for (uint32_t i = 0; i < node.num_faces; i++) {
               const Pair& p = s.pairs[node.offset + i];
               test_face(s.meshes[p.first], p.first, p.second, ray, is);
               is.tri_isects++;
               is.tests_after_hit++;
            }
            continue;
         }
         if (neg[node.flags]) { stack.push_back(ni + 1); stack.push_back(ni + node.offset); }
         else { stack.push_back(ni + node.offset); stack.push_back(ni + 1); }
      } else {
         is.bb_misses++;
      }
   }
   if (is.hits) hit_frame(s, ray, is);
}

// Shadow ray for next-event estimation (no reference counterpart: the
// reference's shadow test is commented out, PathTraceIntegrator.cpp:97-130).
// Closest-hit traversal bounded by tmax (ray.min_t starts at tmax, so boxes and
// faces beyond it are culled exactly as a closer hit would cull them) that
// stops at the first accepted face.  No debug-pixel quirk on shadow rays.
static bool occluded(const Scene& s, const V3& o, const V3& d, float tmax) {
   Ray ray{o, d, tmax};
   Isect is;
   is.px = -1; is.py = -1;
   const V3 inv = {1.f / ray.d.x, 1.f / ray.d.y, 1.f / ray.d.z};
   const bool neg[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
   thread_local std::vector<uint32_t> stack;
   stack.clear();
   stack.push_back(0);
   while (!stack.empty()) {
      const uint32_t ni = stack.back();
      stack.pop_back();
      const CompactNode& node = s.nodes[ni];
      if (!box_hits(node.bb, ray.o, inv, ray.min_t)) continue;
      if (node.num_faces > 0) {
         for (uint32_t i = 0; i < node.num_faces; i++) {
            const Pair& p = s.pairs[node.offset + i];
            test_face(s.meshes[p.first], p.first, p.second, ray, is);
            if (is.hits) return true;
         }
         continue;
      }
      if (neg[node.flags]) { stack.push_back(ni + 1); stack.push_back(ni + node.offset); }
      else { stack.push_back(ni + node.offset); stack.push_back(ni + 1); }
   }
   return false;
}

// ----------------------------------------------------------------------------- RNG, sincos

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC 2011).
static inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
   uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
   uint32_t k0 = key_in[0], k1 = key_in[1];
   for (int r = 0; r < 10; r++) {
      const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
      const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
      const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
      const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
      const uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
      c0 = n0; c1 = n1; c2 = n2; c3 = n3;
      k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
   }
   out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
constexpr uint32_t kRngKeyHi = 0x706F6C79u;   // second key word: ASCII "poly"
static inline float u01(uint32_t r) { return (float)(r >> 8) * 0x1p-24f; }

// sin and cos of theta in [0, 2pi] (larger |theta| are reduced less accurately):
// quadrant k = floor(theta*2/pi + 1/2), Cody-Waite reduction r = theta - k*pi/2
// with a three-part pi/2, then Taylor polynomials on [-pi/4, pi/4] in Horner form
// with every step an explicit fma.
static inline void ptx_sincos(float theta, float* s_out, float* c_out) {
   const float q = theta * 0.636619772367581343f;
   const int k = (int)std::floor(q + 0.5f);
   const float kf = (float)k;
   float r = std::fmaf(-kf, 1.57079625129699707031f, theta);
   r = std::fmaf(-kf, 7.54978941586159635335e-08f, r);
   r = std::fmaf(-kf, 5.39030252995776476554e-15f, r);
   const float r2 = r * r;
   float ps = std::fmaf(r2, 2.75573192239858906526e-06f, -1.98412698412698412698e-04f);
   ps = std::fmaf(r2, ps, 8.33333333333333333333e-03f);
   ps = std::fmaf(r2, ps, -1.66666666666666666667e-01f);
   const float sn = std::fmaf(r * r2, ps, r);
   float pc = std::fmaf(r2, -2.75573192239858906526e-07f, 2.48015873015873015873e-05f);
   pc = std::fmaf(r2, pc, -1.38888888888888888889e-03f);
   pc = std::fmaf(r2, pc, 4.16666666666666666667e-02f);
   pc = std::fmaf(r2, pc, -0.5f);
   const float cs = std::fmaf(r2, pc, 1.0f);
   switch (k & 3) {
      case 0: *s_out = sn; *c_out = cs; break;
      case 1: *s_out = cs; *c_out = -sn; break;
      case 2: *s_out = -sn; *c_out = -cs; break;
      default: *s_out = -cs; *c_out = sn; break;
   }
}

// acos for the glossy lobe, project-owned so that CPU and GPU agree bitwise:
// acos(x) = pi/2 - asin(x) for |x| <= 1/2, else 2 asin(sqrt((1 - |x|)/2)), taken
// from pi for x < 0; asin(t) = t + t^3 P(t^2) with P the Maclaurin series of
// (asin t - t) / t^3, coefficients (2n)! / (4^n (n!)^2 (2n + 1)) for n = 1..12,
// in Horner form with every step an explicit fma.  |x| > 1 gives NaN, as acosf.
static inline float asin_series(float t, float z) {
   float p = std::fmaf(z, 0.006447210311889649f, 0.0073125258735988454f);
   p = std::fmaf(z, p, 0.008390335809616815f);
   p = std::fmaf(z, p, 0.009761609529194078f);
   p = std::fmaf(z, p, 0.011551800896139705f);
   p = std::fmaf(z, p, 0.01396484375f);
   p = std::fmaf(z, p, 0.017352764423076924f);
   p = std::fmaf(z, p, 0.022372159090909092f);
   p = std::fmaf(z, p, 0.030381944444444444f);
   p = std::fmaf(z, p, 0.044642857142857144f);
   p = std::fmaf(z, p, 0.075f);
   p = std::fmaf(z, p, 0.16666666666666666f);
   return std::fmaf(t * z, p, t);
}

static inline float ptx_acos(float x) {
   const float a = std::fabs(x);
   if (a <= 0.5f) return 1.57079637050628662109f - asin_series(x, x * x);
   const float z = (1.0f - a) * 0.5f;
   const float r = 2.0f * asin_series(std::sqrt(z), z);
   return x < 0.0f ? 3.14159274101257324219f - r : r;
}

// GlossyBRDF::sample (src/cpu/shading/brdf/glossy_brdf.h:27-60): a cosine-
// hemisphere sample rotated about (mirror direction x normal) by -acos(mirror . n),
// then mixed with the mirror direction by the roughness; repeated with new
// uniforms while the result points below the surface.  The reference loops
// without bound; here attempt a draws uniforms draw(a) and the lobe gives up
// (returns false) after kGlossyTries attempts.  The result is not normalised, as
// in the reference (LocalToWorld normalises it).  Trig supplies acos and
// sin/cos: the project-owned ones for rendering, libm for pinning against the
// reference's own function.
constexpr uint32_t kGlossyTries = 64;

struct PtxTrig {
   static float acos_(float x) { return ptx_acos(x); }
   static void sincos_(float t, float* s, float* c) { ptx_sincos(t, s, c); }
};
struct LibmTrig {
   static float acos_(float x) { return std::acos(x); }
   static void sincos_(float t, float* s, float* c) { *s = std::sin(t); *c = std::cos(t); }
};

template <class Trig, class Draw>
static bool glossy_sample(V3 in, float rough, Draw draw, V3& out, uint32_t* used = nullptr) {
   for (uint32_t a = 0; a < kGlossyTries; a++) {
      if (used) *used = a + 1;
      float u0, u1;
      draw(a, u0, u1);
      const float r = std::sqrt(u0);                          // CosineSampleHemisphere
      const float theta = kTwoPI * u1;
      float sn, cs;
      Trig::sincos_(theta, &sn, &cs);
      const V3 lam = {r * cs, std::sqrt(std::max(0.0f, 1.0f - u0)), r * sn};
      const float factor = (0.f * in.x + 1.f * in.y + 0.f * in.z) * 2;   // incoming . n, n = (0, 1, 0)
      V3 spec = {in.x - 0.f * factor, in.y - 1.f * factor, in.z - 0.f * factor};
      spec = normalize(spec);
      const float angle = -Trig::acos_(0.f * spec.x + 1.f * spec.y + 0.f * spec.z);
      const float x = spec.y * 0.f - spec.z * 1.f;            // spec x n (Vector::Cross(Normal))
      const float y = spec.z * 0.f - spec.x * 0.f;
      const float z = spec.x * 1.f - spec.y * 0.f;
      float s, c;                                              // Transform::Rotate (Transform.cpp:214-228)
      Trig::sincos_(angle, &s, &c);
      const float omc = 1 - c;
      const float m00 = c + x * x * omc, m01 = x * y * omc - z * s, m02 = x * z * omc + y * s;
      const float m10 = y * x * omc + z * s, m11 = c + y * y * omc, m12 = y * z * omc - x * s;
      const float m20 = z * x * omc - y * s, m21 = z * y * omc + x * s, m22 = c + z * z * omc;
      const V3 rot = {lam.x * m00 + lam.y * m01 + lam.z * m02,   // Transform::Apply(Vector)
                      lam.x * m10 + lam.y * m11 + lam.z * m12,
                      lam.x * m20 + lam.y * m21 + lam.z * m22};
      const float k = 1 - rough;
      const V3 o = {spec.x * k + rot.x * rough, spec.y * k + rot.y * rough, spec.z * k + rot.z * rough};
      if (o.y < 0) continue;
      out = o;
      return true;
   }
   return false;
}

// ----------------------------------------------------------------------------- render

// A mesh's shading data.  Glossy meshes keep the reflectance GlossyBRDF::sample
// returns, Ks (1 - r) + Kd r per channel (glossy_brdf.h:52), which is constant.
// A textured Lambert mesh keeps its faces' corner texture coordinates.
static void set_material(Mesh& m, const MeshIn& in) {
   if (in.texture > 0 && in.brdf == 1 && !in.is_light && in.u && in.v) {
      m.tex = (int)in.texture - 1;
      m.uv.resize(size_t(6) * in.num_faces);
      for (uint32_t f = 0; f < in.num_faces; f++) {
         const uint32_t ix[3] = {in.fv0[f], in.fv1[f], in.fv2[f]};
         for (int c = 0; c < 3; c++) {
            m.uv[6 * f + 2 * c] = in.u[ix[c]];
            m.uv[6 * f + 2 * c + 1] = in.v[ix[c]];
         }
      }
   }
   m.brdf = in.brdf;
   std::memcpy(m.refl, in.refl, sizeof(m.refl));
   if (in.brdf == 3) {
      const float r = in.roughness;
      for (int c = 0; c < 3; c++) m.refl[c] = in.specular[c] * (1 - r) + in.refl[c] * r;
      m.rough = r;
   }
   m.light = in.is_light != 0;
   std::memcpy(m.L, in.emission, sizeof(m.L));
}

struct CameraIn {                 // mirrors ptx_camera_desc
   float camera_to_world[16];
   float fov;                     // degrees
   uint32_t width, height;
   float one_over_width, one_over_height, aspect_ratio, tan_fov_over_2;  // host-precomputed
};

struct JobIn {                    // mirrors ptx_job_desc
   uint32_t width, height, spp, max_depth;
   uint32_t seed;
   uint32_t flags;                // bit 0: next-event estimation to the point lights
};

struct Camera {
   const float* m;
   float oow, ooh, aspect, tanf2;
   // PerspectiveCamera::get_ray_for_pixel (src/cpu/cameras/PerspectiveCamera.cpp:27-49)
   Ray ray_for_pixel(float px, float py) const {
      const float nx = px * oow;
      const float ny = py * ooh;
      const float cx = (2 * nx - 1) * aspect * tanf2;
      const float cy = (1 - 2 * ny) * tanf2;
      const float z = 1;
      V3 dir = normalize({cx, cy, z});
      V3 o = apply_point(m, {0, 0, 0});
      V3 d = apply_vector(m, dir);
      d = normalize(d);
      return {o, d, kInf};
   }
};

struct Spd { float r, g, b; };

// Mesh::intersect's interpolated texture coordinates (mesh.cpp:342-353): u, v,
// w are the hit's barycentric weights of the face's corners 0, 1, 2.
static inline void tex_lerp(const float uv[6], float u, float v, float w, float& tu, float& tv) {
   tu = uv[0] * u + uv[2] * v + uv[4] * w;
   tv = uv[1] * u + uv[3] * v + uv[5] * w;
}

// (unsigned int)f where C++ defines it (0 <= f < 2^32); 0 below, 0xFFFFFFFF above.
static inline uint32_t texel_trunc(float f) {
   if (!(f > 0.f)) return 0u;
   if (!(f < 4294967296.f)) return 0xFFFFFFFFu;
   return (uint32_t)f;
}

// texture::evaluate_rgba's texel (texture.h:25-31): flat index row * W + column
// of column (unsigned)(u * W), row (unsigned)((1 - v) * H).  Where that index
// is outside the image (the reference reads outside its data vector) the
// column and row are clamped to the image instead.
static inline uint32_t texel_index(float tu, float tv, uint32_t W, uint32_t H) {
   const uint32_t col = texel_trunc(tu * (float)W);
   const uint32_t row = texel_trunc((1.f - tv) * (float)H);
   const uint64_t flat = (uint64_t)row * W + col;
   if (flat < (uint64_t)W * H) return (uint32_t)flat;
   return std::min(row, H - 1) * W + std::min(col, W - 1);
}

// LambertBRDF::sample's reflectance with a texture (lambert_brdf.h:49-56):
// ReflectanceSpectrum(int, int, int) of the texel's R, G, B (spectrum.h:17-18).
static inline void textured_refl(const Scene& s, const Mesh& m, const Isect& is, float refl[3]) {
   float tu, tv;
   tex_lerp(&m.uv[6 * is.face], is.u, is.v, is.w, tu, tv);
   const Texture& t = s.textures[m.tex];
   const uint8_t* px = &t.rgba[4 * size_t(texel_index(tu, tv, t.w, t.h))];
   const float one_over_255 = 1.0f / 255.0f;
   refl[0] = (float)px[0] * one_over_255;
   refl[1] = (float)px[1] * one_over_255;
   refl[2] = (float)px[2] * one_over_255;
}

// PathTraceIntegrator::GetSample (src/cpu/integrators/PathTraceIntegrator.cpp:10-135)
// Next-event estimation at a Lambert vertex (opt-in; no reference
// counterpart): for each point light in order, l = normalize(light - P) with
// d2 = |light - P|^2, cos = n . l; if cos > 0 and the shadow ray (P, l, t < d)
// is unoccluded, direct += (src * Kd) * (I * (cos / (pi * d2))).  P is the
// offset hit Location, n the shading normal facing the incoming ray, src the
// throughput before this vertex's reflectance.
static void add_direct(const Scene& s, const Isect& is, const float refl[3], float sr, float sg, float sb, Spd& direct) {
   for (const PointLight& L : s.lights) {
      const float wx = L.p.x - is.loc.x, wy = L.p.y - is.loc.y, wz = L.p.z - is.loc.z;
      const float d2 = wx * wx + wy * wy + wz * wz;
      const float d = std::sqrt(d2);
      const float inv = 1.0f / d;
      const V3 l = {wx * inv, wy * inv, wz * inv};
      const float cosv = is.n.x * l.x + is.n.y * l.y + is.n.z * l.z;
      if (!(cosv > 0.f)) continue;
      if (occluded(s, is.loc, l, d)) continue;
      const float g = cosv / (kPI * d2);
      direct.r = direct.r + (sr * refl[0]) * (L.I[0] * g);
      direct.g = direct.g + (sg * refl[1]) * (L.I[1] * g);
      direct.b = direct.b + (sb * refl[2]) * (L.I[2] * g);
   }
}

// Work of a render, for the algorithmic-byte accounting: rays traced (every
// bounce's closest-hit query) and, summed over them, the reference's per-ray
// counters (num_bb_hits + num_bb_misses, num_triangle_isects).
struct RayStats {
   uint64_t rays = 0, boxes = 0, tris = 0;
};

static Spd get_sample(const Scene& s, const Ray& primary, int x, int y, uint32_t pixel_index,
                      uint32_t sample_index, uint32_t max_depth, uint32_t seed, bool nee, RayStats* stats = nullptr) {
   Ray cur = primary;
   uint32_t bounces = 0;
   float sr = 1.f, sg = 1.f, sb = 1.f;
   Spd direct = {0, 0, 0};
   // terminal radiance, plus the next-event sum when it is on
   auto done = [&](Spd t) -> Spd {
      if (nee) { t.r = t.r + direct.r; t.g = t.g + direct.g; t.b = t.b + direct.b; }
      return t;
   };
   while (true) {
      if (sr < 0.0039f && sg < 0.0039f && sb < 0.0039f) return done({0, 0, 0});
      Isect is;
      is.px = x; is.py = y;
      cur.min_t = kInf;
      intersect(s, cur, is);
      if (stats) {
         stats->rays++;
         stats->boxes += is.bb_hits + is.bb_misses;
         stats->tris += is.tri_isects;
      }
      if (!is.hits) {
         Spd spd = {0, 0, 0};
         if (s.has_sky) spd = {s.sky[0] * sr, s.sky[1] * sg, s.sky[2] * sb};
         spd.r += 0.f; spd.g += 0.f; spd.b += 0.f;
         return done(spd);
      }
      const Mesh& m = s.meshes[is.mesh];
      if (m.light) return done({(m.L[0] + 0.f) * sr, (m.L[1] + 0.f) * sg, (m.L[2] + 0.f) * sb});
      if (bounces >= max_depth) return done({0, 0, 0});
      float kd[3] = {m.refl[0], m.refl[1], m.refl[2]};
      if (m.tex >= 0) textured_refl(s, m, is, kd);
      if (nee && m.brdf == 1) add_direct(s, is, kd, sr, sg, sb, direct);
      const uint32_t bounce_index = bounces;
      bounces++;
      // Intersection::WorldToLocal (Intersection.cpp:10-14); bent normal == geo normal
      V3 lin = {dot(cur.d, is.t1), is.n.x * cur.d.x + is.n.y * cur.d.y + is.n.z * cur.d.z, dot(cur.d, is.t2)};
      lin = normalize(lin);
      V3 lout;
      float rr, rg, rb;
      if (m.brdf == 2) {
         rr = m.refl[0]; rg = m.refl[1]; rb = m.refl[2];
         lout = {lin.x, -lin.y, lin.z};
      } else if (m.brdf == 3) {
         rr = m.refl[0]; rg = m.refl[1]; rb = m.refl[2];
         const uint32_t key[2] = {seed, kRngKeyHi};
         auto draw = [&](uint32_t a, float& u0, float& u1) {
            const uint32_t ctr[4] = {pixel_index, sample_index, bounce_index, a};
            uint32_t rnd[4];
            philox4x32_10(ctr, key, rnd);
            u0 = u01(rnd[0]);
            u1 = u01(rnd[1]);
         };
         if (!glossy_sample<PtxTrig>(lin, m.rough, draw, lout)) {
            sr = sg = sb = 0.f;                 // no accepted direction: the path ends black
            continue;
         }
      } else {
         rr = kd[0]; rg = kd[1]; rb = kd[2];
         const uint32_t ctr[4] = {pixel_index, sample_index, bounce_index, 0u};
         const uint32_t key[2] = {seed, kRngKeyHi};
         uint32_t rnd[4];
         philox4x32_10(ctr, key, rnd);
         const float u0 = u01(rnd[0]), u1 = u01(rnd[1]);
         // CosineSampleHemisphere (src/cpu/constants.h:106-115)
         const float r = std::sqrt(u0);
         const float theta = kTwoPI * u1;
         float sn, cs;
         ptx_sincos(theta, &sn, &cs);
         lout = {r * cs, std::sqrt(std::max(0.0f, 1.0f - u0)), r * sn};
      }
      // Intersection::LocalToWorld (Intersection.cpp:16-23)
      V3 wout = {is.t1.x * lout.x + is.n.x * lout.y + is.t2.x * lout.z,
                 is.t1.y * lout.x + is.n.y * lout.y + is.t2.y * lout.z,
                 is.t1.z * lout.x + is.n.z * lout.y + is.t2.z * lout.z};
      wout = normalize(wout);
      cur = {is.loc, wout, kInf};
      sr = sr * rr; sg = sg * rg; sb = sb * rb;
   }
}

}  // namespace orc

// ----------------------------------------------------------------------------- C ABI

extern "C" {

struct orc_hit {             // 100-byte record, same layout the reference driver writes
   uint32_t hit, mesh, face;
   float t, u, v, w;
   float loc[3], n[3], t1[3], t2[3], err[3];
   uint32_t bb_hits, bb_misses, tri_isects;
};

orc::Scene* orc_scene_create(const orc::MeshIn* meshes, uint32_t n_meshes, int has_sky, const float* sky) {
   auto* s = new orc::Scene();
   s->meshes.resize(n_meshes);
   for (uint32_t mi = 0; mi < n_meshes; mi++) {
      const orc::MeshIn& in = meshes[mi];
      orc::Mesh& m = s->meshes[mi];
      m.nf = in.num_faces;
      m.wv.resize(size_t(3) * in.num_faces);
      m.has_normals = in.nx != nullptr;
      if (m.has_normals) m.vn.resize(size_t(3) * in.num_faces);
      for (uint32_t f = 0; f < in.num_faces; f++) {
         const uint32_t ix[3] = {in.fv0[f], in.fv1[f], in.fv2[f]};
         for (int c = 0; c < 3; c++) {
            m.wv[3 * f + c] = orc::apply_point_inplace(in.object_to_world, {in.x[ix[c]], in.y[ix[c]], in.z[ix[c]]});
            if (m.has_normals)
               m.vn[3 * f + c] = orc::apply_normal(in.world_to_object, {in.nx[ix[c]], in.ny[ix[c]], in.nz[ix[c]]});
         }
      }
      set_material(m, in);
   }
   s->has_sky = has_sky != 0;
   if (sky) std::memcpy(s->sky, sky, sizeof(s->sky));
   orc::build_bvh(*s);
   return s;
}

// Same scene with a BVH supplied by the caller (reference compact layout:
// 32-byte nodes and (mesh, face) leaf pairs) instead of building one.  Lets a
// test check traversal on scenes whose build would take the single-threaded
// restatement minutes, with the builder's parity checked separately.
orc::Scene* orc_scene_create_prebuilt(const orc::MeshIn* meshes, uint32_t n_meshes, int has_sky, const float* sky,
                                      const void* nodes, uint32_t num_nodes, const uint32_t* pairs,
                                      uint32_t num_pairs) {
   auto* s = new orc::Scene();
   s->meshes.resize(n_meshes);
   for (uint32_t mi = 0; mi < n_meshes; mi++) {
      const orc::MeshIn& in = meshes[mi];
      orc::Mesh& m = s->meshes[mi];
      m.nf = in.num_faces;
      m.wv.resize(size_t(3) * in.num_faces);
      m.has_normals = in.nx != nullptr;
      if (m.has_normals) m.vn.resize(size_t(3) * in.num_faces);
      for (uint32_t f = 0; f < in.num_faces; f++) {
         const uint32_t ix[3] = {in.fv0[f], in.fv1[f], in.fv2[f]};
         for (int c = 0; c < 3; c++) {
            m.wv[3 * f + c] = orc::apply_point_inplace(in.object_to_world, {in.x[ix[c]], in.y[ix[c]], in.z[ix[c]]});
            if (m.has_normals)
               m.vn[3 * f + c] = orc::apply_normal(in.world_to_object, {in.nx[ix[c]], in.ny[ix[c]], in.nz[ix[c]]});
         }
      }
      set_material(m, in);
   }
   s->has_sky = has_sky != 0;
   if (sky) std::memcpy(s->sky, sky, sizeof(s->sky));
   s->nodes.resize(num_nodes);
   std::memcpy(s->nodes.data(), nodes, size_t(num_nodes) * sizeof(orc::CompactNode));
   s->pairs.resize(num_pairs);
   for (uint32_t i = 0; i < num_pairs; i++) s->pairs[i] = {pairs[2 * i], pairs[2 * i + 1]};
   return s;
}

void orc_scene_destroy(orc::Scene* s) { delete s; }

// Point lights for next-event estimation: 6 floats each (position, intensity).
void orc_scene_set_lights(orc::Scene* s, const float* lights, uint32_t n) {
   s->lights.resize(n);
   for (uint32_t i = 0; i < n; i++)
      s->lights[i] = {{lights[6 * i], lights[6 * i + 1], lights[6 * i + 2]},
                      {lights[6 * i + 3], lights[6 * i + 4], lights[6 * i + 5]}};
}
// Textures that meshes' MeshIn::texture indexes (copied).
void orc_scene_set_textures(orc::Scene* s, const orc::TexIn* tex, uint32_t n) {
   s->textures.resize(n);
   for (uint32_t i = 0; i < n; i++) {
      s->textures[i].w = tex[i].width;
      s->textures[i].h = tex[i].height;
      s->textures[i].rgba.assign(tex[i].rgba, tex[i].rgba + 4 * size_t(tex[i].width) * tex[i].height);
   }
}

// The texture coordinates of a hit and the texel they select (for unit checks).
void orc_tex_lerp(const float uv[6], float u, float v, float w, float out[2]) {
   orc::tex_lerp(uv, u, v, w, out[0], out[1]);
}
uint32_t orc_texel_index(float tu, float tv, uint32_t width, uint32_t height) {
   return orc::texel_index(tu, tv, width, height);
}

uint32_t orc_num_nodes(const orc::Scene* s) { return (uint32_t)s->nodes.size(); }
uint32_t orc_num_pairs(const orc::Scene* s) { return (uint32_t)s->pairs.size(); }
void orc_copy_bvh(const orc::Scene* s, void* nodes, uint32_t* pairs) {
   if (nodes) std::memcpy(nodes, s->nodes.data(), s->nodes.size() * sizeof(orc::CompactNode));
   if (pairs)
      for (size_t i = 0; i < s->pairs.size(); i++) {
         pairs[2 * i] = s->pairs[i].first;
         pairs[2 * i + 1] = s->pairs[i].second;
      }
}

static void run_parallel(uint32_t n, int threads, const std::function<void(uint32_t, uint32_t)>& fn) {
   if (threads <= 1) { fn(0, n); return; }
   std::vector<std::thread> pool;
   for (int t = 0; t < threads; t++) {
      const uint32_t lo = (uint32_t)((uint64_t)n * t / threads), hi = (uint32_t)((uint64_t)n * (t + 1) / threads);
      pool.emplace_back([=, &fn]() { fn(lo, hi); });
   }
   for (auto& th : pool) th.join();
}

void orc_intersect(const orc::Scene* s, const float* rays, uint32_t n, orc_hit* out, int threads) {
   run_parallel(n, threads, [&](uint32_t lo, uint32_t hi) {
      for (uint32_t i = lo; i < hi; i++) {
         orc::Ray ray = {{rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]},
                         {rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]}, orc::kInf};
         orc::Isect is;
         orc::intersect(*s, ray, is);
         orc_hit& h = out[i];
         std::memset(&h, 0, sizeof(h));
         h.hit = is.hits;
         h.mesh = is.hits ? is.mesh : 0xFFFFFFFFu;
         h.face = is.hits ? is.face : 0xFFFFFFFFu;
         if (is.hits) {
            h.t = ray.min_t; h.u = is.u; h.v = is.v; h.w = is.w;
            h.loc[0] = is.loc.x; h.loc[1] = is.loc.y; h.loc[2] = is.loc.z;
            h.n[0] = is.n.x; h.n[1] = is.n.y; h.n[2] = is.n.z;
            h.t1[0] = is.t1.x; h.t1[1] = is.t1.y; h.t1[2] = is.t1.z;
            h.t2[0] = is.t2.x; h.t2[1] = is.t2.y; h.t2[2] = is.t2.z;
            h.err[0] = is.err.x; h.err[1] = is.err.y; h.err[2] = is.err.z;
         }
         h.bb_hits = is.bb_hits; h.bb_misses = is.bb_misses; h.tri_isects = is.tri_isects;
      }
   });
}

// Closest-hit timing loop (CPU baseline): hit ids only, checksum returned.
uint64_t orc_intersect_ids(const orc::Scene* s, const float* rays, uint32_t n, uint32_t* face_out, int threads) {
   std::atomic<uint64_t> sum{0};
   run_parallel(n, threads, [&](uint32_t lo, uint32_t hi) {
      uint64_t local = 0;
      for (uint32_t i = lo; i < hi; i++) {
         orc::Ray ray = {{rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]},
                         {rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]}, orc::kInf};
         orc::Isect is;
         orc::intersect(*s, ray, is);
         const uint32_t id = is.hits ? is.face : 0xFFFFFFFFu;
         if (face_out) face_out[i] = id;
         local += id;
      }
      sum += local;
   });
   return sum.load();
}

// Full render: TileRunner's 2x2 tile dispenser (src/cpu/runners/TileRunner.cpp:61-117)
// over `threads` workers, BoxFilter accumulation in sample order, then x 1/spp.
// If `row_step` > 1 only rows y % row_step == 0 are rendered (a deterministic
// subset for timing); other rows are left as zeros.
// stats (optional): {rays, boxes, face tests, samples} of the rendered rows.
void orc_render_counted(const orc::Scene* s, const orc::CameraIn* cam, const orc::JobIn* job, float* film,
                        int threads, uint32_t row_step, uint64_t* stats);

void orc_render(const orc::Scene* s, const orc::CameraIn* cam, const orc::JobIn* job, float* film,
                int threads, uint32_t row_step) {
   orc_render_counted(s, cam, job, film, threads, row_step, nullptr);
}

void orc_render_counted(const orc::Scene* s, const orc::CameraIn* cam, const orc::JobIn* job, float* film,
                        int threads, uint32_t row_step, uint64_t* stats) {
   const uint32_t W = job->width, H = job->height;
   orc::Camera c{cam->camera_to_world, cam->one_over_width, cam->one_over_height, cam->aspect_ratio,
                 cam->tan_fov_over_2};
   const uint32_t tiles_x = (W + 1) / 2, tiles_y = (H + 1) / 2;
   std::atomic<uint32_t> next{0};
   std::mutex stats_mu;
   if (row_step == 0) row_step = 1;
   if (stats) stats[0] = stats[1] = stats[2] = stats[3] = 0;
   auto worker = [&]() {
      orc::RayStats mine;
      uint64_t samples = 0;
      orc::RayStats* st = stats ? &mine : nullptr;
      while (true) {
         const uint32_t t = next.fetch_add(1);
         if (t >= tiles_x * tiles_y) break;
         const uint32_t tx = t % tiles_x, ty = t / tiles_x;
         for (uint32_t y = ty * 2; y < std::min(ty * 2 + 2, H); y++) {
            if (y % row_step) continue;
            for (uint32_t x = tx * 2; x < std::min(tx * 2 + 2, W); x++) {
               const orc::Ray primary = c.ray_for_pixel(x + 0.5f, y + 0.5f);
               float ar = 0, ag = 0, ab = 0;
               const uint32_t pix = y * W + x;
               for (uint32_t sidx = 0; sidx < job->spp; sidx++) {
                  orc::Spd smp = orc::get_sample(*s, primary, (int)x, (int)y, pix, sidx, job->max_depth, job->seed,
                                                 (job->flags & 1u) != 0, st);
                  samples++;
                  if (smp.r > 255) smp.r = 255;
                  if (smp.g > 255) smp.g = 255;
                  if (smp.b > 255) smp.b = 255;
                  ar += smp.r; ag += smp.g; ab += smp.b;
               }
               const float div = 1.0f / (unsigned long)job->spp;
               film[3 * size_t(pix)] = ar * div;
               film[3 * size_t(pix) + 1] = ag * div;
               film[3 * size_t(pix) + 2] = ab * div;
            }
         }
      }
      if (stats) {
         std::lock_guard<std::mutex> lock(stats_mu);
         stats[0] += mine.rays;
         stats[1] += mine.boxes;
         stats[2] += mine.tris;
         stats[3] += samples;
      }
   };
   if (threads <= 1) worker();
   else {
      std::vector<std::thread> pool;
      for (int i = 0; i < threads; i++) pool.emplace_back(worker);
      for (auto& th : pool) th.join();
   }
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
   orc::philox4x32_10(ctr, key, out);
}
void orc_sincos(float theta, float* s, float* c) { orc::ptx_sincos(theta, s, c); }

float orc_acos(float x) { return orc::ptx_acos(x); }

// GlossyBRDF::sample on given uniforms (pairs u0, u1 per attempt); use_libm
// selects libm's acos/sin/cos (as the reference calls them) over the
// project-owned ones.  Returns 1 and the direction when an attempt is accepted
// within n_pairs attempts, else 0; *used = attempts drawn.
int orc_glossy_sample(const float in[3], float roughness, const float* uniforms, uint32_t n_pairs, int use_libm,
                      float out[3], uint32_t* used) {
   const orc::V3 v = {in[0], in[1], in[2]};
   auto draw = [&](uint32_t a, float& u0, float& u1) {
      if (a < n_pairs) { u0 = uniforms[2 * a]; u1 = uniforms[2 * a + 1]; }
      else { u0 = 0.f; u1 = 0.f; }
   };
   orc::V3 o = {0, 0, 0};
   uint32_t u = 0;
   const bool ok = use_libm ? orc::glossy_sample<orc::LibmTrig>(v, roughness, draw, o, &u)
                            : orc::glossy_sample<orc::PtxTrig>(v, roughness, draw, o, &u);
   out[0] = o.x; out[1] = o.y; out[2] = o.z;
   *used = u;
   return ok && u <= n_pairs ? 1 : 0;
}
void orc_box_hits(const float box[6], const float ray[6], float min_t, int* hit) {
   const orc::Box b = {{box[0], box[1], box[2]}, {box[3], box[4], box[5]}};
   const orc::V3 o = {ray[0], ray[1], ray[2]};
   const orc::V3 inv = {1.f / ray[3], 1.f / ray[4], 1.f / ray[5]};
   *hit = orc::box_hits(b, o, inv, min_t) ? 1 : 0;
}
// Primary ray of a pixel centre, for tests of the camera.
void orc_camera_ray(const orc::CameraIn* cam, float px, float py, float out[6]) {
   orc::Camera c{cam->camera_to_world, cam->one_over_width, cam->one_over_height, cam->aspect_ratio,
                 cam->tan_fov_over_2};
   const orc::Ray r = c.ray_for_pixel(px, py);
   out[0] = r.o.x; out[1] = r.o.y; out[2] = r.o.z; out[3] = r.d.x; out[4] = r.d.y; out[5] = r.d.z;
}

}  // extern "C"
