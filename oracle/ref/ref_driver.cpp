// Test-infrastructure driver for the reference polytope CPU path.
//
// This file is part of the checker, not of the product: it links the
// reference's own sources (compiled in place from the reference tree by
// oracle/ref/Makefile) and exposes them through a small command-line tool that
// reads scenes and rays from files and writes the reference's outputs
// (compact BVH nodes, leaf pairs, per-ray hits with counters, rendered film)
// as raw little-endian binaries.  Those outputs become the golden fixtures in
// tests/golden, and on a machine where the tool is built it also serves as the
// "reference" CPU baseline.
//
// It defines the globals that the reference's main.cpp defines
// (src/cpu/main.cpp:27, 50-51) and otherwise only calls reference APIs:
//   bvh::bound / compact / intersect_compact   (src/cpu/acceleration/bvh.cpp:36, 695, 476)
//   pbrt_parser::parse_file                     (src/common/parsers/pbrt_parser.cpp:275)
//   AbstractRunner::Trace, BoxFilter::Output    (src/cpu/runners/AbstractRunner.cpp:11,
//                                                src/cpu/filters/BoxFilter.cpp:54)

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>
#include <atomic>

#include <pthread.h>
#include <sched.h>

#include "common/utilities/Common.h"
#include "common/parsers/pbrt_parser.h"
#include "cpu/acceleration/bvh.h"
#include "cpu/cameras/PerspectiveCamera.h"
#include "cpu/filters/BoxFilter.h"
#include "cpu/films/PNGFilm.h"
#include "cpu/integrators/PathTraceIntegrator.h"
#include "cpu/runners/TileRunner.h"
#include "cpu/samplers/samplers.h"
#include "cpu/scenes/Scene.h"
#include "cpu/scenes/skyboxes/ColorSkybox.h"
#include <cassert>                       // glossy_brdf.h uses assert without including it
#include <cmath>
#include <random>
#include "cpu/shading/brdf/glossy_brdf.h"
#include "cpu/shading/brdf/lambert_brdf.h"
#include "cpu/shading/brdf/mirror_brdf.h"
#include "cpu/shading/texture.h"
#include "cpu/shapes/mesh.h"
#include "cpu/structures/stats.h"
#include "ptxm_scene.h"

poly::Logger Log;
struct poly::stats main_stats;
thread_local struct poly::stats thread_stats;

namespace {

using ptxm::Job;
using ptxm::Reader;
using ptxm::load_ptxm;
using ptxm::write_file;

// Per-ray record written by "hits": 25 four-byte words.
struct HitRecord {
   uint32_t hit, mesh, face;
   float t, u, v, w;
   float loc[3], geo_n[3], t1[3], t2[3], err[3];
   uint32_t bb_hits, bb_misses, tri_isects;
};
static_assert(sizeof(HitRecord) == 100, "hit record layout");

// Everything the reference's parser produced, as JSON with each float as its
// 32-bit pattern: film size, samples, integrator depth, sampler, camera,
// skybox and, per Scene::Shapes entry in order, its transforms, light or
// material and packed geometry (pbrt_parser.cpp:907-1721, mesh_parsers.cpp).
void describe(const poly::AbstractRunner& runner, const std::string& path) {
   std::FILE* f = std::fopen(path.c_str(), "w");
   if (!f) { std::fprintf(stderr, "ref_driver: cannot write %s\n", path.c_str()); std::exit(1); }
   auto bits = [](float v) { uint32_t u; std::memcpy(&u, &v, 4); return u; };
   auto farr = [&](const float* v, size_t n) {
      std::fputc('[', f);
      for (size_t i = 0; i < n; i++) std::fprintf(f, "%s%u", i ? "," : "", bits(v[i]));
      std::fputc(']', f);
   };
   auto uarr = [&](const std::vector<unsigned int>& v) {
      std::fputc('[', f);
      for (size_t i = 0; i < v.size(); i++) std::fprintf(f, "%s%u", i ? "," : "", v[i]);
      std::fputc(']', f);
   };
   auto mat = [&](const poly::Transform& t) { farr(&t.Matrix.Matrix[0][0], 16); };
   const auto* integrator = dynamic_cast<const poly::PathTraceIntegrator*>(runner.Integrator.get());
   const auto* camera = dynamic_cast<const poly::PerspectiveCamera*>(runner.Scene->Camera.get());
   std::fprintf(f, "{\"width\": %u, \"height\": %u, \"spp\": %u, \"max_depth\": %d, \"center_sampler\": %s,\n",
                runner.Bounds.x, runner.Bounds.y, runner.NumSamples, integrator ? (int)integrator->MaxDepth : -1,
                dynamic_cast<const poly::CenterSampler*>(runner.Sampler.get()) ? "true" : "false");
   std::fprintf(f, " \"camera\": {\"c2w\": ");
   mat(camera->CameraToWorld);
   std::fprintf(f, ", \"fov\": %u},\n \"sky\": ", bits(camera->Settings.FieldOfView));
   if (const auto* sky = dynamic_cast<const poly::ColorSkybox*>(runner.Scene->Skybox.get())) {
      const poly::SpectralPowerDistribution s = sky->GetSpd(poly::Vector(0.f, 0.f, 1.f));
      const float rgb[3] = {s.r, s.g, s.b};
      farr(rgb, 3);
   } else {
      std::fprintf(f, "null");
   }
   std::fprintf(f, ",\n \"meshes\": [");
   for (size_t i = 0; i < runner.Scene->Shapes.size(); i++) {
      const poly::Mesh& m = *runner.Scene->Shapes[i];
      const poly::mesh_geometry& g = *m.mesh_geometry;
      std::fprintf(f, "%s\n  {\"o2w\": ", i ? "," : "");
      mat(*m.object_to_world);
      std::fprintf(f, ", \"w2o\": ");
      mat(*m.world_to_object);
      std::fprintf(f, ", \"light\": %s", m.spd ? "true" : "false");
      if (m.spd) {
         const float e[3] = {m.spd->r, m.spd->g, m.spd->b};
         std::fprintf(f, ", \"spd\": ");
         farr(e, 3);
      }
      const poly::AbstractBRDF* brdf = m.material ? m.material->BRDF.get() : nullptr;
      const char* kind = "none";
      const poly::ReflectanceSpectrum* refl = nullptr;
      const poly::texture* tex = nullptr;
      if (const auto* lam = dynamic_cast<const poly::LambertBRDF*>(brdf)) {
         kind = "lambert";
         if (lam->texture) tex = lam->texture.get();
         else refl = &lam->refl;
      }
      else if (const auto* mir = dynamic_cast<const poly::MirrorBRDF*>(brdf)) { kind = "mirror"; refl = &mir->refl; }
      else if (dynamic_cast<const poly::GlossyBRDF*>(brdf)) kind = "glossy";     // its parameters are private
      else if (brdf) kind = "other";
      std::fprintf(f, ", \"brdf\": \"%s\"", kind);
      if (refl) {
         const float r[3] = {refl->r, refl->g, refl->b};
         std::fprintf(f, ", \"refl\": ");
         farr(r, 3);
      }
      if (tex) {
         // FNV-1a 64 of the decoded texel bytes
         uint64_t h = 1469598103934665603ull;
         for (unsigned char c : tex->data) { h ^= c; h *= 1099511628211ull; }
         std::fprintf(f, ", \"texture\": {\"name\": \"%s\", \"width\": %u, \"height\": %u, \"format\": \"%s\", "
                         "\"bytes\": %zu, \"fnv1a\": \"%016llx\"}",
                      tex->name.c_str(), tex->width, tex->height,
                      tex->data_format == poly::texture::format::RGBA ? "rgba" : "grey", tex->data.size(),
                      (unsigned long long)h);
      }
      std::fprintf(f, ", \"has_uv\": %s, \"u\": ", g.has_vertex_uvs ? "true" : "false");
      farr(g.u_packed.data(), g.u_packed.size());
      std::fprintf(f, ", \"v\": ");
      farr(g.v_packed.data(), g.v_packed.size());
      std::fprintf(f, ", \"has_normals\": %s, \"x\": ", g.has_vertex_normals ? "true" : "false");
      farr(g.x_packed.data(), g.x_packed.size());
      std::fprintf(f, ", \"y\": ");
      farr(g.y_packed.data(), g.y_packed.size());
      std::fprintf(f, ", \"z\": ");
      farr(g.z_packed.data(), g.z_packed.size());
      std::fprintf(f, ", \"nx\": ");
      farr(g.nx_packed.data(), g.nx_packed.size());
      std::fprintf(f, ", \"ny\": ");
      farr(g.ny_packed.data(), g.ny_packed.size());
      std::fprintf(f, ", \"nz\": ");
      farr(g.nz_packed.data(), g.nz_packed.size());
      std::fprintf(f, ", \"fv0\": ");
      uarr(g.fv0);
      std::fprintf(f, ", \"fv1\": ");
      uarr(g.fv1);
      std::fprintf(f, ", \"fv2\": ");
      uarr(g.fv2);
      std::fputc('}', f);
   }
   std::fprintf(f, "\n ]}\n");
   std::fclose(f);
}

// GlossyBRDF::sample (glossy_brdf.h:27-60) on recorded randomness: for each
// case (incoming xyz, roughness, seed) this translation unit's generator
// (constants.h:86-104, a per-translation-unit static) is seeded, the lobe is
// sampled, and a copy of the generator is replayed through the same
// distribution until it reaches the same state, which lists the uniforms the
// sample consumed.  Output per case: outgoing xyz, count, up to 128 uniforms.
void glossy_cases(const std::string& in_path, const std::string& out_path) {
   Reader r(in_path.c_str());
   const uint32_t n = r.u32();
   std::vector<float> out;
   for (uint32_t i = 0; i < n; i++) {
      float v[4];
      r.raw(v, sizeof(v));
      const uint32_t seed = r.u32();
      poly::Generator.seed(seed);
      std::mt19937 replay = poly::Generator;
      const poly::ReflectanceSpectrum one(1.f, 1.f, 1.f);
      poly::GlossyBRDF brdf(one, one, v[3]);
      poly::ReflectanceSpectrum refl;
      float pdf = 0;
      const poly::Vector o = brdf.sample(poly::Vector(v[0], v[1], v[2]), 0.f, 0.f, refl, pdf);
      std::vector<float> used;
      while (!(replay == poly::Generator) && used.size() < 4096) {
         std::uniform_real_distribution<float> d(0.0f, 1.0f);
         used.push_back(d(replay));
      }
      float rec[3 + 1 + 128] = {};
      rec[0] = o.x; rec[1] = o.y; rec[2] = o.z;
      const uint32_t cnt = (uint32_t)used.size();
      std::memcpy(&rec[3], &cnt, 4);
      for (size_t k = 0; k < used.size() && k < 128; k++) rec[4 + k] = used[k];
      out.insert(out.end(), rec, rec + 132);
   }
   write_file(out_path, out.data(), out.size() * 4);
}

// Texture lookups on the reference's own path: for each ray, Scene::intersect,
// then, when the hit mesh is a textured Lambert, LambertBRDF::sample with the
// intersection's u/v_tex_lerp (PathTraceIntegrator.cpp:78, lambert_brdf.h:49-56)
// for the reflectance.  That call is made only where texture::evaluate_rgba's
// index lies inside the texel array (texture.h:25-31); elsewhere the reference
// reads outside it and the record says so.  Record per ray (36 bytes):
//   flags (1 hit, 2 mesh has uvs, 4 textured Lambert, 8 index inside, 16
//   column/row conversion defined), mesh, face, u_tex, v_tex, index, r, g, b.
struct TexProbe {
   uint32_t flags, mesh, face;
   float tu, tv;
   uint32_t index;
   float r, g, b;
};
static_assert(sizeof(TexProbe) == 36, "tex probe record");

void tex_probe(poly::Scene* scene, const std::vector<float>& rays, const std::string& out_path) {
   const size_t n = rays.size() / 6;
   std::vector<TexProbe> out(n);
   for (size_t i = 0; i < n; i++) {
      poly::Ray ray(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2], rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
      poly::Intersection is = scene->intersect(ray, -1, -1);
      TexProbe& p = out[i];
      std::memset(&p, 0, sizeof(p));
      if (!is.Hits) { p.mesh = p.face = 0xFFFFFFFFu; continue; }
      p.flags = 1;
      uint32_t mi = 0;
      for (; mi < scene->Shapes.size(); mi++) if (scene->Shapes[mi] == is.Shape) break;
      p.mesh = mi; p.face = is.faceIndex;
      if (!is.Shape->mesh_geometry->has_vertex_uvs) continue;
      p.flags |= 2;
      p.tu = is.u_tex_lerp; p.tv = is.v_tex_lerp;
      const poly::AbstractBRDF* brdf = is.Shape->material ? is.Shape->material->BRDF.get() : nullptr;
      const auto* lam = dynamic_cast<const poly::LambertBRDF*>(brdf);
      if (!lam || !lam->texture) continue;
      p.flags |= 4;
      const poly::texture& t = *lam->texture;
      const float fu = p.tu * (float)t.width, fv = (1.f - p.tv) * (float)t.height;
      if (!(fu >= 0.f && fu < 4294967296.f && fv >= 0.f && fv < 4294967296.f)) continue;
      p.flags |= 16;
      const unsigned int u_texel = (unsigned int)fu, v_texel = (unsigned int)fv;
      p.index = v_texel * t.width + u_texel;
      if (size_t(p.index) * 4 + 2 >= t.data.size()) continue;
      p.flags |= 8;
      poly::ReflectanceSpectrum refl;
      float pdf = 0;
      (void)lam->sample(poly::Vector(0.f, -1.f, 0.f), p.tu, p.tv, refl, pdf);
      p.r = refl.r; p.g = refl.g; p.b = refl.b;
   }
   write_file(out_path, out.data(), out.size() * sizeof(TexProbe));
}

// Pins worker t to one core, as src/cpu/main.cpp:197-205 pins thread i to
// CPU i; here the t-th CPU of the process's own affinity set, so that the
// pinning also holds where the process may use only some of the host's CPUs.
static void pin_thread(std::thread& th, unsigned t) {
   cpu_set_t allowed;
   CPU_ZERO(&allowed);
   if (sched_getaffinity(0, sizeof(allowed), &allowed) != 0) return;
   const int n = CPU_COUNT(&allowed);
   if (n <= 0) return;
   int want = (int)(t % (unsigned)n), cpu = -1;
   for (int c = 0; c < CPU_SETSIZE; c++)
      if (CPU_ISSET(c, &allowed) && want-- == 0) { cpu = c; break; }
   if (cpu < 0) return;
   cpu_set_t one;
   CPU_ZERO(&one);
   CPU_SET(cpu, &one);
   (void)pthread_setaffinity_np(th.native_handle(), sizeof(one), &one);
}

void usage() {
   std::fprintf(stderr,
      "usage: ref_driver (--ptxm scene.bin | --pbrt scene.pbrt) [--nodes out] [--pairs out] [--describe out.json]\n"
      "                  [--rays rays.bin --hits out] [--rays rays.bin --tex-probe out] [--render out] [--threads N]\n"
      "                  [--time-rays rays.bin --threads N --repeat R] [--time-render --threads N]\n");
   std::exit(2);
}

}  // namespace

int main(int argc, char** argv) {
   std::string ptxm, pbrt, nodes_out, pairs_out, rays_in, hits_out, render_out, time_rays, describe_out;
   std::string glossy_in, glossy_out, tex_out;
   bool png = false;
   bool time_render = false;
   unsigned threads = 1, repeat = 1;
   for (int i = 1; i < argc; i++) {
      std::string a = argv[i];
      auto next = [&]() -> std::string { if (i + 1 >= argc) usage(); return argv[++i]; };
      if (a == "--ptxm") ptxm = next();
      else if (a == "--pbrt") pbrt = next();
      else if (a == "--nodes") nodes_out = next();
      else if (a == "--pairs") pairs_out = next();
      else if (a == "--rays") rays_in = next();
      else if (a == "--hits") hits_out = next();
      else if (a == "--tex-probe") tex_out = next();
      else if (a == "--render") render_out = next();
      else if (a == "--time-rays") time_rays = next();
      else if (a == "--threads") threads = (unsigned)std::atoi(next().c_str());
      else if (a == "--repeat") repeat = (unsigned)std::atoi(next().c_str());
      else if (a == "--time-render") time_render = true;
      else if (a == "--png") png = true;   // after --render: PNGFilm::Output into the working directory
      else if (a == "--describe") describe_out = next();
      else if (a == "--glossy-cases") { glossy_in = next(); glossy_out = next(); }
      else usage();
   }
   if (!glossy_in.empty()) {
      glossy_cases(glossy_in, glossy_out);
      return 0;
   }
   Job job;
   std::unique_ptr<poly::AbstractRunner> runner;
   if (!ptxm.empty()) {
      runner = load_ptxm(ptxm.c_str(), job);
   } else if (!pbrt.empty()) {
      poly::pbrt_parser parser;
      runner = parser.parse_file(pbrt);
      runner->Scene->bvh_root.num_nodes = 0;
      job.width = runner->Bounds.x; job.height = runner->Bounds.y; job.spp = runner->NumSamples;
   } else {
      usage();
   }
   poly::Scene* scene = runner->Scene;
   if (!describe_out.empty()) {
      describe(*runner, describe_out);
      if (nodes_out.empty() && pairs_out.empty() && rays_in.empty() && render_out.empty() && time_rays.empty() &&
          !time_render)
         return 0;
   }

   const auto t0 = std::chrono::steady_clock::now();
   scene->bvh_root.bound(scene->Shapes);
   scene->bvh_root.compact();
   const auto t1 = std::chrono::steady_clock::now();
   std::fprintf(stderr, "ref_driver: bvh %u nodes, %zu pairs, build %.3f s\n",
                scene->bvh_root.compact_root->num_nodes,
                scene->bvh_root.compact_root->leaf_ordered_indices.size(),
                std::chrono::duration<double>(t1 - t0).count());

   const poly::compact_bvh* cb = scene->bvh_root.compact_root;
   if (!nodes_out.empty())
      write_file(nodes_out, cb->nodes, sizeof(poly::compact_bvh_node) * cb->num_nodes);
   if (!pairs_out.empty()) {
      std::vector<uint32_t> flat;
      flat.reserve(cb->leaf_ordered_indices.size() * 2);
      for (const auto& p : cb->leaf_ordered_indices) { flat.push_back(p.first); flat.push_back(p.second); }
      write_file(pairs_out, flat.data(), flat.size() * 4);
   }

   auto load_rays = [](const std::string& path) {
      Reader r(path.c_str());
      const uint32_t n = r.u32();
      std::vector<float> rays(size_t(n) * 6);
      r.raw(rays.data(), rays.size() * 4);
      return rays;
   };

   if (!rays_in.empty() && !hits_out.empty()) {
      std::vector<float> rays = load_rays(rays_in);
      const size_t n = rays.size() / 6;
      std::vector<HitRecord> out(n);
      for (size_t i = 0; i < n; i++) {
         poly::Ray ray(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2],
                       rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
         poly::Intersection is = scene->intersect(ray, -1, -1);
         HitRecord& h = out[i];
         std::memset(&h, 0, sizeof(h));
         h.hit = is.Hits ? 1u : 0u;
         if (is.Hits) {
            uint32_t mi = 0;
            for (; mi < scene->Shapes.size(); mi++) if (scene->Shapes[mi] == is.Shape) break;
            h.mesh = mi; h.face = is.faceIndex;
            h.t = ray.MinT; h.u = is.u; h.v = is.v; h.w = is.w;
            h.loc[0] = is.Location.x; h.loc[1] = is.Location.y; h.loc[2] = is.Location.z;
            h.geo_n[0] = is.geo_normal.x; h.geo_n[1] = is.geo_normal.y; h.geo_n[2] = is.geo_normal.z;
            h.t1[0] = is.Tangent1.x; h.t1[1] = is.Tangent1.y; h.t1[2] = is.Tangent1.z;
            h.t2[0] = is.Tangent2.x; h.t2[1] = is.Tangent2.y; h.t2[2] = is.Tangent2.z;
            h.err[0] = is.error.x; h.err[1] = is.error.y; h.err[2] = is.error.z;
         } else {
            h.mesh = 0xFFFFFFFFu; h.face = 0xFFFFFFFFu;
         }
         h.bb_hits = is.num_bb_hits; h.bb_misses = is.num_bb_misses; h.tri_isects = is.num_triangle_isects;
      }
      write_file(hits_out, out.data(), out.size() * sizeof(HitRecord));
   }

   if (!rays_in.empty() && !tex_out.empty()) tex_probe(scene, load_rays(rays_in), tex_out);

   if (!time_rays.empty()) {
      // CPU baseline: contiguous chunks of the ray list per thread, as a ray-batch
      // analogue of the reference's per-thread tile loop (src/cpu/main.cpp:187-211).
      std::vector<float> rays = load_rays(time_rays);
      const size_t n = rays.size() / 6;
      std::atomic<uint64_t> checksum{0};
      const auto s = std::chrono::steady_clock::now();
      for (unsigned rep = 0; rep < repeat; rep++) {
         std::vector<std::thread> pool;
         for (unsigned t = 0; t < threads; t++) {
            pool.emplace_back([&, t]() {
               const size_t lo = n * t / threads, hi = n * (t + 1) / threads;
               uint64_t local = 0;
               for (size_t i = lo; i < hi; i++) {
                  poly::Ray ray(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2],
                                rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
                  poly::Intersection is = scene->intersect(ray, -1, -1);
                  local += is.Hits ? is.faceIndex : 0xFFFFFFFFu;
               }
               checksum += local;
            });
            pin_thread(pool.back(), t);
         }
         for (auto& th : pool) th.join();
      }
      const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - s).count();
      std::printf("{\"rays\": %zu, \"repeat\": %u, \"threads\": %u, \"seconds\": %.6f, \"mrays_per_s\": %.4f, \"checksum\": %llu}\n",
                  n, repeat, threads, secs, double(n) * repeat / secs / 1e6,
                  (unsigned long long)checksum.load());
   }

   if (time_render) {
      // CPU render baseline: the reference's own runner (TileRunner::Run, one
      // thread per core over its 2x2-pixel tile cursor, src/cpu/runners/TileRunner.cpp:61-117),
      // spawned as src/cpu/main.cpp:187-211 does.
      const auto s = std::chrono::steady_clock::now();
      std::vector<std::thread> pool;
      for (unsigned t = 0; t < threads; t++) {
         pool.push_back(runner->Spawn((int)t));
         pin_thread(pool.back(), t);
      }
      for (auto& th : pool) th.join();
      const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - s).count();
      const double samples = double(runner->Bounds.x) * runner->Bounds.y * runner->NumSamples;
      std::printf("{\"samples\": %.0f, \"threads\": %u, \"seconds\": %.6f, \"msamples_per_s\": %.4f}\n",
                  samples, threads, secs, samples / secs / 1e6);
   }

   if (!render_out.empty()) {
      // All samples of a pixel run in sequence (AbstractRunner::Trace); pixels are
      // split over threads by rows.  Lambertian scenes use the reference's
      // unseeded thread-local RNG, so only mirror/emitter/skybox scenes are
      // reproducible.
      const unsigned W = runner->Bounds.x, H = runner->Bounds.y;
      std::vector<std::thread> pool;
      for (unsigned t = 0; t < threads; t++) {
         pool.emplace_back([&, t]() {
            for (unsigned y = t; y < H; y += threads)
               for (unsigned x = 0; x < W; x++) runner->Trace((int)x, (int)y);
         });
      }
      for (auto& th : pool) th.join();
      std::vector<float> film(size_t(W) * H * 3);
      for (unsigned y = 0; y < H; y++)
         for (unsigned x = 0; x < W; x++) {
            poly::SpectralPowerDistribution spd = runner->Film->Filter->Output(poly::Point2i((int)x, (int)y));
            float* p = &film[(size_t(y) * W + x) * 3];
            p[0] = spd.r; p[1] = spd.g; p[2] = spd.b;
         }
      write_file(render_out, film.data(), film.size() * 4);
      if (png) runner->Output();
   }
   return 0;
}
