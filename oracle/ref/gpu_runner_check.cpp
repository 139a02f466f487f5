// Test-infrastructure driver: renders a PTXM scene through the reference's
// runner API with poly::GpuRunner (polytope_amd/binding/GpuRunner.cpp), the
// drop-in GPU runner, and writes the film as the runner's BoxFilter outputs it
// (W*H*3 raw little-endian floats).  The GPU tests compare that film with the
// reference's own CPU film (deterministic scenes) and with the oracle's.
//
//   gpu_runner_check (--ptxm scene.bin | --pbrt scene.pbrt) --out film.bin [--seed S] [--scene-bvh]
//
// --pbrt reads the scene with the reference's own parser (pbrt_parser::parse_file,
// src/common/parsers/pbrt_parser.cpp:275), PLY meshes and PNG textures included,
// and hands the parsed runner's sampler, scene, integrator and film to GpuRunner.
//
// --scene-bvh builds the tree with the reference's bvh::bound + bvh::compact
// (src/cpu/acceleration/bvh.cpp:36, 695) and hands it to the library instead
// of letting the library build its own.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

#include "GpuRunner.h"
#include "common/parsers/pbrt_parser.h"
#include "common/utilities/Common.h"
#include "cpu/structures/stats.h"
#include "ptxm_scene.h"

poly::Logger Log;
struct poly::stats main_stats;
thread_local struct poly::stats thread_stats;

int main(int argc, char **argv) {
   std::string scene_path, pbrt_path, out_path;
   uint32_t seed = 1;
   bool scene_bvh = false;
   for (int i = 1; i < argc; i++) {
      const std::string a = argv[i];
      if (a == "--ptxm" && i + 1 < argc) scene_path = argv[++i];
      else if (a == "--pbrt" && i + 1 < argc) pbrt_path = argv[++i];
      else if (a == "--out" && i + 1 < argc) out_path = argv[++i];
      else if (a == "--seed" && i + 1 < argc) seed = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
      else if (a == "--scene-bvh") scene_bvh = true;
      else {
         std::fprintf(stderr, "usage: gpu_runner_check (--ptxm scene.bin | --pbrt scene.pbrt) --out film.bin "
                              "[--seed S] [--scene-bvh]\n");
         return 2;
      }
   }
   if ((scene_path.empty() == pbrt_path.empty()) || out_path.empty()) return 2;
   ptxm::Job job;
   ptxm::Parts p;
   std::unique_ptr<poly::AbstractRunner> parsed;
   if (!pbrt_path.empty()) {
      poly::pbrt_parser parser;
      parsed = parser.parse_file(pbrt_path);
      parsed->Scene->bvh_root.num_nodes = 0;
      p.sampler = std::move(parsed->Sampler);
      p.scene = parsed->Scene;
      p.integrator = std::move(parsed->Integrator);
      p.film = std::move(parsed->Film);
      p.bounds = parsed->Bounds;
      job.spp = parsed->NumSamples;
   } else {
      p = ptxm::load_parts(scene_path.c_str(), job);
   }
   if (scene_bvh) {
      p.scene->bvh_root.bound(p.scene->Shapes);
      p.scene->bvh_root.compact();
   }
   poly::GpuRunner runner(std::move(p.sampler), p.scene, std::move(p.integrator), std::move(p.film), job.spp,
                          p.bounds, seed, 0, scene_bvh);
   // the reference's main spawns one Run per thread (src/cpu/main.cpp:187-211)
   std::vector<std::thread> threads;
   for (int t = 0; t < 4; t++) threads.push_back(runner.Spawn(t));
   for (auto &t : threads) t.join();
   if (runner.status() != 0) {
      std::fprintf(stderr, "gpu_runner_check: %s\n", runner.error().c_str());
      return 1;
   }
   const unsigned W = runner.Bounds.x, H = runner.Bounds.y;
   std::vector<float> film((size_t)W * H * 3);
   for (unsigned y = 0; y < H; y++)
      for (unsigned x = 0; x < W; x++) {
         const poly::SpectralPowerDistribution s = runner.Film->Filter->Output(poly::Point2i((int)x, (int)y));
         float *q = &film[((size_t)y * W + x) * 3];
         q[0] = s.r; q[1] = s.g; q[2] = s.b;
      }
   ptxm::write_file(out_path, film.data(), film.size() * sizeof(float));
   std::printf("gpu_runner_check: %ux%u, %u spp\n", W, H, job.spp);
   return 0;
}
