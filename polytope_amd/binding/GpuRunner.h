// GpuRunner: the MI355X render path as a drop-in runner for polytope.
//
// A poly::AbstractRunner (src/cpu/runners/AbstractRunner.h:18-58) whose Run()
// renders the whole frame on the GPU through libptx_hip.so (include/ptx.h) and
// fills the runner's film, so that AbstractRunner::Output() writes the image
// the CPU runners would, from the same Scene, PerspectiveCamera, materials,
// ColorSkybox and PathTraceIntegrator{MaxDepth}.  It is compiled inside the
// reference tree (it includes the reference's own headers) and linked against
// libptx_hip.so; nothing else in the reference changes:
//
//   auto runner = std::make_unique<poly::GpuRunner>(std::move(sampler), scene,
//        std::move(integrator), std::move(film), spp, bounds);
//   runner->Run(0);      // or Spawn(i) for every thread id, as main.cpp does
//   runner->Output();
//
// Supported descriptions are the ones the reference's CPU path renders:
// CenterSampler, PathTraceIntegrator, LambertBRDF (constant or with an RGBA
// texture and the mesh's u, v) / MirrorBRDF materials, emitters (Mesh::spd),
// an optional ColorSkybox and the
// left-handed PerspectiveCamera the pbrt parser builds.  Anything else makes
// Run() fail with status() != 0 and a message in error(); the film is then
// left empty.  Lambert bounces draw from the library's counter-based RNG
// (Philox4x32-10 keyed by `seed`), not the reference's unseeded mt19937.
#pragma once

#include <cstdint>
#include <memory>
#include <string>

#include "cpu/runners/AbstractRunner.h"
#include "ptx.h"

namespace poly {

class GpuRunner : public AbstractRunner {
public:
   GpuRunner(std::unique_ptr<AbstractSampler> sampler, poly::Scene *scene,
             std::unique_ptr<AbstractIntegrator> integrator, std::unique_ptr<AbstractFilm> film,
             unsigned int numSamples, const poly::Bounds bounds, uint32_t seed = 1, int device = 0,
             bool use_scene_bvh = false)
         : AbstractRunner(std::move(sampler), scene, std::move(integrator), std::move(film), numSamples, bounds),
           seed_(seed), device_(device), use_scene_bvh_(use_scene_bvh) {}

   // Thread 0 renders the frame; every other thread id returns at once, so a
   // main that spawns one Run per core drives it unchanged.
   void Run(int threadId) override;

   int status() const { return status_; }
   const std::string &error() const { return error_; }

private:
   int fail(int code, const std::string &msg);

   uint32_t seed_;
   int device_;
   bool use_scene_bvh_;   // hand Scene::bvh_root's compact tree to the library
   int status_ = PTX_OK;
   std::string error_;
};

}  // namespace poly
