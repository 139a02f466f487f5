// Path-tracing kernels for gfx950: the bounce loop of
// PathTraceIntegrator::GetSample (src/cpu/integrators/PathTraceIntegrator.cpp:10-135)
// with CenterSampler pixel centres (src/cpu/samplers/samplers.cpp:13-18), the
// pinhole camera (src/cpu/cameras/PerspectiveCamera.cpp:27-49), Lambert and mirror
// BRDFs (src/cpu/shading/brdf/*.h) and BoxFilter accumulation
// (src/cpu/filters/BoxFilter.cpp:14-71), in three forms:
//
//   plain     k_mega: one launch per sample, one thread per pixel path with every
//             bounce inline — the launch shape of the reference CUDA path
//             (PathTracerKernel::Trace, src/cuda/kernels/path_tracer.cu:899-1018);
//   regen     k_primary traces each pixel's primary ray once (CenterSampler gives
//             every sample of a pixel the same ray); then one persistent k_paths
//             launch per round of path slots (many jobs' samples) in which every
//             lane carries a path from its cached primary hit through all its
//             bounces and takes the next slot as soon as it ends;
//   wavefront k_primary as above, k_round_start shades the primary hit per
//             sample, then per bounce k_extend (the persistent traversal of
//             traverse.h fed from a queue) and k_shade (frame, BRDF, compaction
//             into the next queue).
// Batched rounds end with k_accumulate, which adds each pixel's samples to its
// film in sample order; k_finalize scales by 1/spp.
//
// All forms run the same arithmetic in the same order, so their films are
// identical to each other and to the CPU restatement.
#include <hip/hip_runtime.h>

#include "launch.h"
#include "divmagic.h"
#include "traverse.h"

#pragma clang fp contract(off)

namespace ptx {
namespace dev {

constexpr float kPI = 3.14159265358979323846f;
constexpr float kTwoPI = 2.0f * kPI;
constexpr uint32_t kRngKeyHi = 0x706F6C79u;   // second Philox key word: ASCII "poly"
constexpr uint32_t kNone = 0xFFFFFFFFu;

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC 2011): ten rounds of the
// multiply-xor bijection with the Weyl key schedule.
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                              uint32_t k1, uint32_t &o0, uint32_t &o1) {
#pragma unroll
   for (int r = 0; r < 10; r++) {
      const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
      const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
      const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
      c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
      k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
   }
   o0 = c0;
   o1 = c1;
}

__device__ __forceinline__ float u01(uint32_t r) { return (float)(r >> 8) * 0x1p-24f; }

// sin and cos for theta in [0, 2pi]: quadrant k = floor(theta*2/pi + 1/2),
// three-part Cody-Waite reduction, Taylor polynomials on [-pi/4, pi/4] in Horner
// form with explicit fma (the same operation sequence as the CPU restatement).
__device__ __forceinline__ void sincos_poly(float theta, float &s_out, float &c_out) {
   const float q = theta * 0.636619772367581343f;
   const int k = (int)__builtin_floorf(q + 0.5f);
   const float kf = (float)k;
   float r = fmaf_(-kf, 1.57079625129699707031f, theta);
   r = fmaf_(-kf, 7.54978941586159635335e-08f, r);
   r = fmaf_(-kf, 5.39030252995776476554e-15f, r);
   const float r2 = r * r;
   float ps = fmaf_(r2, 2.75573192239858906526e-06f, -1.98412698412698412698e-04f);
   ps = fmaf_(r2, ps, 8.33333333333333333333e-03f);
   ps = fmaf_(r2, ps, -1.66666666666666666667e-01f);
   const float sn = fmaf_(r * r2, ps, r);
   float pc = fmaf_(r2, -2.75573192239858906526e-07f, 2.48015873015873015873e-05f);
   pc = fmaf_(r2, pc, -1.38888888888888888889e-03f);
   pc = fmaf_(r2, pc, 4.16666666666666666667e-02f);
   pc = fmaf_(r2, pc, -0.5f);
   const float cs = fmaf_(r2, pc, 1.0f);
   switch (k & 3) {
      case 0: s_out = sn; c_out = cs; break;
      case 1: s_out = cs; c_out = -sn; break;
      case 2: s_out = -sn; c_out = -cs; break;
      default: s_out = -cs; c_out = sn; break;
   }
}

// acos for the glossy lobe, project-owned so that the GPU and the oracle agree
// bitwise: acos(x) = pi/2 - asin(x) for |x| <= 1/2, else 2 asin(sqrt((1 - |x|)/2)),
// taken from pi for x < 0; asin(t) = t + t^3 P(t^2), P the Maclaurin series of
// (asin t - t) / t^3 (coefficients (2n)! / (4^n (n!)^2 (2n + 1)), n = 1..12) in
// Horner form with explicit fma.  |x| > 1 gives NaN, as acosf does.
__device__ __forceinline__ float asin_series(float t, float z) {
   float p = fmaf_(z, 0.006447210311889649f, 0.0073125258735988454f);
   p = fmaf_(z, p, 0.008390335809616815f);
   p = fmaf_(z, p, 0.009761609529194078f);
   p = fmaf_(z, p, 0.011551800896139705f);
   p = fmaf_(z, p, 0.01396484375f);
   p = fmaf_(z, p, 0.017352764423076924f);
   p = fmaf_(z, p, 0.022372159090909092f);
   p = fmaf_(z, p, 0.030381944444444444f);
   p = fmaf_(z, p, 0.044642857142857144f);
   p = fmaf_(z, p, 0.075f);
   p = fmaf_(z, p, 0.16666666666666666f);
   return fmaf_(t * z, p, t);
}

__device__ __forceinline__ float acos_poly(float x) {
   const float a = __builtin_fabsf(x);
   if (a <= 0.5f) return 1.57079637050628662109f - asin_series(x, x * x);
   const float z = (1.0f - a) * 0.5f;
   const float r = 2.0f * asin_series(__builtin_sqrtf(z), z);
   return x < 0.0f ? 3.14159274101257324219f - r : r;
}

__device__ __forceinline__ void normalize3(float &x, float &y, float &z) {   // Vector::Normalize
   const float ld = 1.0f / __builtin_sqrtf(x * x + y * y + z * z);
   x *= ld; y *= ld; z *= ld;
}

// PerspectiveCamera::get_ray_for_pixel at the pixel centre (CenterSampler).
__device__ __forceinline__ void camera_ray(const JobDev &J, uint32_t x, uint32_t y, float o[3], float d[3]) {
   const float px = (float)x + 0.5f, py = (float)y + 0.5f;
   const float nx = px * J.oow, ny = py * J.ooh;
   const float cx = (2 * nx - 1) * J.aspect * J.tanf2;
   const float cy = (1 - 2 * ny) * J.tanf2;
   float dx = cx, dy = cy, dz = 1.0f;
   normalize3(dx, dy, dz);
   const float *m = J.cam;
   // Transform::Apply(Point) of the origin and Apply(Vector) (Transform.cpp:93-133)
   const float ox = 0.f * m[0] + 0.f * m[1] + 0.f * m[2] + m[3];
   const float oy = 0.f * m[4] + 0.f * m[5] + 0.f * m[6] + m[7];
   const float oz = 0.f * m[8] + 0.f * m[9] + 0.f * m[10] + m[11];
   const float w = 0.f * m[12] + 0.f * m[13] + 0.f * m[14] + m[15];
   if (w == 1) { o[0] = ox; o[1] = oy; o[2] = oz; }
   else {
      const float dv = 1.f / w;
      o[0] = ox * dv; o[1] = oy * dv; o[2] = oz * dv;
   }
   float wx = dx * m[0] + dy * m[1] + dz * m[2];
   float wy = dx * m[4] + dy * m[5] + dz * m[6];
   float wz = dx * m[8] + dy * m[9] + dz * m[10];
   normalize3(wx, wy, wz);
   d[0] = wx; d[1] = wy; d[2] = wz;
}

struct Radiance {
   float r, g, b;
};

// Terminal radiance of a path whose ray misses (PathTraceIntegrator.cpp:46-54).
__device__ __forceinline__ Radiance miss_radiance(const DevScene &S, const float src[3]) {
   Radiance L = {0.f, 0.f, 0.f};
   if (S.has_sky) L = {S.sky[0] * src[0], S.sky[1] * src[1], S.sky[2] * src[2]};
   L.r += 0.f; L.g += 0.f; L.b += 0.f;
   return L;
}

// GlossyBRDF::sample (glossy_brdf.h:27-60): a cosine-hemisphere sample rotated
// about (mirror direction x n) by -acos(mirror . n), n = (0, 1, 0), then mixed
// with the mirror direction by the roughness; drawn again (uniforms of counter
// word 3 = attempt) while it points below the surface.  The reference loops
// without bound; after kGlossyTries attempts this returns false and the path
// ends black.  The result is not normalised (LocalToWorld does that).
constexpr uint32_t kGlossyTries = 64;

__device__ __forceinline__ bool glossy_sample(const float in[3], float rough, uint32_t pixel, uint32_t sample,
                                              uint32_t bounce_index, uint32_t seed, float &lx, float &ly,
                                              float &lz) {
   for (uint32_t a = 0; a < kGlossyTries; a++) {
      // the mirror frame below depends only on `in`; recomputing it per attempt
      // keeps four values live across the loop instead of the whole frame
      float i0 = in[0], i1 = in[1], i2 = in[2];
      asm volatile("" : "+v"(i0), "+v"(i1), "+v"(i2));
      uint32_t r0, r1;
      philox4x32_10(pixel, sample, bounce_index, a, seed, kRngKeyHi, r0, r1);
      const float u0 = u01(r0), u1 = u01(r1);
      const float r = __builtin_sqrtf(u0);                   // CosineSampleHemisphere
      const float theta = kTwoPI * u1;
      float sn, cs;
      sincos_poly(theta, sn, cs);
      const float om = 1.0f - u0;
      const float hx = r * cs, hy = __builtin_sqrtf(0.0f < om ? om : 0.0f), hz = r * sn;
      const float factor = (0.f * i0 + 1.f * i1 + 0.f * i2) * 2;
      float sx = i0 - 0.f * factor, sy = i1 - 1.f * factor, sz = i2 - 0.f * factor;
      normalize3(sx, sy, sz);
      const float angle = -acos_poly(0.f * sx + 1.f * sy + 0.f * sz);
      const float x = sy * 0.f - sz * 1.f, y = sz * 0.f - sx * 0.f, z = sx * 1.f - sy * 0.f;
      float s, c;                                              // Transform::Rotate (Transform.cpp:214-228)
      sincos_poly(angle, s, c);
      const float omc = 1 - c;
      const float m00 = c + x * x * omc, m01 = x * y * omc - z * s, m02 = x * z * omc + y * s;
      const float m10 = y * x * omc + z * s, m11 = c + y * y * omc, m12 = y * z * omc - x * s;
      const float m20 = z * x * omc - y * s, m21 = z * y * omc + x * s, m22 = c + z * z * omc;
      const float rx = hx * m00 + hy * m01 + hz * m02;        // Transform::Apply(Vector)
      const float ry = hx * m10 + hy * m11 + hz * m12;
      const float rz = hx * m20 + hy * m21 + hz * m22;
      const float k = 1 - rough;
      const float ox = sx * k + rx * rough, oy = sy * k + ry * rough, oz = sz * k + rz * rough;
      if (oy < 0) continue;
      lx = ox; ly = oy; lz = oz;
      return true;
   }
   return false;
}

// The bounce (PathTraceIntegrator.cpp:71-95): local incoming direction,
// BRDF sample, world outgoing direction; updates src.  For glossy meshes
// M.refl is the lobe's constant reflectance and M.emission[0] its roughness.
template <bool kGlossy = true>
__device__ __forceinline__ void bounce(const MeshRec &M, const Frame &fr, const float lin[3], uint32_t pixel,
                                       uint32_t sample, uint32_t bounce_index, uint32_t seed, float wout[3],
                                       float src[3]) {
   float lx, ly, lz;
   bool live = true;
   if (M.brdf == PTX_BRDF_MIRROR) {            // MirrorBRDF::sample (mirror_brdf.h:23-35)
      lx = lin[0]; ly = -lin[1]; lz = lin[2];
   } else if (kGlossy && M.brdf == PTX_BRDF_GLOSSY) {
      live = glossy_sample(lin, M.emission[0], pixel, sample, bounce_index, seed, lx, ly, lz);
      if (!live) { lx = 0.f; ly = 1.f; lz = 0.f; }
   } else {                                     // LambertBRDF -> AbstractBRDF::sample -> CosineSampleHemisphere
      uint32_t r0, r1;
      philox4x32_10(pixel, sample, bounce_index, 0u, seed, kRngKeyHi, r0, r1);
      const float u0 = u01(r0), u1 = u01(r1);
      const float r = __builtin_sqrtf(u0);
      const float theta = kTwoPI * u1;
      float sn, cs;
      sincos_poly(theta, sn, cs);
      const float om = 1.0f - u0;
      lx = r * cs;
      ly = __builtin_sqrtf(0.0f < om ? om : 0.0f);
      lz = r * sn;
   }
   // Intersection::LocalToWorld (Intersection.cpp:16-23)
   float wx = fr.t1x * lx + fr.nx * ly + fr.t2x * lz;
   float wy = fr.t1y * lx + fr.ny * ly + fr.t2y * lz;
   float wz = fr.t1z * lx + fr.nz * ly + fr.t2z * lz;
   normalize3(wx, wy, wz);
   wout[0] = wx; wout[1] = wy; wout[2] = wz;
   src[0] = live ? src[0] * M.refl[0] : 0.f;
   src[1] = live ? src[1] * M.refl[1] : 0.f;
   src[2] = live ? src[2] * M.refl[2] : 0.f;
}

// Intersection::WorldToLocal (Intersection.cpp:10-14)
__device__ __forceinline__ void world_to_local(const Frame &fr, const float d[3], float lin[3]) {
   float x = fr.t1x * d[0] + fr.t1y * d[1] + fr.t1z * d[2];
   float y = fr.nx * d[0] + fr.ny * d[1] + fr.nz * d[2];
   float z = fr.t2x * d[0] + fr.t2y * d[1] + fr.t2z * d[2];
   normalize3(x, y, z);
   lin[0] = x; lin[1] = y; lin[2] = z;
}

__device__ __forceinline__ bool src_is_zero(const float src[3]) {        // spectrum.h:65-67
   return src[0] < 0.0039f && src[1] < 0.0039f && src[2] < 0.0039f;
}

// ---------------------------------------------------------------- next-event estimation

// One point light seen from P with shading normal n: direction l, distance d and
// the geometry factor g = cos / (pi * d^2); returns false when cos <= 0.
__device__ __forceinline__ bool light_setup(const float *light, float px, float py, float pz, float nx, float ny,
                                            float nz, float l[3], float &d, float &g) {
   const float wx = light[0] - px, wy = light[1] - py, wz = light[2] - pz;
   const float d2 = wx * wx + wy * wy + wz * wz;
   d = __builtin_sqrtf(d2);
   const float inv = 1.0f / d;
   l[0] = wx * inv; l[1] = wy * inv; l[2] = wz * inv;
   const float cosv = nx * l[0] + ny * l[1] + nz * l[2];
   if (!(cosv > 0.f)) return false;
   g = cosv / (kPI * d2);
   return true;
}

// Shadow ray: closest-hit traversal bounded by tmax that stops at the first
// accepted face.  No debug-pixel quirk on shadow rays.
template <int kLds>
__device__ __forceinline__ bool occluded(const DevScene &S, const Stack<kLds> &st, float ox, float oy, float oz,
                                         const float l[3], float tmax) {
   Trav s;
   make_ray(s.r, ox, oy, oz, l[0], l[1], l[2]);
   begin(s, S.root_box, S.root_ref, S.root_count, tmax);
   while (s.cur != kNone) {
      Fetched f;
      fetch(s, S.nodes, S.tris, f);
      const bool more = advance<kLds>(s, f, st, false);
      if (s.hit != kNone) return true;
      if (!more) break;
   }
   return false;
}

__device__ __forceinline__ Frame frame_of(const DevScene &S, uint32_t tri, const float d[3], float u, float v,
                                          float w, uint32_t after, uint32_t mesh) {
   Ray r;
   r.dx = d[0]; r.dy = d[1]; r.dz = d[2];
   const float *cn = nullptr;
   if (S.normals && (S.meshes[mesh].flags & kRecNormals)) cn = S.normals + 9 * (size_t)tri;
   Frame fr;
   hit_frame(S.tris + 9 * (size_t)tri, cn, r, u, v, w, after, fr);
   return fr;
}

// ---------------------------------------------------------------- textures

// (unsigned int)f where the conversion is defined (0 <= f < 2^32); 0 for a
// negative or NaN f, 0xFFFFFFFF above.
__device__ __forceinline__ uint32_t texel_trunc(float f) {
   return f > 0.f ? (f < 4294967296.f ? (uint32_t)f : 0xFFFFFFFFu) : 0u;
}

// Texel of a hit on a textured Lambert mesh (LambertBRDF::sample,
// lambert_brdf.h:49-56): the texture coordinates Mesh::intersect interpolates
// (mesh.cpp:342-353; u, v, w weight the face's corners 0, 1, 2), then
// texture::evaluate_rgba's flat index row * W + column (texture.h:25-31).
// Where that index falls outside the image (the reference reads outside its
// texel array there) the column and the row clamp to the image.
__device__ __forceinline__ uint32_t texel_index(const float uv[6], float u, float v, float w, uint32_t wh) {
   const float tu = (uv[0] * u + uv[2] * v) + uv[4] * w;
   const float tv = (uv[1] * u + uv[3] * v) + uv[5] * w;
   const uint32_t W = wh & 0x7FFFu, H = (wh >> 15) & 0x7FFFu;
   const uint32_t col = texel_trunc(tu * (float)W), row = texel_trunc((1.0f - tv) * (float)H);
   const uint64_t flat = (uint64_t)row * W + col;
   if (flat < (uint64_t)W * H) return (uint32_t)flat;
   return min(row, H - 1u) * W + min(col, W - 1u);
}

// ReflectanceSpectrum(int r, int g, int b) (spectrum.h:17-18) of an RGBA8 texel.
__device__ __forceinline__ void texel_refl(uint32_t t, float refl[3]) {
   refl[0] = (float)(t & 255u) * (1.0f / 255.0f);
   refl[1] = (float)((t >> 8) & 255u) * (1.0f / 255.0f);
   refl[2] = (float)((t >> 16) & 255u) * (1.0f / 255.0f);
}

// The texel pool of a scene with texture records (after its num_tris records).
__device__ __forceinline__ const uint32_t *texel_pool(const DevScene &S) {
   return reinterpret_cast<const uint32_t *>(S.tex_tris + S.num_tris);
}

// The RGBA8 texel under a hit on leaf-ordered triangle `tri` of a textured mesh.
__device__ __forceinline__ uint32_t tri_texel(const DevScene &S, uint32_t tri, float u, float v, float w) {
   const TexTri &t = S.tex_tris[tri];
   const float uv[6] = {t.uv[0], t.uv[1], t.uv[2], t.uv[3], t.uv[4], t.uv[5]};
   return texel_pool(S)[t.texel0 + texel_index(uv, u, v, w, t.wh)];
}

// ---------------------------------------------------------------- primary hits

// Record per pixel (16 words): w0 = kind | mesh << 2 (kind 0 = the path
// bounces, 1 = it ends here with radiance w1..w3 for src = 1), or, for a
// textured Lambert hit, 2 | RGB texel << 8; for kinds 0 and 2: w1..3 Location,
// w4..6 tangent1, w7..9 normal, w10..12 tangent2, w13..15 the normalised local
// incoming direction.  Written from the pixel's finished primary traversal s.
__device__ __forceinline__ void write_primary_record(const DevScene &S, const JobDev &J, uint32_t pixel,
                                                     const Trav &s) {
   float rec[16];
#pragma unroll
   for (int k = 0; k < 16; k++) rec[k] = 0.f;
   const float one[3] = {1.f, 1.f, 1.f};
   if (s.hit == kNone) {
      const Radiance L = miss_radiance(S, one);
      rec[0] = __uint_as_float(1u);
      rec[1] = L.r; rec[2] = L.g; rec[3] = L.b;
   } else {
      const uint32_t mesh = S.pairs[2 * (size_t)s.hit];
      const MeshRec &M = S.meshes[mesh];
      if (M.is_light) {
         rec[0] = __uint_as_float(1u | (mesh << 2));
         rec[1] = (M.emission[0] + 0.f) * 1.f;
         rec[2] = (M.emission[1] + 0.f) * 1.f;
         rec[3] = (M.emission[2] + 0.f) * 1.f;
      } else if (J.max_depth == 0) {
         rec[0] = __uint_as_float(1u | (mesh << 2));
      } else {
         const float d[3] = {s.r.dx, s.r.dy, s.r.dz};
         const Frame fr = frame_of(S, s.hit, d, s.hu, s.hv, s.hw, s.after, mesh);
         float lin[3];
         world_to_local(fr, d, lin);
         uint32_t w0 = mesh << 2;
         if (M.flags & kRecTextured) w0 = 2u | (tri_texel(S, s.hit, s.hu, s.hv, s.hw) & 0xFFFFFFu) << 8;
         rec[0] = __uint_as_float(w0);
         rec[1] = fr.lx; rec[2] = fr.ly; rec[3] = fr.lz;
         rec[4] = fr.t1x; rec[5] = fr.t1y; rec[6] = fr.t1z;
         rec[7] = fr.nx; rec[8] = fr.ny; rec[9] = fr.nz;
         rec[10] = fr.t2x; rec[11] = fr.t2y; rec[12] = fr.t2z;
         rec[13] = lin[0]; rec[14] = lin[1]; rec[15] = lin[2];
      }
   }
   float4 *dst = reinterpret_cast<float4 *>(J.primary + 16 * (size_t)pixel);
   dst[0] = make_float4(rec[0], rec[1], rec[2], rec[3]);
   dst[1] = make_float4(rec[4], rec[5], rec[6], rec[7]);
   dst[2] = make_float4(rec[8], rec[9], rec[10], rec[11]);
   dst[3] = make_float4(rec[12], rec[13], rec[14], rec[15]);
}

struct PrimaryPolicy {
   struct Ctx {
      uint32_t dbg;
   };
   const JobDev *job;          // device pointer (per launch: one job)
   DevScene S;
   JobDev J;
   __device__ __forceinline__ void load(uint32_t, Fetched &) const {}
   __device__ __forceinline__ int start(uint32_t id, const Fetched &, Ctx &ctx, Trav &s) const {
      const uint32_t x = id % J.width, y = id / J.width;
      float o[3], d[3];
      camera_ray(J, x, y, o, d);
      make_ray(s.r, o[0], o[1], o[2], d[0], d[1], d[2]);
      ctx.dbg = (x == 128 && y == 128) ? 1u : 0u;
      return 0;
   }
   __device__ __forceinline__ void load2(Ctx &, Fetched &) const {}
   __device__ __forceinline__ void start2(Ctx &, const Fetched &, Trav &) const {}
   __device__ __forceinline__ void begin_root(Ctx &, Trav &s) const { begin(s, S.root_box, S.root_ref, S.root_count); }
   __device__ __forceinline__ const Node64 *nodes(const Ctx &) const { return S.nodes; }
   __device__ __forceinline__ const float *tris(const Ctx &) const { return S.tris; }
   __device__ __forceinline__ bool debug_pixel(const Ctx &ctx) const { return ctx.dbg != 0; }
   __device__ __forceinline__ void finish(uint32_t id, const Ctx &, const Trav &s) const {
      write_primary_record(S, J, id, s);
   }
};

// Primary rays of every job of a batch in one launch: ids are the jobs' pixels
// back to back (offsets[j] = the first id of job j, offsets[njobs] = the
// total), each lane finds its job by binary search and traces in that job's
// scene.
struct PrimariesPolicy {
   struct Ctx {
      uint32_t dbg, job;
      const Node64 *nodes;
      const float *tris;
   };
   const JobDev *jobs;
   const DevScene *scenes;
   const uint32_t *offsets;
   uint32_t njobs;
   __device__ __forceinline__ void load(uint32_t, Fetched &) const {}
   __device__ __forceinline__ int start(uint32_t id, const Fetched &, Ctx &ctx, Trav &s) const {
      uint32_t lo = 0, hi = njobs;
      while (hi - lo > 1) {
         const uint32_t mid = (lo + hi) / 2;
         if (offsets[mid] <= id) lo = mid; else hi = mid;
      }
      const JobDev &J = jobs[lo];
      const uint32_t p = id - offsets[lo];
      const uint32_t x = p % J.width, y = p / J.width;
      float o[3], d[3];
      camera_ray(J, x, y, o, d);
      make_ray(s.r, o[0], o[1], o[2], d[0], d[1], d[2]);
      ctx.dbg = (x == 128 && y == 128) ? 1u : 0u;
      ctx.job = lo;
      const DevScene &S = scenes[J.scene];
      ctx.nodes = S.nodes;
      ctx.tris = S.tris;
      return 0;
   }
   __device__ __forceinline__ void load2(Ctx &, Fetched &) const {}
   __device__ __forceinline__ void start2(Ctx &, const Fetched &, Trav &) const {}
   __device__ __forceinline__ void begin_root(Ctx &ctx, Trav &s) const {
      const DevScene &S = scenes[jobs[ctx.job].scene];
      begin(s, S.root_box, S.root_ref, S.root_count);
   }
   __device__ __forceinline__ const Node64 *nodes(const Ctx &ctx) const { return ctx.nodes; }
   __device__ __forceinline__ const float *tris(const Ctx &ctx) const { return ctx.tris; }
   __device__ __forceinline__ bool debug_pixel(const Ctx &ctx) const { return ctx.dbg != 0; }
   __device__ __forceinline__ void finish(uint32_t id, const Ctx &ctx, const Trav &s) const {
      const JobDev &J = jobs[ctx.job];
      write_primary_record(scenes[J.scene], J, id - offsets[ctx.job], s);
   }
};

template <int kLds>
__global__ void __launch_bounds__(256) k_primary(DevScene S, JobDev J, uint32_t *work, uint4 *spill, uint32_t leaf_min) {
   const PrimaryPolicy P{nullptr, S, J};
   persistent_trace<kLds>(P, J.width * J.height, work, spill, 1u, leaf_min, S.max_stack);
}

template <int kLds>
__global__ void __launch_bounds__(256) k_primaries(const JobDev *jobs, const DevScene *scenes, const uint32_t *offsets,
                                                   uint32_t njobs, uint32_t *work, uint4 *spill, uint32_t leaf_min,
                                                   uint32_t max_stack) {
   const PrimariesPolicy P{jobs, scenes, offsets, njobs};
   persistent_trace<kLds>(P, offsets[njobs], work, spill, 1u, leaf_min, max_stack);
}

// ---------------------------------------------------------------- path queues

// Queue entry, 16 words: path slot, job, pixel, sample, origin xyz, direction
// xyz, throughput rgb, flags (bit 0: the reference's debug pixel (128, 128)),
// scene, unused.
constexpr int kQWords = 16;

struct QueueRefs {
   float *q;                   // [capacity][16]
   uint32_t *count;            // appended entries
};

// Appends one entry per calling lane with `want`; one atomic per wave.
__device__ __forceinline__ void enqueue(const QueueRefs &Q, bool want, const float e[kQWords]) {
   const uint64_t m = __ballot(want);
   if (!m) return;
   const uint32_t lane = __lane_id();
   const uint32_t leader = __ffsll((unsigned long long)m) - 1;
   uint32_t base = 0;
   if (lane == leader) base = atomicAdd(Q.count, (uint32_t)__popcll(m));
   base = __shfl(base, leader);
   if (want) {
      const uint32_t slot = base + __popcll(m & ((1ull << lane) - 1ull));
      float4 *dst = reinterpret_cast<float4 *>(Q.q + (size_t)kQWords * slot);
      dst[0] = make_float4(e[0], e[1], e[2], e[3]);
      dst[1] = make_float4(e[4], e[5], e[6], e[7]);
      dst[2] = make_float4(e[8], e[9], e[10], e[11]);
      dst[3] = make_float4(e[12], e[13], e[14], e[15]);
   }
}

// Shadow-connect queue of the wavefront form (next-event estimation): one
// record per Lambert vertex of a next-event job, 12 words: the vertex P and the
// path slot, the shading normal and the scene, throughput x Kd.
struct ShadowRefs {
   float4 *q;                  // [capacity][3], or null: no next-event job in the round
   uint32_t *count;
};

__device__ __forceinline__ void enqueue_shadow(const ShadowRefs &Q, bool want, const Frame &fr, uint32_t slot,
                                               uint32_t scene, const float k[3]) {
   const uint64_t m = __ballot(want);
   if (!m) return;
   const uint32_t lane = __lane_id();
   const uint32_t leader = __ffsll((unsigned long long)m) - 1;
   uint32_t base = 0;
   if (lane == leader) base = atomicAdd(Q.count, (uint32_t)__popcll(m));
   base = __shfl(base, leader);
   if (want) {
      float4 *dst = Q.q + 3 * (size_t)(base + __popcll(m & ((1ull << lane) - 1ull)));
      dst[0] = make_float4(fr.lx, fr.ly, fr.lz, __uint_as_float(slot));
      dst[1] = make_float4(fr.nx, fr.ny, fr.nz, __uint_as_float(scene));
      dst[2] = make_float4(k[0], k[1], k[2], 0.f);
   }
}

__device__ __forceinline__ void write_res(float *res, uint32_t path, Radiance L) {
   res[3 * (size_t)path] = L.r;
   res[3 * (size_t)path + 1] = L.g;
   res[3 * (size_t)path + 2] = L.b;
}

__device__ __forceinline__ uint32_t find_chunk(const Chunk *chunks, uint32_t nchunks, uint32_t slot) {
   uint32_t lo = 0, hi = nchunks;
   while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (chunks[mid].base <= slot) lo = mid; else hi = mid;
   }
   return lo;
}

// Shades every path slot of a round at its primary hit (bounce 0).
__global__ void __launch_bounds__(256) k_round_start(const Chunk *chunks, uint32_t nchunks, uint32_t nslots,
                                                     const JobDev *jobs, const DevScene *scenes, float *res,
                                                     QueueRefs out, ShadowRefs shadow, float *direct) {
   for (uint32_t slot0 = blockIdx.x * blockDim.x; slot0 < nslots; slot0 += gridDim.x * blockDim.x) {
      const uint32_t slot = slot0 + threadIdx.x;
      bool want = false, want_shadow = false;
      float e[kQWords];
      Frame fr;
      float kd[3] = {0.f, 0.f, 0.f};
      uint32_t scene = 0;
      if (slot < nslots && shadow.q) {
         direct[3 * (size_t)slot] = 0.f;
         direct[3 * (size_t)slot + 1] = 0.f;
         direct[3 * (size_t)slot + 2] = 0.f;
      }
      if (slot < nslots) {
         const Chunk c = chunks[find_chunk(chunks, nchunks, slot)];
         const uint32_t local = slot - c.base;
         const uint32_t sample = c.s0 + local / c.npix, pixel = local % c.npix;
         const JobDev &J = jobs[c.job];
         const float4 *pr = reinterpret_cast<const float4 *>(J.primary + 16 * (size_t)pixel);
         const float4 a = pr[0], b = pr[1], cc = pr[2], d = pr[3];
         const uint32_t w0 = __float_as_uint(a.x);
         if (w0 & 1u) {
            write_res(res, slot, Radiance{a.y, a.z, a.w});
         } else {
            const DevScene &S = scenes[J.scene];
            MeshRec M;
            if (w0 & 2u) {                      // textured Lambert: the texel rides in the record
               M.brdf = PTX_BRDF_LAMBERT;
               texel_refl(w0 >> 8, M.refl);
            } else {
               M = S.meshes[w0 >> 2];
            }
            fr.lx = a.y; fr.ly = a.z; fr.lz = a.w;
            fr.t1x = b.x; fr.t1y = b.y; fr.t1z = b.z;
            fr.nx = b.w; fr.ny = cc.x; fr.nz = cc.y;
            fr.t2x = cc.z; fr.t2y = cc.w; fr.t2z = d.x;
            const float lin[3] = {d.y, d.z, d.w};
            float src[3] = {1.f, 1.f, 1.f};
            if (shadow.q && (J.flags & PTX_JOB_NEE) && M.brdf == PTX_BRDF_LAMBERT && S.num_lights > 0) {
               want_shadow = true;
               kd[0] = src[0] * M.refl[0]; kd[1] = src[1] * M.refl[1]; kd[2] = src[2] * M.refl[2];
               scene = J.scene;
            }
            float wout[3];
            bounce(M, fr, lin, pixel, sample, 0u, J.seed, wout, src);
            if (src_is_zero(src)) {
               write_res(res, slot, Radiance{0.f, 0.f, 0.f});
            } else {
               want = true;
               const uint32_t x = pixel % J.width, y = pixel / J.width;
               e[0] = __uint_as_float(slot); e[1] = __uint_as_float(c.job);
               e[2] = __uint_as_float(pixel); e[3] = __uint_as_float(sample);
               e[4] = fr.lx; e[5] = fr.ly; e[6] = fr.lz;
               e[7] = wout[0]; e[8] = wout[1]; e[9] = wout[2];
               e[10] = src[0]; e[11] = src[1]; e[12] = src[2];
               e[13] = __uint_as_float((x == 128 && y == 128) ? 1u : 0u);
               e[14] = __uint_as_float(J.scene); e[15] = 0.f;
            }
         }
      }
      enqueue(out, want, e);
      if (shadow.q) enqueue_shadow(shadow, want_shadow, fr, slot, scene, kd);
   }
}

// Extend stage: closest hit of every queued path (persistent traversal).  Path
// hit record (8 words): leaf-ordered triangle (kNone = miss), t, u, v, w, face
// tests since the accepting one.
struct ExtendPolicy {
   struct Ctx {
      uint32_t path, dbg, scene;
      const Node64 *nodes;
      const float *tris;
   };
   const float *queue;
   const DevScene *scenes;
   float *hits;
   __device__ __forceinline__ void load(uint32_t id, Fetched &f) const {
      const float4 *src = reinterpret_cast<const float4 *>(queue + (size_t)kQWords * id);
      const float4 a = src[0], b = src[1], c = src[2], d = src[3];
      f.q[0] = a.x; f.q[1] = a.y; f.q[2] = a.z; f.q[3] = a.w;
      f.q[4] = b.x; f.q[5] = b.y; f.q[6] = b.z; f.q[7] = b.w;
      f.q[8] = c.x; f.q[9] = c.y; f.q[10] = c.z; f.q[11] = c.w;
      f.q[12] = d.x; f.q[13] = d.y; f.q[14] = d.z; f.q[15] = d.w;
   }
   __device__ __forceinline__ int start(uint32_t, const Fetched &f, Ctx &ctx, Trav &s) const {
      make_ray(s.r, f.q[4], f.q[5], f.q[6], f.q[7], f.q[8], f.q[9]);
      ctx.path = __float_as_uint(f.q[0]);
      ctx.dbg = __float_as_uint(f.q[13]) & 1u;
      ctx.scene = __float_as_uint(f.q[14]);
      return 1;
   }
   __device__ __forceinline__ void load2(Ctx &ctx, Fetched &f) const {
      const float4 *src = reinterpret_cast<const float4 *>(scenes + ctx.scene);
      const float4 a = src[0], b = src[1], c = src[2];
      f.q[0] = a.x; f.q[1] = a.y; f.q[2] = a.z; f.q[3] = a.w;
      f.q[4] = b.x; f.q[5] = b.y; f.q[6] = b.z; f.q[7] = b.w;
      f.q[8] = c.x; f.q[9] = c.y; f.q[10] = c.z; f.q[11] = c.w;
   }
   __device__ __forceinline__ void start2(Ctx &ctx, const Fetched &f, Trav &s) const {
      const uint64_t np = (uint64_t)__float_as_uint(f.q[0]) | ((uint64_t)__float_as_uint(f.q[1]) << 32);
      const uint64_t tp = (uint64_t)__float_as_uint(f.q[2]) | ((uint64_t)__float_as_uint(f.q[3]) << 32);
      ctx.nodes = reinterpret_cast<const Node64 *>(np);
      ctx.tris = reinterpret_cast<const float *>(tp);
      begin(s, &f.q[4], __float_as_uint(f.q[10]), __float_as_uint(f.q[11]));
   }
   __device__ __forceinline__ void begin_root(Ctx &, Trav &) const {}
   __device__ __forceinline__ const Node64 *nodes(const Ctx &ctx) const { return ctx.nodes; }
   __device__ __forceinline__ const float *tris(const Ctx &ctx) const { return ctx.tris; }
   __device__ __forceinline__ bool debug_pixel(const Ctx &ctx) const { return ctx.dbg != 0; }
   __device__ __forceinline__ void finish(uint32_t, const Ctx &ctx, const Trav &s) const {
      float4 *dst = reinterpret_cast<float4 *>(hits + 8 * (size_t)ctx.path);
      dst[0] = make_float4(__uint_as_float(s.hit), s.min_t, s.hu, s.hv);
      dst[1] = make_float4(s.hw, __uint_as_float(s.after), 0.f, 0.f);
   }
};

template <int kLds>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8)))
k_extend(const float *queue, const uint32_t *count, const DevScene *scenes, float *hits, uint32_t *work, uint4 *spill,
         uint32_t max_stack, uint32_t leaf_min) {
   const ExtendPolicy P{queue, scenes, hits};
   const uint32_t n = __builtin_amdgcn_readfirstlane(*count);
   persistent_trace<kLds>(P, n, work, spill, 1u, leaf_min, max_stack);
}

// Shade stage at bounce b >= 1 (PathTraceIntegrator.cpp:23-95).
__global__ void __launch_bounds__(256) k_shade(const float *queue_in, const uint32_t *count_in, const float *hits,
                                               const JobDev *jobs, const DevScene *scenes, float *res,
                                               QueueRefs out, const uint32_t *bounce_ptr, ShadowRefs shadow) {
   const uint32_t n = *count_in;
   const uint32_t b = *bounce_ptr;
   for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
      const uint32_t i = i0 + threadIdx.x;
      bool want = false, want_shadow = false;
      float e[kQWords];
      Frame fr;
      float kd[3] = {0.f, 0.f, 0.f};
      uint32_t path = 0, scene = 0;
      if (i < n) {
         const float4 *qe = reinterpret_cast<const float4 *>(queue_in + (size_t)kQWords * i);
         const float4 a = qe[0], bq = qe[1], c = qe[2], d = qe[3];
         path = __float_as_uint(a.x);
         const uint32_t job = __float_as_uint(a.y);
         const uint32_t pixel = __float_as_uint(a.z), sample = __float_as_uint(a.w);
         const float dir[3] = {bq.w, c.x, c.y};
         float src[3] = {c.z, c.w, d.x};
         scene = __float_as_uint(d.z);
         const float4 *hp = reinterpret_cast<const float4 *>(hits + 8 * (size_t)path);
         const float4 h0 = hp[0], h1 = hp[1];
         const uint32_t tri = __float_as_uint(h0.x);
         const DevScene &S = scenes[scene];
         const JobDev &J = jobs[job];
         if (tri == kNone) {
            write_res(res, path, miss_radiance(S, src));
         } else {
            const uint32_t mesh = S.pairs[2 * (size_t)tri];
            const MeshRec &Mg = S.meshes[mesh];
            if (Mg.is_light) {
               write_res(res, path, Radiance{(Mg.emission[0] + 0.f) * src[0], (Mg.emission[1] + 0.f) * src[1],
                                             (Mg.emission[2] + 0.f) * src[2]});
            } else if (b >= J.max_depth) {
               write_res(res, path, Radiance{0.f, 0.f, 0.f});
            } else {
               MeshRec M = Mg;
               if (M.flags & kRecTextured) texel_refl(tri_texel(S, tri, h0.z, h0.w, h1.x), M.refl);
               fr = frame_of(S, tri, dir, h0.z, h0.w, h1.x, __float_as_uint(h1.y), mesh);
               if (shadow.q && (J.flags & PTX_JOB_NEE) && M.brdf == PTX_BRDF_LAMBERT && S.num_lights > 0) {
                  want_shadow = true;
                  kd[0] = src[0] * M.refl[0]; kd[1] = src[1] * M.refl[1]; kd[2] = src[2] * M.refl[2];
               }
               float lin[3], wout[3];
               world_to_local(fr, dir, lin);
               bounce(M, fr, lin, pixel, sample, b, J.seed, wout, src);
               if (src_is_zero(src)) {
                  write_res(res, path, Radiance{0.f, 0.f, 0.f});
               } else {
                  want = true;
                  e[0] = a.x; e[1] = a.y; e[2] = a.z; e[3] = a.w;
                  e[4] = fr.lx; e[5] = fr.ly; e[6] = fr.lz;
                  e[7] = wout[0]; e[8] = wout[1]; e[9] = wout[2];
                  e[10] = src[0]; e[11] = src[1]; e[12] = src[2];
                  e[13] = d.y; e[14] = d.z; e[15] = 0.f;
               }
            }
         }
      }
      enqueue(out, want, e);
      if (shadow.q) enqueue_shadow(shadow, want_shadow, fr, path, scene, kd);
   }
}

// Shadow-connect stage of the wavefront form: for each queued vertex, the
// point lights in order (light_setup, a bounded shadow traversal), adding
// (throughput x Kd) x (I x cos / (pi d^2)) of each unoccluded light to the
// path's next-event sum.  One vertex per thread; a path has at most one
// vertex per bounce, so its sum takes the terms in the order the oracle does.
template <int kLds>
__global__ void __launch_bounds__(256) k_connect(const float4 *queue, const uint32_t *count, const DevScene *scenes,
                                                 float *direct, uint4 *spill, uint32_t max_stack) {
   __shared__ uint32_t lds_ref[kLds * 256];
   __shared__ float lds_near[kLds * 256];
   __shared__ uint16_t lds_cnt[kLds * 256];
   const Stack<kLds> st{lds_ref, lds_near, lds_cnt, reinterpret_cast<u32x4 *>(spill), threadIdx.x,
                        blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, max_stack};
   const uint32_t n = *count;
   for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
      const float4 a = queue[3 * (size_t)i], b = queue[3 * (size_t)i + 1], c = queue[3 * (size_t)i + 2];
      const uint32_t slot = __float_as_uint(a.w);
      const DevScene &S = scenes[__float_as_uint(b.w)];
      float *dsum = direct + 3 * (size_t)slot;
      float dr = dsum[0], dg = dsum[1], db = dsum[2];
      for (uint32_t j = 0; j < S.num_lights; j++) {
         const float *light = S.lights + 6 * (size_t)j;
         float l[3], dist, g;
         if (!light_setup(light, a.x, a.y, a.z, b.x, b.y, b.z, l, dist, g)) continue;
         if (occluded<kLds>(S, st, a.x, a.y, a.z, l, dist)) continue;
         dr = dr + c.x * (light[3] * g);
         dg = dg + c.y * (light[4] * g);
         db = db + c.z * (light[5] * g);
      }
      dsum[0] = dr; dsum[1] = dg; dsum[2] = db;
   }
}

// Adds each chunk's samples to its job's film, per pixel in sample order, each
// sample clamped to 255 per channel (BoxFilter::AddSample/Output).
__global__ void __launch_bounds__(256) k_accumulate(const Chunk *chunks, uint32_t nchunks, const JobDev *jobs,
                                                    const float *res, const float *direct) {
   for (uint32_t ci = 0; ci < nchunks; ci++) {
      const Chunk c = chunks[ci];
      const JobDev &J = jobs[c.job];
      // wavefront form: a next-event job's sample is its terminal radiance plus its next-event sum
      const bool add = direct && (J.flags & PTX_JOB_NEE);
      for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < c.npix; p += gridDim.x * blockDim.x) {
         float *film = J.film + 3 * (size_t)p;
         float ar = film[0], ag = film[1], ab = film[2];
         for (uint32_t s = 0; s < c.k; s++) {
            const size_t slot = (size_t)c.base + (size_t)s * c.npix + p;
            const float *r = res + 3 * slot;
            float sr = r[0], sg = r[1], sb = r[2];
            if (add) {
               sr = sr + direct[3 * slot];
               sg = sg + direct[3 * slot + 1];
               sb = sb + direct[3 * slot + 2];
            }
            if (sr > 255) sr = 255;
            if (sg > 255) sg = 255;
            if (sb > 255) sb = 255;
            ar += sr; ag += sg; ab += sb;
         }
         film[0] = ar; film[1] = ag; film[2] = ab;
      }
   }
}

__global__ void k_finalize(float *film, uint32_t npix, uint32_t spp) {
   const float div = 1.0f / (float)spp;
   for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
      film[3 * (size_t)p] *= div;
      film[3 * (size_t)p + 1] *= div;
      film[3 * (size_t)p + 2] *= div;
   }
}

// Start of a bounce: clears the traversal work counters and the next queue's
// count, and advances the device-resident bounce index.
__global__ void k_bounce_prologue(uint32_t *work, uint32_t *count_out, uint32_t *bounce, uint32_t *count_shadow) {
   const uint32_t i = threadIdx.x;
   if (i < kWorkWords) work[i] = 0u;
   if (i == 8) *count_out = 0u;
   if (i == 9) *bounce += 1u;
   if (i == 10 && count_shadow) *count_shadow = 0u;
}

__global__ void k_zero_words(uint32_t *a, uint32_t na, uint32_t *b, uint32_t nb) {
   const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
   if (i < na) a[i] = 0u;
   if (b && i < nb) b[i] = 0u;
}

__global__ void k_zero_floats(float *a, size_t n) {
   for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
      a[i] = 0.f;
}

// ---------------------------------------------------------------- plain path

// One sample of every pixel: one thread per pixel path, all bounces inline.
// kTex: the scene has textured meshes.  kCached: the path starts from the
// pixel's primary record (k_primary: CenterSampler gives every sample of a
// pixel the same primary ray) and traces from bounce 1; otherwise it traces
// its primary ray itself, as the reference's CUDA path does for every sample.
template <int kLds, bool kNee, bool kTex, bool kCached>
__global__ void __launch_bounds__(256) k_mega(DevScene S, JobDev J, uint32_t sample, uint4 *spill) {
   __shared__ uint32_t lds_ref[kLds * 256];
   __shared__ float lds_near[kLds * 256];
   __shared__ uint16_t lds_cnt[kLds * 256];
   const Stack<kLds> st{lds_ref, lds_near, lds_cnt, reinterpret_cast<u32x4 *>(spill), threadIdx.x,
                        blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, S.max_stack};
   const uint32_t npix = J.width * J.height;
   for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
      const uint32_t x = p % J.width, y = p / J.width;
      const bool dbg = x == 128 && y == 128;
      float o[3], d[3];
      float src[3] = {1.f, 1.f, 1.f};
      Radiance L = {0.f, 0.f, 0.f};
      const bool nee = kNee && (J.flags & PTX_JOB_NEE) != 0;
      float direct[3] = {0.f, 0.f, 0.f};
      uint32_t b0 = 0;
      if constexpr (kCached) {
         // bounce 0 from the primary record: terminal radiance, or the hit frame
         const float4 *pr = reinterpret_cast<const float4 *>(J.primary + 16 * (size_t)p);
         const float4 a = pr[0], bq = pr[1], cq = pr[2], dq = pr[3];
         const uint32_t w0 = __float_as_uint(a.x);
         if (w0 & 1u) {
            L = Radiance{a.y, a.z, a.w};
            b0 = kNone;
         } else {
            MeshRec M;
            if (w0 & 2u) {                      // textured Lambert: the texel rides in the record
               M.brdf = PTX_BRDF_LAMBERT;
               texel_refl(w0 >> 8, M.refl);
            } else {
               M = S.meshes[w0 >> 2];
            }
            Frame fr;
            fr.lx = a.y; fr.ly = a.z; fr.lz = a.w;
            fr.t1x = bq.x; fr.t1y = bq.y; fr.t1z = bq.z;
            fr.nx = bq.w; fr.ny = cq.x; fr.nz = cq.y;
            fr.t2x = cq.z; fr.t2y = cq.w; fr.t2z = dq.x;
            const float lin[3] = {dq.y, dq.z, dq.w};
            if (nee && M.brdf == PTX_BRDF_LAMBERT) {
               for (uint32_t j = 0; j < S.num_lights; j++) {
                  const float *light = S.lights + 6 * (size_t)j;
                  float l[3], dist, g;
                  if (!light_setup(light, fr.lx, fr.ly, fr.lz, fr.nx, fr.ny, fr.nz, l, dist, g)) continue;
                  if (occluded<kLds>(S, st, fr.lx, fr.ly, fr.lz, l, dist)) continue;
                  direct[0] = direct[0] + (src[0] * M.refl[0]) * (light[3] * g);
                  direct[1] = direct[1] + (src[1] * M.refl[1]) * (light[4] * g);
                  direct[2] = direct[2] + (src[2] * M.refl[2]) * (light[5] * g);
               }
            }
            float wout[3];
            bounce(M, fr, lin, p, sample, 0u, J.seed, wout, src);
            o[0] = fr.lx; o[1] = fr.ly; o[2] = fr.lz;
            d[0] = wout[0]; d[1] = wout[1]; d[2] = wout[2];
            b0 = 1;
         }
      } else {
         camera_ray(J, x, y, o, d);
      }
      for (uint32_t b = b0; b != kNone; b++) {
         if (src_is_zero(src)) break;
         Trav s;
         make_ray(s.r, o[0], o[1], o[2], d[0], d[1], d[2]);
         begin(s, S.root_box, S.root_ref, S.root_count);
         while (s.cur != kNone) {
            Fetched f;
            fetch(s, S.nodes, S.tris, f);
            if (!advance<kLds>(s, f, st, dbg)) break;
         }
         if (s.hit == kNone) {
            L = miss_radiance(S, src);
            break;
         }
         const uint32_t mesh = S.pairs[2 * (size_t)s.hit];
         const MeshRec &Mg = S.meshes[mesh];
         if (Mg.is_light) {
            L = Radiance{(Mg.emission[0] + 0.f) * src[0], (Mg.emission[1] + 0.f) * src[1],
                         (Mg.emission[2] + 0.f) * src[2]};
            break;
         }
         if (b >= J.max_depth) break;
         MeshRec M = Mg;
         if (kTex && (M.flags & kRecTextured)) texel_refl(tri_texel(S, s.hit, s.hu, s.hv, s.hw), M.refl);
         const Frame fr = frame_of(S, s.hit, d, s.hu, s.hv, s.hw, s.after, mesh);
         if (nee && M.brdf == PTX_BRDF_LAMBERT) {
            for (uint32_t j = 0; j < S.num_lights; j++) {
               const float *light = S.lights + 6 * (size_t)j;
               float l[3], dist, g;
               if (!light_setup(light, fr.lx, fr.ly, fr.lz, fr.nx, fr.ny, fr.nz, l, dist, g)) continue;
               if (occluded<kLds>(S, st, fr.lx, fr.ly, fr.lz, l, dist)) continue;
               direct[0] = direct[0] + (src[0] * M.refl[0]) * (light[3] * g);
               direct[1] = direct[1] + (src[1] * M.refl[1]) * (light[4] * g);
               direct[2] = direct[2] + (src[2] * M.refl[2]) * (light[5] * g);
            }
         }
         float lin[3], wout[3];
         world_to_local(fr, d, lin);
         bounce(M, fr, lin, p, sample, b, J.seed, wout, src);
         o[0] = fr.lx; o[1] = fr.ly; o[2] = fr.lz;
         d[0] = wout[0]; d[1] = wout[1]; d[2] = wout[2];
      }
      if (nee) {
         L.r = L.r + direct[0];
         L.g = L.g + direct[1];
         L.b = L.b + direct[2];
      }
      float *film = J.film + 3 * (size_t)p;
      float sr = L.r, sg = L.g, sb = L.b;
      if (sr > 255) sr = 255;
      if (sg > 255) sg = 255;
      if (sb > 255) sb = 255;
      film[0] += sr;
      film[1] += sg;
      film[2] += sb;
   }
}


// ---------------------------------------------------------------- path regeneration

// One persistent launch per round: each lane carries a whole path — the cached
// primary hit, then trace and shade for every bounce — and takes the next path
// slot of the round as soon as its path ends, so no lane waits for the longest
// path of its wave and small jobs share the GPU with large ones.  Each loop
// step issues one mode's loads for every lane before any compute (the traversal
// step's node or faces, the primary record, the hit's mesh index, its mesh
// record, its faces), as in the closest-hit loop.  The round's chunk table and
// scene table sit in LDS.  The shading steps (primary bounce, hit frame and
// BRDF sample) cost several traversal steps and run for the whole wave when any
// lane takes them, so a wave shades only when at least shade_min of its lanes
// wait for it, or when none is tracing; until then those lanes hold.
enum : uint32_t {
   kPIdle = 0, kPPrim = 1, kPTrace = 2, kPShade1 = 3, kPShade2 = 4, kPShade3 = 5,
   kPLight = 6,    // next-event estimation: add the last light's term, set up the next shadow ray
   kPShadow = 7,   // next-event estimation: shadow-ray traversal
   kPHold = 8      // this iteration only: a lane that waits for its wave to gather
};
// Lane word: bits 0-3 state, bit 4 the reference's debug pixel, bits 5-10 the
// path's chunk in the round, bits 11-30 its bounce index, bit 31 set when the
// path's job uses next-event estimation.
constexpr uint32_t kPModeMask = 15u, kPDbg = 16u, kPChunkShift = 5, kPBounceShift = 11, kPNee = 1u << 31;
constexpr uint32_t kMaxRoundScenes = 64;
static_assert(kMaxRoundChunks <= 64, "chunk index has 6 bits in the lane word");

__device__ __forceinline__ uint32_t lw_chunk(uint32_t w) { return (w >> kPChunkShift) & 63u; }
__device__ __forceinline__ uint32_t lw_bounce(uint32_t w) { return (w & ~kPNee) >> kPBounceShift; }
__device__ __forceinline__ uint32_t lw_with_state(uint32_t w, uint32_t state) { return (w & ~kPModeMask) | state; }

struct LdsChunk {
   uint32_t base, npix, s0, width;
   uint32_t seed, max_depth, scene, flags;
   const uint32_t *primary;
   DivU32 by_npix;            // local slot -> (sample, pixel) without a run-time division
};

// (sample offset, pixel) of a chunk's local slot: local = sample offset * npix + pixel.
__device__ __forceinline__ void chunk_split(const LdsChunk &c, uint32_t local, uint32_t &sample_off,
                                            uint32_t &pixel) {
   sample_off = div_u32(local, c.by_npix);
   pixel = local - sample_off * c.npix;
}

__device__ __forceinline__ uint32_t lds_find_chunk(const LdsChunk *c, uint32_t n, uint32_t slot) {
   uint32_t lo = 0, hi = n;
   while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (c[mid].base <= slot) lo = mid; else hi = mid;
   }
   return lo;
}

// Terminal radiance of a path plus its next-event sum (paths of next-event jobs).
template <bool kNee>
__device__ __forceinline__ Radiance fin(Radiance R, const float direct[3], uint32_t lane_word) {
   if (kNee && (lane_word & kPNee)) {
      R.r = R.r + direct[0];
      R.g = R.g + direct[1];
      R.b = R.b + direct[2];
   }
   return R;
}

// Starts the traversal of ray s.r in scene S; returns the next state, or kPIdle
// after writing the miss result when the root box is missed.
template <bool kNee>
__device__ __forceinline__ uint32_t p_root(const DevScene &S, Trav &s, float *res, uint32_t slot,
                                           const float src[3], const float direct[3], uint32_t lane_word) {
   begin(s, S.root_box, S.root_ref, S.root_count);
   if (s.cur != kNone) return kPTrace;
   write_res(res, slot, fin<kNee>(miss_radiance(S, src), direct, lane_word));
   return kPIdle;
}

// Per-lane record of a next-event vertex (3 x 16 bytes in nee_scratch): the
// vertex P, the next ray's direction, throughput x Kd, the shading normal.
struct NeeRecord {
   float4 a, b, c;   // a = P.xyz, wout.x; b = wout.yz, k.xy; c = k.z, n.xyz
};

// kGlossy: some scene of the round has glossy meshes; kTex: some has textured ones.
template <int kLds, bool kNee, bool kGlossy, bool kTex>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_paths(const Chunk *chunks, uint32_t nchunks, uint32_t nslots, const JobDev *jobs, const DevScene *scenes,
        uint32_t nscenes, float *res, uint32_t *work, uint4 *spill, uint32_t max_stack, uint32_t shade_min,
        uint32_t leaf_min, float4 *nee_scratch) {
   __shared__ uint32_t lds_ref[kLds * 256];
   __shared__ float lds_near[kLds * 256];
   __shared__ uint16_t lds_cnt[kLds * 256];
   __shared__ LdsChunk sch[kMaxRoundChunks];
   __shared__ DevScene ssc[kMaxRoundScenes];
   for (uint32_t i = threadIdx.x; i < nchunks; i += blockDim.x) {
      const Chunk c = chunks[i];
      const JobDev &J = jobs[c.job];
      LdsChunk l;
      l.base = c.base; l.npix = c.npix; l.s0 = c.s0; l.width = J.width;
      l.seed = J.seed; l.max_depth = J.max_depth; l.scene = J.scene; l.flags = J.flags;
      l.primary = J.primary; l.by_npix = div_u32_setup(c.npix);
      sch[i] = l;
   }
   for (uint32_t i = threadIdx.x; i < nscenes; i += blockDim.x) ssc[i] = scenes[i];
   __syncthreads();
   const uint32_t lane = __lane_id();
   const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
   const Stack<kLds> st{lds_ref, lds_near, lds_cnt, reinterpret_cast<u32x4 *>(spill), threadIdx.x,
                        gtid, gridDim.x * blockDim.x, max_stack};
   RaySource srcq;
   srcq.init(nslots, work, lane);
   Trav s;
   s.hit = kNone;
   s.after = 0;
   s.cur = kNone;
   s.cnt = 0;
   s.sp = 0;
   uint32_t mode = kPIdle;       // lane word
   uint32_t slot = 0;
   uint32_t t0 = 0;              // Shade1 -> Shade2: mesh; Shade2 -> Shade3: brdf; Light/Shadow: next light
   float t1 = 0.f, t2 = 0.f, t3 = 0.f;   // Shade2 -> Shade3: reflectance; Light/Shadow: visible, last vertex
   // Shade1 -> Shade2 of a textured hit keeps the texel's address and the texture
   // record's flags in ray words a found hit no longer needs (origin, 1 / d)
   float src[3] = {1.f, 1.f, 1.f};
   float direct[3] = {0.f, 0.f, 0.f};    // next-event sum of the path (kNee)
   float pend[3] = {0.f, 0.f, 0.f};      // the light term a running shadow ray decides (kNee)
   const Node64 *nodes = nullptr;
   const float *tris = nullptr;
   // At a Lambert vertex of a next-event job: keep the vertex in this lane's
   // record and go through the lights before tracing the bounce ray.
   auto nee_vertex = [&](const Frame &fr, const float refl[3], const float wout[3], uint32_t next_mode) {
      const float k0 = src[0] * refl[0], k1 = src[1] * refl[1], k2 = src[2] * refl[2];
      (void)k0;
      NeeRecord r;
      r.a = make_float4(fr.lx, fr.ly, fr.lz, wout[0]);
      r.b = make_float4(wout[1], wout[2], k0, k1);
      r.c = make_float4(k2, fr.nx, fr.ny, fr.nz);
      float4 *dst = nee_scratch + 3 * (size_t)gtid;
      dst[0] = r.a; dst[1] = r.b; dst[2] = r.c;
      t0 = 0u;
      t1 = 0.f;
      mode = lw_with_state(next_mode, kPLight);
   };
   while (true) {
      uint64_t idle = __ballot((mode & kPModeMask) == kPIdle);
      if (idle && !srcq.done) {
         while (idle) {
            srcq.refill(work, lane);
            if (srcq.done) break;
            const uint32_t take = min((uint32_t)__popcll(idle), srcq.pool_left);
            const uint32_t rank = __popcll(idle & ((1ull << lane) - 1ull));
            const bool mine = ((idle >> lane) & 1ull) && rank < take;
            if (mine) {
               slot = srcq.pool_base + rank;
               mode = kPPrim;
            }
            idle &= ~__ballot(mine);
            srcq.pool_base += take;
            srcq.pool_left -= take;
         }
      }
      const uint32_t m0 = mode & kPModeMask;
      if (!__any(m0 != kPIdle)) break;
      const bool heavy = m0 == kPPrim || m0 == kPShade2 || m0 == kPShade3;
      const bool tracing = m0 == kPTrace || m0 == kPShadow;
      // single-load steps that never wait for the wave
      const bool light_step = m0 == kPShade1 || m0 == kPLight;
      const bool run_heavy = (uint32_t)__popcll(__ballot(heavy)) >= shade_min || !__any(tracing || light_step);
      // leaf steps gather the same way; lanes at interior nodes, in Shade1 or in
      // Light always step, so some lane progresses whenever any lane holds
      const bool at_leaf = tracing && (s.cur & kLeafBit) != 0;
      const bool run_leaf = (uint32_t)__popcll(__ballot(at_leaf)) >= leaf_min ||
                            !__any((tracing && !at_leaf) || light_step);
      const uint32_t m = (heavy && !run_heavy) || (at_leaf && !run_leaf) ? kPHold : m0;
      // ------------------------------------------------------------ fetch
      Fetched f;
      if (m == kPTrace || m == kPShadow) {
         fetch(s, nodes, tris, f);
      } else if (m == kPPrim) {
         const uint32_t cf = lds_find_chunk(sch, nchunks, slot);
         mode = kPPrim | (cf << kPChunkShift);
         const LdsChunk &c = sch[cf];
         uint32_t sample_off, pixel;
         chunk_split(c, slot - c.base, sample_off, pixel);
         const PTX_GLOBAL f32x4u *pr = as_global(reinterpret_cast<const f32x4u *>(c.primary + 16 * (size_t)pixel));
         const f32x4u q0 = pr[0], q1 = pr[1], q2 = pr[2], q3 = pr[3];
         f.q[0] = q0.x; f.q[1] = q0.y; f.q[2] = q0.z; f.q[3] = q0.w;
         f.q[4] = q1.x; f.q[5] = q1.y; f.q[6] = q1.z; f.q[7] = q1.w;
         f.q[8] = q2.x; f.q[9] = q2.y; f.q[10] = q2.z; f.q[11] = q2.w;
         f.q[12] = q3.x; f.q[13] = q3.y; f.q[14] = q3.z; f.q[15] = q3.w;
      } else if (m == kPShade1) {
         const DevScene &S = ssc[sch[lw_chunk(mode)].scene];
         f.q[0] = __uint_as_float(as_global(S.pairs)[2 * (size_t)s.hit]);
         if (kTex) {                                // the triangle's texture record (flags 0 without one)
            f.q[7] = 0.f;
            if (S.tex_tris) {
               const PTX_GLOBAL f32x4u *tr = as_global(reinterpret_cast<const f32x4u *>(S.tex_tris + s.hit));
               const f32x4u a = tr[0], b = tr[1];
               f.q[1] = a.x; f.q[2] = a.y; f.q[3] = a.z; f.q[4] = a.w;
               f.q[5] = b.x; f.q[6] = b.y; f.q[7] = b.z; f.q[8] = b.w;
            }
         }
      } else if (m == kPShade2) {
         if (kTex && (__float_as_uint(s.r.oy) & kTexTextured)) {      // the texel instead of the mesh record
            f.q[0] = __uint_as_float(*as_global(reinterpret_cast<const uint32_t *>(
                  (uint64_t)__float_as_uint(s.r.ix) | ((uint64_t)__float_as_uint(s.r.iy) << 32))));
         } else {
            const PTX_GLOBAL uint32_t *mr =
                  as_global(reinterpret_cast<const uint32_t *>(ssc[sch[lw_chunk(mode)].scene].meshes + t0));
#pragma unroll
            for (int k = 0; k < 9; k++) f.q[k] = __uint_as_float(mr[k]);
         }
         const PTX_GLOBAL float *v = as_global(tris + 9 * (size_t)s.hit);
#pragma unroll
         for (int k = 0; k < 9; k++) f.q[9 + k] = v[k];
      } else if (m == kPShade3) {
         const PTX_GLOBAL float *v = as_global(tris + 9 * (size_t)s.hit);
         const PTX_GLOBAL float *nrm = as_global(ssc[sch[lw_chunk(mode)].scene].normals + 9 * (size_t)s.hit);
#pragma unroll
         for (int k = 0; k < 9; k++) { f.q[k] = v[k]; f.q[9 + k] = nrm[k]; }
      } else if (kNee && m == kPLight) {
         const PTX_GLOBAL f32x4u *rec = as_global(reinterpret_cast<const f32x4u *>(nee_scratch + 3 * (size_t)gtid));
         const f32x4u a = rec[0], b = rec[1], c = rec[2];
         f.q[0] = a.x; f.q[1] = a.y; f.q[2] = a.z; f.q[3] = a.w;
         f.q[4] = b.x; f.q[5] = b.y; f.q[6] = b.z; f.q[7] = b.w;
         f.q[8] = c.x; f.q[9] = c.y; f.q[10] = c.z; f.q[11] = c.w;
         const DevScene &S = ssc[sch[lw_chunk(mode)].scene];
         if (t0 < S.num_lights) {
            const PTX_GLOBAL float *lp = as_global(S.lights + 6 * (size_t)t0);
            const f32x4u la = *reinterpret_cast<const PTX_GLOBAL f32x4u *>(lp);
            const f32x2u lb = *reinterpret_cast<const PTX_GLOBAL f32x2u *>(lp + 4);
            f.q[12] = la.x; f.q[13] = la.y; f.q[14] = la.z; f.q[15] = la.w;
            f.q[16] = lb.x; f.q[17] = lb.y;
         }
      }
      // ------------------------------------------------------------ compute
      if (m == kPTrace) {
         if (!advance<kLds>(s, f, st, (mode & kPDbg) != 0)) {
            if (s.hit == kNone) {
               write_res(res, slot, fin<kNee>(miss_radiance(ssc[sch[lw_chunk(mode)].scene], src), direct, mode));
               mode = kPIdle;
            } else {
               mode = lw_with_state(mode, kPShade1);
            }
         }
      } else if (kNee && m == kPShadow) {
         const bool more = advance<kLds>(s, f, st, false);
         if (s.hit != kNone) {                    // occluded
            t1 = 0.f;
            mode = lw_with_state(mode, kPLight);
         } else if (!more) {                      // nothing before the light
            t1 = 1.f;
            mode = lw_with_state(mode, kPLight);
         }
      } else if (kNee && m == kPLight) {
         const DevScene &S = ssc[sch[lw_chunk(mode)].scene];
         if (t1 != 0.f) {
            direct[0] = direct[0] + pend[0];
            direct[1] = direct[1] + pend[1];
            direct[2] = direct[2] + pend[2];
         }
         t1 = 0.f;
         if (t0 < S.num_lights) {
            float l[3], dist, g;
            const uint32_t j = t0;
            t0 = j + 1;
            if (light_setup(&f.q[12], f.q[0], f.q[1], f.q[2], f.q[9], f.q[10], f.q[11], l, dist, g)) {
               pend[0] = f.q[6] * (f.q[15] * g);
               pend[1] = f.q[7] * (f.q[16] * g);
               pend[2] = f.q[8] * (f.q[17] * g);
               make_ray(s.r, f.q[0], f.q[1], f.q[2], l[0], l[1], l[2]);
               begin(s, S.root_box, S.root_ref, S.root_count, dist);
               if (s.cur == kNone) t1 = 1.f;       // root box missed: nothing in the way
               else mode = lw_with_state(mode, kPShadow);
            }
         } else if (t2 != 0.f) {                  // the path ends at this vertex
            write_res(res, slot, fin<kNee>(Radiance{0.f, 0.f, 0.f}, direct, mode));
            mode = kPIdle;
         } else {                                 // trace the bounce ray
            make_ray(s.r, f.q[0], f.q[1], f.q[2], f.q[3], f.q[4], f.q[5]);
            const uint32_t next = p_root<kNee>(S, s, res, slot, src, direct, mode);
            mode = next == kPIdle ? kPIdle : lw_with_state(mode, next);
         }
      } else {
         // A path's start (from the pixel's primary record) and a hit's
         // shading (Shade2 / Shade3) end in the same bounce: the branches
         // below set up the vertex, and one copy of the bounce, the
         // next-event set-up and the ray start then serves every lane that
         // shades in this step, whichever branch it came through.  The
         // vertex travels in registers the lane holds anyway: the frame and
         // the incoming direction in f.q[1..15] (the primary record's
         // layout), the material in t0 (BRDF), t1..t3 (reflectance) and s.cnt
         // (glossy roughness).
         bool shade = false;
         uint32_t pixel = 0, sample = 0, b = 0, seed = 0;
         uint32_t nb = 0;                         // the lane word after the bounce, state bits clear
         uint32_t scene = 0;                      // index into the round's scene table
         if (m == kPPrim) {
            const LdsChunk &c = sch[lw_chunk(mode)];
            uint32_t sample_off;
            chunk_split(c, slot - c.base, sample_off, pixel);
            sample = c.s0 + sample_off;
            const uint32_t w0 = __float_as_uint(f.q[0]);
            if (kNee) {
               direct[0] = direct[1] = direct[2] = 0.f;
               if (c.flags & PTX_JOB_NEE) mode |= kPNee;
            }
            if (w0 & 1u) {
               write_res(res, slot, fin<kNee>(Radiance{f.q[1], f.q[2], f.q[3]}, direct, mode));
               mode = kPIdle;
            } else {
               if (kTex && (w0 & 2u)) {            // textured Lambert: the texel rides in the record
                  float refl[3];
                  texel_refl(w0 >> 8, refl);
                  t0 = (uint32_t)PTX_BRDF_LAMBERT;
                  t1 = refl[0]; t2 = refl[1]; t3 = refl[2];
               } else {
                  const PTX_GLOBAL uint32_t *mw =
                        as_global(reinterpret_cast<const uint32_t *>(ssc[c.scene].meshes + (w0 >> 2)));
                  t0 = mw[0];
                  t1 = __uint_as_float(mw[1]);
                  t2 = __uint_as_float(mw[2]);
                  t3 = __uint_as_float(mw[3]);
                  if (kGlossy) s.cnt = mw[5];      // glossy roughness
               }
               src[0] = src[1] = src[2] = 1.f;
               // pixel (128, 128) (mesh.cpp:192-194): x == 128 needs width > 128, and
               // then y * width + x == 128 * width + 128 holds for that pixel only
               const uint32_t dbg =
                     c.width > 128u && (uint64_t)pixel == 128ull * c.width + 128ull ? kPDbg : 0u;
               nb = (mode & ~kPModeMask) | dbg | (1u << kPBounceShift);
               b = 0u;
               seed = c.seed;
               scene = c.scene;
               shade = true;
            }
         } else if (m == kPShade1) {
            t0 = __float_as_uint(f.q[0]);
            if (kTex) {
               const uint32_t wh = __float_as_uint(f.q[7]);
               s.r.oy = f.q[7];                    // the texture record's flags
               if (wh & kTexTextured) {            // the texel's address, for Shade2's load
                  const DevScene &S = ssc[sch[lw_chunk(mode)].scene];
                  const uint32_t k = __float_as_uint(f.q[8]) + texel_index(&f.q[1], s.hu, s.hv, s.hw, wh);
                  const uint64_t a = reinterpret_cast<uint64_t>(texel_pool(S) + k);
                  s.r.ix = __uint_as_float((uint32_t)a);
                  s.r.iy = __uint_as_float((uint32_t)(a >> 32));
               }
            }
            mode = lw_with_state(mode, kPShade2);
         } else if (m == kPShade2 || m == kPShade3) {
            const LdsChunk &c = sch[lw_chunk(mode)];
            const uint32_t bi = lw_bounce(mode);
            bool go = true;
            if (kTex && m == kPShade2 && (__float_as_uint(s.r.oy) & kTexTextured)) {
               // textured Lambert (never an emitter): the texel is the reflectance
               if (bi >= c.max_depth) {
                  write_res(res, slot, fin<kNee>(Radiance{0.f, 0.f, 0.f}, direct, mode));
                  mode = kPIdle;
                  go = false;
               } else {
                  float refl[3];
                  texel_refl(__float_as_uint(f.q[0]), refl);
                  t0 = (uint32_t)PTX_BRDF_LAMBERT;
                  t1 = refl[0]; t2 = refl[1]; t3 = refl[2];
                  if ((__float_as_uint(s.r.oy) & kTexNormals) && ssc[c.scene].normals) {
                     mode = lw_with_state(mode, kPShade3);
                     go = false;
                  } else {
#pragma unroll
                     for (int k = 0; k < 9; k++) f.q[k] = f.q[9 + k];
                  }
               }
            } else if (m == kPShade2) {
               const int32_t brdf = __float_as_int(f.q[0]);
               const int32_t is_light = __float_as_int(f.q[4]);
               const uint32_t mflags = __float_as_uint(f.q[8]);
               if (kGlossy) s.cnt = __float_as_uint(f.q[5]);          // glossy roughness, kept until the bounce
               if (is_light) {
                  write_res(res, slot, fin<kNee>(Radiance{(f.q[5] + 0.f) * src[0], (f.q[6] + 0.f) * src[1],
                                                          (f.q[7] + 0.f) * src[2]}, direct, mode));
                  mode = kPIdle;
                  go = false;
               } else if (bi >= c.max_depth) {
                  write_res(res, slot, fin<kNee>(Radiance{0.f, 0.f, 0.f}, direct, mode));
                  mode = kPIdle;
                  go = false;
               } else if ((mflags & kRecNormals) && ssc[c.scene].normals) {
                  t0 = (uint32_t)brdf;
                  t1 = f.q[1]; t2 = f.q[2]; t3 = f.q[3];
                  mode = lw_with_state(mode, kPShade3);
                  go = false;
               } else {
                  t0 = (uint32_t)brdf;
                  t1 = f.q[1]; t2 = f.q[2]; t3 = f.q[3];
#pragma unroll
                  for (int k = 0; k < 9; k++) f.q[k] = f.q[9 + k];
               }
            }
            if (go) {
               Frame fr;
               hit_frame_sel(&f.q[0], &f.q[9], m == kPShade3, s.r, s.hu, s.hv, s.hw, s.after, fr);
               const float d[3] = {s.r.dx, s.r.dy, s.r.dz};
               float lin[3];
               world_to_local(fr, d, lin);
               f.q[1] = fr.lx; f.q[2] = fr.ly; f.q[3] = fr.lz;
               f.q[4] = fr.t1x; f.q[5] = fr.t1y; f.q[6] = fr.t1z;
               f.q[7] = fr.nx; f.q[8] = fr.ny; f.q[9] = fr.nz;
               f.q[10] = fr.t2x; f.q[11] = fr.t2y; f.q[12] = fr.t2z;
               f.q[13] = lin[0]; f.q[14] = lin[1]; f.q[15] = lin[2];
               uint32_t sample_off;
               chunk_split(c, slot - c.base, sample_off, pixel);
               sample = c.s0 + sample_off;
               b = bi;
               seed = c.seed;
               nb = (mode & ~kPModeMask) + (1u << kPBounceShift);
               scene = c.scene;
               shade = true;
            }
         }
         if (shade) {
            const DevScene &S = ssc[scene];
            MeshRec M;
            M.brdf = (int32_t)t0;
            M.refl[0] = t1; M.refl[1] = t2; M.refl[2] = t3;
            if (kGlossy) M.emission[0] = __uint_as_float(s.cnt);
            Frame fr;
            fr.lx = f.q[1]; fr.ly = f.q[2]; fr.lz = f.q[3];
            fr.t1x = f.q[4]; fr.t1y = f.q[5]; fr.t1z = f.q[6];
            fr.nx = f.q[7]; fr.ny = f.q[8]; fr.nz = f.q[9];
            fr.t2x = f.q[10]; fr.t2y = f.q[11]; fr.t2z = f.q[12];
            const float lin[3] = {f.q[13], f.q[14], f.q[15]};
            const bool nee = kNee && (mode & kPNee) && M.brdf == PTX_BRDF_LAMBERT && S.num_lights > 0;
            const float s_before[3] = {src[0], src[1], src[2]};
            float wout[3];
            bounce<kGlossy>(M, fr, lin, pixel, sample, b, seed, wout, src);
            nodes = S.nodes;
            tris = S.tris;
            if (nee) {
               const float keep[3] = {src[0], src[1], src[2]};
               src[0] = s_before[0]; src[1] = s_before[1]; src[2] = s_before[2];
               nee_vertex(fr, M.refl, wout, nb);
               src[0] = keep[0]; src[1] = keep[1]; src[2] = keep[2];
               t2 = src_is_zero(src) ? 1.f : 0.f;
            } else if (src_is_zero(src)) {
               write_res(res, slot, fin<kNee>(Radiance{0.f, 0.f, 0.f}, direct, mode));
               mode = kPIdle;
            } else {
               make_ray(s.r, fr.lx, fr.ly, fr.lz, wout[0], wout[1], wout[2]);
               const uint32_t next = p_root<kNee>(S, s, res, slot, src, direct, mode);
               mode = next == kPIdle ? kPIdle : nb | next;
            }
         }
      }
   }
}
}  // namespace dev

// ---------------------------------------------------------------- launchers

namespace {
constexpr uint32_t kRenderLds = 12;
}

uint32_t render_lds_depth() { return kRenderLds; }

int render_blocks_per_cu(int which) {
   int b = 0;
   hipError_t e = hipSuccess;
   if (which == 0) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, dev::k_extend<kRenderLds>, 256, 0);
   else if (which == 1) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, dev::k_primary<kRenderLds>, 256, 0);
   else if (which == 2) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, dev::k_mega<kRenderLds, false, false, false>, 256, 0);
   else if (which == 3) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, dev::k_mega<kRenderLds, true, false, false>, 256, 0);
   else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, dev::k_connect<kRenderLds>, 256, 0);
   return e == hipSuccess && b > 0 ? b : 1;
}

void launch_primary(const DevScene &S, const JobDev &J, uint32_t *work, uint4 *spill, uint32_t leaf_min, uint32_t grid,
                    hipStream_t st) {
   hipLaunchKernelGGL(dev::k_zero_words, dim3(1), dim3(64), 0, st, work, 8u, (uint32_t *)nullptr, 0u);
   hipLaunchKernelGGL(dev::k_primary<kRenderLds>, dim3(grid), dim3(256), 0, st, S, J, work, spill, leaf_min);
}

void launch_primaries(const JobDev *jobs, const DevScene *scenes, const uint32_t *offsets, uint32_t njobs,
                      uint32_t *work, uint4 *spill, uint32_t leaf_min, uint32_t max_stack, uint32_t grid,
                      hipStream_t st) {
   hipLaunchKernelGGL(dev::k_zero_words, dim3(1), dim3(64), 0, st, work, 8u, (uint32_t *)nullptr, 0u);
   hipLaunchKernelGGL(dev::k_primaries<kRenderLds>, dim3(grid), dim3(256), 0, st, jobs, scenes, offsets, njobs, work,
                      spill, leaf_min, max_stack);
}

void launch_round_start(const Chunk *chunks, uint32_t nchunks, uint32_t nslots, const JobDev *jobs,
                        const DevScene *scenes, float *res, float *q_out, uint32_t *count_out, uint32_t *bounce,
                        const ShadowArgs &sh, uint4 *spill, uint32_t max_stack, uint32_t grid, hipStream_t st) {
   // the bounce index and the shadow count are adjacent words
   hipLaunchKernelGGL(dev::k_zero_words, dim3(1), dim3(64), 0, st, count_out, 1u, bounce, sh.q ? 2u : 1u);
   hipLaunchKernelGGL(dev::k_round_start, dim3(grid), dim3(256), 0, st, chunks, nchunks, nslots, jobs, scenes, res,
                      dev::QueueRefs{q_out, count_out}, dev::ShadowRefs{sh.q, sh.count}, sh.direct);
   if (sh.q)
      hipLaunchKernelGGL(dev::k_connect<kRenderLds>, dim3(sh.grid), dim3(256), 0, st, sh.q, sh.count, scenes,
                         sh.direct, spill, max_stack);
}

// One bounce: prologue, extend, shade, and with next-event jobs the shadow
// connection.  Captured once per queue direction and replayed for every bounce.
void launch_bounce(const float *q_in, const uint32_t *count_in, float *q_out, uint32_t *count_out, float *hits,
                   const JobDev *jobs, const DevScene *scenes, float *res, uint32_t *work, uint4 *spill,
                   uint32_t max_stack, uint32_t leaf_min, uint32_t *bounce, const ShadowArgs &sh, uint32_t ext_grid,
                   uint32_t shade_grid, hipStream_t st) {
   hipLaunchKernelGGL(dev::k_bounce_prologue, dim3(1), dim3(64), 0, st, work, count_out, bounce, sh.count);
   hipLaunchKernelGGL(dev::k_extend<kRenderLds>, dim3(ext_grid), dim3(256), 0, st, q_in, count_in, scenes, hits, work,
                      spill, max_stack, leaf_min);
   hipLaunchKernelGGL(dev::k_shade, dim3(shade_grid), dim3(256), 0, st, q_in, count_in, hits, jobs, scenes, res,
                      dev::QueueRefs{q_out, count_out}, bounce, dev::ShadowRefs{sh.q, sh.count});
   if (sh.q)
      hipLaunchKernelGGL(dev::k_connect<kRenderLds>, dim3(sh.grid), dim3(256), 0, st, sh.q, sh.count, scenes,
                         sh.direct, spill, max_stack);
}

void launch_accumulate(const Chunk *chunks, uint32_t nchunks, const JobDev *jobs, const float *res,
                       const float *direct, uint32_t grid, hipStream_t st) {
   hipLaunchKernelGGL(dev::k_accumulate, dim3(grid), dim3(256), 0, st, chunks, nchunks, jobs, res, direct);
}

void launch_finalize(float *film, uint32_t npix, uint32_t spp, hipStream_t st) {
   const uint32_t grid = std::min<uint32_t>(4096, (npix + 255) / 256);
   hipLaunchKernelGGL(dev::k_finalize, dim3(std::max(1u, grid)), dim3(256), 0, st, film, npix, spp);
}

void launch_zero_floats(float *a, size_t n, hipStream_t st) {
   const uint32_t grid = (uint32_t)std::min<size_t>(4096, (n + 255) / 256);
   hipLaunchKernelGGL(dev::k_zero_floats, dim3(std::max(1u, grid)), dim3(256), 0, st, a, n);
}

int paths_blocks_per_cu() {
   int b[8] = {0, 0, 0, 0, 0, 0, 0, 0};
   hipError_t e[8];
#define PTX_PATHS_OCC(I, N, G, T) \
   e[I] = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b[I], dev::k_paths<kRenderLds, N, G, T>, 256, 0)
   PTX_PATHS_OCC(0, false, false, false);
   PTX_PATHS_OCC(1, true, false, false);
   PTX_PATHS_OCC(2, false, true, false);
   PTX_PATHS_OCC(3, true, true, false);
   PTX_PATHS_OCC(4, false, false, true);
   PTX_PATHS_OCC(5, true, false, true);
   PTX_PATHS_OCC(6, false, true, true);
   PTX_PATHS_OCC(7, true, true, true);
#undef PTX_PATHS_OCC
   int m = b[0];
   for (int i = 0; i < 8; i++) {
      if (e[i] != hipSuccess) return 1;
      m = std::min(m, b[i]);
   }
   return std::max(1, m);   // one grid and one scratch area serve every variant
}

uint32_t paths_max_chunks() { return kMaxRoundChunks; }
uint32_t paths_max_scenes() { return dev::kMaxRoundScenes; }

void launch_paths(const Chunk *chunks, uint32_t nchunks, uint32_t nslots, const JobDev *jobs, const DevScene *scenes,
                  uint32_t nscenes, float *res, uint32_t *work, uint4 *spill, uint32_t max_stack, uint32_t shade_min,
                  uint32_t leaf_min, float4 *nee_scratch, bool glossy, bool textured, uint32_t grid, hipStream_t st) {
   hipLaunchKernelGGL(dev::k_zero_words, dim3(1), dim3(64), 0, st, work, 8u, (uint32_t *)nullptr, 0u);
#define PTX_LAUNCH_PATHS(N, G, T)                                                                                     \
   hipLaunchKernelGGL((dev::k_paths<kRenderLds, N, G, T>), dim3(grid), dim3(256), 0, st, chunks, nchunks, nslots, jobs, \
                      scenes, nscenes, res, work, spill, max_stack, shade_min, leaf_min, nee_scratch)
   const int v = (nee_scratch ? 1 : 0) | (glossy ? 2 : 0) | (textured ? 4 : 0);
   switch (v) {
      case 0: PTX_LAUNCH_PATHS(false, false, false); break;
      case 1: PTX_LAUNCH_PATHS(true, false, false); break;
      case 2: PTX_LAUNCH_PATHS(false, true, false); break;
      case 3: PTX_LAUNCH_PATHS(true, true, false); break;
      case 4: PTX_LAUNCH_PATHS(false, false, true); break;
      case 5: PTX_LAUNCH_PATHS(true, false, true); break;
      case 6: PTX_LAUNCH_PATHS(false, true, true); break;
      default: PTX_LAUNCH_PATHS(true, true, true); break;
   }
#undef PTX_LAUNCH_PATHS
}

namespace {
template <bool T, bool C>
void launch_mega_t(const DevScene &S, const JobDev &J, uint32_t sample, uint4 *spill, uint32_t grid, hipStream_t st) {
   if (J.flags & PTX_JOB_NEE)
      hipLaunchKernelGGL((dev::k_mega<kRenderLds, true, T, C>), dim3(grid), dim3(256), 0, st, S, J, sample, spill);
   else
      hipLaunchKernelGGL((dev::k_mega<kRenderLds, false, T, C>), dim3(grid), dim3(256), 0, st, S, J, sample, spill);
}
}  // namespace

void launch_mega(const DevScene &S, const JobDev &J, uint32_t sample, uint4 *spill, bool textured, bool cached,
                 uint32_t grid, hipStream_t st) {
   if (cached) {
      if (textured) launch_mega_t<true, true>(S, J, sample, spill, grid, st);
      else launch_mega_t<false, true>(S, J, sample, spill, grid, st);
   } else {
      if (textured) launch_mega_t<true, false>(S, J, sample, spill, grid, st);
      else launch_mega_t<false, false>(S, J, sample, spill, grid, st);
   }
}

}  // namespace ptx
