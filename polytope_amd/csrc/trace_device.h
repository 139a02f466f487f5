// Device-side building blocks of the closest-hit path: the box test, the Wald
// watertight triangle test, the post-hit shading frame and the per-lane
// traversal state machine.  Included by the HIP translation units only.
//
// Arithmetic follows the reference operation by operation (fused only where it
// calls std::fma): this code is compiled with -ffp-contract=off and clang's
// correctly rounded fp32 division and square root.
#pragma once

#include <hip/hip_runtime.h>

#include "ptx_internal.h"

#pragma clang fp contract(off)

namespace ptx {
namespace dev {

constexpr float kInf = __builtin_huge_valf();
// src/cpu/constants.h:247, 320-333
constexpr float kMachineEpsilon = 5.9604644775390625e-8f;   // numeric_limits<float>::epsilon() * 0.5f
constexpr float gamma_n(int n) { return (n * kMachineEpsilon) / (1.f - n * kMachineEpsilon); }
constexpr float kGamma2 = (2 * kMachineEpsilon) / (1 - 2 * kMachineEpsilon);
constexpr float kGamma3 = (3 * kMachineEpsilon) / (1 - 3 * kMachineEpsilon);
constexpr float kGamma5 = (5 * kMachineEpsilon) / (1 - 5 * kMachineEpsilon);
constexpr float kGamma7 = (7 * kMachineEpsilon) / (1 - 7 * kMachineEpsilon);
constexpr float kSlabScale = 1 + 2 * gamma_n(3);              // BoundingBox.cpp:27

__device__ __forceinline__ float fmaf_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// src/cpu/constants.h:121-133
__device__ __forceinline__ float diff_of_products(float a, float b, float c, float d) {
   const float cd = c * d;
   const float dop = fmaf_(a, b, -cd);
   const float err = fmaf_(-c, d, cd);
   return dop + err;
}
__device__ __forceinline__ float sum_of_products(float a, float b, float c, float d) {
   const float cd = c * d;
   const float sop = fmaf_(a, b, cd);
   const float err = fmaf_(c, d, -cd);
   return sop + err;
}

// std::nextafter(x, +/-inf) for float, by bit stepping.
__device__ __forceinline__ float next_up(float x) {
   if (x != x || x == kInf) return x;
   if (x == 0.f) return __uint_as_float(1u);
   const uint32_t b = __float_as_uint(x);
   return __uint_as_float(x > 0.f ? b + 1u : b - 1u);
}
__device__ __forceinline__ float next_down(float x) {
   if (x != x || x == -kInf) return x;
   if (x == 0.f) return __uint_as_float(0x80000001u);
   const uint32_t b = __float_as_uint(x);
   return __uint_as_float(x > 0.f ? b - 1u : b + 1u);
}

// Per-ray constants.  The reference recomputes the permutation and shear for
// every face (mesh.cpp:196-236); they depend on the ray only, so computing them
// once gives the same values.
struct Ray {
   float ox, oy, oz, dx, dy, dz;
   float ix, iy, iz;          // 1 / d  (bvh.cpp:477-481)
   uint32_t nk;               // bit a (0-2): inverse_direction[a] < 0 (bvh.cpp:483); bits 3-4: the
                              // Wald dominant axis kz, with kx = kz+1, ky = kz+2 (mod 3)
   float Sx, Sy, Sz;
};

__device__ __forceinline__ float pick3(uint32_t k, float a, float b, float c) {
   return k == 0 ? a : (k == 1 ? b : c);
}

__device__ __forceinline__ void make_ray(Ray &r, float ox, float oy, float oz, float dx, float dy, float dz) {
   r.ox = ox; r.oy = oy; r.oz = oz; r.dx = dx; r.dy = dy; r.dz = dz;
   r.ix = 1.f / dx; r.iy = 1.f / dy; r.iz = 1.f / dz;
   const uint32_t neg = (r.ix < 0 ? 1u : 0u) | (r.iy < 0 ? 2u : 0u) | (r.iz < 0 ? 4u : 0u);
   const float ax = __builtin_fabsf(dx), ay = __builtin_fabsf(dy), az = __builtin_fabsf(dz);
   uint32_t kz = 0;                                            // mesh.cpp:200-204
   if (ay > ax && ay > az) kz = 1;
   else if (az > ax && az > ay) kz = 2;
   r.nk = neg | kz << 3;
   const uint32_t kx = kz == 2 ? 0 : kz + 1;
   const uint32_t ky = kx == 2 ? 0 : kx + 1;
   const float pdx = pick3(kx, dx, dy, dz), pdy = pick3(ky, dx, dy, dz), pdz = pick3(kz, dx, dy, dz);
   r.Sx = -pdx / pdz;                                          // mesh.cpp:234-236
   r.Sy = -pdy / pdz;
   r.Sz = pick3(kz, r.ix, r.iy, r.iz);                         // 1 / pdz, the same division
}

// BoundingBox::Hits (BoundingBox.cpp:12-85) as the slab interval [near, far]
// folded from near = 0 and far = +inf with the reference's selections.  The
// reference result for MinT m is
//   inside || (near <= far && near <= m),
// because its far bound is the same fold started from m instead of +inf and
// its running near/far are monotone, so its early exits agree with the final
// comparison.  Its `inside` term (origin strictly inside the box) never
// decides: then b - o is exact and nonzero on every axis with opposite signs
// at the two faces, so every axis gives one negative and one positive slab
// value (no NaN, even with 1/d infinite), near stays 0 and far > 0, and
// 0 <= m because every MinT is positive.  So the test is
// near <= far && near <= m, and a deferred child's near (0 for an origin
// inside it) is tested at pop time as near <= MinT.  The per-axis swap keeps
// the reference's selections (with a NaN slab value, from 0 * inf, it
// decides which value goes on).  The folds are fmaxf/fminf, which the
// compiler merges into three-input max/min: the reference's
// `near = tn > near ? tn : near` keeps near when tn is NaN, and so does fmaxf
// because near starts at 0 and is never NaN (likewise far from +inf).  The
// two can differ only in the sign of a zero result, and near and far are only
// ever compared, where -0 == +0.
__device__ __forceinline__ void slab(float b0, float b1, float b2, float b3, float b4, float b5,
                                     const Ray &r, float &near, float &far) {
   near = 0.f;
   far = kInf;
   float tn, tf, sw;
   tn = (b0 - r.ox) * r.ix; tf = (b3 - r.ox) * r.ix;
   sw = tn; tn = tn > tf ? tf : tn; tf = sw > tf ? sw : tf; tf *= kSlabScale;
   near = fmaxf(near, tn); far = fminf(far, tf);
   tn = (b1 - r.oy) * r.iy; tf = (b4 - r.oy) * r.iy;
   sw = tn; tn = tn > tf ? tf : tn; tf = sw > tf ? sw : tf; tf *= kSlabScale;
   near = fmaxf(near, tn); far = fminf(far, tf);
   tn = (b2 - r.oz) * r.iz; tf = (b5 - r.oz) * r.iz;
   sw = tn; tn = tn > tf ? tf : tn; tf = sw > tf ? sw : tf; tf *= kSlabScale;
   near = fmaxf(near, tn); far = fminf(far, tf);
}

// MaxComponent (mesh.cpp:149-155) of three absolute values, as a three-input
// max.  The reference's comparison chain returns the same value whenever no
// input is NaN.  Its results feed only deltaT, which matters only through
// `t <= deltaT`, and a NaN input (a NaN coordinate, U, V or W) makes det or T,
// and so t, NaN: the comparison is then false whatever deltaT is.
__device__ __forceinline__ float max_component(float x, float y, float z) {
   return fmaxf(fmaxf(x, y), z);
}

// Mesh::intersect test part for one face (mesh.cpp:196-340).  On acceptance
// updates min_t and the hit's (u, v, w) and returns true.
__device__ __forceinline__ bool test_triangle(const float *v, const Ray &r, float &min_t,
                                              float &hu, float &hv, float &hw, bool debug_pixel) {
   if (debug_pixel && r.dx != r.dx) return false;              // mesh.cpp:192-194
   const float v0x = v[0], v0y = v[1], v0z = v[2];
   const float v1x = v[3], v1y = v[4], v1z = v[5];
   const float v2x = v[6], v2y = v[7], v2z = v[8];
   // p_i = v_i - origin, then permuted (kx, ky, kz) (mesh.cpp:214-231)
   const float a0 = v0x - r.ox, a1 = v0y - r.oy, a2 = v0z - r.oz;
   const float b0 = v1x - r.ox, b1 = v1y - r.oy, b2 = v1z - r.oz;
   const float c0 = v2x - r.ox, c1 = v2y - r.oy, c2 = v2z - r.oz;
   const uint32_t kz = r.nk >> 3;
   // kz = 0 -> (y, z, x); kz = 1 -> (z, x, y); kz = 2 -> (x, y, z)
   float p0x = pick3(kz, a1, a2, a0), p0y = pick3(kz, a2, a0, a1), p0z = pick3(kz, a0, a1, a2);
   float p1x = pick3(kz, b1, b2, b0), p1y = pick3(kz, b2, b0, b1), p1z = pick3(kz, b0, b1, b2);
   float p2x = pick3(kz, c1, c2, c0), p2y = pick3(kz, c2, c0, c1), p2z = pick3(kz, c0, c1, c2);
   p0x = fmaf_(r.Sx, p0z, p0x); p0y = fmaf_(r.Sy, p0z, p0y);   // mesh.cpp:239-244
   p1x = fmaf_(r.Sx, p1z, p1x); p1y = fmaf_(r.Sy, p1z, p1y);
   p2x = fmaf_(r.Sx, p2z, p2x); p2y = fmaf_(r.Sy, p2z, p2y);
   float U = diff_of_products(p1x, p2y, p1y, p2x);             // mesh.cpp:247-249
   float V = diff_of_products(p2x, p0y, p2y, p0x);
   float W = diff_of_products(p0x, p1y, p0y, p1x);
   if (U == 0.0f || V == 0.0f || W == 0.0f) {                  // mesh.cpp:252-264
      const double CxBy = (double)p2x * (double)p1y;
      const double CyBx = (double)p2y * (double)p1x;
      U = (float)(CyBx - CxBy);
      const double AxCy = (double)p0x * (double)p2y;
      const double AyCx = (double)p0y * (double)p2x;
      V = (float)(AyCx - AxCy);
      const double BxAy = (double)p1x * (double)p0y;
      const double ByAx = (double)p1y * (double)p0x;
      W = (float)(ByAx - BxAy);
   }
   // mesh.cpp:272: (U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0).
   // fminf/fmaxf skip a NaN operand (and give NaN only when all three are), so
   // the smallest and largest of the three decide exactly what the two
   // disjunctions decide, zeros of either sign included.
   if (fminf(fminf(U, V), W) < 0.0f && fmaxf(fmaxf(U, V), W) > 0.0f) return false;
   const float det = U + V + W;
   if (det == 0.0f) return false;
   p0z *= r.Sz; p1z *= r.Sz; p2z *= r.Sz;
   const float T = fmaf_(U, p0z, sum_of_products(V, p1z, W, p2z));
   if (det < 0.f && (T >= 0 || T < min_t * det)) return false; // mesh.cpp:293-298
   if (det > 0.f && (T <= 0 || T > min_t * det)) return false;
   const float rcp = 1.0f / det;
   const float u = U * rcp, vv = V * rcp, w = W * rcp;
   const float t = T * rcp;
   if (debug_pixel && t != t) return false;                    // mesh.cpp:307-310
   const float maxZt = max_component(__builtin_fabsf(p0z), __builtin_fabsf(p1z), __builtin_fabsf(p2z));
   const float deltaZ = kGamma3 * maxZt;
   const float maxXt = max_component(__builtin_fabsf(p0x), __builtin_fabsf(p1x), __builtin_fabsf(p2x));
   const float maxYt = max_component(__builtin_fabsf(p0y), __builtin_fabsf(p1y), __builtin_fabsf(p2y));
   const float deltaX = kGamma5 * (maxXt + maxZt);
   const float deltaY = kGamma5 * (maxYt + maxZt);
   const float deltaE = 2 * (kGamma2 * maxXt * maxYt + deltaY * maxXt + deltaX * maxYt);
   const float maxE = max_component(__builtin_fabsf(U), __builtin_fabsf(V), __builtin_fabsf(W));
   const float deltaT = 3 * (kGamma3 * maxE * maxZt + deltaE * maxZt + deltaZ * maxE) * __builtin_fabsf(rcp);
   if (t <= deltaT) return false;                              // mesh.cpp:321-322
   min_t = t;
   hu = u; hv = vv; hw = w;
   return true;
}

// Post-hit frame (mesh.cpp:325-340 for the error bound and Location, 366-461
// for normal, offset and tangents), with the origin offset applied `repeats`
// times: the reference re-runs this block after every face test that follows
// the accepting one.
struct Frame {
   float lx, ly, lz, nx, ny, nz, t1x, t1y, t1z, t2x, t2y, t2z, ex, ey, ez;
};

// use_cn selects the interpolated corner normals cn (which must be readable
// either way) over the geometric normal.
__device__ __forceinline__ void hit_frame_sel(const float *__restrict__ v, const float *__restrict__ cn, bool use_cn,
                                              const Ray &r, float u, float vv, float w, uint32_t repeats,
                                              Frame &f) {
   const float v0x = v[0], v0y = v[1], v0z = v[2];
   const float v1x = v[3], v1y = v[4], v1z = v[5];
   const float v2x = v[6], v2y = v[7], v2z = v[8];
   const float xe = __builtin_fabsf(u * v0x) + __builtin_fabsf(vv * v1x) + __builtin_fabsf(w * v2x);
   const float ye = __builtin_fabsf(u * v0y) + __builtin_fabsf(vv * v1y) + __builtin_fabsf(w * v2y);
   const float ze = __builtin_fabsf(u * v0z) + __builtin_fabsf(vv * v1z) + __builtin_fabsf(w * v2z);
   f.ex = xe * kGamma7; f.ey = ye * kGamma7; f.ez = ze * kGamma7;
   float px = fmaf_(v0x, u, sum_of_products(v1x, vv, v2x, w));
   float py = fmaf_(v0y, u, sum_of_products(v1y, vv, v2y, w));
   float pz = fmaf_(v0z, u, sum_of_products(v1z, vv, v2z, w));
   const float e0x = v1x - v0x, e0y = v1y - v0y, e0z = v1z - v0z;
   const float e1x = v2x - v1x, e1y = v2y - v1y, e1z = v2z - v1z;
   const float e2x = v0x - v2x, e2y = v0y - v2y, e2z = v0z - v2z;
   float nx, ny, nz;
   if (use_cn) {
      nx = cn[0] * u + cn[3] * vv + cn[6] * w;
      ny = cn[1] * u + cn[4] * vv + cn[7] * w;
      nz = cn[2] * u + cn[5] * vv + cn[8] * w;
   } else {                                                   // Vector::Cross
      nx = e0y * e1z - e0z * e1y;
      ny = e0z * e1x - e0x * e1z;
      nz = e0x * e1y - e0y * e1x;
   }
   const float rdn = nx * r.dx + ny * r.dy + nz * r.dz;
   const float flip = rdn > 0 ? -1.f : 1.f;
   nx = nx * flip; ny = ny * flip; nz = nz * flip;
   {
      const float ld = 1.0f / __builtin_sqrtf(nx * nx + ny * ny + nz * nz);
      nx *= ld; ny *= ld; nz *= ld;
   }
   const float an_x = __builtin_fabsf(nx), an_y = __builtin_fabsf(ny), an_z = __builtin_fabsf(nz);
   const float d = an_x * f.ex + an_y * f.ey + an_z * f.ez;
   float ox = nx * d, oy = ny * d, oz = nz * d;
   if ((nx * r.dx + ny * r.dy + nz * r.dz) > 0) { ox = -ox; oy = -oy; oz = -oz; }
   for (uint32_t k = 0; k < repeats; k++) {
      px = px + ox; py = py + oy; pz = pz + oz;
      if (ox > 0) px = next_up(px); else if (ox < 0) px = next_down(px);
      if (oy > 0) py = next_up(py); else if (oy < 0) py = next_down(py);
      if (oz > 0) pz = next_up(pz); else if (oz < 0) pz = next_down(pz);
   }
   f.lx = px; f.ly = py; f.lz = pz;
   f.nx = nx; f.ny = ny; f.nz = nz;
   const float ed0 = __builtin_fabsf(e1x * e0x + e1y * e0y + e1z * e0z);
   const float ed1 = __builtin_fabsf(e2x * e1x + e2y * e1y + e2z * e1z);
   const float ed2 = __builtin_fabsf(e0x * e2x + e0y * e2y + e0z * e2z);
   float tx, ty, tz;
   if (ed0 > ed1 && ed0 > ed2) { tx = e0x; ty = e0y; tz = e0z; }
   else if (ed1 > ed0 && ed1 > ed2) { tx = e1x; ty = e1y; tz = e1z; }
   else { tx = e2x; ty = e2y; tz = e2z; }
   float sx = ty * nz - tz * ny, sy = tz * nx - tx * nz, sz = tx * ny - ty * nx;
   {
      const float ld = 1.0f / __builtin_sqrtf(tx * tx + ty * ty + tz * tz);
      f.t1x = tx * ld; f.t1y = ty * ld; f.t1z = tz * ld;
   }
   {
      const float ld = 1.0f / __builtin_sqrtf(sx * sx + sy * sy + sz * sz);
      f.t2x = sx * ld; f.t2y = sy * ld; f.t2z = sz * ld;
   }
}

__device__ __forceinline__ void hit_frame(const float *__restrict__ v, const float *__restrict__ cn,
                                          const Ray &r, float u, float vv, float w, uint32_t repeats,
                                          Frame &f) {
   hit_frame_sel(v, cn ? cn : v, cn != nullptr, r, u, vv, w, repeats, f);
}

// ---------------------------------------------------------------- traversal
//
// bvh::intersect_compact (bvh.cpp:476-523) over the Node64 layout.  The
// reference pops a node, tests its box against the current MinT and, for an
// interior node, pushes both children so that the child on the near side of
// the split axis is popped next.  Here a visited interior node's record holds
// both child boxes, tested at once:
//   * the child popped next is tested with the same MinT the reference would
//     use, so its result is the reference's;
//   * the other child is popped only after the first child's subtree; its slab
//     interval is kept on the stack and re-checked at pop time against the
//     MinT of that moment (near <= MinT).  A child that misses now
//     would miss later too (MinT only decreases), so it is not pushed.
// Leaves test their triangles in leaf order, as the reference does.  The
// counters reproduce the reference's num_bb_hits / num_bb_misses /
// num_triangle_isects exactly: every child the reference pushes is popped and
// counted once.

struct Trav {
   Ray r;
   float min_t, hu, hv, hw;
   uint32_t hit;          // leaf-ordered triangle index, 0xFFFFFFFF = none
   uint32_t after;        // face tests since (and including) the accepting one
   uint32_t cur;          // node to process next: kLeafBit | first triangle, or interior index
   uint32_t cnt;          // faces left in the leaf `cur` (0 for an interior node)
   uint32_t sp;           // deferred stack depth
   uint32_t bb_hits, bb_misses, tris;
};

}  // namespace dev
}  // namespace ptx
