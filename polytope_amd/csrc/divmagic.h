// Exact unsigned 32-bit division by a divisor that stays fixed for many
// dividends (a chunk's pixel count), as a multiply-high and shifts instead of
// the ~20-instruction integer division the compiler emits for a run-time
// divisor.  Granlund and Montgomery, "Division by invariant integers using
// multiplication" (PLDI 1994), Figure 4.1: with l = ceil(log2 d) and
// m = floor(2^32 (2^l - d) / d) + 1 (which fits 32 bits),
//   n / d = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0),  t = mulhi(m, n),
// for every 32-bit n and every d >= 1.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ptx {

struct DivU32 {
   uint32_t m;    // the multiplier
   uint32_t sh;   // min(l, 1) | max(l - 1, 0) << 8
};

__host__ __device__ __forceinline__ DivU32 div_u32_setup(uint32_t d) {
   uint32_t l = 0;                         // ceil(log2 d)
   while (l < 32 && (uint64_t(1) << l) < d) l++;
   const uint64_t m = ((((uint64_t(1) << l) - d) << 32) / d) + 1;
   const uint32_t sh1 = l < 1 ? l : 1u, sh2 = l > 1 ? l - 1u : 0u;
   return DivU32{uint32_t(m), sh1 | sh2 << 8};
}

__host__ __device__ __forceinline__ uint32_t div_u32(uint32_t n, DivU32 q) {
   const uint32_t t = uint32_t((uint64_t(q.m) * n) >> 32);
   return (t + ((n - t) >> (q.sh & 255u))) >> (q.sh >> 8);
}

}  // namespace ptx
