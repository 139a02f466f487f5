// Host check of the division by invariant integers in
// polytope_amd/csrc/divmagic.h (the header the render kernel uses), against
// the `/` operator: every dividend below 2^16 for divisors below 3000, the
// boundary dividends of every such divisor, divisors near powers of two, and
// random 32-bit pairs.  Prints the number of cases and exits 1 on a mismatch.
#include <cstdint>
#include <cstdio>
#include <random>

#include "divmagic.h"

int main() {
   using ptx::div_u32;
   using ptx::div_u32_setup;
   long cases = 0, bad = 0;
   auto check = [&](uint32_t n, uint32_t d, ptx::DivU32 q) {
      cases++;
      if (div_u32(n, q) != n / d) {
         if (bad < 5) std::printf("mismatch n=%u d=%u\n", n, d);
         bad++;
      }
   };
   for (uint32_t d = 1; d < 3000; d++) {
      const ptx::DivU32 q = div_u32_setup(d);
      for (uint32_t n = 0; n < (1u << 16); n++) check(n, d, q);
      for (uint32_t n : {0xFFFFFFFFu, 0xFFFFFFFEu, 0x80000000u, 0x7FFFFFFFu, d - 1, d, d + 1, 2 * d - 1, 2 * d})
         check(n, d, q);
   }
   for (uint32_t k = 1; k < 32; k++)
      for (uint32_t d : {(1u << k) - 1, 1u << k, (1u << k) + 1}) {
         const ptx::DivU32 q = div_u32_setup(d);
         for (uint32_t n : {0u, 1u, d - 1, d, d + 1, 0x7FFFFFFFu, 0x80000000u, 0xFFFFFFFEu, 0xFFFFFFFFu})
            check(n, d, q);
      }
   std::mt19937_64 rng(7);
   for (int i = 0; i < 4000000; i++) {
      uint32_t d = uint32_t(rng()) >> (rng() % 32);
      if (d == 0) d = 1;
      const uint32_t n = uint32_t(rng()) >> (i % 3 == 0 ? rng() % 32 : 0);
      check(n, d, div_u32_setup(d));
   }
   std::printf("%ld cases, %ld mismatches\n", cases, bad);
   return bad ? 1 : 0;
}
