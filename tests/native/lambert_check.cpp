// Exhaustive check of the float32 arctangent restated in ctemsoft_amd/csrc/ctem_lambert.h against
// the host C library's atanf, which the reference's compiled Lambert inverse calls
// (Lambert.f90:405-433): every float in [-1, 1] (the arguments y/x with |y| <= |x|), and every
// float outside it with exponent below 26 as a wider check. Prints the mismatch counts; exit
// status 0 when both are zero.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../ctemsoft_amd/csrc/ctem_lambert.h"

static float from_bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

int main() {
  long long bad_in = 0, n_in = 0, bad_out = 0, n_out = 0;
  const uint32_t one = bits(1.0f), big = bits(0x1p26f);
#pragma omp parallel for reduction(+ : bad_in, n_in, bad_out, n_out) schedule(static, 1 << 16)
  for (long long u = 0; u < (long long)big; u++) {
    for (int sgn = 0; sgn < 2; sgn++) {
      const float x = from_bits((uint32_t)u | (sgn ? 0x80000000u : 0u));
      const bool in = (uint32_t)u <= one;
      const bool bad = bits(ctem::atanf_ref(x)) != bits(atanf(x));
      if (in) { n_in++; bad_in += bad; } else { n_out++; bad_out += bad; }
    }
  }
  printf("atanf in [-1,1]: %lld of %lld differ\n", bad_in, n_in);
  printf("atanf 1 < |x| < 2^26: %lld of %lld differ\n", bad_out, n_out);
  return (bad_in || bad_out) ? 1 : 0;
}
