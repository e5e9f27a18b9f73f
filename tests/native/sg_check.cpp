// Exhaustive check of the float32 math restated in ctemsoft_amd/csrc/ctem_sg.h against the
// host C library that the reference's compiled CalcsgSingle calls (glibc acosf, cosf):
//   acosf_ref  on every float in [-1, 1];
//   cosf_ref   on every float in [0, pi/4) (the branch the excitation errors use);
//   sqrtf_cr   on every float in [2^-10, 2^10) against the correctly rounded sqrtf.
// Prints the number of mismatches per function; exit status 0 when all are zero.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../ctemsoft_amd/csrc/ctem_sg.h"

static float from_bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

int main() {
  long long bad_acos = 0, bad_cos = 0, bad_sqrt = 0, n_acos = 0, n_cos = 0, n_sqrt = 0;
  const uint32_t one = bits(1.0f);
  // acosf: +0 .. 1 and -0 .. -1
#pragma omp parallel for reduction(+ : bad_acos, n_acos) schedule(static, 1 << 16)
  for (long long u = 0; u <= (long long)one; u++) {
    for (int sgn = 0; sgn < 2; sgn++) {
      const float x = from_bits((uint32_t)u | (sgn ? 0x80000000u : 0u));
      n_acos++;
      if (bits(ctem::acosf_ref(x)) != bits(acosf(x))) bad_acos++;
    }
  }
  const uint32_t qpi = 0x3f490fdau;  // first float >= pi/4
#pragma omp parallel for reduction(+ : bad_cos, n_cos) schedule(static, 1 << 16)
  for (long long u = 0; u < (long long)qpi; u++) {
    const float x = from_bits((uint32_t)u);
    n_cos++;
    if (bits(ctem::cosf_ref(x)) != bits(cosf(x))) bad_cos++;
  }
  const uint32_t lo = bits(0x1p-10f), hi = bits(0x1p10f);
#pragma omp parallel for reduction(+ : bad_sqrt, n_sqrt) schedule(static, 1 << 16)
  for (long long u = lo; u < (long long)hi; u++) {
    const float x = from_bits((uint32_t)u);
    n_sqrt++;
    if (bits(ctem::sqrtf_cr(x)) != bits(sqrtf(x))) bad_sqrt++;
  }
  printf("acosf %lld of %lld differ\ncosf %lld of %lld differ\nsqrtf %lld of %lld differ\n", bad_acos,
         n_acos, bad_cos, n_cos, bad_sqrt, n_sqrt);
  return (bad_acos || bad_cos || bad_sqrt) ? 1 : 0;
}
