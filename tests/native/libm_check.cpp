// Exhaustive check of the float32 functions restated in ctemsoft_amd/csrc/ctem_libm.h against the
// host C library that the reference's compiled CalcUcg and Weickenmeier-Kohl routines call:
//   expf     on every float (NaNs aside);
//   logf     on every float >= +0 (negative arguments give NaN in both);
//   sincosf  on every finite float (both results), and cexpf(0 + iy) = (cos y, sin y) of it
//            (the form of CalcUcg's phase factors) on every float |y| < 4096;
//   cabsf    on 2^28 pairs of floats with exponents spread over the float range, and on
//            every 7th non-negative float paired with each of a fixed set of second components.
// Every entry of the restated tables is reached by these inputs. Prints the number of
// mismatches per function; exit status 0 when all are zero.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../ctemsoft_amd/csrc/ctem_libm.h"

extern "C" float cabsf(float _Complex);
extern "C" float _Complex cexpf(float _Complex);

static float from_bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static bool same(float a, float b) { return bits(a) == bits(b) || (isnan(a) && isnan(b)); }

static uint64_t splitmix(uint64_t* s) {
  uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

int main() {
  long long bad_exp = 0, bad_log = 0, bad_sc = 0, bad_abs = 0, bad_cx = 0, n_exp = 0, n_log = 0, n_sc = 0,
            n_abs = 0, n_cx = 0;
#pragma omp parallel for reduction(+ : bad_exp, n_exp, bad_sc, n_sc) schedule(static, 1 << 20)
  for (long long u = 0; u < (1ll << 32); u++) {
    const float x = from_bits((uint32_t)u);
    if (isnan(x)) continue;
    n_exp++;
    if (!same(ctem::libm::expf(x), expf(x))) bad_exp++;
    if (isinf(x)) continue;
    float s, c, rs, rc;
    ctem::libm::sincosf(x, &s, &c);
    sincosf(x, &rs, &rc);
    n_sc++;
    if (!same(s, rs) || !same(c, rc)) bad_sc++;
  }
  const uint32_t ylim = bits(4096.0f);
#pragma omp parallel for reduction(+ : bad_cx, n_cx) schedule(static, 1 << 20)
  for (long long u = 0; u < (long long)ylim; u++) {
    for (int sgn = 0; sgn < 2; sgn++) {
      const float y = from_bits((uint32_t)u | (sgn ? 0x80000000u : 0u));
      float s, c;
      ctem::libm::sincosf(y, &s, &c);
      float _Complex z;
      __real__ z = 0.0f;
      __imag__ z = y;
      const float _Complex e = cexpf(z);
      n_cx++;
      if (!same(__real__ e, c) || !same(__imag__ e, s)) bad_cx++;
    }
  }
#pragma omp parallel for reduction(+ : bad_log, n_log) schedule(static, 1 << 20)
  for (long long u = 0; u <= 0x7f800000ll; u++) {
    const float x = from_bits((uint32_t)u);
    n_log++;
    if (!same(ctem::libm::logf(x), logf(x))) bad_log++;
  }
  // cabsf: random pairs (exponent and mantissa bits uniform over the finite floats' range)
#pragma omp parallel for reduction(+ : bad_abs, n_abs) schedule(static, 1 << 16)
  for (long long i = 0; i < (1ll << 28); i++) {
    uint64_t st = (uint64_t)i * 0x2545f4914f6cdd1dull;
    const uint64_t r = splitmix(&st);
    uint32_t a = (uint32_t)r & 0x7fffffffu, b = (uint32_t)(r >> 32) & 0x7fffffffu;
    if (a >= 0x7f800000u || b >= 0x7f800000u) continue;
    // half the pairs within 2^12 of each other's magnitude, where both squares matter
    if (i & 1) b = (a & 0x807fffffu) | ((((a >> 23) + ((b >> 23) & 31) - 16) & 0xff) << 23);
    if ((b & 0x7f800000u) == 0x7f800000u) continue;
    const float x = from_bits(a | ((uint32_t)(r >> 16) & 0x80000000u)), y = from_bits(b);
    float _Complex z;
    __real__ z = x;
    __imag__ z = y;
    n_abs++;
    if (!same(ctem::libm::cabsf(x, y), cabsf(z))) bad_abs++;
  }
  const float second[6] = {0.0f, 1.0e-10f, 3.0e-11f, 1.0f, 7.5e-6f, 1.0e20f};
#pragma omp parallel for reduction(+ : bad_abs, n_abs) schedule(static, 1 << 20)
  for (long long u = 0; u < 0x7f800000ll; u += 7) {
    const float x = from_bits((uint32_t)u);
    for (float y : second) {
      float _Complex z;
      __real__ z = x;
      __imag__ z = y;
      n_abs++;
      if (!same(ctem::libm::cabsf(x, y), cabsf(z))) bad_abs++;
    }
  }
  printf("expf %lld of %lld differ\nlogf %lld of %lld differ\nsincosf %lld of %lld differ\n"
         "cabsf %lld of %lld differ\ncexpf %lld of %lld differ\n",
         bad_exp, n_exp, bad_log, n_log, bad_sc, n_sc, bad_abs, n_abs, bad_cx, n_cx);
  return (bad_exp || bad_log || bad_sc || bad_abs || bad_cx) ? 1 : 0;
}
