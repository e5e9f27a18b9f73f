// The float32 expf, logf, sincosf and cabsf of the host C library, restated so that device code
// returns the same bits.
//
// The reference's compiled CalcUcg and Weickenmeier-Kohl routines (diffraction.f90:515-626,
// others.f90:68-778) call the C library for these functions. On x86-64 the C library (glibc 2.28
// and later) uses ARM's optimized-routines algorithms, built with fused multiply-add where the
// processor has it. None of them is correctly rounded, so another implementation differs from
// them in the last bit for some inputs. The functions below restate the published algorithms:
//   - expf: x/ln2 = k/32 + r, 2^(k/32) from a 32-entry table of 2^(i/32) (bit pattern minus
//     i << 47) and 2^(r/32) from a cubic, all in double, rounded once to float;
//   - logf: x = 2^k z with z in [0x3f330000, 2*0x3f330000), a 16-entry table of (1/c, log c)
//     for subintervals of z, log1p(z/c - 1) from a cubic, all in double, rounded once;
//   - sincosf: the even degree-8 and odd degree-7 polynomials in double on [-pi/4, pi/4], with
//     the quadrant from a double multiply by 2/pi for |x| < 120 and from a fixed-point product
//     with 2/pi to 192 bits beyond;
//   - cabsf (hypotf): sqrt((double)x*x + (double)y*y) rounded to float.
// Each multiply-add that the FMA build of the C library fuses is written as an explicit fma();
// every other step is a single IEEE operation in the written order (contraction off).
// tests/test_sg_math.py checks every function against the C library on every float input of the
// domains the setup uses.
//
// The 2^(i/32) table and the 2/pi words follow from their definitions; the polynomial
// coefficients and the logf table are the published constants of those algorithms. The
// exhaustive comparison reaches every entry of every table.
#pragma once
#include <stdint.h>

#include "ctem_sg.h"

#pragma clang fp contract(off)

namespace ctem {
namespace libm {

CTEM_HD inline uint64_t d2u(double d) { uint64_t u; __builtin_memcpy(&u, &d, 8); return u; }
CTEM_HD inline double u2d(uint64_t u) { double d; __builtin_memcpy(&d, &u, 8); return d; }
CTEM_HD inline uint32_t top12(float x) { return f2u(x) >> 20; }

// ---- expf ------------------------------------------------------------------------------------
CTEM_HD inline float expf(float x) {
#pragma clang fp contract(off)
  // bit pattern of 2^(i/32) minus i << 47: adding k << 47 gives 2^(k/32) for the whole range
  static constexpr uint64_t kTab[32] = {
      0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,
      0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,
      0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
      0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,
      0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,
      0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
      0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,
      0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull};
  constexpr double kInvLn2N = 0x1.71547652b82fep+0 * 32, kShift = 0x1.8p+52;
  constexpr double kC0 = 0x1.c6af84b912394p-5 / 32 / 32 / 32, kC1 = 0x1.ebfce50fac4f3p-3 / 32 / 32,
                   kC2 = 0x1.62e42ff0c52d6p-1 / 32;
  const uint32_t abstop = top12(x) & 0x7ff;
  if (abstop >= top12(88.0f)) {
    if (f2u(x) == 0xff800000u) return 0.0f;          // -inf
    if (abstop >= top12(__builtin_inff())) return x + x;
    if (x > 0x1.62e42ep6f) return __builtin_inff();  // above log(2^128)
    if (x < -0x1.9fe368p6f) return 0.0f;             // below log(2^-150)
  }
  const double xd = (double)x;
  double kd = __builtin_fma(kInvLn2N, xd, kShift);
  const uint64_t ki = d2u(kd);
  kd = kd - kShift;
  const double r = __builtin_fma(kInvLn2N, xd, -kd);
  const double s = u2d(kTab[ki % 32] + (ki << 47));
  const double z = __builtin_fma(kC0, r, kC1);
  const double r2 = r * r;
  double y = __builtin_fma(kC2, r, 1.0);
  y = __builtin_fma(z, r2, y);
  y = y * s;
  return (float)y;
}

// ---- logf ------------------------------------------------------------------------------------
CTEM_HD inline float logf(float x) {
#pragma clang fp contract(off)
  static constexpr double kInvc[16] = {
      0x1.661ec79f8f3bep+0, 0x1.571ed4aaf883dp+0, 0x1.49539f0f010bp+0,  0x1.3c995b0b80385p+0,
      0x1.30d190c8864a5p+0, 0x1.25e227b0b8eap+0,  0x1.1bb4a4a1a343fp+0, 0x1.12358f08ae5bap+0,
      0x1.0953f419900a7p+0, 0x1p+0,               0x1.e608cfd9a47acp-1, 0x1.ca4b31f026aap-1,
      0x1.b2036576afce6p-1, 0x1.9c2d163a1aa2dp-1, 0x1.886e6037841edp-1, 0x1.767dcf5534862p-1};
  static constexpr double kLogc[16] = {
      -0x1.57bf7808caadep-2, -0x1.2bef0a7c06ddbp-2, -0x1.01eae7f513a67p-2, -0x1.b31d8a68224e9p-3,
      -0x1.6574f0ac07758p-3, -0x1.1aa2bc79c81p-3,   -0x1.a4e76ce8c0e5ep-4, -0x1.1973c5a611ccp-4,
      -0x1.252f438e10c1ep-5, 0x0p+0,                0x1.aa5aa5df25984p-5,  0x1.c5e53aa362eb4p-4,
      0x1.526e57720db08p-3,  0x1.bc2860d22477p-3,   0x1.1058bc8a07ee1p-2,  0x1.4043057b6ee09p-2};
  constexpr double kLn2 = 0x1.62e42fefa39efp-1;
  constexpr double kA0 = -0x1.00ea348b88334p-2, kA1 = 0x1.5575b0be00b6ap-2, kA2 = -0x1.ffffef20a4123p-2;
  uint32_t ix = f2u(x);
  if (ix == 0x3f800000u) return 0.0f;
  if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u) {
    if (ix * 2 == 0) return -__builtin_inff();
    if (ix == 0x7f800000u) return x;
    if ((ix & 0x80000000u) || ix * 2 >= 0xff000000u) return __builtin_nanf("");
    ix = f2u(x * 0x1p23f);  // subnormal: scale to normal
    ix -= 23u << 23;
  }
  const uint32_t tmp = ix - 0x3f330000u;
  const int i = (int)((tmp >> 19) % 16);
  const int k = (int32_t)tmp >> 23;
  const uint32_t iz = ix - (tmp & (0x1ffu << 23));
  const double z = (double)u2f(iz);
  const double r = __builtin_fma(z, kInvc[i], -1.0);
  const double y0 = __builtin_fma((double)k, kLn2, kLogc[i]);
  const double r2 = r * r;
  double y = __builtin_fma(kA1, r, kA2);
  y = __builtin_fma(kA0, r2, y);
  y = __builtin_fma(y, r2, y0 + r);
  return (float)y;
}

// ---- sincosf ---------------------------------------------------------------------------------
struct SinCosPoly {
  double c0, c1, c2, c3, c4, s1, s2, s3;
};

CTEM_HD inline void sincos_poly(double x, double x2, const SinCosPoly& p, int n, float* sinp, float* cosp) {
#pragma clang fp contract(off)
  const double x4 = x2 * x2, x3 = x2 * x;
  const double c2 = __builtin_fma(x2, p.c4, p.c3);
  const double s1 = __builtin_fma(x2, p.s3, p.s2);
  if (n & 1) {
    float* t = sinp;
    sinp = cosp;
    cosp = t;
  }
  const double c1 = __builtin_fma(x2, p.c1, p.c0);
  const double x5 = x3 * x2, x6 = x4 * x2;
  const double s = __builtin_fma(x3, p.s1, x);
  const double c = __builtin_fma(x4, p.c2, c1);
  *sinp = (float)__builtin_fma(x5, s1, s);
  *cosp = (float)__builtin_fma(x6, c2, c);
}

CTEM_HD inline void sincosf(float y, float* sinp, float* cosp) {
#pragma clang fp contract(off)
  // the cosine polynomial, and its negation for quadrants 2 and 3
  constexpr SinCosPoly kP0 = {0x1p0,  -0x1.ffffffd0c621cp-2, 0x1.55553e1068f19p-5, -0x1.6c087e89a359dp-10,
                              0x1.99343027bf8c3p-16, -0x1.555545995a603p-3, 0x1.1107605230bc4p-7,
                              -0x1.994eb3774cf24p-13};
  constexpr SinCosPoly kP1 = {-0x1p0, 0x1.ffffffd0c621cp-2, -0x1.55553e1068f19p-5, 0x1.6c087e89a359dp-10,
                              -0x1.99343027bf8c3p-16, -0x1.555545995a603p-3, 0x1.1107605230bc4p-7,
                              -0x1.994eb3774cf24p-13};
  // 2/pi to 192 bits, as 32-bit words that advance 8 bits per entry
  static constexpr uint32_t kInvPio4[24] = {
      0xa2,       0xa2f9,     0xa2f983,   0xa2f9836e, 0xf9836e4e, 0x836e4e44, 0x6e4e4415, 0x4e441529,
      0x441529fc, 0x1529fc27, 0x29fc2757, 0xfc2757d1, 0x2757d1f5, 0x57d1f534, 0xd1f534dd, 0xf534ddc0,
      0x34ddc0db, 0xddc0db62, 0xc0db6295, 0xdb629599, 0x6295993c, 0x95993c43, 0x993c4390, 0x3c439041};
  constexpr double kHpiInv = 0x1.45F306DC9C883p+23;  // 2/pi * 2^24
  constexpr double kHpi = 0x1.921FB54442D18p0;
  constexpr double kPi63 = 0x1.921FB54442D18p-62;    // 2 pi * 2^-64
  const uint32_t at = top12(y) & 0x7ff;
  double x = (double)y;
  if (at < (top12((float)0x1.921FB54442D18p-1) & 0x7ff)) {
    if (at < (top12(0x1p-12f) & 0x7ff)) {
      *sinp = y;
      *cosp = 1.0f;
      return;
    }
    sincos_poly(x, x * x, kP0, 0, sinp, cosp);
  } else if (at < (top12(120.0f) & 0x7ff)) {
    const double r = x * kHpiInv;
    const int n = (((int32_t)r) + 0x800000) >> 24;
    x = __builtin_fma(-(double)n, kHpi, x);
    const double s = (n & 3) == 0 || (n & 3) == 3 ? 1.0 : -1.0;
    sincos_poly(x * s, x * x, (n & 2) ? kP1 : kP0, n, sinp, cosp);
  } else if (at < (top12(__builtin_inff()) & 0x7ff)) {
    uint32_t xi = f2u(y);
    const int sign = (int)(xi >> 31);
    const uint32_t* arr = &kInvPio4[(xi >> 26) & 15];
    const int shift = (int)((xi >> 23) & 7);
    xi = (xi & 0xffffffu) | 0x800000u;
    xi <<= shift;
    uint64_t res0 = (uint64_t)(uint32_t)(xi * arr[0]);
    const uint64_t res1 = (uint64_t)xi * arr[4];
    const uint64_t res2 = (uint64_t)xi * arr[8];
    res0 = (res2 >> 32) | (res0 << 32);
    res0 += res1;
    const uint64_t nn = (res0 + (1ull << 61)) >> 62;
    res0 -= nn << 62;
    const int n = (int)nn;
    x = (double)(int64_t)res0 * kPi63;
    const int q = (n + sign) & 3;
    const double s = (q == 0 || q == 3) ? 1.0 : -1.0;
    sincos_poly(x * s, x * x, (q & 2) ? kP1 : kP0, n, sinp, cosp);
  } else {
    *sinp = *cosp = y - y;
  }
}

// ---- cabsf -----------------------------------------------------------------------------------
CTEM_HD inline float cabsf(float re, float im) {
#pragma clang fp contract(off)
  if (__builtin_isinf(re) || __builtin_isinf(im)) return __builtin_inff();
  const double a = (double)re, b = (double)im;
  return (float)__builtin_sqrt(a * a + b * b);
}

}  // namespace libm
}  // namespace ctem
