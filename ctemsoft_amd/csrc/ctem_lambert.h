// Float32 pieces of the EBSD detector-pattern program (CTEMEBSD.f90), bit-exact on host and
// device with the reference's compiled code:
//   - atanf_ref: the single-precision arctangent the reference's compiled Lambert inverse calls
//     (the C library's atanf, the Sun fdlibm algorithm: reduction to |x| < 7/16 with atan(0.5),
//     atan(1), atan(1.5), atan(inf) split into high and low parts, then an odd polynomial of
//     degree 23 evaluated as two interleaved polynomials in x^4);
//   - lambert_sq_inverse: Lambert2DSquareInverseSingle (Lambert.f90:405-433) times the grid
//     scale;
//   - quat_mult_f: quat_mult (quaternions.f90:255-265);
//   - sum3_kahan: SUM of a three-element REAL(4) array as the Fortran runtime evaluates it,
//     compensated (Kahan) summation in single precision in element order (the normalisation of
//     the rotated direction, CTEMEBSD.f90:540).
// tests/test_lambert_math.py checks atanf_ref against the C library on every float in [-1, 1]
// (the arguments the Lambert inverse passes, |y| <= |x|).
//
// Every step is a single IEEE operation in the written order (contraction off; division and
// square root correctly rounded on the device through ctem_sg.h).
#pragma once
#include "ctem_sg.h"

#pragma clang fp contract(off)

namespace ctem {

CTEM_HD inline float atanf_ref(float x) {
#pragma clang fp contract(off)
  const float atanhi[4] = {4.6364760399e-01f, 7.8539812565e-01f, 9.8279368877e-01f, 1.5707962513e+00f};
  const float atanlo[4] = {5.0121582440e-09f, 3.7748947079e-08f, 3.4473217170e-08f, 7.5497894159e-08f};
  const float aT[11] = {3.3333334327e-01f,  -2.0000000298e-01f, 1.4285714924e-01f, -1.1111110449e-01f,
                        9.0908870101e-02f,  -7.6918758452e-02f, 6.6610731184e-02f, -5.8335702866e-02f,
                        4.9768779427e-02f,  -3.6531571299e-02f, 1.6285819933e-02f};
  const uint32_t hx = f2u(x), ix = hx & 0x7fffffffu;
  int id;
  if (ix >= 0x4c000000u) {  // |x| >= 2^25
    if (ix > 0x7f800000u) return x + x;
    return (hx >> 31) ? -atanhi[3] - atanlo[3] : atanhi[3] + atanlo[3];
  }
  if (ix < 0x3ee00000u) {          // |x| < 0.4375
    if (ix < 0x31000000u) return x;  // |x| < 2^-29
    id = -1;
  } else {
    x = __builtin_fabsf(x);
    if (ix < 0x3f980000u) {    // |x| < 1.1875
      if (ix < 0x3f300000u) {  // 7/16 <= |x| < 11/16
        id = 0;
        x = CTEM_FDIV(2.0f * x - 1.0f, 2.0f + x);
      } else {  // 11/16 <= |x| < 19/16
        id = 1;
        x = CTEM_FDIV(x - 1.0f, x + 1.0f);
      }
    } else if (ix < 0x401c0000u) {  // |x| < 2.4375
      id = 2;
      x = CTEM_FDIV(x - 1.5f, 1.0f + 1.5f * x);
    } else {  // 2.4375 <= |x| < 2^25
      id = 3;
      x = CTEM_FDIV(-1.0f, x);
    }
  }
  const float z = x * x;
  const float w = z * z;
  const float s1 = z * (aT[0] + w * (aT[2] + w * (aT[4] + w * (aT[6] + w * (aT[8] + w * aT[10])))));
  const float s2 = w * (aT[1] + w * (aT[3] + w * (aT[5] + w * (aT[7] + w * aT[9]))));
  if (id < 0) return x - x * (s1 + s2);
  const float r = atanhi[id] - ((x * (s1 + s2) - atanlo[id]) - x);
  return (hx >> 31) ? -r : r;
}

// ixy = scl * Lambert2DSquareInverseSingle(xyz); returns ierr (1: |1 - |xyz|^2| > 1e-6, ixy = 0).
// spi2 is sngl(LPs%sPi2), sqrt(pi)/2 of InitLambertParameters (Lambert.f90:239-244) rounded to
// float; the host computes it with the C library as the reference does (sqrt(4 atan(1)) / 2).
CTEM_HD inline int lambert_sq_inverse(const float xyz[3], float scl, float spi2, float ixy[2]) {
#pragma clang fp contract(off)
  float r0, r1;
  int ierr = 0;
  const float ss = xyz[0] * xyz[0] + xyz[1] * xyz[1] + xyz[2] * xyz[2];
  if (__builtin_fabsf(1.0f - ss) > 1.0e-6f) {
    r0 = r1 = 0.0f;
    ierr = 1;
  } else if (__builtin_fabsf(xyz[2]) == 1.0f) {
    r0 = r1 = 0.0f;
  } else if (__builtin_fabsf(xyz[1]) <= __builtin_fabsf(xyz[0])) {
    const float q = CTEM_FDIV(__builtin_fabsf(xyz[0]), xyz[0]) * CTEM_FSQRT(2.0f * (1.0f - xyz[2]));
    r0 = q * spi2;
    r1 = CTEM_FDIV(q * atanf_ref(CTEM_FDIV(xyz[1], xyz[0])), spi2);
  } else {
    const float q = CTEM_FDIV(__builtin_fabsf(xyz[1]), xyz[1]) * CTEM_FSQRT(2.0f * (1.0f - xyz[2]));
    r0 = CTEM_FDIV(q * atanf_ref(CTEM_FDIV(xyz[0], xyz[1])), spi2);
    r1 = q * spi2;
  }
  ixy[0] = scl * r0;
  ixy[1] = scl * r1;
  return ierr;
}

CTEM_HD inline void quat_mult_f(const float x[4], const float y[4], float r[4]) {
#pragma clang fp contract(off)
  r[0] = x[0] * y[0] - x[1] * y[1] - x[2] * y[2] - x[3] * y[3];
  r[1] = x[0] * y[1] + x[1] * y[0] + x[2] * y[3] - x[3] * y[2];
  r[2] = x[0] * y[2] - x[1] * y[3] + x[2] * y[0] + x[3] * y[1];
  r[3] = x[0] * y[3] + x[1] * y[2] - x[2] * y[1] + x[3] * y[0];
}

CTEM_HD inline float sum3_kahan(float a, float b, float c) {
#pragma clang fp contract(off)
  const float v[3] = {a, b, c};
  float s = 0.0f, comp = 0.0f;
  for (int i = 0; i < 3; i++) {
    const float y = v[i] - comp;
    const float t = s + y;
    comp = (t - s) - y;
    s = t;
  }
  return s;
}

}  // namespace ctem
