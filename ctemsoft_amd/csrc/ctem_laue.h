// Laue-symmetry equivalents of a Lambert pixel (Apply2DLaueSymmetry, Lambert.f90:1942-2046),
// usable on host and device. The hexagonal-grid cases keep the reference's float32 floor
// and truncation steps.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define CTEM_LHD __host__ __device__
#else
#define CTEM_LHD
#endif

namespace ctem {

CTEM_LHD inline int fl2(int v) {  // floor(float(v)/2.0)
  float f = (float)v / 2.0f;
  int i = (int)f;
  return (f < 0.0f && (float)i != f) ? i - 1 : i;
}

// Writes up to 12 (i,j) pairs into q; returns their count.
CTEM_LHD inline int laue_equiv(int ipx, int ipy, int isym, int32_t* q) {
  int ne = 1;
  q[0] = ipx; q[1] = ipy;
  if (ipx == 0 && ipy == 0) return 1;
  int fly = fl2(ipy);
  float my;
#define SETQ(n, a, b) do { q[2 * (n)] = (a); q[2 * (n) + 1] = (b); } while (0)
  switch (isym) {
    case 2: SETQ(1, ipx, -ipy); ne = 2; break;
    case 4: SETQ(1, -ipy, ipx); SETQ(2, ipy, -ipx); SETQ(3, -ipx, -ipy); ne = 4; break;
    case 3: case 10: SETQ(1, -ipx, ipy); SETQ(2, ipx, -ipy); SETQ(3, -ipx, -ipy); ne = 4; break;
    case 5: case 11:
      SETQ(1, -ipx, ipy); SETQ(2, ipx, -ipy); SETQ(3, -ipx, -ipy); SETQ(4, ipy, ipx);
      SETQ(5, -ipy, ipx); SETQ(6, ipy, -ipx); SETQ(7, -ipy, -ipx); ne = 8; break;
    case 6: case 7:
      SETQ(1, ipy - 2 * fly - fl2(ipx + 3 * ipy - 3 * fly), ipx - ipy + fly);
      SETQ(2, ipy - fly - fl2(ipx - fly), -ipx - fly);
      ne = 3; break;
    case 13:
      my = (float)(ipy % 2) / 2.0f;
      SETQ(1, ipy - fl2(ipx + fly), (int)((float)(ipx + ipy / 2) - my));
      SETQ(2, ipy - 2 * fly - fl2(ipx + 3 * ipy - 3 * fly), ipx - ipy + fly);
      SETQ(3, ipx, -ipy);
      SETQ(4, ipy - fly - fl2(ipx - fly), -ipx - fly);
      SETQ(5, -fly - fl2(ipx + ipy - fly), (int)((float)(-ipx + ipy / 2) + my));
      ne = 6; break;
    case 8:
      my = (float)(ipy % 2);
      SETQ(1, ipx - fly - fl2(ipx + 2 * ipy - 3 * fly), ipx + fly);
      SETQ(2, ipy - 2 * fly - fl2(ipx + 3 * ipy - 3 * fly), ipx - ipy + fly);
      SETQ(3, (int)((float)(-ipx) + my), -ipy);
      SETQ(4, ipy - fly - fl2(ipx - fly), -ipx - fly);
      SETQ(5, ipx + ipy - fl2(ipx + ipy - fly), (int)((float)(-ipx + ipy / 2) + my / 2.0f));
      ne = 6; break;
    case 9:
      my = (float)(ipy % 2) / 2.0f;
      SETQ(1, fly + fl2(ipx + ipy - fly), (int)((float)(ipx - ipy / 2) - my));
      SETQ(2, ipx - fly - fl2(ipx + 2 * ipy - 3 * fly), ipx + fly);
      SETQ(3, ipx, -ipy);
      SETQ(4, ipy - 2 * fly - fl2(ipx + 3 * ipy - 3 * fly), ipx - ipy + fly);
      SETQ(5, ipx - fl2(ipx + 2 * ipy - fly), -ipx - fly);
      SETQ(6, (int)((float)(-ipx) + 2.0f * my), -ipy);
      SETQ(7, -fly - fl2(ipx + ipy - fly), (int)((float)(-ipx + ipy / 2) + my));
      SETQ(8, ipy - fly - fl2(ipx - fly), -ipx - fly);
      SETQ(9, (int)((float)(-ipx) + 2.0f * my), ipy);
      SETQ(10, ipx + ipy - fl2(ipx + ipy - fly), (int)(((float)(-2 * ipx + ipy) + 2.0f * my) / 2.0f));
      SETQ(11, ipy - 2 * fly - fl2(ipx - 3 * fly), ipx + fly);
      ne = 12; break;
    case 12:
      my = (float)(ipy % 2);
      SETQ(1, ipx + fly - fl2(ipx - ipy + fly), ipx - ipy + fly);
      SETQ(2, ipy - 2 * fly - fl2(ipx + 3 * ipy - 3 * fly), ipx - ipy + fly);
      SETQ(3, ipx - fl2(ipx + 2 * ipy - fly), -ipx - fly);
      SETQ(4, ipy - fly - fl2(ipx - fly), -ipx - fly);
      SETQ(5, (int)((float)(-ipx) + my), ipy);
      ne = 6; break;
    default: break;
  }
#undef SETQ
  return ne;
}

}  // namespace ctem
