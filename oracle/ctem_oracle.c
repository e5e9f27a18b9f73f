/* CPU restatement of CTEMsoft's EBSD master-pattern beam loop: the oracle.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, the package's smoke check and bench.py's CPU
 * baseline load this file's library, and only as the checker. The product (the HIP library
 * under ctemsoft_amd/) never links or calls it.
 *
 * Float32 steps are written with float variables in the reference's operation order so that
 * results match the reference bit for bit where it computes in single precision; this file is
 * compiled with -ffp-contract=off so that no multiply-add is fused. acosf/cosf/cos/sin/exp
 * are the C library's, which is what the reference's compiled code calls.
 */
#include "ctem_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define CPI 3.141592653589793238 /* constants.f90:70 */

/* CalcDotSingle 'r' (crystal.f90:411-429): dot_product(p, matmul(rmt, q)) with the double
 * metric, result stored in single precision */
static float dot_r_single(const float p[3], const float q[3], const double rmt[9]) {
  double v[3];
  for (int i = 0; i < 3; i++) {
    double s = 0.0;
    for (int j = 0; j < 3; j++) s += rmt[i * 3 + j] * (double)q[j];
    v[i] = s;
  }
  double c = 0.0;
  for (int i = 0; i < 3; i++) c += (double)p[i] * v[i];
  return (float)c;
}

/* CalcLengthSingle 'r' (crystal.f90:564-577) */
static float len_r_single(const float p[3], const double rmt[9]) {
  return sqrtf(dot_r_single(p, p, rmt));
}

/* CalcAngleSingle 'r' (crystal.f90:632-666) */
static float angle_r_single(const float p[3], const float q[3], const double rmt[9]) {
  float x = dot_r_single(p, q, rmt);
  float y = len_r_single(p, rmt);
  float z = len_r_single(q, rmt);
  float yz = y * z;
  float t = x / yz;
  if (t >= 1.0f) return 0.0f;
  if (t <= -1.0f) return (float)CPI;
  return acosf(t);
}

/* CalcsgSingle (diffraction.f90:649-674) */
float orc_calcsg(const float g[3], const float k[3], const float fn[3], const double rmt[9]) {
  float kpg[3], tkpg[3];
  for (int i = 0; i < 3; i++) {
    kpg[i] = k[i] + g[i];
    float t2 = 2.0f * k[i];
    tkpg[i] = t2 + g[i];
  }
  float xnom = -dot_r_single(g, tkpg, rmt);
  float q1 = len_r_single(kpg, rmt);
  float q2 = angle_r_single(kpg, fn, rmt);
  float t = 2.0f * q1;
  float xden = t * cosf(q2);
  return xnom / xden;
}

/* Compute_DynMat 'BLOCHBETHE' beam selection and ordering (gvectors.f90:988-1120):
 * |s_g| against weakcutoff*lambda|U_g| (strong) and cutoff*lambda|U_g| (weak), with the
 * sg cutoff for double-diffraction reflections; 000 is always strong and first. */
int orc_select_beams_fn(int nref, const int32_t* hkl, const double* ucg, const uint8_t* dbdiff,
                        double mlambda, float weakcutoff, float cutoff, float sgcutoff,
                        const double rmt[9], const double kvec[3], const float* fnv, int32_t* strong,
                        float* strongsg, int32_t* nns, int32_t* weak, float* weaksg,
                        int32_t* nnw);

int orc_select_beams(int nref, const int32_t* hkl, const double* ucg, const uint8_t* dbdiff,
                     double mlambda, float weakcutoff, float cutoff, float sgcutoff,
                     const double rmt[9], const double kvec[3], int32_t* strong,
                     float* strongsg, int32_t* nns, int32_t* weak, float* weaksg,
                     int32_t* nnw) {
  return orc_select_beams_fn(nref, hkl, ucg, dbdiff, mlambda, weakcutoff, cutoff, sgcutoff, rmt, kvec, NULL,
                             strong, strongsg, nns, weak, weaksg, nnw);
}

/* the same with a fixed foil normal fnv (DynFN of the channeling-pattern program,
 * CTEMECP.f90:204-205; gvectors.f90:1014); fnv = NULL: DynFN = k of each direction
 * (CTEMEBSDmaster.f90:414) */
int orc_select_beams_fn(int nref, const int32_t* hkl, const double* ucg, const uint8_t* dbdiff,
                        double mlambda, float weakcutoff, float cutoff, float sgcutoff,
                        const double rmt[9], const double kvec[3], const float* fnv, int32_t* strong,
                        float* strongsg, int32_t* nns, int32_t* weak, float* weaksg,
                        int32_t* nnw) {
  float k[3] = {(float)kvec[0], (float)kvec[1], (float)kvec[2]}; /* sngl(kk) */
  const float* fn = fnv ? fnv : k;
  int ns = 0, nw = 0;
  strong[ns] = 0;
  strongsg[ns] = 0.0f;
  ns++;
  for (int ig = 1; ig < nref; ig++) {
    float g[3] = {(float)hkl[3 * ig], (float)hkl[3 * ig + 1], (float)hkl[3 * ig + 2]};
    float sg = orc_calcsg(g, k, fn, rmt);
    float sgp = fabsf(sg);
    float lug = (float)(cabs(ucg[2 * ig] + I * ucg[2 * ig + 1]) * mlambda);
    float cut1 = cutoff * lug;
    float cut2 = weakcutoff * lug;
    if (dbdiff[ig]) {
      if (sgp <= sgcutoff) {
        strong[ns] = ig;
        strongsg[ns] = sg;
        ns++;
      }
    } else if (sgp <= cut1) {
      if (sgp <= cut2) {
        strong[ns] = ig;
        strongsg[ns] = sg;
        ns++;
      } else {
        weak[nw] = ig;
        weaksg[nw] = sg;
        nw++;
      }
    }
  }
  *nns = ns;
  *nnw = nw;
  return 0;
}

/* LUT(h,k,l) at the column-major offset of its Fortran bounds, with no per-axis check, as the
 * compiled reference addresses it (gvectors.f90:1133-1160): for a hexagonal list whose beams
 * span more than the LUT box, an offset inside the array reads the aliased entry. An offset
 * outside the array (where the reference reads neighbouring heap memory) gives NaN. */
static inline double complex lut_at(const double* lut, const int32_t lo[3], const int32_t dim[3],
                                    int h, int k, int l) {
  const long i = (long)(h - lo[0]) + (long)dim[0] * ((long)(k - lo[1]) + (long)dim[1] * (long)(l - lo[2]));
  if (i < 0 || i >= (long)dim[0] * dim[1] * dim[2]) return NAN;
  return lut[2 * i] + I * lut[2 * i + 1];
}

/* Bethe-corrected DynMat (gvectors.f90:1123-1173) */
int orc_dynmat(int nns, const int32_t* strong, const float* strongsg, int nnw,
               const int32_t* weak, const float* weaksg, const int32_t* hkl,
               const double* lut, const int32_t lo[3], const int32_t dim[3], double mlambda,
               float dynupz, double* dynmat) {
  for (int ir = 0; ir < nns; ir++) {
    const int32_t* gr = hkl + 3 * strong[ir];
    for (int ic = 0; ic < nns; ic++) {
      const int32_t* gc = hkl + 3 * strong[ic];
      double complex v;
      if (ic != ir) {
        v = lut_at(lut, lo, dim, gr[0] - gc[0], gr[1] - gc[1], gr[2] - gc[2]);
        double wr = 0.0, wi = 0.0;
        for (int iw = 0; iw < nnw; iw++) {
          const int32_t* gw = hkl + 3 * weak[iw];
          double complex a = lut_at(lut, lo, dim, gr[0] - gw[0], gr[1] - gw[1], gr[2] - gw[2]);
          double complex b = lut_at(lut, lo, dim, gw[0] - gc[0], gw[1] - gc[1], gw[2] - gc[2]);
          double pr = creal(a) * creal(b) - cimag(a) * cimag(b);
          double pi = creal(a) * cimag(b) + cimag(a) * creal(b);
          double inv = 1.0 / (double)weaksg[iw];
          wr += pr * inv;
          wi += pi * inv;
        }
        double c = 0.5 * mlambda;
        v = (creal(v) - c * wr) + I * (cimag(v) - c * wi);
      } else {
        double s = 0.0;
        for (int iw = 0; iw < nnw; iw++) {
          const int32_t* gw = hkl + 3 * weak[iw];
          double complex a = lut_at(lut, lo, dim, gr[0] - gw[0], gr[1] - gw[1], gr[2] - gw[2]);
          double m = cabs(a);
          s += m * m / (double)weaksg[iw];
        }
        s = s * mlambda / 2.0;
        v = (2.0 * (double)strongsg[ir] / mlambda - s) + I * (double)dynupz;
      }
      dynmat[2 * (ir * nns + ic)] = creal(v);
      dynmat[2 * (ir * nns + ic) + 1] = cimag(v);
    }
  }
  return 0;
}

/* Sgh for one site (CTEMEBSDmaster.f90:418-449): sum over the orbit of
 * cmplx(Z^2 DBWF) * cmplx(cos, sin), each factor and the product in float32, summed in
 * double. No occupancy factor, as in the reference. */
int orc_sgh(int nns, const int32_t* strong, const int32_t* hkl, const double rmt[9], int32_t z,
            float b_nm2, int norbit, const double* orbit, double* sgh) {
  const double tpi = 2.0 * CPI;
  float zf = (float)z;
  double znsq = (double)(zf * zf);
  for (int ir = 0; ir < nns; ir++) {
    for (int ic = 0; ic < nns; ic++) {
      int kk[3];
      for (int t = 0; t < 3; t++) kk[t] = hkl[3 * strong[ir] + t] - hkl[3 * strong[ic] + t];
      float kf[3] = {(float)kk[0], (float)kk[1], (float)kk[2]};
      float L = len_r_single(kf, rmt);
      float L2 = L * L;
      double kkl = (double)(0.25f * L2);
      double sr = 0.0, si = 0.0;
      for (int a = 0; a < norbit; a++) {
        double s = 0.0;
        for (int t = 0; t < 3; t++) s += (double)kk[t] * orbit[3 * a + t];
        double arg = tpi * s;
        double dbwf = exp(-((double)b_nm2 * kkl));
        float pre = (float)(znsq * dbwf);
        float c = (float)cos(arg), sn = (float)sin(arg);
        float re = pre * c;
        float im = pre * sn;
        sr += (double)re;
        si += (double)im;
      }
      sgh[2 * (ir * nns + ic)] = sr;
      sgh[2 * (ir * nns + ic) + 1] = si;
    }
  }
  return 0;
}

/* The channeling structure matrix of the channeling-pattern program (CTEMECP.f90:318-350), all
 * sites in one matrix: per site Znsq = float(Z)**2 * occ, kkl = 0.25 CalcLength(float(g-h))**2
 * and DBWF = Znsq * exp(-B kkl), all float32 (expf); per orbit position the float32 product
 * cmplx(DBWF, 0) * cmplx(cos(arg), sin(arg)) with arg = fl32(2 pi) * sum((g-h) * r) in double,
 * added in double site after site. sgh is row-major [ir][ic]. */
int orc_sgh_ecp(int nns, const int32_t* strong, const int32_t* hkl, const double rmt[9], int nsite,
                const int32_t* z, const float* b_nm2, const float* occ, const int32_t* norbit,
                const double* orbit, double* sgh) {
  const float tpi = (float)(2.0 * CPI);
  for (int ir = 0; ir < nns; ir++) {
    for (int ic = 0; ic < nns; ic++) {
      int kk[3];
      for (int t = 0; t < 3; t++) kk[t] = hkl[3 * strong[ir] + t] - hkl[3 * strong[ic] + t];
      float kf[3] = {(float)kk[0], (float)kk[1], (float)kk[2]};
      float L = len_r_single(kf, rmt);
      float kkl = 0.25f * (L * L);
      double sr = 0.0, si = 0.0;
      const double* orb = orbit;
      for (int p = 0; p < nsite; p++) {
        float zf = (float)z[p];
        float znsq = (zf * zf) * occ[p];
        float dbwf = znsq * expf(-b_nm2[p] * kkl);
        for (int a = 0; a < norbit[p]; a++) {
          double sd = 0.0;
          for (int t = 0; t < 3; t++) sd += (double)kk[t] * orb[3 * a + t];
          double arg = (double)tpi * sd;
          float c = (float)cos(arg), sn = (float)sin(arg);
          sr += (double)(dbwf * c);
          si += (double)(dbwf * sn);
        }
        orb += 3 * norbit[p];
      }
      sgh[2 * (ir * nns + ic)] = sr;
      sgh[2 * (ir * nns + ic) + 1] = si;
    }
  }
  return 0;
}

/* ---------------- extended-precision Lgh (scaling and squaring) ---------------------- */
typedef long double complex lcx;

static void lmatmul(int n, const lcx* a, const lcx* b, lcx* c) {
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++) c[i * n + j] = 0;
    for (int k = 0; k < n; k++) {
      lcx aik = a[i * n + k];
      if (aik == 0) continue;
      const lcx* bk = b + k * n;
      lcx* ci = c + i * n;
      for (int j = 0; j < n; j++) ci[j] += aik * bk[j];
    }
  }
}

int orc_lgh(int n, const double* dynmat, double kn, double dz, double thick, int nz,
            const float* lam, double* lgh) {
  size_t nn = (size_t)n * n;
  lcx* A = malloc(nn * sizeof(lcx));
  lcx* E = malloc(nn * sizeof(lcx));
  lcx* T = malloc(nn * sizeof(lcx));
  lcx* L = calloc(nn, sizeof(lcx));
  lcx* psi = malloc(n * sizeof(lcx));
  lcx* tmp = malloc(n * sizeof(lcx));
  if (!A || !E || !T || !L || !psi || !tmp) return -1;
  /* A = 2 pi i dz DynMat / fl32(2 kn)  (CTEMEBSDmaster.f90:693,703-716) */
  long double d2kn = (long double)(float)(2.0 * kn);
  lcx pre = (2.0L * 3.141592653589793238462643383279L * (long double)dz / d2kn) * I;
  long double norm1 = 0;
  for (int j = 0; j < n; j++) {
    long double cs = 0;
    for (int i = 0; i < n; i++) {
      A[i * n + j] = pre * ((long double)dynmat[2 * (i * n + j)] + I * (long double)dynmat[2 * (i * n + j) + 1]);
      cs += cabsl(A[i * n + j]);
    }
    if (cs > norm1) norm1 = cs;
  }
  int s = 0;
  while (norm1 > 0.25L) {
    norm1 *= 0.5L;
    s++;
  }
  long double sc = ldexpl(1.0L, -s);
  for (size_t i = 0; i < nn; i++) A[i] *= sc;
  /* Taylor series, Horner form, 18 terms: E = I + A(I + A/2(I + A/3(...))) */
  for (size_t i = 0; i < nn; i++) E[i] = 0;
  for (int i = 0; i < n; i++) E[i * n + i] = 1;
  for (int m = 18; m >= 1; m--) {
    lmatmul(n, A, E, T);
    for (size_t i = 0; i < nn; i++) E[i] = T[i] / (long double)m;
    for (int i = 0; i < n; i++) E[i * n + i] += 1;
  }
  for (int q = 0; q < s; q++) {
    lmatmul(n, E, E, T);
    memcpy(E, T, nn * sizeof(lcx));
  }
  /* depth recursion psi_{k+1} = E psi_k, psi_0 = e_1; accumulate lam_k psi_k psi_k^H */
  for (int i = 0; i < n; i++) psi[i] = (i == 0) ? 1 : 0;
  for (int k = 0; k < nz; k++) {
    for (int i = 0; i < n; i++) {
      lcx acc = 0;
      for (int j = 0; j < n; j++) acc += E[i * n + j] * psi[j];
      tmp[i] = acc;
    }
    memcpy(psi, tmp, n * sizeof(lcx));
    long double w = (long double)lam[k];
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) L[i * n + j] += w * psi[i] * conjl(psi[j]);
  }
  long double dzt = (long double)dz / (long double)thick;
  for (size_t i = 0; i < nn; i++) {
    lcx v = L[i] * dzt;
    lgh[2 * i] = (double)creall(v);
    lgh[2 * i + 1] = (double)cimagl(v);
  }
  free(A); free(E); free(T); free(L); free(psi); free(tmp);
  return 0;
}

double orc_lgh_dot_sgh(int n, const double* lgh, const double* sgh) {
  double s = 0.0;
  for (size_t i = 0; i < (size_t)n * n; i++) s += lgh[2 * i] * sgh[2 * i] - lgh[2 * i + 1] * sgh[2 * i + 1];
  return s;
}

/* ---------------- RoscaLambert incident directions ----------------------------------- */
static int get_sextant(double x, double y) { /* kvectors.f90:781-817 */
  const double srt = (double)1.732050808f; /* a default-real literal in the reference */
  double xx = fabs(x * srt);
  if (y >= 0) {
    if (y >= xx) return 0;
    return (x > 0.0) ? 1 : 5;
  }
  if (fabs(y) >= xx) return 3;
  return (x > 0.0) ? 2 : 4;
}

typedef struct {
  int cap, n;
  int32_t* ij;
  double* k;
  double* kn;
  const double* dsm;
  double mlambda;
} klist;

static void push_k(klist* L, int i, int j, const double kstar_in[3]) {
  if (L->n >= L->cap) {
    L->n++;
    return;
  }
  double ks[3] = {kstar_in[0], kstar_in[1], kstar_in[2]};
  double x = sqrt(ks[0] * ks[0] + ks[1] * ks[1] + ks[2] * ks[2]); /* NormVec 'c' */
  if (x != 0.0)
    for (int t = 0; t < 3; t++) ks[t] = ks[t] / x;
  for (int t = 0; t < 3; t++) ks[t] = ks[t] / L->mlambda;
  for (int a = 0; a < 3; a++) { /* matmul(transpose(dsm), kstar) */
    double s = 0.0;
    for (int b = 0; b < 3; b++) s += L->dsm[b * 3 + a] * ks[b];
    L->k[3 * L->n + a] = s;
  }
  L->ij[2 * L->n] = i;
  L->ij[2 * L->n + 1] = j;
  L->kn[L->n] = 1.0 / L->mlambda;
  L->n++;
}

/* AddkVector (kvectors.f90:839-947) */
static void add_k(klist* L, double delta, int i, int j, int usehex) {
  const double srt = 0.86602540, rtt = 1.7320508076, prea = 0.525037568, preb = 1.050075136,
               prec = 0.90689968, pred = 2.09439510;
  const double ipi = 1.0 / CPI;
  double x, y;
  if (usehex) {
    float jh = (float)j * 0.5f;
    x = (double)((float)i - jh) * delta;
    y = (double)j * delta * srt;
  } else {
    x = i * delta;
    y = j * delta;
  }
  if (i == 0 && j == 0) return;
  double ks[3];
  if (usehex) {
    double XX = 0, YY = 0, xp, yp;
    switch (get_sextant(x, y)) {
      case 0: case 3:
        XX = preb * y * cos(x * prec / y);
        YY = preb * y * sin(x * prec / y);
        break;
      case 1: case 4:
        xp = y + rtt * x;
        yp = y * pred / xp;
        XX = prea * xp * sin(yp);
        YY = prea * xp * cos(yp);
        break;
      default:
        xp = y - rtt * x;
        yp = y * pred / xp;
        XX = prea * xp * sin(yp);
        YY = -prea * xp * cos(yp);
        break;
    }
    double q = XX * XX + YY * YY;
    ks[0] = 0.5 * XX * sqrt(4.0 - q);
    ks[1] = 0.5 * YY * sqrt(4.0 - q);
    ks[2] = 1.0 - 0.5 * q;
    /* hexagonal -> square-array index; integer division as in the reference (:934) */
    push_k(L, i - j / 2 + (j % 2) / 2, j, ks);
  } else {
    if (fabs(x) <= fabs(y)) {
      double q = 2.0 * y * ipi * sqrt(CPI - y * y);
      double ang = x * CPI * 0.25 / y;
      ks[0] = q * sin(ang);
      ks[1] = q * cos(ang);
      ks[2] = 1.0 - 2.0 * y * y * ipi;
    } else {
      double q = 2.0 * x * ipi * sqrt(CPI - x * x);
      double ang = y * CPI * 0.25 / x;
      ks[0] = q * cos(ang);
      ks[1] = q * sin(ang);
      ks[2] = 1.0 - 2.0 * x * x * ipi;
    }
    push_k(L, i, j, ks);
  }
}

/* Calckvectors 'RoscaLambert' (kvectors.f90:421-542) */
int orc_kvectors(int npx, int npy, int isym, int usehex, const double dsm[9], double mlambda,
                 int cap, int32_t* ij, double* k, double* kn) {
  klist L = {cap, 0, ij, k, kn, dsm, mlambda};
  double delta = usehex ? 2.0 * sqrt(CPI) / pow(3.0, 0.75) / (double)npx
                        : sqrt(CPI * 0.5) / (double)npx;
  double head[3] = {0.0, 0.0, 1.0};
  push_k(&L, 0, 0, head);
  switch (isym) {
    case 1:
      for (int j = -npy; j <= npy; j++)
        for (int i = -npx; i <= npx; i++) add_k(&L, delta, i, j, usehex);
      break;
    case 2:
      for (int j = 0; j <= npy; j++)
        for (int i = -npx; i <= npx; i++) add_k(&L, delta, i, j, usehex);
      break;
    case 3: case 4: case 10:
      for (int j = 0; j <= npy; j++)
        for (int i = 0; i <= npx; i++) add_k(&L, delta, i, j, usehex);
      break;
    case 5: case 11:
      for (int i = 0; i <= npx; i++)
        for (int j = 0; j <= i; j++) add_k(&L, delta, i, j, usehex);
      break;
    case 6: case 7:
      for (int j = 0; j <= npy; j++)
        for (int i = 0; i <= npy; i++) add_k(&L, delta, i, j, usehex);
      break;
    case 8:
      for (int j = 0; j <= npy; j++)
        for (int i = j; i <= npy; i++) add_k(&L, delta, i, j, usehex);
      break;
    case 9:
      for (int j = 0; j <= npx; j++)
        for (int i = 2 * j; i <= npx; i++) add_k(&L, delta, i, j, usehex);
      break;
    case 12:
      for (int j = 0; j <= npy / 2; j++)
        for (int i = 2 * j; i <= npy; i++) {
          add_k(&L, delta, i, j, usehex);
          add_k(&L, delta, i - j, -j, usehex);
        }
      break;
    default:
      return -2;
  }
  return (L.n > cap) ? -1 : L.n;
}

/* Apply2DLaueSymmetry (Lambert.f90:1942-2046) */
int orc_laue_equiv(int ipx, int ipy, int isym, int32_t* q) {
#define SET(n, a, b) do { q[2 * (n)] = (a); q[2 * (n) + 1] = (b); } while (0)
  memset(q, 0, 24 * sizeof(int32_t));
  SET(0, ipx, ipy);
  int ne = 1;
  if (ipx == 0 && ipy == 0) return 1;
  int fly = (int)floorf((float)ipy / 2.0f);
#define FL(v) ((int)floorf((float)(v) / 2.0f))
  float my;
  switch (isym) {
    case 2: SET(1, ipx, -ipy); ne = 2; break;
    case 4: SET(1, -ipy, ipx); SET(2, ipy, -ipx); SET(3, -ipx, -ipy); ne = 4; break;
    case 3: case 10: SET(1, -ipx, ipy); SET(2, ipx, -ipy); SET(3, -ipx, -ipy); ne = 4; break;
    case 5: case 11:
      SET(1, -ipx, ipy); SET(2, ipx, -ipy); SET(3, -ipx, -ipy); SET(4, ipy, ipx);
      SET(5, -ipy, ipx); SET(6, ipy, -ipx); SET(7, -ipy, -ipx); ne = 8; break;
    case 6: case 7:
      SET(1, ipy - 2 * fly - FL(ipx + 3 * ipy - 3 * fly), ipx - ipy + fly);
      SET(2, ipy - fly - FL(ipx - fly), -ipx - fly);
      ne = 3; break;
    case 13:
      my = (float)(ipy % 2) / 2.0f;
      SET(1, ipy - FL(ipx + fly), (int)((float)(ipx + ipy / 2) - my));
      SET(2, ipy - 2 * fly - FL(ipx + 3 * ipy - 3 * fly), ipx - ipy + fly);
      SET(3, ipx, -ipy);
      SET(4, ipy - fly - FL(ipx - fly), -ipx - fly);
      SET(5, -fly - FL(ipx + ipy - fly), (int)((float)(-ipx + ipy / 2) + my));
      ne = 6; break;
    case 8:
      my = (float)(ipy % 2);
      SET(1, ipx - fly - FL(ipx + 2 * ipy - 3 * fly), ipx + fly);
      SET(2, ipy - 2 * fly - FL(ipx + 3 * ipy - 3 * fly), ipx - ipy + fly);
      SET(3, (int)((float)(-ipx) + my), -ipy);
      SET(4, ipy - fly - FL(ipx - fly), -ipx - fly);
      SET(5, ipx + ipy - FL(ipx + ipy - fly), (int)((float)(-ipx + ipy / 2) + my / 2.0f));
      ne = 6; break;
    case 9:
      my = (float)(ipy % 2) / 2.0f;
      SET(1, fly + FL(ipx + ipy - fly), (int)((float)(ipx - ipy / 2) - my));
      SET(2, ipx - fly - FL(ipx + 2 * ipy - 3 * fly), ipx + fly);
      SET(3, ipx, -ipy);
      SET(4, ipy - 2 * fly - FL(ipx + 3 * ipy - 3 * fly), ipx - ipy + fly);
      SET(5, ipx - FL(ipx + 2 * ipy - fly), -ipx - fly);
      SET(6, (int)((float)(-ipx) + 2.0f * my), -ipy);
      SET(7, -fly - FL(ipx + ipy - fly), (int)((float)(-ipx + ipy / 2) + my));
      SET(8, ipy - fly - FL(ipx - fly), -ipx - fly);
      SET(9, (int)((float)(-ipx) + 2.0f * my), ipy);
      SET(10, ipx + ipy - FL(ipx + ipy - fly), (int)(((float)(-2 * ipx + ipy) + 2.0f * my) / 2.0f));
      SET(11, ipy - 2 * fly - FL(ipx - 3 * fly), ipx + fly);
      ne = 12; break;
    case 12:
      my = (float)(ipy % 2);
      SET(1, ipx + fly - FL(ipx - ipy + fly), ipx - ipy + fly);
      SET(2, ipy - 2 * fly - FL(ipx + 3 * ipy - 3 * fly), ipx - ipy + fly);
      SET(3, ipx - FL(ipx + 2 * ipy - fly), -ipx - fly);
      SET(4, ipy - fly - FL(ipx - fly), -ipx - fly);
      SET(5, (int)((float)(-ipx) + my), ipy);
      ne = 6; break;
    default: break;
  }
#undef SET
#undef FL
  return ne;
}

/* ---- EBSD detector patterns (CTEMEBSD.f90) ---------------------------------------------- */

/* InitLambertParameters (Lambert.f90:231-262): sqrt(pi/2) and sqrt(pi)/2 with pi = 4 atan(1) */
static double lps_spio2(void) { return sqrt(4.0 * atan(1.0) * 0.5); }
static double lps_spi2(void) { return sqrt(4.0 * atan(1.0)) * 0.5; }

/* Lambert2DSquareInverseSingle (Lambert.f90:405-433) in float32, times scl (the caller's
 * ixy = scl * LambertSphereToSquare(dc)) */
int orc_lambert_inverse_single(const float xyz[3], float scl, float xy[2]) {
  const float spi2 = (float)lps_spi2();
  float r0, r1;
  int ierr = 0;
  const float ss = xyz[0] * xyz[0] + xyz[1] * xyz[1] + xyz[2] * xyz[2];
  if (fabsf(1.0f - ss) > 1.0e-6f) {
    r0 = r1 = 0.0f;
    ierr = 1;
  } else if (fabsf(xyz[2]) == 1.0f) {
    r0 = r1 = 0.0f;
  } else if (fabsf(xyz[1]) <= fabsf(xyz[0])) {
    const float q = fabsf(xyz[0]) / xyz[0] * sqrtf(2.0f * (1.0f - xyz[2]));
    r0 = q * spi2;
    r1 = q * atanf(xyz[1] / xyz[0]) / spi2;
  } else {
    const float q = fabsf(xyz[1]) / xyz[1] * sqrtf(2.0f * (1.0f - xyz[2]));
    r0 = q * atanf(xyz[0] / xyz[1]) / spi2;
    r1 = q * spi2;
  }
  xy[0] = scl * r0;
  xy[1] = scl * r1;
  return ierr;
}

/* element (k, ix, iy) of a Fortran array a(nk, -nx:nx, -ny:ny) at its column-major offset
 * without a bounds check, as the compiled program reads it (an index one past a row reads the
 * next row); offsets outside the array give 0 */
static float at3_i(const int32_t* a, int nk, int nx, int ny, int k, int ix, int iy) {
  const int64_t o = (int64_t)k + (int64_t)nk * ((ix + nx) + (int64_t)(2 * nx + 1) * (iy + ny));
  const int64_t tot = (int64_t)nk * (2 * nx + 1) * (2 * ny + 1);
  return (o >= 0 && o < tot) ? (float)a[o] : 0.0f;
}
static float at3_f(const float* a, int nx, int ny, int nk, int ix, int iy, int k) {
  const int64_t o = (int64_t)(ix + nx) + (int64_t)(2 * nx + 1) * ((iy + ny) + (int64_t)(2 * ny + 1) * k);
  const int64_t tot = (int64_t)nk * (2 * nx + 1) * (2 * ny + 1);
  return (o >= 0 && o < tot) ? a[o] : 0.0f;
}

int orc_ebsd_detector(float L, float thetac, float delta, int numsx, int numsy, float xpc, float ypc,
                      float sig, int numEbins, int nsx, int nsy, const int32_t* accum_e, int Emin,
                      int Emax, float* rg, float* wdet, int64_t* num_el) {
  const float dtor = 0.0174533f; /* CTEMEBSD.f90:148 */
  const int64_t np = (int64_t)numsx * numsy;
  float* sx_ = (float*)malloc(sizeof(float) * numsx);
  float* sy_ = (float*)malloc(sizeof(float) * numsy);
  if (!sx_ || !sy_) return -1;
  /* :414-415 */
  for (int i = 0; i < numsx; i++) sx_[i] = -((xpc - (1.0f - (float)numsx) * 0.5f) - (float)i) * delta;
  for (int j = 0; j < numsy; j++) sy_[j] = ((ypc - (1.0f - (float)numsy) * 0.5f) - (float)j) * delta;
  /* :421-423 */
  const float alp = (float)(0.5 * CPI - (double)((sig - thetac) * dtor));
  const float ca = cosf(alp), sa = sinf(alp);
  /* :431-442 */
  const float L2 = L * L;
  float* rgx = rg;
  float* rgy = rg + np;
  float* rgz = rg + 2 * np;
  for (int j = 0; j < numsy; j++) {
    const float sx = L2 + sy_[j] * sy_[j];
    const float Ls = ca * sy_[j] + L * sa;
    const float Lc = -sa * sy_[j] + L * ca;
    for (int i = 0; i < numsx; i++) {
      const float rhos = 1.0f / sqrtf(sx + sx_[i] * sx_[i]);
      const int64_t p = (int64_t)j * numsx + i;
      rgx[p] = Ls * rhos;
      rgy[p] = sx_[i] * rhos;
      rgz[p] = Lc * rhos;
    }
  }
  /* :445-450 */
  for (int64_t p = 0; p < np; p++) {
    const float z = sqrtf(rgx[p] * rgx[p] + rgy[p] * rgy[p] + rgz[p] * rgz[p]);
    rgx[p] = rgx[p] / z;
    rgy[p] = rgy[p] / z;
    rgz[p] = rgz[p] / z;
  }
  /* :469-502 */
  const float scl = (float)((double)(float)nsx / lps_spio2());
  memset(wdet, 0, sizeof(float) * (size_t)np * numEbins);
  for (int j = 0; j < numsy; j++)
    for (int i = 0; i < numsx; i++) {
      const int64_t p = (int64_t)j * numsx + i;
      float dc[3] = {rgx[p], rgy[p], rgz[p]};
      if (dc[2] < 0.0f) { dc[0] = -dc[0]; dc[1] = -dc[1]; dc[2] = -dc[2]; }
      float ixy[2];
      orc_lambert_inverse_single(dc, scl, ixy);
      const float x = ixy[0];
      ixy[0] = ixy[1];
      ixy[1] = -x;
      const int nix = (int)((float)nsx + ixy[0]) - nsx;
      const int niy = (int)((float)nsy + ixy[1]) - nsy;
      const float dx = ixy[0] - (float)nix, dy = ixy[1] - (float)niy;
      const float dxm = 1.0f - dx, dym = 1.0f - dy;
      for (int k = Emin - 1; k < Emax; k++)
        wdet[p * numEbins + k] = at3_i(accum_e, numEbins, nsx, nsy, k, nix, niy) * dxm * dym +
                                 at3_i(accum_e, numEbins, nsx, nsy, k, nix + 1, niy) * dx * dym +
                                 at3_i(accum_e, numEbins, nsx, nsy, k, nix, niy + 1) * dxm * dy +
                                 at3_i(accum_e, numEbins, nsx, nsy, k, nix + 1, niy + 1) * dx * dy;
    }
  /* :503, :508: num_el = nint(sum(accum_e_detector)) over the whole array (every bin), as the
   * Fortran runtime's SUM of a REAL(4) array evaluates it: compensated (Kahan) summation in single
   * precision, in array element order; nint rounds half away from zero */
  for (int64_t e = 0; e < np * numEbins; e++) wdet[e] = wdet[e] * 0.25f;
  float tot = 0.0f, comp = 0.0f;
  for (int64_t e = 0; e < np * numEbins; e++) {
    const float y = wdet[e] - comp;
    const float t = tot + y;
    comp = (t - tot) - y;
    tot = t;
  }
  *num_el = (int64_t)lroundf(tot);
  free(sx_);
  free(sy_);
  return 0;
}

/* SUM of a REAL(4) array as the compiled program evaluates it (sum(dc**2), CTEMEBSD.f90:540):
 * the Fortran runtime's compensated (Kahan) summation in single precision, in element order */
static float sum3_kahan(float a, float b, float c) {
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

int orc_ebsd_pattern(int numsx, int numsy, const float* rg, const float* wdet, int numEbins, int Emin,
                     int Emax, int npx, int npy, const float* sr, const float q[4], double prefactor,
                     float* out) {
  const int64_t np = (int64_t)numsx * numsy;
  const float scl = (float)((double)(float)npx / lps_spio2());
  const float q1[4] = {q[0], -q[1], -q[2], -q[3]}; /* conjg (quaternions.f90:306-313) */
  for (int j = 0; j < numsy; j++)
    for (int i = 0; i < numsx; i++) {
      const int64_t p = (int64_t)j * numsx + i;
      const float qq[4] = {0.0f, rg[p], rg[np + p], rg[2 * np + p]};
      float t[4], r[4];
      /* quat_mult (quaternions.f90:255-265): t = qq * conjg(q), r = q * t */
      t[0] = qq[0] * q1[0] - qq[1] * q1[1] - qq[2] * q1[2] - qq[3] * q1[3];
      t[1] = qq[0] * q1[1] + qq[1] * q1[0] + qq[2] * q1[3] - qq[3] * q1[2];
      t[2] = qq[0] * q1[2] - qq[1] * q1[3] + qq[2] * q1[0] + qq[3] * q1[1];
      t[3] = qq[0] * q1[3] + qq[1] * q1[2] - qq[2] * q1[1] + qq[3] * q1[0];
      r[0] = q[0] * t[0] - q[1] * t[1] - q[2] * t[2] - q[3] * t[3];
      r[1] = q[0] * t[1] + q[1] * t[0] + q[2] * t[3] - q[3] * t[2];
      r[2] = q[0] * t[2] - q[1] * t[3] + q[2] * t[0] + q[3] * t[1];
      r[3] = q[0] * t[3] + q[1] * t[2] - q[2] * t[1] + q[3] * t[0];
      (void)r[0];
      float dc[3] = {r[1], r[2], r[3]};
      const float nrm = sqrtf(sum3_kahan(dc[0] * dc[0], dc[1] * dc[1], dc[2] * dc[2]));
      dc[0] = dc[0] / nrm;
      dc[1] = dc[1] / nrm;
      dc[2] = dc[2] / nrm;
      if (dc[2] < 0.0f) { dc[0] = -dc[0]; dc[1] = -dc[1]; dc[2] = -dc[2]; }
      float ixy[2];
      orc_lambert_inverse_single(dc, scl, ixy);
      const int nix = (int)((float)npx + ixy[0]) - npx;
      const int niy = (int)((float)npy + ixy[1]) - npy;
      const float dx = ixy[0] - (float)nix, dy = ixy[1] - (float)niy;
      const float dxm = 1.0f - dx, dym = 1.0f - dy;
      float acc = 0.0f;
      for (int k = Emin - 1; k < Emax; k++)
        acc = acc + wdet[p * numEbins + k] * (at3_f(sr, npx, npy, numEbins, nix, niy, k) * dxm * dym +
                                              at3_f(sr, npx, npy, numEbins, nix + 1, niy, k) * dx * dym +
                                              at3_f(sr, npx, npy, numEbins, nix, niy + 1, k) * dxm * dy +
                                              at3_f(sr, npx, npy, numEbins, nix + 1, niy + 1, k) * dx * dy);
      out[p] = (float)(prefactor * (double)acc);
    }
  return 0;
}

/* noise.f90: ran (:136-165) and POIDEV (:167-228) with their saved variables */
struct orc_noise {
  int32_t idum, ix, iy, started;
  float am, oldm, g, sq, alxm;
};
static float orc_ran(struct orc_noise* s) {
  const int32_t IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836;
  if (!s->started) { s->ix = -1; s->iy = -1; s->started = 1; }
  if (s->idum <= 0 || s->iy < 0) {
    s->am = nextafterf(1.0f, 0.0f) / (float)IM;
    const int32_t a = s->idum < 0 ? -s->idum : s->idum;
    s->iy = (888889999 ^ a) | 1;
    s->ix = 777755555 ^ a;
    s->idum = a + 1;
  }
  uint32_t x = (uint32_t)s->ix;
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  s->ix = (int32_t)x;
  const int32_t k = s->iy / IQ;
  s->iy = IA * (s->iy - k * IQ) - IR * k;
  if (s->iy < 0) s->iy += IM;
  return s->am * (float)((IM & (s->ix ^ s->iy)) | 1);
}
/* alogam (noise.f90:47-133; Pike and Hill, CACM algorithm 291) */
static double orc_alogam(double x) {
  if (x <= 0.0) return 0.0;
  double y = x, f, z;
  if (x < 7.0) {
    f = 1.0;
    z = y;
    while (z < 7.0) {
      f = f * z;
      z = z + 1.0;
    }
    y = z;
    f = -log(f);
  } else {
    f = 0.0;
  }
  z = 1.0 / y / y;
  return f + (y - 0.5) * log(y) - y + 0.918938533204673 +
         (((-0.000595238095238 * z + 0.000793650793651) * z - 0.002777777777778) * z + 0.083333333333333) / y;
}
static float orc_poidev(float xm, struct orc_noise* s) {
  const float PI = 3.141592654f;
  float em, t, y;
  if (xm < 20.0f) {
    if (xm != s->oldm) {
      s->oldm = xm;
      s->g = expf(-xm);
    }
    em = -1.0f;
    t = 1.0f;
    for (;;) {
      em = em + 1.0f;
      t = t * orc_ran(s);
      if (t <= s->g) break;
    }
  } else {
    if (xm != s->oldm) {
      s->oldm = xm;
      s->sq = sqrtf(2.0f * xm);
      s->alxm = logf(xm);
      s->g = (float)((double)(xm * s->alxm) - orc_alogam((double)(xm + 1.0f)));
    }
    for (;;) {
      for (;;) {
        y = tanf(PI * orc_ran(s));
        em = s->sq * y + xm;
        if (em >= 0.0f) break;
      }
      em = (float)(int32_t)em;
      t = (float)((double)(0.9f * (1.0f + y * y)) *
                  exp((double)(em * s->alxm) - orc_alogam((double)(em + 1.0f)) - (double)s->g));
      if (orc_ran(s) <= t) break;
    }
  }
  return em;
}
void orc_poisson_noise(float* pattern, int numsx, int numsy, int64_t* st) {
  struct orc_noise s;
  if (st[1] == 0) { /* first use: the saved variables' initial values */
    memset(&s, 0, sizeof(s));
    s.idum = (int32_t)st[0];
    s.oldm = -1.0f;
  } else {
    memcpy(&s, &st[2], sizeof(s));
    s.idum = (int32_t)st[0];
  }
  for (int i = 0; i < numsx; i++)
    for (int j = 0; j < numsy; j++) {
      float* v = &pattern[(int64_t)j * numsx + i];
      *v = orc_poidev(*v, &s);
    }
  st[0] = s.idum;
  st[1] = 1;
  memcpy(&st[2], &s, sizeof(s));
}

/* eu2qu (rotations.f90:1322-1351), float32; e in radians */
void orc_eu2qu(const float e[3], float q[4]) {
  const float ee[3] = {0.5f * e[0], 0.5f * e[1], 0.5f * e[2]};
  const float cPhi = cosf(ee[1]), sPhi = sinf(ee[1]);
  const float cm = cosf(ee[0] - ee[2]), sm = sinf(ee[0] - ee[2]);
  const float cp = cosf(ee[0] + ee[2]), sp = sinf(ee[0] + ee[2]);
  const float qq = cPhi * cp;
  if (qq < 0.0f) {
    q[0] = -qq; q[1] = sPhi * cm; q[2] = sPhi * sm; q[3] = cPhi * sp;
  } else {
    q[0] = qq; q[1] = -sPhi * cm; q[2] = -sPhi * sm; q[3] = -cPhi * sp;
  }
}

/* axisangle rotation of the program (CTEMEBSD.f90:264-270): qax = ax2qu(a) with a(4) in degrees
 * times dtor (ax2om, om2qu: rotations.f90; quat_norm's sum in element order, quaternions.f90),
 * then q = quat_mult(qax, q) for each of the n quaternions */
void orc_apply_axisangle(int n, const float ax[4], float* quats) {
  if (ax[3] == 0.0f) return;
  const float a3 = ax[3] * 0.0174533f;
  const float c = cosf(a3), s = sinf(a3), omc = 1.0f - c;
  float m[3][3];
  for (int i = 0; i < 3; i++) m[i][i] = ax[i] * ax[i] * omc + c;
  float q = omc * ax[0] * ax[1];
  m[0][1] = q + s * ax[2];
  m[1][0] = q - s * ax[2];
  q = omc * ax[1] * ax[2];
  m[1][2] = q + s * ax[0];
  m[2][1] = q - s * ax[0];
  q = omc * ax[2] * ax[0];
  m[2][0] = q + s * ax[1];
  m[0][2] = q - s * ax[1];
  float r[4];
  r[0] = sqrtf(m[0][0] + m[1][1] + m[2][2] + 1.0f) / 2.0f;
  r[1] = sqrtf(fabsf(m[0][0] - m[1][1] - m[2][2] + 1.0f)) / 2.0f;
  r[2] = sqrtf(fabsf(-m[0][0] + m[1][1] - m[2][2] + 1.0f)) / 2.0f;
  r[3] = sqrtf(fabsf(-m[0][0] - m[1][1] + m[2][2] + 1.0f)) / 2.0f;
  if (m[2][1] - m[1][2] < 0.0f) r[1] = -r[1];
  if (m[0][2] - m[2][0] < 0.0f) r[2] = -r[2];
  if (m[1][0] - m[0][1] < 0.0f) r[3] = -r[3];
  const float sc = 1.0f / sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
  for (int i = 0; i < 4; i++) r[i] = sc * r[i];
  for (int k = 0; k < n; k++) {
    float* x = quats + 4 * k;
    const float t0 = r[0] * x[0] - r[1] * x[1] - r[2] * x[2] - r[3] * x[3];
    const float t1 = r[0] * x[1] + r[1] * x[0] + r[2] * x[3] - r[3] * x[2];
    const float t2 = r[0] * x[2] - r[1] * x[3] + r[2] * x[0] + r[3] * x[1];
    const float t3 = r[0] * x[3] + r[1] * x[2] - r[2] * x[1] + r[3] * x[0];
    x[0] = t0; x[1] = t1; x[2] = t2; x[3] = t3;
  }
}

/* Kahan sum in element order (the Fortran runtime's SUM of REAL(4) values) */
static float kahan_add(float* s, float* c, float x) {
  const float y = x - *c;
  const float t = *s + y;
  *c = (t - *s) - y;
  *s = t;
  return t;
}

/* sr = sum(srtmp, 4) (CTEMEBSD.f90:404): per element over the nset sites */
void orc_site_sum(int64_t per, int nset, const float* srtmp, float* sr) {
  for (int64_t e = 0; e < per; e++) {
    float s = 0.0f, c = 0.0f;
    for (int q = 0; q < nset; q++) kahan_add(&s, &c, srtmp[e + per * q]);
    sr[e] = s;
  }
}

/* binning and gamma (CTEMEBSD.f90:575-590); binning divides numsx and numsy */
void orc_bin_pattern(int numsx, int numsy, int binning, int gamma, float gammavalue, const float* pat,
                     float* binned) {
  const int bx = numsx / binning, by = numsy / binning;
  const float bindx = 1.0f / ((float)binning * (float)binning);
  for (int jb = 0; jb < by; jb++)
    for (int ib = 0; ib < bx; ib++) {
      float s = 0.0f, c = 0.0f;
      for (int j = jb * binning; j < (jb + 1) * binning; j++)
        for (int i = ib * binning; i < (ib + 1) * binning; i++) kahan_add(&s, &c, pat[(int64_t)j * numsx + i]);
      binned[(int64_t)jb * bx + ib] = binning == 1 ? pat[(int64_t)jb * numsx + ib] : s * bindx;
    }
  if (gamma)
    for (int64_t e = 0; e < (int64_t)bx * by; e++) binned[e] = powf(binned[e], gammavalue);
}

/* ---------------- Monte Carlo electron trajectories (CTEMMC.f90) --------------------------- */

/* the generator of rng.f95: a lag-3 subtraction modulo 2147483579 (s1 - s3) plus the LCG
 * 69069 s4 + 1013904243 (32-bit wrap), u = fl32(0.5 + 0.23283064e-9 (imz + s4)) */

void orc_mcrng_seed(orc_mcrng* r, int32_t seed) {
  r->s[0] = seed; r->s[1] = 362436069; r->s[2] = 16163801; r->s[3] = 1131199299;
}

float orc_mcrng_uniform(orc_mcrng* r) {
  int32_t imz = r->s[0] - r->s[2];
  if (imz < 0) imz += 2147483579;
  r->s[0] = r->s[1];
  r->s[1] = r->s[2];
  r->s[2] = imz;
  r->s[3] = (int32_t)(69069u * (uint32_t)r->s[3] + 1013904243u);
  imz = (int32_t)((uint32_t)imz + (uint32_t)r->s[3]);
  return (float)(0.5 + 0.23283064e-9 * (double)imz);
}

/* jump of the generator by n draws: the lag-3 recurrence x_k = x_{k-3} - x_{k-1} (mod m) as a 3x3
 * matrix power, the LCG as an affine map power */
#define ORC_MCM 2147483579LL
static void m3mul(const int64_t a[9], const int64_t b[9], int64_t c[9]) {
  int64_t t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      int64_t s = 0;
      for (int k = 0; k < 3; k++) s = (s + a[3 * i + k] * b[3 * k + j]) % ORC_MCM;
      t[3 * i + j] = s;
    }
  memcpy(c, t, sizeof t);
}
void orc_mcrng_jump(orc_mcrng* r, uint64_t n) {
  /* state vector (s1, s2, s3) = (x_{k-2}, x_{k-1}, x_k): next = (s2, s3, s1 - s3) */
  int64_t P[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, B[9] = {0, 1, 0, 0, 0, 1, 1, 0, ORC_MCM - 1};
  uint32_t A = 1, C = 0, a = 69069u, c = 1013904243u;
  while (n) {
    if (n & 1) {
      m3mul(B, P, P);
      C = a * C + c;
      A = a * A;
    }
    m3mul(B, B, B);
    c = a * c + c;
    a = a * a;
    n >>= 1;
  }
  int64_t v[3] = {r->s[0], r->s[1], r->s[2]}, w[3];
  for (int i = 0; i < 3; i++) w[i] = ((P[3 * i] * v[0]) % ORC_MCM + (P[3 * i + 1] * v[1]) % ORC_MCM + (P[3 * i + 2] * v[2]) % ORC_MCM) % ORC_MCM;
  for (int i = 0; i < 3; i++) r->s[i] = (int32_t)w[i];
  r->s[3] = (int32_t)(A * (uint32_t)r->s[3] + C);
}

static double orc_sum3_kahan_d(double a, double b, double c) {
  const double v[3] = {a, b, c};
  double s = 0.0, comp = 0.0;
  for (int i = 0; i < 3; i++) {
    const double y = v[i] - comp;
    const double t = s + y;
    comp = (t - s) - y;
    s = t;
  }
  return s;
}

int64_t orc_mc_block = 16384; /* draws per electron in the blocked mode (>= 1 + 3 * 5000) */

/* single_run (CTEMMC.f90:304-526) for num_el electrons. mode 0: one stream seeded primeseed + 1,
 * electrons in order (the program); mode 1: electron e draws from that stream's block starting at
 * e * 16384. kahan: the normalisation's SUM compensated (the Fortran runtime's). accum_e
 * [numsy][numsx][numEbins], accum_z [nz][nz][numzbins][numEbins] (energy fastest), added to. Returns
 * the number of hits that fell outside the energy bins or the depth map's pixel box (the program
 * writes past its arrays there: an energy window that is not a whole number of bins, or
 * (numsx - 1) / 2 mod 10 >= 5). */
int64_t orc_mc_run(double sig, double omega, int32_t numsx, int32_t primeseed, int64_t num_el, double EkeV,
                   double Ehistmin, double Ebinsize, double depthmax, double depthstep, double Ze, double density,
                   double at_wt, int mode, int kahan, int32_t* accum_e, int32_t* accum_z) {
  const int numEbins = (int)((EkeV - Ehistmin) / Ebinsize) + 1;
  const int numzbins = (int)(depthmax / depthstep) + 1;
  const int nx = (numsx - 1) / 2, nzx = nx / 10, nzw = 2 * nzx + 1;
  const double scaled = 1.0e8, min_energy = 1.0, presig = 1.5273987e19;
  const double cPi = 3.141592653589, cAvogadro = 6.0221415e23, cDtoR = 0.017453293;
  const double dpi = 4.0 * atan(1.0);
  const double sPio2 = sqrt(dpi * 0.5), sPi2 = sqrt(dpi) * 0.5;
  const double Emin = Ehistmin - Ebinsize / 2.0;
  const double pre = at_wt / cAvogadro / density;
  const double prealpha = 3.4e-3 * pow(Ze, (double)0.67f);
  double J = (9.76 * Ze + 58.5 / pow(Ze, 0.19)) * 1.0e-3 / (double)1.166f;
  J = 1.0 / J;
  const double predEds = -78500.0 * density * Ze / at_wt;
  const int num = 5000;
  const double tpi = 2.0 * cPi;
  const double tano = tan(omega * cDtoR);
  const double delta = (double)nx / sPio2;
  const double cxstart = cos((90.0 - sig) * cDtoR);
  const double czstart = -sin((90.0 - sig) * cDtoR);
  int64_t dropped = 0;
  orc_mcrng r0, r;
  orc_mcrng_seed(&r0, primeseed + 1);
  r = r0;
  for (int64_t el = 0; el < num_el; el++) {
    if (mode == 1) {
      r = r0;
      orc_mcrng_jump(&r, (uint64_t)el * (uint64_t)orc_mc_block);
    }
    double Ec = EkeV;
    double xyz[3] = {0.0, 0.0, 0.0};
    double alpha = prealpha / Ec;
    double step = Ec * (Ec + 1024.0) / Ze / (Ec + 511.0);
    double sige = presig * step * step * alpha * (1.0 + alpha);
    double lambda = pre * sige;
    double rr = (double)orc_mcrng_uniform(&r);
    step = -lambda * log(rr);
    double cxyz[3] = {cxstart, 0.0, czstart};
    for (int i = 0; i < 3; i++) xyz[i] = xyz[i] + step * scaled * cxyz[i];
    int traj = 0;
    while (traj < num) {
      const double dEds = predEds * log(Ec * J + 1.0) / Ec;
      const double dE = dEds * step;
      Ec = Ec + dE;
      if (Ec < min_energy) break;
      rr = (double)orc_mcrng_uniform(&r);
      const double cphi = 1.0 - 2.0 * alpha * rr / (1.0 + alpha - rr);
      const double sphi = sin(acos(cphi));
      rr = (double)orc_mcrng_uniform(&r);
      const double psi = tpi * rr;
      const double spsi = sin(psi), cpsi = cos(psi);
      double cxyzp[3];
      if (fabs(cxyz[2]) > 0.99999) {
        cxyzp[0] = sphi * cpsi;
        cxyzp[1] = sphi * spsi;
        cxyzp[2] = (cxyz[2] / fabs(cxyz[2])) * cphi;
      } else {
        const double dsq = sqrt(1.0 - cxyz[2] * cxyz[2]);
        const double dsqi = 1.0 / dsq;
        cxyzp[0] = sphi * (cxyz[0] * cxyz[2] * cpsi - cxyz[1] * spsi) * dsqi + cxyz[0] * cphi;
        cxyzp[1] = sphi * (cxyz[1] * cxyz[2] * cpsi + cxyz[0] * spsi) * dsqi + cxyz[1] * cphi;
        cxyzp[2] = -sphi * cpsi * dsq + cxyz[2] * cphi;
      }
      const double q0 = cxyzp[0] * cxyzp[0], q1 = cxyzp[1] * cxyzp[1], q2 = cxyzp[2] * cxyzp[2];
      const double ss = kahan ? orc_sum3_kahan_d(q0, q1, q2) : (q0 + q1) + q2;
      const double dd = 1.0 / sqrt(ss);
      for (int i = 0; i < 3; i++) cxyzp[i] = cxyzp[i] * dd;
      alpha = prealpha / Ec;
      step = Ec * (Ec + 1024.0) / Ze / (Ec + 511.0);
      sige = presig * step * step * alpha * (1.0 + alpha);
      lambda = pre * sige;
      rr = (double)orc_mcrng_uniform(&r);
      step = -lambda * log(rr);
      double xyzn[3];
      for (int i = 0; i < 3; i++) xyzn[i] = xyz[i] + step * scaled * cxyzp[i];
      const double znmax = xyzn[1] * tano;
      if (xyzn[2] > znmax) {
        /* Lambert2DSquareInverseDouble (Lambert.f90:452-483) times delta */
        double res0 = 0.0, res1 = 0.0;
        const double s2 = (cxyzp[0] * cxyzp[0] + cxyzp[1] * cxyzp[1]) + cxyzp[2] * cxyzp[2];
        if (!(fabs(1.0 - s2) > 1.0e-12) && fabs(cxyzp[2]) != 1.0) {
          if (fabs(cxyzp[1]) <= fabs(cxyzp[0])) {
            const double q = fabs(cxyzp[0]) / cxyzp[0] * sqrt(2.0 * (1.0 - cxyzp[2]));
            res0 = q * sPi2;
            res1 = q * atan(cxyzp[1] / cxyzp[0]) / sPi2;
          } else {
            const double q = fabs(cxyzp[1]) / cxyzp[1] * sqrt(2.0 * (1.0 - cxyzp[2]));
            res0 = q * atan(cxyzp[0] / cxyzp[1]) / sPi2;
            res1 = q * sPi2;
          }
        }
        const double dx0 = delta * res0, dx1 = delta * res1;
        const long ix = lround(dx1), iy = lround(-dx0);
        if (labs(ix) <= nx && labs(iy) <= nx && Ec > Emin) {
          const long iE = lround((Ec - Ehistmin) / Ebinsize) + 1;
          if (iE < 1 || iE > numEbins) {
            dropped++;
          } else {
            const double edis = fabs(xyz[2] / cxyzp[2]);
            const long iz = lround(edis * 0.1 / depthstep) + 1;
            if (iz > 0 && iz <= numzbins) {
              const long px = lroundf((float)ix / 10.0f), py = lroundf((float)iy / 10.0f);
              if (labs(px) <= nzx && labs(py) <= nzx)
                accum_z[(((py + nzx) * nzw + (px + nzx)) * numzbins + (iz - 1)) * numEbins + (iE - 1)] += 1;
              else
                dropped++;
            }
            accum_e[((iy + nx) * numsx + (ix + nx)) * numEbins + (iE - 1)] += 1;
          }
        }
        break;
      }
      for (int i = 0; i < 3; i++) { cxyz[i] = cxyzp[i]; xyz[i] = xyzn[i]; }
      traj++;
    }
  }
  return dropped;
}
