// Native per-energy setup of the master-pattern program: what the reference's host computes
// before the beam loop (CTEMEBSDmaster.f90:238-401), restated in C++ so that a host without the
// reference's modules can produce the `ctem_crystal` the plans consume:
//   CalcMatrices (crystal.f90:102-184)            metric and structure matrices
//   GenerateSymmetry / MakeGenerators / matrixmult / isitnew (symmetry.f90:72-450), with the
//     point-group pass of ReadDataFile (files.f90:354-367)
//   CalcOrbit / CalcPositions (symmetry.f90:614-676,773-863)
//   CalcUcg 'WK' with absorption and the Weickenmeier-Kohl routines FSCATT, WEKO, FPHON, RI1,
//     RI2, RIH1, RIH2, EI, FCORE (diffraction.f90:391-626, others.f90:68-778)
//   CalcWaveLength (diffraction.f90:229-302)
//   Compute_ReflectionList 'ALL' with CalcFamily and IsGAllowed (gvectors.f90:1262-1444,
//     symmetry.f90:544-594,1329-1356)
// The reference computes these in float32 (and double where declared so) and calls the C
// library's expf, logf, powf, cexpf and cabsf; this file makes the same calls, in the same order
// and with the same roundings, so the results are bit-equal. Every step is a single IEEE
// operation in the written order (contraction off).
//
// Reference state that carries between calls is kept the same way: the symmetry-operator array
// keeps the entries of the point-group pass when the space group is generated (isitnew compares
// against one slot past the last operator), and CalcFamily's work array keeps the previous
// family (its duplicate test reads one entry past the current count).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex.h>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/ctem_ebsd.h"
#include "../../include/ctem_setup.h"
#include "ctem_libm.h"
#include "ctem_sg.h"
#include "ctem_tables.h"

#pragma clang fp contract(off)

extern "C" {
float cabsf(float _Complex);
float _Complex cexpf(float _Complex);
}

namespace ctem {
namespace su {

// The Weickenmeier-Kohl routines below are written once and evaluated with the host C library
// (HostMath: the reference's own calls) or with the restatement of that library's algorithms
// (DevMath: the LUT build on the GPU). Both return the same bits for every input, so the GPU's
// LUT is bit-equal to the host's. powf is called only as Z^0.333333 for the atomic number of a
// site; the host evaluates it once per site and hands the value in.
struct HostMath {
  static float exp(float x) { return ::expf(x); }
  static float log(float x) { return ::logf(x); }
  static float sqrt(float x) { return ::sqrtf(x); }
};
struct DevMath {
  __device__ static float exp(float x) { return libm::expf(x); }
  __device__ static float log(float x) { return libm::logf(x); }
  __device__ static float sqrt(float x) { return sqrtf_cr(x); }
};
// Z^0.333333 of FCORE (others.f90), from the host C library
inline float zcube_root(int z) { return ::powf((float)z, .333333f); }
#define WK template <class MT> __host__ __device__

constexpr double cPi = 3.141592653589793238, cLight = 299792458.0, cPlanck = 6.626075e-34,
                 cCharge = 1.602177e-19, cRestmass = 9.109389e-31;

// first space group of each of the 32 point groups (symmetryvars.f90:242-244)
constexpr int kSGPG[32] = {1,   2,   3,   6,   10,  16,  25,  47,  75,  81,  83,  89,  99,  111, 123, 143,
                           147, 149, 156, 162, 168, 174, 175, 177, 183, 187, 191, 195, 200, 207, 215, 221};

struct M4 { double m[4][4]; };

struct State {
  // cell
  int xtal_system = 1;
  double a = 0, b = 0, c = 0, alpha = 0, beta = 0, gamma = 0;
  int sgnum = 1, sgset = 1, ntype = 0;
  std::vector<int> atom_type;
  std::vector<float> atom_pos;  // 5 per site: x, y, z, occupancy, B
  double dmt[3][3], rmt[3][3], dsm[3][3], rsm[3][3], vol = 0;
  bool hexset = false;
  // symmetry (SG%SYM_data etc.); data keeps what earlier generations left
  M4 data[193];
  M4 sc;  // SG%SYM_c
  int gennum = 0, matnum = 0, numpt = 0;
  bool centrosym = false;
  double direc[48][3][3], recip[48][3][3];
  char lattice = 'P';
  int itmp[49][3];
  // positions: apos[type][j][3] (float32 values), orbit sizes
  std::vector<std::vector<double>> apos;
  std::vector<int> numat;
  // electron
  double accvol = 0, relcor = 0, psihat = 0, lambda = 0, sigma = 0;
  State() {
    std::memset(data, 0, sizeof data);
    std::memset(&sc, 0, sizeof sc);
    std::memset(itmp, 0, sizeof itmp);
  }
};

// ---- crystal.f90 ----------------------------------------------------------------------------
void calc_matrices(State& S) {
  const double pirad = cPi / 180.0;
  // the reference's compiled code takes the sine and cosine of one angle from one sincos call
  double ca, cb, cg, sa, sb, sg;
  ::sincos(pirad * S.alpha, &sa, &ca);
  ::sincos(pirad * S.beta, &sb, &cb);
  ::sincos(pirad * S.gamma, &sg, &cg);
  const double tg = std::tan(pirad * S.gamma);
  const double a = S.a, b = S.b, c = S.c;
  S.dmt[0][0] = a * a; S.dmt[1][1] = b * b; S.dmt[2][2] = c * c;
  S.dmt[0][1] = a * b * cg; S.dmt[1][0] = S.dmt[0][1];
  S.dmt[0][2] = a * c * cb; S.dmt[2][0] = S.dmt[0][2];
  S.dmt[1][2] = b * c * ca; S.dmt[2][1] = S.dmt[1][2];
  const double abc = a * b * c;
  const double det = (abc * abc) * ((((1.0 - ca * ca) - cb * cb) - cg * cg) + ((2.0 * ca) * cb) * cg);
  S.vol = std::sqrt(det);
  const double bcs = b * c * sa, acs = a * c * sb, abs_ = a * b * sg;
  S.rmt[0][0] = bcs * bcs; S.rmt[1][1] = acs * acs; S.rmt[2][2] = abs_ * abs_;
  S.rmt[0][1] = ((a * b) * (c * c)) * (ca * cb - cg); S.rmt[1][0] = S.rmt[0][1];
  S.rmt[0][2] = ((a * (b * b)) * c) * (cg * ca - cb); S.rmt[2][0] = S.rmt[0][2];
  S.rmt[1][2] = (((a * a) * b) * c) * (cb * cg - ca); S.rmt[2][1] = S.rmt[1][2];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) S.rmt[i][j] = S.rmt[i][j] / det;
  S.dsm[0][0] = a; S.dsm[0][1] = b * cg; S.dsm[0][2] = c * cb;
  S.dsm[1][0] = 0.0; S.dsm[1][1] = b * sg; S.dsm[1][2] = -(c * (cb * cg - ca)) / sg;
  S.dsm[2][0] = 0.0; S.dsm[2][1] = 0.0; S.dsm[2][2] = S.vol / ((a * b) * sg);
  S.rsm[0][0] = 1.0 / a; S.rsm[0][1] = 0.0; S.rsm[0][2] = 0.0;
  S.rsm[1][0] = -1.0 / (a * tg); S.rsm[1][1] = 1.0 / (b * sg); S.rsm[1][2] = 0.0;
  S.rsm[2][0] = ((b * c) * (cg * ca - cb)) / (S.vol * sg);
  S.rsm[2][1] = ((a * c) * (cb * cg - ca)) / (S.vol * sg);
  S.rsm[2][2] = ((a * b) * sg) / S.vol;
}

// CalcDot/CalcLength in reciprocal space, double and single (crystal.f90:411-429,564-577)
double dot_r_double(const double p[3], const double q[3], const double r[3][3]) {
  double v[3];
  for (int i = 0; i < 3; i++) v[i] = ((0.0 + r[i][0] * q[0]) + r[i][1] * q[1]) + r[i][2] * q[2];
  return ((0.0 + p[0] * v[0]) + p[1] * v[1]) + p[2] * v[2];
}
float len_r_single(const float p[3], const double r[3][3]) {
  const double rm[9] = {r[0][0], r[0][1], r[0][2], r[1][0], r[1][1], r[1][2], r[2][0], r[2][1], r[2][2]};
  return std::sqrt(dot_r(p, p, rm));
}

// ---- symmetry.f90 ---------------------------------------------------------------------------
void fillgen(State& S, const char t[4], int isgn) {
  const double sgn = (double)isgn;
  double(&c)[4][4] = S.sc.m;
  std::memset(c, 0, sizeof c);
  c[3][3] = 1.0;
  switch (t[0]) {
    case 'a': c[0][0] = 1; c[1][1] = 1; c[2][2] = 1; break;
    case 'b': c[0][0] = -1; c[1][1] = -1; c[2][2] = 1; break;
    case 'c': c[0][0] = -1; c[1][1] = 1; c[2][2] = -1; break;
    case 'd': c[0][2] = 1; c[1][0] = 1; c[2][1] = 1; break;
    case 'e': c[0][1] = 1; c[1][0] = 1; c[2][2] = -1; break;
    case 'f': c[0][1] = -1; c[1][0] = -1; c[2][2] = -1; break;
    case 'g': c[0][1] = -1; c[1][0] = 1; c[2][2] = 1; break;
    case 'h': c[0][0] = -1; c[1][1] = -1; c[2][2] = -1; break;
    case 'i': c[0][0] = 1; c[1][1] = 1; c[2][2] = -1; break;
    case 'j': c[0][0] = 1; c[1][1] = -1; c[2][2] = 1; break;
    case 'k': c[0][1] = -1; c[1][0] = -1; c[2][2] = 1; break;
    case 'l': c[0][1] = 1; c[1][0] = 1; c[2][2] = 1; break;
    case 'm': c[0][1] = 1; c[1][0] = -1; c[2][2] = -1; break;
    case 'n': c[0][1] = -1; c[1][0] = 1; c[1][1] = -1; c[2][2] = 1; break;
    default: break;
  }
  for (int j = 1; j < 4; j++) {
    double v = c[j - 1][3];
    switch (t[j]) {
      case 'A': v = sgn / 6.0; break;
      case 'B': v = sgn / 4.0; break;
      case 'C': v = sgn / 3.0; break;
      case 'D': v = sgn / 2.0; break;
      case 'E': v = (sgn * 2.0) / 3.0; break;
      case 'F': v = (sgn * 3.0) / 4.0; break;
      case 'G': v = (sgn * 5.0) / 6.0; break;
      case 'O': v = 0.0; break;
      case 'X': v = (-sgn * 3.0) / 8.0; break;
      case 'Y': v = -sgn / 4.0; break;
      case 'Z': v = -sgn / 8.0; break;
      default: break;
    }
    c[j - 1][3] = v;
  }
}

M4 matmul4(const M4& a, const M4& b) {
  M4 r;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double s = 0.0;
      for (int k = 0; k < 4; k++) s = s + a.m[i][k] * b.m[k][j];
      r.m[i][j] = s;
    }
  return r;
}

char gen_char(const char* g, int pos1) {  // 1-based character of a blank-padded 40-character string
  const int n = (int)std::strlen(g);
  return (pos1 >= 1 && pos1 <= n) ? g[pos1 - 1] : ' ';
}

int make_generators(State& S) {
  const double eps = 0.0005;
  const char* genst = kGenerators[S.sgnum - 1];
  S.lattice = kLattice[S.sgnum - 1];
  fillgen(S, "aOOO", 1);
  S.data[1] = S.sc;
  S.gennum = gen_char(genst, 2) - 48;
  for (int i = 2; i <= 2 + S.gennum - 1; i++) {
    char t[4];
    for (int k = 1; k <= 4; k++) t[k - 1] = gen_char(genst, 2 + 4 * (i - 2) + k);
    fillgen(S, t, 1);
    S.data[i] = S.sc;
  }
  const int i = 2 + 4 * S.gennum + 1;
  if (gen_char(genst, 1) == '1') {
    S.centrosym = true;
    fillgen(S, "hOOO", 1);
    S.data[S.gennum + 2] = S.sc;
    S.gennum = S.gennum + 2;
  } else {
    S.gennum = S.gennum + 1;
  }
  if (gen_char(genst, i) != '0') {
    if (S.sgset != 1 && S.sgset != 2) return -1;  // the reference asks the user here
    if (S.sgset == 2) {
      const char t[4] = {'a', gen_char(genst, i + 1), gen_char(genst, i + 2), gen_char(genst, i + 3)};
      for (int l = 2; l <= S.gennum; l++) {
        fillgen(S, t, -1);
        M4 d = S.data[l];
        const M4 e = matmul4(d, S.sc);
        fillgen(S, t, 1);
        d = matmul4(S.sc, e);
        for (int k = 0; k < 3; k++) {
          double& v = d.m[k][3];
          if (std::fabs(v) < eps) v = 0.0;
          if (v < 0.0) while (v < 0.0) v = v + 1.0;
          if (v >= 1.0) while (v >= 1.0) v = v - 1.0;
          if (std::fabs(v - 1.0) < eps) v = 0.0;
        }
        S.data[l] = d;
      }
    }
  }
  return 0;
}

void matrixmult(State& S, int k1, int k2) {
  const double eps = 0.0005;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double s = 0.0;
      for (int k = 0; k < 4; k++) s = s + S.data[k1].m[i][k] * S.data[k2].m[k][j];
      S.sc.m[i][j] = s;
    }
  for (int k = 0; k < 3; k++) {
    double& v = S.sc.m[k][3];
    if (std::fabs(v) < eps) v = 0.0;
    if (v < 0.0) while (v < 0.0) v = v + 1.0;
    if (v > 1.0) while (v > 1.0) v = v - 1.0;
    if (std::fabs(v - 1.0) < eps) v = 0.0;
  }
}

bool isitnew(const State& S, int nsym) {
  const double eps = 0.0005;
  int k = 0, n = 0;
  while (k <= nsym && n != 12) {
    n = 0;
    k = k + 1;
    if (k > 192) break;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 4; j++)
        if (std::fabs(S.sc.m[i][j] - S.data[k].m[i][j]) < eps) n = n + 1;
  }
  return n != 12;
}

int generate_symmetry(State& S, bool dopg) {
  if (make_generators(S)) return -1;
  int nsym = S.gennum;
  for (int k = 1; k <= S.gennum; k++) {
    matrixmult(S, k, k);
    if (isitnew(S, nsym)) { nsym = nsym + 1; S.data[nsym] = S.sc; }
  }
  int k1 = 1;
  while (k1 <= nsym) {
    int k2 = k1 + 1;
    while (k2 <= nsym) {
      matrixmult(S, k2, k1);
      if (isitnew(S, nsym)) {
        nsym = nsym + 1;
        S.data[nsym] = S.sc;
        if (nsym >= 192) { k2 = nsym; k1 = nsym; }
      }
      k2 = k2 + 1;
    }
    k1 = k1 + 1;
  }
  S.matnum = nsym;
  for (int i = 1; i <= S.matnum; i++)
    for (int j = 0; j < 3; j++) S.data[i].m[j][3] = std::fmod(S.data[i].m[j][3], 1.0);
  if (dopg) {
    S.numpt = 0;
    for (int i = 1; i <= S.matnum; i++) {
      const double sm = (S.data[i].m[0][3] * S.data[i].m[0][3] + S.data[i].m[1][3] * S.data[i].m[1][3]) +
                        S.data[i].m[2][3] * S.data[i].m[2][3];
      if (sm < 0.1) {
        const int np = S.numpt;
        S.numpt = S.numpt + 1;
        for (int j = 0; j < 3; j++)
          for (int k = 0; k < 3; k++) S.direc[np][j][k] = S.data[i].m[j][k];
        for (int j = 0; j < 3; j++)
          for (int k = 0; k < 3; k++) {
            double q = 0.0;
            for (int l1 = 0; l1 < 3; l1++)
              for (int l2 = 0; l2 < 3; l2++) q = q + (S.dmt[j][l1] * S.data[i].m[l1][l2]) * S.rmt[l2][k];
            S.recip[np][j][k] = q;
          }
      }
    }
  }
  return 0;
}

// ReadDataFile after reading the file: settings, matrices, the point-group pass, the space group
int read_data(State& S) {
  S.hexset = (S.xtal_system == 4) || (S.xtal_system == 5 && S.sgset != 2);
  calc_matrices(S);
  int ipg = 0;
  for (int i = 1; i <= 32; i++)
    if (kSGPG[i - 1] <= S.sgnum) ipg = i;
  if (kSGPG[ipg - 1] == S.sgnum) return generate_symmetry(S, true);
  const int isave = S.sgnum;
  S.sgnum = kSGPG[ipg - 1];
  if (generate_symmetry(S, true)) return -1;
  S.sgnum = isave;
  return generate_symmetry(S, false);
}

// CalcOrbit with SYM_reduce: positions of site m (0-based) under every operator
std::vector<double> calc_orbit(const State& S, int m) {
  const double eps = 1.0e-4;
  std::vector<double> ct;
  double r[3];
  for (int i = 0; i < 3; i++) r[i] = (double)S.atom_pos[5 * m + i];
  ct.insert(ct.end(), r, r + 3);
  int n = 1;
  for (int i = 2; i <= S.matnum; i++) {
    double s[3];
    for (int j = 0; j < 3; j++) {
      s[j] = S.data[i].m[j][3];
      for (int k = 0; k < 3; k++) s[j] = s[j] + S.data[i].m[j][k] * r[k];
    }
    for (int j = 0; j < 3; j++) s[j] = std::fmod(s[j] + 100.0, 1.0);
    bool isnew = true;
    for (int mm = 0; mm < n; mm++) {
      double diff = 0.0;
      for (int j = 0; j < 3; j++) diff = diff + std::fabs(ct[3 * mm + j] - s[j]);
      if (diff < eps) isnew = false;
    }
    if (isnew) {
      n = n + 1;
      ct.insert(ct.end(), s, s + 3);
    }
  }
  return ct;
}

// CalcPositions('v'): the orbit of every site, each coordinate stored through float32
void calc_positions(State& S) {
  S.apos.assign(S.ntype, {});
  S.numat.assign(S.ntype, 0);
  for (int i = 0; i < S.ntype; i++) {
    const std::vector<double> ct = calc_orbit(S, i);
    const int n = (int)ct.size() / 3;
    for (int kk = 0; kk < n; kk++)
      for (int mm = 0; mm < 3; mm++) {
        const float r = (float)((ct[3 * kk + mm] + 1.0) - 1.0);
        S.apos[i].push_back((double)r);
      }
    S.numat[i] = n;
  }
}

bool is_g_allowed(const State& S, const int g[3]) {
  switch (S.lattice) {
    case 'F': {
      const int seo = (g[0] + 100) % 2 + (g[1] + 100) % 2 + (g[2] + 100) % 2;
      return !(seo == 1 || seo == 2);
    }
    case 'I': return (g[0] + g[1] + g[2] + 100) % 2 != 1;
    case 'A': return (g[1] + g[2] + 100) % 2 != 1;
    case 'B': return (g[0] + g[2] + 100) % 2 != 1;
    case 'C': return (g[0] + g[1] + 100) % 2 != 1;
    case 'R':
      if (S.hexset) return ((-g[0] + g[1] + g[2] + 90) % 3) == 0;
      return true;
    default: return true;
  }
}

// CalcFamily in reciprocal space; returns the count, members in S.itmp[0..num-1]
int calc_family(State& S, const int ind[3]) {
  const float eps = 0.0001f;
  int j = 1;
  for (int t = 0; t < 3; t++) S.itmp[0][t] = ind[t];
  const float h = (float)ind[0], k = (float)ind[1], l = (float)ind[2];
  for (int i = 2; i <= S.numpt; i++) {
    const double(&R)[3][3] = S.recip[i - 1];
    const float ih = (float)((R[0][0] * (double)h + R[0][1] * (double)k) + R[0][2] * (double)l);
    const float ik = (float)((R[1][0] * (double)h + R[1][1] * (double)k) + R[1][2] * (double)l);
    const float il = (float)((R[2][0] * (double)h + R[2][1] * (double)k) + R[2][2] * (double)l);
    bool newpoint = true;
    for (int m = 1; m <= j + 1; m++) {
      const float d0 = (float)S.itmp[m - 1][0] - ih, d1 = (float)S.itmp[m - 1][1] - ik,
                  d2 = (float)S.itmp[m - 1][2] - il;
      const float idiff = (d0 * d0 + d1 * d1) + d2 * d2;
      if (idiff < eps) newpoint = false;
    }
    if (newpoint) {
      j = j + 1;
      S.itmp[j - 1][0] = (int)std::lround(ih);
      S.itmp[j - 1][1] = (int)std::lround(ik);
      S.itmp[j - 1][2] = (int)std::lround(il);
    }
  }
  return j;
}

// ---- others.f90: Weickenmeier-Kohl scattering amplitudes (float32) ---------------------------
WK float ei(float x) {
  const float A1 = 8.57332f, A2 = 18.05901f, A3 = 8.63476f, A4 = .26777f, B1 = 9.57332f, B2 = 25.63295f,
              B3 = 21.09965f, B4 = 3.95849f;
  if (x < -60.f) return 0.f;
  if (x < -1.f) {
    const float xp = fabsf(x);
    const float num = A4 + xp * (A3 + xp * (A2 + xp * (A1 + xp)));
    const float den = B4 + xp * (B3 + xp * (B2 + xp * (B1 + xp)));
    return -(((num / den) * MT::exp(-xp)) / xp);
  }
  float e = .577216f + MT::log(fabsf(x));
  int i = 1;
  float si = x, summ = si;
  while (fabsf(si / x) > 1.e-6f) {
    const float ri = (float)i, rii = ri + 1.f;
    si = ((si * x) * ri) / (rii * rii);
    summ = summ + si;
    i = i + 1;
  }
  return e + summ;
}

WK float rih2(float x) {
  const float F[21] = {1.000000f, 1.005051f, 1.010206f, 1.015472f, 1.020852f, 1.026355f, 1.031985f,
                              1.037751f, 1.043662f, 1.049726f, 1.055956f, 1.062364f, 1.068965f, 1.075780f,
                              1.082830f, 1.090140f, 1.097737f, 1.105647f, 1.113894f, 1.122497f, 1.131470f};
  const float x1 = 1.f / x;
  const int i = (int)(200.f * x1);
  const int i1 = i + 1;
  return F[i] + (200.f * (F[i1] - F[i])) * (x1 - .5e-3f * (float)i);
}

WK float rih1(float x1, float x2, float x3) {
  if (x2 <= 20.f && x3 <= 20.f) return MT::exp(-x1) * (ei<MT>(x2) - ei<MT>(x3));
  float r;
  if (x2 > 20.f) r = (MT::exp(x2 - x1) * rih2<MT>(x2)) / x2;
  else r = MT::exp(-x1) * ei<MT>(x2);
  if (x3 > 20.f) r = r - (MT::exp(x3 - x1) * rih2<MT>(x3)) / x3;
  else r = r - MT::exp(-x1) * ei<MT>(x3);
  return r;
}

WK float ri1(float bi, float bj, float g) {
  const float PI = 3.1415927f, C = 0.5772157f;
  const float g2 = g * g;
  float eps = fmaxf(bi, bj);
  eps = eps * g2;
  float r;
  if (eps <= 0.1f) {
    r = bi * MT::log((bi + bj) / bi) + bj * MT::log((bi + bj) / bj);
    r = r * PI;
    if (g == 0.f) return r;
    const float bi2 = bi * bi, bj2 = bj * bj;
    float temp = (.5f * bi2) * MT::log(bi / (bi + bj)) + (.5f * bj2) * MT::log(bj / (bi + bj));
    temp = (temp + .75f * (bi2 + bj2)) - (.25f * (bi + bj)) * (bi + bj);
    temp = temp - (.5f * (bi - bj)) * (bi - bj);
    r = r + (PI * g2) * temp;
    return r;
  }
  const float big2 = bi * g2, bjg2 = bj * g2;
  r = ((2.f * C + MT::log(big2)) + MT::log(bjg2)) - 2.f * ei<MT>(-(((bi * bj) * g2) / (bi + bj)));
  r = r + rih1<MT>(big2, (big2 * bi) / (bi + bj), big2);
  r = r + rih1<MT>(bjg2, (bjg2 * bj) / (bi + bj), bjg2);
  r = (r * PI) / g2;
  return r;
}

WK float ri2(float bi, float bj, float g, float u) {
  const float PI = 3.1415927f;
  const float u2 = u * u, u22 = .5f * u2, g2 = g * g;
  const float biuh = bi + .5f * u2, bjuh = bj + .5f * u2, biu = bi + u2, bju = bj + u2;
  float eps = fmaxf(fmaxf(bi, bj), u2);
  eps = eps * g2;
  float r;
  if (eps <= 0.1f) {
    r = (bi + u2) * MT::log(((bi + bj) + u2) / (bi + u2));
    r = r + bj * MT::log(((bi + bj) + u2) / (bj + u2));
    r = r + u2 * MT::log(u2 / (bj + u2));
    r = r * PI;
    if (g == 0.f) return r;
    float temp = ((.5f * u22) * u22) * MT::log((biu * bju) / (u2 * u2));
    temp = temp + ((.5f * biuh) * biuh) * MT::log(biu / (biuh + bjuh));
    temp = temp + ((.5f * bjuh) * bjuh) * MT::log(bju / (biuh + bjuh));
    temp = (temp + (.25f * biu) * biu) + (.5f * bi) * bi;
    temp = (temp + (.25f * bju) * bju) + (.5f * bj) * bj;
    temp = temp - (.25f * (biuh + bjuh)) * (biuh + bjuh);
    const float q = (bi * biu - bj * bju) / (biuh + bjuh);
    temp = temp - .5f * (q * q);
    temp = temp - u22 * u22;
    r = r + (PI * g2) * temp;
    return r;
  }
  r = ei<MT>(-((((.5f * u2) * g2) * biuh) / biu)) + ei<MT>(-((((.5f * u2) * g2) * bjuh) / bju));
  r = (r - ei<MT>(-(((biuh * bjuh) * g2) / (biuh + bjuh)))) - ei<MT>(-((.25f * u2) * g2));
  r = 2.f * r;
  r = r + rih1<MT>((.5f * u2) * g2, (.25f * u2) * g2, (((.25f * u2) * u2) * g2) / biu);
  r = r + rih1<MT>((.5f * u2) * g2, (.25f * u2) * g2, (((.25f * u2) * u2) * g2) / bju);
  r = r + rih1<MT>(biuh * g2, ((biuh * biuh) * g2) / (biuh + bjuh), ((biuh * biuh) * g2) / biu);
  r = r + rih1<MT>(bjuh * g2, ((bjuh * bjuh) * g2) / (biuh + bjuh), ((bjuh * bjuh) * g2) / bju);
  r = (r * PI) / g2;
  return r;
}

WK float weko(const float A[4], const float B[4], float s) {
  float w = 0.f, s2 = 0.f;
  if (s > 0.f) s2 = 1.f / (s * s);
  for (int i = 0; i < 4; i++) {
    const float argu = (B[i] * s) * s;
    if (argu < .1f) w = w + (A[i] * B[i]) * (1.f - .5f * argu);
    else if (argu > 20.f) w = w + A[i] * s2;
    else w = w + (A[i] * (1.f - MT::exp(-argu))) * s2;
  }
  return w;
}

WK float fphon(float g, float ul, const float A[4], const float B[4]) {
  const float FOURPI = 12.56636f, FP2 = FOURPI * FOURPI;
  const float u2 = ul * ul;
  float a1[4], b1[4];
  for (int i = 0; i < 4; i++) a1[i] = A[i] * FP2;
  for (int i = 0; i < 4; i++) b1[i] = B[i] / FP2;
  float f = 0.f;
  const float g2 = g * g;
  const float dewa = MT::exp(-((.5f * u2) * g2));
  for (int j = 0; j < 4; j++) {
    f = f + (a1[j] * a1[j]) * (dewa * ri1<MT>(b1[j], b1[j], g) - ri2<MT>(b1[j], b1[j], g, ul));
    for (int i = 0; i < j; i++)
      f = f + ((2.f * a1[j]) * a1[i]) * (dewa * ri1<MT>(b1[i], b1[j], g) - ri2<MT>(b1[i], b1[j], g, ul));
  }
  return f;
}

WK float fcore(float g, int z, float zr, float accvlt) {
  const float PI = 3.1415927f, A0 = .5289f;
  const float k0 = .5068f * MT::sqrt(1022.f * accvlt + accvlt * accvlt);
  const float de = (float)(6 * z) * 1.e-3f;
  const float thetae = ((de / (2.f * accvlt)) * (2.f * accvlt + 1022.f)) / (accvlt + 1022.f);
  const float r = (0.885f * A0) / zr;
  const float ta = 1.f / (k0 * r);
  const float tb = g / (2.f * k0);
  const float omega = (2.f * tb) / ta;
  const float kappa = thetae / ta;
  const float k2 = kappa * kappa, o2 = omega * omega;
  const float s4 = MT::sqrt(o2 + 4.f * k2);
  const float x1 = (omega / ((1.f + o2) * s4)) * MT::log((omega + s4) / (2.f * kappa));
  const float q = (1.f + o2) * (1.f + o2) + (4.f * k2) * o2;
  const float x2 = (1.f / MT::sqrt(q)) *
                   MT::log((((1.f + 2.f * k2) + o2) + MT::sqrt(q)) / ((2.f * kappa) * MT::sqrt(1.f + k2)));
  float x3;
  if (omega > 1e-2f) {
    const float s5 = MT::sqrt(o2 + 4.f * (1.f + k2));
    x3 = (1.f / (omega * s5)) * MT::log((omega + s5) / (2.f * MT::sqrt(1.f + k2)));
  } else {
    x3 = 1.f / (4.f * (1.f + k2));
  }
  const float hi = ((float)(2 * z) / (ta * ta)) * ((-x1 + x2) - x3);
  return ((((4.f / (A0 * A0)) * 2.f) * PI) / (k0 * k0)) * hi;
}

// FSCATT with absorption (ABSFLG = 3), ACCFLG and DWFLG set; returns (real, imaginary).
// zr = Z^0.333333 (zcube_root).
WK void fscatt(float g, float ul, int z, float zr, const float* A, const float* B, float accvlt, float* fre,
               float* fim) {
  const float FOURPI = 12.56637062f;
  const float dewa = MT::exp(-((((.5f * ul) * ul) * g) * g));
  const float s = g / FOURPI;
  float freal = (FOURPI * dewa) * weko<MT>(A, B, s);
  float fima = fcore<MT>(g, z, zr, accvlt) * dewa + fphon<MT>(g, ul, A, B);
  const float k0 = .5068f * MT::sqrt(1022.f * accvlt + accvlt * accvlt);
  const float gamma = (accvlt + 511.f) / 511.f;
  freal = freal * gamma;
  fima = ((fima * gamma) * gamma) / k0;
  *fre = freal;
  *fim = fima;
}

// CalcUcg, method 'WK' with absorption: Ucg (float32 complex) and Vmod, Vpmod
struct Ucg { float re, im, vmod, vpmod; };

Ucg calc_ucg(const State& S, const int hkl[3]) {
  const float twopi = (float)(2.0 * cPi);
  float g;
  if (hkl[0] * hkl[0] + hkl[1] * hkl[1] + hkl[2] * hkl[2] == 0) {
    g = 0.f;
  } else {
    const double hd[3] = {(double)hkl[0], (double)hkl[1], (double)hkl[2]};
    g = (float)std::sqrt(dot_r_double(hd, hd, S.rmt));
  }
  const float swk = 0.1f * twopi;
  const float dwwk = (float)(100.0 / (8.0 * (cPi * cPi)));
  const float s = g * swk;
  const float pref = (float)(((double)0.04787801f / S.vol) / (4.0 * cPi));
  const float preg = (2.0f * (float)((cRestmass * cCharge) / (cPlanck * cPlanck))) * 1.0e-18f;
  const float pre = pref * preg;
  float ffr = 0.f, ffi = 0.f, ggr = 0.f, ggi = 0.f;
  const float accvlt = (float)S.accvol / 1000.0f;
  for (int m = 0; m < S.ntype; m++) {
    const float ul = ::sqrtf(S.atom_pos[5 * m + 4] * dwwk);
    float sfr, sfi;
    fscatt<HostMath>(s, ul, S.atom_type[m], zcube_root(S.atom_type[m]), kWKA[S.atom_type[m] - 1],
                     kWKB[S.atom_type[m] - 1], accvlt, &sfr, &sfi);
    const float occ = S.atom_pos[5 * m + 3];
    sfr = sfr * occ;
    sfi = sfi * occ;
    float p1r = 0.f, p1i = 0.f;
    const std::vector<double>& ap = S.apos[m];
    for (int j = 0; j < S.numat[m]; j++) {
      const double sum = ((0.0 + (double)(float)hkl[0] * ap[3 * j]) + (double)(float)hkl[1] * ap[3 * j + 1]) +
                         (double)(float)hkl[2] * ap[3 * j + 2];
      const float arg = (float)((double)twopi * sum);
      float _Complex zc;
      __real__ zc = 0.0f;
      __imag__ zc = -arg;
      const float _Complex e = cexpf(zc);
      p1r = p1r + __real__ e;
      p1i = p1i + __imag__ e;
    }
    ffr = ffr + p1r * sfr;
    ffi = ffi + p1i * sfr;
    ggr = ggr + p1r * sfi;
    ggi = ggi + p1i * sfi;
  }
  float _Complex ff, gg;
  __real__ ff = ffr; __imag__ ff = ffi;
  __real__ gg = ggr; __imag__ gg = ggi;
  Ucg u;
  u.vmod = pref * cabsf(ff);
  u.vpmod = pref * cabsf(gg);
  u.re = pre * (ffr - ggi);
  u.im = pre * (ffi + ggr);
  return u;
}

// CalcWaveLength with the crystal's mean inner potential (scattering set 3)
void calc_wavelength(State& S, double voltage) {
  S.accvol = voltage;
  const double temp1 = ((1.0e9 * cPlanck) / std::sqrt((2.0 * cRestmass) * cCharge));
  const double temp2 = (((cCharge * 0.5) * voltage) / cRestmass) / (cLight * cLight);
  S.relcor = 1.0 + 2.0 * temp2;
  S.psihat = voltage * (1.0 + temp2);
  calc_positions(S);
  const int zero[3] = {0, 0, 0};
  const Ucg u = calc_ucg(S, zero);
  S.psihat = S.psihat + (double)u.vmod;
  S.lambda = temp1 / std::sqrt(S.psihat);
  const double sig = ((((2.0 * cPi) * cRestmass) * S.relcor) * cCharge) * S.lambda;
  S.sigma = (1.0e-18 * sig) / (cPlanck * cPlanck);
}

// ---- the LUT box on the GPU ----------------------------------------------------------------
// Inputs of CalcUcg that do not depend on g, per site: Z, Z^0.333333, ul, occupancy, the fit
// coefficients and the float32-valued orbit positions.
struct DevSites {
  int ntype;
  const int* z; const float* zr; const float* ul; const float* occ; const float* wka; const float* wkb;
  const int* numat; const int* aoff; const double* apos;
};
struct DevBox {
  int lo[3], dim[3];
  char lattice; int hexset;
  double rmt[9];
  float pref, pre, swk, accvlt;
};

__device__ bool dev_allowed(char lattice, int hexset, int h, int k, int l) {
  switch (lattice) {
    case 'F': { const int seo = (h + 100) % 2 + (k + 100) % 2 + (l + 100) % 2; return !(seo == 1 || seo == 2); }
    case 'I': return (h + k + l + 100) % 2 != 1;
    case 'A': return (k + l + 100) % 2 != 1;
    case 'B': return (h + l + 100) % 2 != 1;
    case 'C': return (h + k + 100) % 2 != 1;
    case 'R': return hexset ? ((-h + k + l + 90) % 3) == 0 : true;
    default: return true;
  }
}

// One thread per cell of the box: Ucg of every lattice-allowed g (0 elsewhere), in the order of
// operations of CalcUcg 'WK' with absorption, with the C library's float32 functions restated
// (bit-equal to the host's build).
__global__ __launch_bounds__(256) void k_ucg_box(DevBox B, DevSites S, float2* __restrict__ lut,
                                                 uint8_t* __restrict__ dbd) {
#pragma clang fp contract(off)
  const int64_t n = (int64_t)B.dim[0] * B.dim[1] * B.dim[2];
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const int h = (int)(c % B.dim[0]) + B.lo[0];
  const int k = (int)((c / B.dim[0]) % B.dim[1]) + B.lo[1];
  const int l = (int)(c / ((int64_t)B.dim[0] * B.dim[1])) + B.lo[2];
  float2 out = make_float2(0.f, 0.f);
  uint8_t small = 0;
  if (dev_allowed(B.lattice, B.hexset, h, k, l)) {
    float g = 0.f;
    if (h * h + k * k + l * l != 0) {
      const double hd[3] = {(double)h, (double)k, (double)l};
      double v[3];
      for (int i = 0; i < 3; i++) v[i] = ((0.0 + B.rmt[3 * i] * hd[0]) + B.rmt[3 * i + 1] * hd[1]) + B.rmt[3 * i + 2] * hd[2];
      g = (float)sqrt(((0.0 + hd[0] * v[0]) + hd[1] * v[1]) + hd[2] * v[2]);
    }
    const float s = g * B.swk;
    const float twopi = (float)(2.0 * 3.141592653589793238);
    float ffr = 0.f, ffi = 0.f, ggr = 0.f, ggi = 0.f;
    for (int m = 0; m < S.ntype; m++) {
      float sfr, sfi;
      fscatt<DevMath>(s, S.ul[m], S.z[m], S.zr[m], S.wka + 4 * m, S.wkb + 4 * m, B.accvlt, &sfr, &sfi);
      sfr = sfr * S.occ[m];
      sfi = sfi * S.occ[m];
      float p1r = 0.f, p1i = 0.f;
      const double* ap = S.apos + 3 * S.aoff[m];
      for (int j = 0; j < S.numat[m]; j++) {
        const double sum = ((0.0 + (double)(float)h * ap[3 * j]) + (double)(float)k * ap[3 * j + 1]) +
                           (double)(float)l * ap[3 * j + 2];
        const float arg = (float)((double)twopi * sum);
        float sn, cs;
        libm::sincosf(-arg, &sn, &cs);
        p1r = p1r + cs;
        p1i = p1i + sn;
      }
      ffr = ffr + p1r * sfr;
      ffi = ffi + p1i * sfr;
      ggr = ggr + p1r * sfi;
      ggi = ggi + p1i * sfi;
    }
    out = make_float2(B.pre * (ffr - ggi), B.pre * (ffi + ggr));
    small = (h != 0 || k != 0 || l != 0) && libm::cabsf(out.x, out.y) <= 1.0e-10f;
  }
  lut[c] = out;
  dbd[c] = small;
}

}  // namespace su
}  // namespace ctem

using namespace ctem;
using namespace ctem::su;

// The setup object of the C-ABI: the state above plus the arrays a ctem_crystal points into.
struct ctem_setup {
  State S;
  float ekev = 0, dmin = 0, weakcutoff = 0, cutoff = 0, sgcutoff = 0;
  int isym = 0, usehex = 0;
  int lut_lo[3] = {0, 0, 0}, lut_dim[3] = {0, 0, 0}, imhkl[3] = {0, 0, 0};
  float dynupz = 0.0f;
  std::vector<int32_t> hkl;       // list order
  std::vector<double> ucg;        // per list entry (float32 values)
  std::vector<uint8_t> dbd;       // per list entry
  std::vector<double> lut;        // box, Fortran order, complex interleaved
  std::vector<uint8_t> lut_dbd;   // box
  std::vector<int32_t> site_z, site_norbit;
  std::vector<float> site_b;
  std::vector<double> orbit;
  std::string err;
};

namespace {
thread_local std::string g_setup_err;
int32_t setup_fail(int32_t code, const std::string& m) {
  g_setup_err = m;
  return code;
}

int laue_isym(int sgnum, int* usehex) {
  static const int laue_of_pg[32] = {1, 1, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5, 5, 5, 5, 6,
                                     6, 7, 7, 7, 8, 8, 8, 9, 9, 9, 9, 10, 10, 11, 11, 11};
  static const int second[11] = {149, 151, 153, 156, 158, 160, 161, 164, 165, 166, 167};
  int j = 0;
  for (int i = 1; i <= 32; i++)
    if (kSGPG[i - 1] <= sgnum) j = i;
  int isym = laue_of_pg[j - 1];
  if (isym == 7)
    for (int s : second)
      if (sgnum == s) isym = 12;
  *usehex = ((isym >= 6 && isym <= 9) || isym == 12) ? 1 : 0;
  return isym;
}
}  // namespace

extern "C" const char* ctem_setup_last_error(void) { return g_setup_err.c_str(); }

// Fill the LUT box with CalcUcg of every lattice-allowed g (the reference's outer-box loop and its
// list traversal compute CalcUcg for each of them; the traversal adds the family members the
// centring condition forbids) and the double-diffraction flags, on the host (bit-equal) or on a
// GPU.
static int32_t fill_box_host(ctem_setup* T) {
  State& S = T->S;
  const float ddt = 1.0e-10f;
  for (int l = 0; l < T->lut_dim[2]; l++)
    for (int k = 0; k < T->lut_dim[1]; k++)
      for (int h = 0; h < T->lut_dim[0]; h++) {
        const int g[3] = {h + T->lut_lo[0], k + T->lut_lo[1], l + T->lut_lo[2]};
        if (!is_g_allowed(S, g)) continue;
        const size_t c = (size_t)h + (size_t)T->lut_dim[0] * ((size_t)k + (size_t)T->lut_dim[1] * l);
        const Ucg u = calc_ucg(S, g);
        T->lut[2 * c] = (double)u.re;
        T->lut[2 * c + 1] = (double)u.im;
        float _Complex z;
        __real__ z = u.re;
        __imag__ z = u.im;
        if ((g[0] || g[1] || g[2]) && cabsf(z) <= ddt) T->lut_dbd[c] = 1;
      }
  return CTEM_OK;
}

static int32_t fill_box_gpu(ctem_setup* T, int device, double* ms) {
  State& S = T->S;
  const size_t ncell = (size_t)T->lut_dim[0] * T->lut_dim[1] * T->lut_dim[2];
  std::vector<int> z, numat, aoff;
  std::vector<float> ul, occ, wa, wb, zr;
  std::vector<double> ap;
  const float dwwk = (float)(100.0 / (8.0 * (cPi * cPi)));
  for (int m = 0; m < S.ntype; m++) {
    z.push_back(S.atom_type[m]);
    zr.push_back(zcube_root(S.atom_type[m]));
    ul.push_back(::sqrtf(S.atom_pos[5 * m + 4] * dwwk));
    occ.push_back(S.atom_pos[5 * m + 3]);
    for (int i = 0; i < 4; i++) { wa.push_back(kWKA[S.atom_type[m] - 1][i]); wb.push_back(kWKB[S.atom_type[m] - 1][i]); }
    numat.push_back(S.numat[m]);
    aoff.push_back((int)(ap.size() / 3));
    ap.insert(ap.end(), S.apos[m].begin(), S.apos[m].end());
  }
  DevBox B;
  for (int d = 0; d < 3; d++) { B.lo[d] = T->lut_lo[d]; B.dim[d] = T->lut_dim[d]; }
  B.lattice = S.lattice;
  B.hexset = S.hexset ? 1 : 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) B.rmt[3 * i + j] = S.rmt[i][j];
  const float twopi = (float)(2.0 * cPi);
  B.swk = 0.1f * twopi;
  B.pref = (float)(((double)0.04787801f / S.vol) / (4.0 * cPi));
  B.pre = B.pref * ((2.0f * (float)((cRestmass * cCharge) / (cPlanck * cPlanck))) * 1.0e-18f);
  B.accvlt = (float)S.accvol / 1000.0f;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return setup_fail(CTEM_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  // one allocation for every input and output
  const size_t nb_i = sizeof(int) * (3 * (size_t)S.ntype), nb_f = sizeof(float) * (11 * (size_t)S.ntype),
               nb_d = sizeof(double) * ap.size(), nb_l = sizeof(float2) * ncell, nb_b = ncell;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t total = up(nb_i) + up(nb_f) + up(nb_d) + up(nb_l) + up(nb_b);
  char* buf = nullptr;
  if ((e = hipMalloc(&buf, total)) != hipSuccess) return setup_fail(CTEM_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  char* p = buf;
  int* d_i = (int*)p; p += up(nb_i);
  float* d_f = (float*)p; p += up(nb_f);
  double* d_ap = (double*)p; p += up(nb_d);
  float2* d_lut = (float2*)p; p += up(nb_l);
  uint8_t* d_dbd = (uint8_t*)p;
  std::vector<int> hi(z);
  hi.insert(hi.end(), numat.begin(), numat.end());
  hi.insert(hi.end(), aoff.begin(), aoff.end());
  std::vector<float> hf(ul);
  hf.insert(hf.end(), occ.begin(), occ.end());
  hf.insert(hf.end(), wa.begin(), wa.end());
  hf.insert(hf.end(), wb.begin(), wb.end());
  hf.insert(hf.end(), zr.begin(), zr.end());
  DevSites D;
  D.ntype = S.ntype;
  D.z = d_i; D.numat = d_i + S.ntype; D.aoff = d_i + 2 * S.ntype;
  D.ul = d_f; D.occ = d_f + S.ntype; D.wka = d_f + 2 * S.ntype; D.wkb = d_f + 6 * S.ntype;
  D.zr = d_f + 10 * S.ntype;
  D.apos = d_ap;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  e = hipMemcpy(d_i, hi.data(), sizeof(int) * hi.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_f, hf.data(), sizeof(float) * hf.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && nb_d) e = hipMemcpy(d_ap, ap.data(), nb_d, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_ucg_box, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, nullptr, B, D, d_lut, d_dbd);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
  std::vector<float2> hl(ncell);
  if (e == hipSuccess) e = hipMemcpy(hl.data(), d_lut, nb_l, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(T->lut_dbd.data(), d_dbd, nb_b, hipMemcpyDeviceToHost);
  float kms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&kms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(buf);
  if (e != hipSuccess) return setup_fail(CTEM_E_HIP, std::string("LUT build on the GPU: ") + hipGetErrorString(e));
  for (size_t c = 0; c < ncell; c++) {
    T->lut[2 * c] = (double)hl[c].x;
    T->lut[2 * c + 1] = (double)hl[c].y;
  }
  if (ms) *ms = kms;
  return CTEM_OK;
}

// voltage [V] enters CalcWaveLength as the programs pass it (the master program 1000 ekev in
// float32, the channeling-pattern program its float32 `voltage`). kzone (may be null) and
// convang >= 0: the convergence-angle window of Compute_ReflectionList (gvectors.f90:1295-1301,
// 1349-1395) for zone axis kzone; null: no window (every angle lies in 0..pi).
static int32_t setup_build(const ctem_xtal* x, float ekev, double voltage, const int32_t* kzone, float convang,
                           float dmin, float weakcutoff, float cutoff, float sgcutoff, int device, double* gpu_ms,
                           ctem_setup** out) {
#pragma clang fp contract(off)
  if (!x || !out) return setup_fail(CTEM_E_ARG, "ctem_setup_create: null argument");
  *out = nullptr;
  if (x->ntype < 1 || x->ntype > 250 || !x->atom_type || !x->atom_pos)
    return setup_fail(CTEM_E_ARG, "ctem_setup_create: 1..250 sites needed");
  if (x->sgnum < 1 || x->sgnum > 237) return setup_fail(CTEM_E_ARG, "ctem_setup_create: space group not in 1..237");
  if (!(voltage > 0.0) || !(dmin > 0.0f)) return setup_fail(CTEM_E_ARG, "ctem_setup_create: energy and dmin must be positive");
  if (kzone && (kzone[0] == 0 && kzone[1] == 0 && kzone[2] == 0))
    return setup_fail(CTEM_E_ARG, "ctem_setup_create_ecp: zone axis [000]");
  if (kzone && !(convang >= 0.0f)) return setup_fail(CTEM_E_ARG, "ctem_setup_create_ecp: convergence angle must be >= 0");
  for (int i = 0; i < x->ntype; i++)
    if (x->atom_type[i] < 1 || x->atom_type[i] > 98 || !(x->atom_pos[5 * i + 4] >= 0.0f))
      return setup_fail(CTEM_E_ARG, "ctem_setup_create: atomic numbers 1..98 and B >= 0 needed");
  ctem_setup* T = new ctem_setup();
  State& S = T->S;
  S.xtal_system = x->xtal_system;
  S.a = x->a; S.b = x->b; S.c = x->c; S.alpha = x->alpha; S.beta = x->beta; S.gamma = x->gamma;
  S.sgnum = x->sgnum; S.sgset = x->sgset; S.ntype = x->ntype;
  S.atom_type.assign(x->atom_type, x->atom_type + x->ntype);
  S.atom_pos.assign(x->atom_pos, x->atom_pos + 5 * x->ntype);
  T->ekev = ekev; T->dmin = dmin; T->weakcutoff = weakcutoff; T->cutoff = cutoff; T->sgcutoff = sgcutoff;
  if (read_data(S)) {
    delete T;
    return setup_fail(CTEM_E_ARG, "ctem_setup_create: this space group has two origin settings; sgset must be 1 or 2");
  }
  calc_positions(S);
  T->isym = laue_isym(S.sgnum, &T->usehex);
  calc_wavelength(S, voltage);
  // Compute_ReflectionList 'ALL': the angular window of the family members (gvectors.f90:1295-1301):
  // minang = pi/2 - 3 convang .. maxang = pi/2 + 3 convang about the zone axis kr (reciprocal
  // components of k, float32), or 0..pi without a window
  float kr[3] = {0.0f, 0.0f, 1.0f}, minang = 0.0f, maxang = (float)M_PI;
  if (kzone) {
    for (int j = 0; j < 3; j++) {
      double d = 0.0;
      for (int i = 0; i < 3; i++) d = d + (double)(float)kzone[i] * S.dmt[i][j];  // TransSpace d -> r
      kr[j] = (float)d;
    }
    const float c3 = 3.0f * convang;
    minang = (float)(M_PI * 0.5 - (double)c3);
    maxang = (float)(M_PI * 0.5 + (double)c3);
  }
  const double rm9[9] = {S.rmt[0][0], S.rmt[0][1], S.rmt[0][2], S.rmt[1][0], S.rmt[1][1], S.rmt[1][2],
                         S.rmt[2][0], S.rmt[2][1], S.rmt[2][2]};
  // CalcAngle(float(g), kr, 'r') (crystal.f90:632-666)
  auto angle_r = [&](const int g[3]) -> float {
    const float p[3] = {(float)g[0], (float)g[1], (float)g[2]};
    const float xx = dot_r(p, kr, rm9);
    const float yy = std::sqrt(dot_r(p, p, rm9));
    const float zz = std::sqrt(dot_r(kr, kr, rm9));
    const float t = xx / (yy * zz);
    if (t >= 1.0f) return 0.0f;
    if (t <= -1.0f) return (float)M_PI;
    return ::acosf(t);
  };
  int im[3];
  for (int d = 0; d < 3; d++) {
    im[d] = 1;
    for (;;) {
      im[d] = im[d] + 1;
      float p[3] = {0.f, 0.f, 0.f};
      p[d] = (float)im[d];
      const float dhkl = 1.0f / len_r_single(p, S.rmt);
      if (dhkl < dmin) break;
    }
    T->imhkl[d] = im[d];
    T->lut_lo[d] = -2 * im[d];
    T->lut_dim[d] = 4 * im[d] + 1;
  }
  const size_t ncell = (size_t)T->lut_dim[0] * T->lut_dim[1] * T->lut_dim[2];
  T->lut.assign(2 * ncell, 0.0);
  T->lut_dbd.assign(ncell, 0);
  int32_t rc = device < 0 ? fill_box_host(T) : fill_box_gpu(T, device, gpu_ms);
  if (rc) { delete T; return rc; }
  auto cell_of = [&](const int g[3]) -> size_t {
    return (size_t)(g[0] - T->lut_lo[0]) +
           (size_t)T->lut_dim[0] * ((size_t)(g[1] - T->lut_lo[1]) + (size_t)T->lut_dim[1] * (size_t)(g[2] - T->lut_lo[2]));
  };
  auto in_box = [&](const int g[3]) {
    for (int d = 0; d < 3; d++)
      if (g[d] < T->lut_lo[d] || g[d] >= T->lut_lo[d] + T->lut_dim[d]) return false;
    return true;
  };
  auto add_reflection = [&](const int g[3]) {
    const size_t c = cell_of(g);
    T->hkl.insert(T->hkl.end(), g, g + 3);
    T->ucg.push_back(T->lut[2 * c]);
    T->ucg.push_back(T->lut[2 * c + 1]);
  };
  std::vector<uint8_t> refdone(ncell, 0);
  {
    // 000 first, from the host (its Vpmod is DynUpz and its Vmod entered the wavelength)
    const int g0[3] = {0, 0, 0};
    const Ucg u = calc_ucg(S, g0);
    T->dynupz = u.vpmod;
    const size_t c = cell_of(g0);
    T->lut[2 * c] = (double)u.re;
    T->lut[2 * c + 1] = (double)u.im;
    add_reflection(g0);
    refdone[c] = 2;
  }
  for (int ix = -2 * im[0]; ix <= 2 * im[0]; ix++)
    for (int iy = -2 * im[1]; iy <= 2 * im[1]; iy++)
      for (int iz = -2 * im[2]; iz <= 2 * im[2]; iz++) {
        const int gg[3] = {ix, iy, iz};
        if (!is_g_allowed(S, gg)) continue;
        if (std::abs(ix) > im[0] || std::abs(iy) > im[1] || std::abs(iz) > im[2]) continue;
        if (refdone[cell_of(gg)] != 0) continue;
        const int numr = calc_family(S, gg);
        for (int ir = 0; ir < numr; ir++)
          if (!in_box(S.itmp[ir])) {
            delete T;
            return setup_fail(CTEM_E_RANGE, "ctem_setup_create: a family member lies outside the LUT box");
          }
        // members inside the angular window are added in family order; then the zone's own
        // members (g.k = 0) that the window missed (gvectors.f90:1364-1410)
        std::vector<uint8_t> inrange(numr, 1);
        if (kzone)
          for (int ir = 0; ir < numr; ir++) {
            const float sa = angle_r(S.itmp[ir]);
            inrange[ir] = (sa >= minang) && (sa <= maxang);
          }
        for (int ir = 0; ir < numr; ir++) refdone[cell_of(S.itmp[ir])] = 1;
        std::vector<int> order;
        for (int ir = 0; ir < numr; ir++)
          if (inrange[ir]) order.push_back(ir);
        for (int ir = 0; ir < numr; ir++) {
          const int* gi = S.itmp[ir];
          const int nz = kzone ? gi[0] * kzone[0] + gi[1] * kzone[1] + gi[2] * kzone[2] : 1;
          if (!inrange[ir] && nz == 0) order.push_back(ir);
        }
        for (int ir : order) {
          const int g[3] = {S.itmp[ir][0], S.itmp[ir][1], S.itmp[ir][2]};
          // the family comes from the point group of the first space group of the crystal's
          // point group, whose operators need not keep the lattice centring condition (R-3c in
          // hexagonal axes takes P-31m's): such members are in the list, and in the LUT, too
          if (!is_g_allowed(S, g)) {
            const size_t c = cell_of(g);
            const Ucg u = calc_ucg(S, g);
            T->lut[2 * c] = (double)u.re;
            T->lut[2 * c + 1] = (double)u.im;
            float _Complex zc;
            __real__ zc = u.re;
            __imag__ zc = u.im;
            if (cabsf(zc) <= 1.0e-10f) T->lut_dbd[c] = 1;
          }
          add_reflection(g);
          refdone[cell_of(g)] = 2;
        }
      }
  // LUT cells of the inner box that no list entry took keep the reference's initial 0
  // (gvectors.f90:1326-1327: only AddReflection's branch writes inner-box cells)
  for (int l = -im[2]; l <= im[2]; l++)
    for (int kk = -im[1]; kk <= im[1]; kk++)
      for (int h = -im[0]; h <= im[0]; h++) {
        const int g[3] = {h, kk, l};
        const size_t c = cell_of(g);
        if (refdone[c] != 2) {
          T->lut[2 * c] = 0.0;
          T->lut[2 * c + 1] = 0.0;
          T->lut_dbd[c] = 0;
        }
      }
  const int nref = (int)T->hkl.size() / 3;
  T->dbd.resize(nref);
  for (int i = 0; i < nref; i++) T->dbd[i] = T->lut_dbd[cell_of(&T->hkl[3 * i])];
  // sites: Z, B and the orbits of CalcOrbit (double) for the structure matrices
  for (int p = 0; p < S.ntype; p++) {
    const std::vector<double> ct = calc_orbit(S, p);
    T->site_z.push_back(S.atom_type[p]);
    T->site_b.push_back(S.atom_pos[5 * p + 4]);
    T->site_norbit.push_back((int32_t)(ct.size() / 3));
    T->orbit.insert(T->orbit.end(), ct.begin(), ct.end());
  }
  *out = T;
  return CTEM_OK;
}

extern "C" int32_t ctem_setup_create(const ctem_xtal* x, float ekev, float dmin, float weakcutoff,
                                     float cutoff, float sgcutoff, ctem_setup** out) {
  return setup_build(x, ekev, (double)(ekev * 1000.0f), nullptr, 0.0f, dmin, weakcutoff, cutoff, sgcutoff, -1,
                     nullptr, out);
}

extern "C" int32_t ctem_setup_create_ecp(const ctem_xtal* x, float voltage, float dmin, const int32_t k[3],
                                         float convang, float weakcutoff, float cutoff, float sgcutoff,
                                         ctem_setup** out) {
  if (!k) return setup_fail(CTEM_E_ARG, "ctem_setup_create_ecp: null zone axis");
  return setup_build(x, voltage / 1000.0f, (double)voltage, k, convang, dmin, weakcutoff, cutoff, sgcutoff, -1,
                     nullptr, out);
}

extern "C" int32_t ctem_setup_create_gpu(int32_t device, const ctem_xtal* x, float ekev, float dmin,
                                         float weakcutoff, float cutoff, float sgcutoff, double* lut_ms,
                                         ctem_setup** out) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return setup_fail(CTEM_E_HIP, "no HIP device available");
  if (device < 0 || device >= ndev) return setup_fail(CTEM_E_ARG, "ctem_setup_create_gpu: device out of range");
  return setup_build(x, ekev, (double)(ekev * 1000.0f), nullptr, 0.0f, dmin, weakcutoff, cutoff, sgcutoff, device,
                     lut_ms, out);
}

extern "C" int32_t ctem_setup_crystal(const ctem_setup* T, ctem_crystal* c) {
  if (!T || !c) return setup_fail(CTEM_E_ARG, "ctem_setup_crystal: null argument");
  c->nref = (int32_t)(T->hkl.size() / 3);
  c->hkl = T->hkl.data();
  c->ucg = T->ucg.data();
  c->dbdiff = T->dbd.data();
  for (int d = 0; d < 3; d++) { c->lut_lo[d] = T->lut_lo[d]; c->lut_dim[d] = T->lut_dim[d]; }
  c->lut = T->lut.data();
  c->mlambda = T->S.lambda;
  c->dynupz = T->dynupz;
  c->weakcutoff = T->weakcutoff;
  c->cutoff = T->cutoff;
  c->sgcutoff = T->sgcutoff;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) c->rmt[3 * i + j] = T->S.rmt[i][j];
  c->nsite = T->S.ntype;
  c->site_z = T->site_z.data();
  c->site_b = T->site_b.data();
  c->site_norbit = T->site_norbit.data();
  c->orbit = T->orbit.data();
  return CTEM_OK;
}

extern "C" int32_t ctem_setup_info(const ctem_setup* T, ctem_setup_values* v) {
  if (!T || !v) return setup_fail(CTEM_E_ARG, "ctem_setup_info: null argument");
  const State& S = T->S;
  v->mlambda = S.lambda; v->mrelcor = S.relcor; v->mpsihat = S.psihat; v->msigma = S.sigma;
  v->vol = S.vol;
  v->dynupz = T->dynupz;
  v->isym = T->isym;
  v->usehex = T->usehex;
  v->nsym = S.matnum;
  v->npt = S.numpt;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      v->dmt[3 * i + j] = S.dmt[i][j];
      v->rmt[3 * i + j] = S.rmt[i][j];
      v->dsm[3 * i + j] = S.dsm[i][j];
      v->rsm[3 * i + j] = S.rsm[i][j];
    }
  return CTEM_OK;
}

extern "C" int32_t ctem_setup_density(const ctem_setup* T, float* dens, float* avZ, float* avA) {
#pragma clang fp contract(off)
  if (!T || !dens || !avZ || !avA) return setup_fail(CTEM_E_ARG, "ctem_setup_density: null argument");
  const State& S = T->S;
  // CalcDensity (crystal.f90:1388-1417): numat are the orbit sizes of CalcPositions, float32 sums
  float aw = 0.0f, z = 0.0f;
  int nat = 0;
  for (int p = 0; p < S.ntype; p++) nat += T->site_norbit[p];
  for (int p = 0; p < S.ntype; p++) {
    const float n = (float)T->site_norbit[p];
    aw = aw + (n * kAtomWeights[S.atom_type[p] - 1]) * S.atom_pos[5 * p + 3];
    z = z + n * (float)S.atom_type[p];
  }
  if (nat == 0) return setup_fail(CTEM_E_ARG, "ctem_setup_density: no atoms");
  *avA = aw / (float)nat;
  *avZ = z / (float)nat;
  *dens = aw / (float)(S.vol * 1.0e-21 * 6.0221413e23);  // cAvogadro of constants.f90:75
  return CTEM_OK;
}

extern "C" int32_t ctem_setup_destroy(ctem_setup* T) {
  delete T;
  return CTEM_OK;
}
