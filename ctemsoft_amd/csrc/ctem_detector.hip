// EBSD detector patterns from a master pattern (include/ctem_detector.h): the computation of the
// reference's detector-pattern program (src_programs/CTEMEBSD.f90:117-673) with the orientation
// loop on the GPU.
//
// Set-up (host, once per detector; CTEMEBSD.f90:299-343,411-518): the pixel direction cosines in
// float32 as the program forms them, the Monte Carlo exit-energy histograms interpolated to the
// pixels, num_el and the prefactor. Device layout: the three direction planes [3][npix]; the
// detector histograms of the energy window as [kw/4][npix] float4 (four consecutive bins of a
// pixel in one 16-byte load, coalesced across the threads; kw = the window rounded up to 4,
// zero-padded); the site-summed master energy-innermost, [2npy+1][2npx+1][kp] with the bins from
// Emin on (kp rounded up to 4), so that the four bins of a corner are one 16-byte load and a
// corner's whole energy run is contiguous (the master file stores an energy plane after the
// other; the transposition is done once, at upload).
//
// Patterns (k_ebsd_patterns; CTEMEBSD.f90:525-556): one thread per pixel and group of P
// orientations. Per orientation the thread rotates its direction (quat_mult twice), normalises
// it, folds it into the northern hemisphere, maps it to the Lambert square and forms the
// bilinear weights; then it walks the energy window four bins at a time, reading the pixel's
// histogram weights once for all P orientations and the four corners' bins of each. Every
// float32 step follows the program's order (contraction off; division, square root and atanf
// bit-equal to the host's, ctem_lambert.h; bins summed in order), so the patterns equal the
// program's bit for bit. On the square's right or top edge a corner lies past the end of a row
// of the master; the compiled program's read then wraps into the next row (or energy plane), and
// such positions take a slower path that finds that element; a read outside the whole array (a
// direction on the equator, the top row of the last bin) gives 0, where the program reads
// whatever memory follows its array.
//
// Where the program calls SUM (the direction normalisation, num_el, the site sum of the master,
// the binning), the compiled program -- built at -O0, where each SUM is a call into the Fortran
// runtime -- sums with compensated (Kahan) summation in single precision; the module routines
// built with optimisation (Lambert.f90's norm check, quat_norm) sum plainly in element order.
// Both are restated as such.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "ctem_detector.h"
#include "ctem_lambert.h"

#pragma clang fp contract(off)

using namespace ctem;

namespace {

thread_local std::string g_derr;

int32_t dfail(int32_t code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_derr = buf;
  return code;
}

#define DHIP(expr)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) return dfail(CTEM_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// SUM of a REAL(4) array by the Fortran runtime: compensated summation in element order
struct KahanF {
  float s = 0.0f, c = 0.0f;
  void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

// InitLambertParameters (Lambert.f90:239-245) with the C library, as the reference computes them
double lps_spio2() { return sqrt(4.0 * atan(1.0) * 0.5); }
float lps_spi2_f() { return (float)(sqrt(4.0 * atan(1.0)) * 0.5); }

// Fortran NINT of a double (to the nearest integer, halves away from zero)
int32_t fnint(double x) { return (int32_t)lround(x); }

}  // namespace

struct ctem_detector {
  int32_t device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int32_t numsx = 0, numsy = 0, numEbins = 0, Emin = 1, Emax = 1;
  int64_t num_el = 0;
  double prefactor = 0.0;
  float* d_rg = nullptr;
  float* d_w = nullptr;
  float* d_sr = nullptr;  // [2npy+1][2npx+1][kp]
  int32_t npx = -1, npy = -1, kp = 0;
  float* d_q = nullptr;
  size_t qcap = 0;
  float* d_out = nullptr;
  size_t outcap = 0;
  bool timed = false;
};

namespace {

template <int P>
__global__ __launch_bounds__(256) void k_ebsd_patterns(int npix, int nwin, int npx, int npy, int nE, int e0, int kp,
                                                        const float* __restrict__ rg, const float4* __restrict__ w4,
                                                        const float* __restrict__ srt, const float* __restrict__ quats,
                                                        int n, float scl, float spi2, double prefactor,
                                                        float* __restrict__ out) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int a0 = blockIdx.y * P;
  if (p >= npix) return;
  const int nxr = 2 * npx + 1, nyr = 2 * npy + 1;
  const float4* sr4 = (const float4*)srt;
  const int kp4 = kp >> 2;
  const float qq[4] = {0.0f, rg[p], rg[npix + p], rg[2 * npix + p]};
  int cell[P];
  bool fast[P];
  float dx[P], dy[P], dxm[P], dym[P];
  bool live[P];
#pragma unroll
  for (int a = 0; a < P; a++) {
    live[a] = a0 + a < n;
    const float* q = quats + 4 * (int64_t)(live[a] ? a0 + a : a0);
    const float q2[4] = {q[0], q[1], q[2], q[3]};
    const float q1[4] = {q[0], -q[1], -q[2], -q[3]};  // conjg (quaternions.f90:306-313)
    float t[4], r[4];
    quat_mult_f(qq, q1, t);
    quat_mult_f(q2, t, r);
    float dc[3] = {r[1], r[2], r[3]};
    const float nrm = CTEM_FSQRT(sum3_kahan(dc[0] * dc[0], dc[1] * dc[1], dc[2] * dc[2]));
    dc[0] = CTEM_FDIV(dc[0], nrm);
    dc[1] = CTEM_FDIV(dc[1], nrm);
    dc[2] = CTEM_FDIV(dc[2], nrm);
    if (dc[2] < 0.0f) {
      dc[0] = -dc[0];
      dc[1] = -dc[1];
      dc[2] = -dc[2];
    }
    float ixy[2];
    lambert_sq_inverse(dc, scl, spi2, ixy);
    const int nix = (int)((float)npx + ixy[0]) - npx;
    const int niy = (int)((float)npy + ixy[1]) - npy;
    dx[a] = ixy[0] - (float)nix;
    dy[a] = ixy[1] - (float)niy;
    dxm[a] = 1.0f - dx[a];
    dym[a] = 1.0f - dy[a];
    // int() truncates toward 0, so nix >= -npx; all four corners inside the plane unless on
    // the right or top edge (or a NaN direction)
    fast[a] = nix >= -npx && nix < npx && niy >= -npy && niy < npy;
    cell[a] = (niy + npy) * nxr + (nix + npx);
  }
  // a corner's bin k of the window through the program's column-major offset (row ends wrap)
  const int64_t plane = (int64_t)nxr * nyr, total = plane * nE;
  auto slow = [&](int a, int ox, int oy, int k) -> float {
    const int cx = cell[a] % nxr, cy = cell[a] / nxr;  // cell of corner (nix, niy)
    const int64_t o = (int64_t)(cx + ox) + (int64_t)nxr * (cy + oy) + plane * (e0 + k);
    if (o < 0 || o >= total) return 0.0f;
    const int64_t x = o % nxr, yk = o / nxr, y = yk % nyr, kb = yk / nyr;
    return srt[(y * nxr + x) * kp + (kb - e0)];
  };
  float acc[P];
#pragma unroll
  for (int a = 0; a < P; a++) acc[a] = 0.0f;
  for (int kk = 0; 4 * kk < nwin; kk++) {
    const float4 w = w4[(int64_t)kk * npix + p];
    const float wv[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int a = 0; a < P; a++) {
      float c[4][4];  // corners (nix,niy), (nix+1,niy), (nix,niy+1), (nix+1,niy+1) x 4 bins
      if (fast[a]) {
        const int64_t b = (int64_t)cell[a] * kp4 + kk;
        const float4 v0 = sr4[b], v1 = sr4[b + kp4], v2 = sr4[b + (int64_t)nxr * kp4], v3 = sr4[b + (int64_t)(nxr + 1) * kp4];
        c[0][0] = v0.x; c[0][1] = v0.y; c[0][2] = v0.z; c[0][3] = v0.w;
        c[1][0] = v1.x; c[1][1] = v1.y; c[1][2] = v1.z; c[1][3] = v1.w;
        c[2][0] = v2.x; c[2][1] = v2.y; c[2][2] = v2.z; c[2][3] = v2.w;
        c[3][0] = v3.x; c[3][1] = v3.y; c[3][2] = v3.z; c[3][3] = v3.w;
      } else {
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int k = 4 * kk + e < nwin ? 4 * kk + e : 0;
          c[0][e] = slow(a, 0, 0, k);
          c[1][e] = slow(a, 1, 0, k);
          c[2][e] = slow(a, 0, 1, k);
          c[3][e] = slow(a, 1, 1, k);
        }
      }
#pragma unroll
      for (int e = 0; e < 4; e++)
        if (4 * kk + e < nwin) {
          const float v = c[0][e] * dxm[a] * dym[a] + c[1][e] * dx[a] * dym[a] + c[2][e] * dxm[a] * dy[a] +
                          c[3][e] * dx[a] * dy[a];
          acc[a] = acc[a] + wv[e] * v;
        }
    }
  }
#pragma unroll
  for (int a = 0; a < P; a++)
    if (live[a]) out[(int64_t)(a0 + a) * npix + p] = (float)(prefactor * (double)acc[a]);
}

constexpr int kP = 4;  // orientations per thread

int32_t launch_patterns(ctem_detector* d, int32_t n, const float* dq, float* dout) {
  if (d->npx < 0) return dfail(CTEM_E_STATE, "ctem_detector_patterns: no master pattern set");
  if (n < 0) return dfail(CTEM_E_ARG, "ctem_detector_patterns: n < 0");
  if (n == 0) return CTEM_OK;
  const int npix = d->numsx * d->numsy;
  const int nwin = d->Emax - d->Emin + 1;
  const float scl = (float)((double)(float)d->npx / lps_spio2());
  DHIP(hipSetDevice(d->device));
  const dim3 grid((npix + 255) / 256, (n + kP - 1) / kP);
  if (grid.y > 65535u) return dfail(CTEM_E_RANGE, "ctem_detector_patterns: more than %d orientations per call", 65535 * kP);
  DHIP(hipEventRecord(d->ev0, d->stream));
  hipLaunchKernelGGL(k_ebsd_patterns<kP>, grid, dim3(256), 0, d->stream, npix, nwin, d->npx, d->npy, d->numEbins,
                     d->Emin - 1, d->kp, d->d_rg, (const float4*)d->d_w, d->d_sr, dq, n, scl, lps_spi2_f(),
                     d->prefactor, dout);
  DHIP(hipGetLastError());
  DHIP(hipEventRecord(d->ev1, d->stream));
  d->timed = true;
  return CTEM_OK;
}

int32_t finish(ctem_detector* d) {
  DHIP(hipStreamSynchronize(d->stream));
  DHIP(hipGetLastError());
  return CTEM_OK;
}

}  // namespace

extern "C" {

CTEM_API const char* ctem_detector_last_error(void) { return g_derr.c_str(); }

CTEM_API int32_t ctem_detector_create(int32_t device, const ctem_ebsd_geometry* g, const ctem_mc_energy* mc,
                                      ctem_detector** out) {
  if (!g || !mc || !out || !mc->accum_e) return dfail(CTEM_E_ARG, "ctem_detector_create: null argument");
  *out = nullptr;
  if (g->numsx <= 0 || g->numsy <= 0 || mc->numEbins <= 0 || mc->nsx < 0 || mc->nsy < 0)
    return dfail(CTEM_E_ARG, "ctem_detector_create: empty detector or histogram grid");
  if (!(mc->Ebinsize != 0.0)) return dfail(CTEM_E_ARG, "ctem_detector_create: Ebinsize is 0");
  const int numsx = g->numsx, numsy = g->numsy, nEb = mc->numEbins, nsx = mc->nsx, nsy = mc->nsy;
  const int64_t npix = (int64_t)numsx * numsy;
  // energy window (CTEMEBSD.f90:336-343)
  int Emin = fnint(((double)g->energymin - mc->Ehistmin) / mc->Ebinsize) + 1;
  int Emax = fnint(((double)g->energymax - mc->Ehistmin) / mc->Ebinsize) + 1;
  Emin = std::min(std::max(Emin, 1), nEb);
  Emax = std::min(std::max(Emax, 1), nEb);
  // scintillator coordinates and direction cosines (:411-450)
  std::vector<float> sx(numsx), sy(numsy), rg(3 * npix);
  for (int i = 0; i < numsx; i++) sx[i] = -((g->xpc - (1.0f - (float)numsx) * 0.5f) - (float)i) * g->delta;
  for (int j = 0; j < numsy; j++) sy[j] = ((g->ypc - (1.0f - (float)numsy) * 0.5f) - (float)j) * g->delta;
  const float dtor = 0.0174533f;  // :148
  const float sig = (float)mc->MCsig;
  const float alp = (float)(0.5 * 3.141592653589793238 - (double)((sig - g->thetac) * dtor));
  const float ca = cosf(alp), sa = sinf(alp);
  const float L = g->L, L2 = L * L;
  float* rgx = rg.data();
  float* rgy = rgx + npix;
  float* rgz = rgy + npix;
  for (int j = 0; j < numsy; j++) {
    const float s2 = L2 + sy[j] * sy[j];
    const float Ls = ca * sy[j] + L * sa;
    const float Lc = -sa * sy[j] + L * ca;
    for (int i = 0; i < numsx; i++) {
      const float rhos = 1.0f / sqrtf(s2 + sx[i] * sx[i]);
      const int64_t p = (int64_t)j * numsx + i;
      rgx[p] = Ls * rhos;
      rgy[p] = sx[i] * rhos;
      rgz[p] = Lc * rhos;
    }
  }
  for (int64_t p = 0; p < npix; p++) {
    const float z = sqrtf(rgx[p] * rgx[p] + rgy[p] * rgy[p] + rgz[p] * rgz[p]);
    rgx[p] = rgx[p] / z;
    rgy[p] = rgy[p] / z;
    rgz[p] = rgz[p] / z;
  }
  // exit-energy histograms at the pixels (:469-503): Lambert position with the axes swapped,
  // bilinear weights, bins Emin..Emax; accum_e read at its Fortran offset (a row's end wraps)
  const int nwin = Emax - Emin + 1, kw = (nwin + 3) & ~3;
  std::vector<float> wdet((size_t)kw * npix, 0.0f);  // [kw/4][npix][4]
  auto widx = [&](int64_t p, int k) { return ((size_t)(k >> 2) * npix + p) * 4 + (k & 3); };
  const float scl = (float)((double)(float)nsx / lps_spio2());
  const float spi2 = lps_spi2_f();
  const int64_t mtot = (int64_t)nEb * (2 * nsx + 1) * (2 * nsy + 1);
  auto acc_e = [&](int k, int ix, int iy) -> float {  // outside the array: 0, as for the master
    const int64_t o = (int64_t)k + (int64_t)nEb * ((ix + nsx) + (int64_t)(2 * nsx + 1) * (iy + nsy));
    return (o >= 0 && o < mtot) ? (float)mc->accum_e[o] : 0.0f;
  };
  for (int j = 0; j < numsy; j++)
    for (int i = 0; i < numsx; i++) {
      const int64_t p = (int64_t)j * numsx + i;
      float dc[3] = {rgx[p], rgy[p], rgz[p]};
      if (dc[2] < 0.0f) {
        dc[0] = -dc[0];
        dc[1] = -dc[1];
        dc[2] = -dc[2];
      }
      float ixy[2];
      lambert_sq_inverse(dc, scl, spi2, ixy);
      const float x = ixy[0];
      ixy[0] = ixy[1];
      ixy[1] = -x;
      const int nix = (int)((float)nsx + ixy[0]) - nsx;
      const int niy = (int)((float)nsy + ixy[1]) - nsy;
      const float dx = ixy[0] - (float)nix, dy = ixy[1] - (float)niy;
      const float dxm = 1.0f - dx, dym = 1.0f - dy;
      for (int k = Emin - 1; k < Emax; k++)
        wdet[widx(p, k - (Emin - 1))] =
            (acc_e(k, nix, niy) * dxm * dym + acc_e(k, nix + 1, niy) * dx * dym + acc_e(k, nix, niy + 1) * dxm * dy +
             acc_e(k, nix + 1, niy + 1) * dx * dy) * 0.25f;
    }
  // num_el = nint(sum(accum_e_detector)) over the program's array (energy fastest, the bins
  // outside the window zero) and the prefactor (:508,525)
  KahanF tot;
  for (int64_t p = 0; p < npix; p++)
    for (int k = 0; k < nEb; k++) tot.add((k >= Emin - 1 && k < Emax) ? wdet[widx(p, k - (Emin - 1))] : 0.0f);
  // (the program's num_el is a default integer: past 2^31 its NINT is undefined; here 64 bits)
  const int64_t num_el = (int64_t)llroundf(tot.s);
  const double nAmpere = 6.241e18;
  const double prefactor = 0.25 * nAmpere * g->beamcurrent * g->dwelltime * 1.0e-15 / (double)num_el;

  ctem_detector* d = new ctem_detector();
  d->device = device;
  d->numsx = numsx;
  d->numsy = numsy;
  d->numEbins = nEb;
  d->Emin = Emin;
  d->Emax = Emax;
  d->num_el = num_el;
  d->prefactor = prefactor;
  auto bad = [&](int32_t rc) {
    ctem_detector_destroy(d);
    return rc;
  };
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return bad(dfail(CTEM_E_HIP, "hipSetDevice(%d): %s", device, hipGetErrorString(e)));
  if ((e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreate(&d->ev0)) != hipSuccess || (e = hipEventCreate(&d->ev1)) != hipSuccess ||
      (e = hipMalloc(&d->d_rg, sizeof(float) * rg.size())) != hipSuccess ||
      (e = hipMalloc(&d->d_w, sizeof(float) * wdet.size())) != hipSuccess ||
      (e = hipMemcpy(d->d_rg, rg.data(), sizeof(float) * rg.size(), hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d->d_w, wdet.data(), sizeof(float) * wdet.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return bad(dfail(CTEM_E_HIP, "ctem_detector_create: %s", hipGetErrorString(e)));
  *out = d;
  return CTEM_OK;
}

CTEM_API int32_t ctem_detector_info(const ctem_detector* d, int32_t* emin, int32_t* emax, int64_t* num_el,
                                    double* prefactor) {
  if (!d) return dfail(CTEM_E_ARG, "ctem_detector_info: null detector");
  if (emin) *emin = d->Emin;
  if (emax) *emax = d->Emax;
  if (num_el) *num_el = d->num_el;
  if (prefactor) *prefactor = d->prefactor;
  return CTEM_OK;
}

CTEM_API int32_t ctem_detector_set_master(ctem_detector* d, int32_t npx, int32_t npy, int32_t nE, int32_t numset,
                                          const float* srtmp) {
  if (!d || !srtmp) return dfail(CTEM_E_ARG, "ctem_detector_set_master: null argument");
  if (npx < 1 || npy < 1 || numset < 1) return dfail(CTEM_E_ARG, "ctem_detector_set_master: empty master");
  if (nE != d->numEbins)
    return dfail(CTEM_E_ARG, "ctem_detector_set_master: %d energy bins, the histograms have %d", nE, d->numEbins);
  const int64_t plane = (int64_t)(2 * npx + 1) * (2 * npy + 1), per = plane * nE;
  // bins Emin..nE (a row's end wraps into later bins only), energy innermost
  const int e0 = d->Emin - 1, kp = (nE - e0 + 3) & ~3;
  std::vector<float> sr((size_t)plane * kp, 0.0f);
  for (int k = e0; k < nE; k++)
    for (int64_t c = 0; c < plane; c++) {  // sr = sum(srtmp, 4), each element over the sites
      const int64_t e = c + plane * k;
      KahanF s;
      for (int q = 0; q < numset; q++) s.add(srtmp[e + per * q]);
      sr[(size_t)c * kp + (k - e0)] = s.s;
    }
  DHIP(hipSetDevice(d->device));
  if (d->d_sr) (void)hipFree(d->d_sr);
  d->d_sr = nullptr;
  d->npx = d->npy = -1;
  DHIP(hipMalloc(&d->d_sr, sizeof(float) * sr.size()));
  DHIP(hipMemcpy(d->d_sr, sr.data(), sizeof(float) * sr.size(), hipMemcpyHostToDevice));
  d->npx = npx;
  d->npy = npy;
  d->kp = kp;
  return CTEM_OK;
}

CTEM_API int32_t ctem_detector_patterns(ctem_detector* d, int32_t n, const float* quats, float* patterns) {
  if (!d || (n > 0 && (!quats || !patterns))) return dfail(CTEM_E_ARG, "ctem_detector_patterns: null argument");
  if (n <= 0) return n < 0 ? dfail(CTEM_E_ARG, "ctem_detector_patterns: n < 0") : CTEM_OK;
  const size_t npix = (size_t)d->numsx * d->numsy;
  DHIP(hipSetDevice(d->device));
  if ((size_t)n * 4 > d->qcap) {
    if (d->d_q) (void)hipFree(d->d_q);
    d->d_q = nullptr;
    d->qcap = 0;
    DHIP(hipMalloc(&d->d_q, sizeof(float) * 4 * (size_t)n));
    d->qcap = 4 * (size_t)n;
  }
  if ((size_t)n * npix > d->outcap) {
    if (d->d_out) (void)hipFree(d->d_out);
    d->d_out = nullptr;
    d->outcap = 0;
    DHIP(hipMalloc(&d->d_out, sizeof(float) * npix * (size_t)n));
    d->outcap = npix * (size_t)n;
  }
  DHIP(hipMemcpyAsync(d->d_q, quats, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice, d->stream));
  int32_t rc = launch_patterns(d, n, d->d_q, d->d_out);
  if (rc) return rc;
  DHIP(hipMemcpyAsync(patterns, d->d_out, sizeof(float) * npix * (size_t)n, hipMemcpyDeviceToHost, d->stream));
  return finish(d);
}

CTEM_API int32_t ctem_detector_patterns_device(ctem_detector* d, int32_t n, const float* d_quats, float* d_patterns) {
  if (!d || (n > 0 && (!d_quats || !d_patterns))) return dfail(CTEM_E_ARG, "ctem_detector_patterns_device: null argument");
  return launch_patterns(d, n, d_quats, d_patterns);
}

CTEM_API int32_t ctem_detector_sync(ctem_detector* d) {
  if (!d) return dfail(CTEM_E_ARG, "ctem_detector_sync: null detector");
  DHIP(hipSetDevice(d->device));
  return finish(d);
}

CTEM_API int32_t ctem_detector_last_ms(const ctem_detector* d, double* ms) {
  if (!d || !ms) return dfail(CTEM_E_ARG, "ctem_detector_last_ms: null argument");
  if (!d->timed) return dfail(CTEM_E_STATE, "ctem_detector_last_ms: no patterns computed yet");
  float t = 0.0f;
  DHIP(hipEventSynchronize(d->ev1));
  DHIP(hipEventElapsedTime(&t, d->ev0, d->ev1));
  *ms = t;
  return CTEM_OK;
}

CTEM_API int32_t ctem_detector_destroy(ctem_detector* d) {
  if (!d) return CTEM_OK;
  (void)hipSetDevice(d->device);
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  for (void* p : {(void*)d->d_rg, (void*)d->d_w, (void*)d->d_sr, (void*)d->d_q, (void*)d->d_out})
    if (p) (void)hipFree(p);
  if (d->ev0) (void)hipEventDestroy(d->ev0);
  if (d->ev1) (void)hipEventDestroy(d->ev1);
  if (d->stream) (void)hipStreamDestroy(d->stream);
  delete d;
  return CTEM_OK;
}

// ---- host routines ----------------------------------------------------------------------

CTEM_API int32_t ctem_eu2qu_deg(int32_t n, const float* eul_deg, float* quats) {
  if (n < 0 || (n > 0 && (!eul_deg || !quats))) return dfail(CTEM_E_ARG, "ctem_eu2qu_deg: bad argument");
  const float dtor = 0.0174533f;
  for (int32_t a = 0; a < n; a++) {  // eu2qu (rotations.f90:1322-1351) of eulang * dtor
    float e[3];
    for (int c = 0; c < 3; c++) e[c] = (eul_deg[3 * a + c] * dtor) * 0.5f;
    const float cPhi = cosf(e[1]), sPhi = sinf(e[1]);
    const float cm = cosf(e[0] - e[2]), sm = sinf(e[0] - e[2]);
    const float cp = cosf(e[0] + e[2]), sp = sinf(e[0] + e[2]);
    const float q = cPhi * cp;
    float* r = quats + 4 * a;
    if (q < 0.0f) {
      r[0] = -q; r[1] = sPhi * cm; r[2] = sPhi * sm; r[3] = cPhi * sp;
    } else {
      r[0] = q; r[1] = -sPhi * cm; r[2] = -sPhi * sm; r[3] = -cPhi * sp;
    }
  }
  return CTEM_OK;
}

CTEM_API int32_t ctem_apply_axisangle(int32_t n, const float* axisangle, float* quats) {
  if (n < 0 || !axisangle || (n > 0 && !quats)) return dfail(CTEM_E_ARG, "ctem_apply_axisangle: bad argument");
  if (axisangle[3] == 0.0f) return CTEM_OK;
  const float a[4] = {axisangle[0], axisangle[1], axisangle[2], axisangle[3] * 0.0174533f};
  // ax2om (rotations.f90), then om2qu with quat_norm's plain sum (quaternions.f90)
  const float c = cosf(a[3]), s = sinf(a[3]), omc = 1.0f - c;
  float m[3][3];
  for (int i = 0; i < 3; i++) m[i][i] = a[i] * a[i] * omc + c;
  float q = omc * a[0] * a[1];
  m[0][1] = q + s * a[2];
  m[1][0] = q - s * a[2];
  q = omc * a[1] * a[2];
  m[1][2] = q + s * a[0];
  m[2][1] = q - s * a[0];
  q = omc * a[2] * a[0];
  m[2][0] = q + s * a[1];
  m[0][2] = q - s * a[1];
  float r[4];
  r[0] = sqrtf(m[0][0] + m[1][1] + m[2][2] + 1.0f) / 2.0f;
  r[1] = sqrtf(fabsf(m[0][0] - m[1][1] - m[2][2] + 1.0f)) / 2.0f;
  r[2] = sqrtf(fabsf(-m[0][0] + m[1][1] - m[2][2] + 1.0f)) / 2.0f;
  r[3] = sqrtf(fabsf(-m[0][0] - m[1][1] + m[2][2] + 1.0f)) / 2.0f;
  if (m[2][1] - m[1][2] < 0.0f) r[1] = -r[1];
  if (m[0][2] - m[2][0] < 0.0f) r[2] = -r[2];
  if (m[1][0] - m[0][1] < 0.0f) r[3] = -r[3];
  const float inv = 1.0f / sqrtf(((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]) + r[3] * r[3]);
  for (int i = 0; i < 4; i++) r[i] = inv * r[i];
  for (int32_t k = 0; k < n; k++) {
    float* x = quats + 4 * k;
    float t[4];
    quat_mult_f(r, x, t);
    memcpy(x, t, sizeof t);
  }
  return CTEM_OK;
}

namespace {
// The saved variables of the program's noise routines (noise.f90:136-228): ran's generator
// state and am, POIDEV's last mean and its derived values
struct NoiseState {
  int32_t idum, ix, iy, started;
  float am, oldm, g, sq, alxm;
};
static_assert(sizeof(NoiseState) <= 14 * sizeof(int64_t), "noise state must fit the caller's array");

// Park and Miller's minimal standard generator combined with Marsaglia's xorshift, as ran
float nr_ran(NoiseState& s) {
  const int32_t IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836;
  if (!s.started) {
    s.ix = -1;
    s.iy = -1;
    s.started = 1;
  }
  if (s.idum <= 0 || s.iy < 0) {
    s.am = nextafterf(1.0f, 0.0f) / (float)IM;
    const int32_t a = s.idum < 0 ? -s.idum : s.idum;
    s.iy = (888889999 ^ a) | 1;
    s.ix = 777755555 ^ a;
    s.idum = a + 1;
  }
  uint32_t x = (uint32_t)s.ix;
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  s.ix = (int32_t)x;
  const int32_t k = s.iy / IQ;
  s.iy = IA * (s.iy - k * IQ) - IR * k;
  if (s.iy < 0) s.iy += IM;
  return s.am * (float)((IM & (s.ix ^ s.iy)) | 1);
}

// ln Gamma(x) by Pike and Hill's algorithm 291 (alogam), x > 0
double lgamma_ph(double x) {
  if (x <= 0.0) return 0.0;
  double y = x, f = 0.0, z;
  if (x < 7.0) {
    f = 1.0;
    for (z = y; z < 7.0; z = z + 1.0) f = f * z;
    y = z;
    f = -log(f);
  }
  z = 1.0 / y / y;
  return f + (y - 0.5) * log(y) - y + 0.918938533204673 +
         (((-0.000595238095238 * z + 0.000793650793651) * z - 0.002777777777778) * z + 0.083333333333333) / y;
}

// Numerical Recipes' Poisson deviate (POIDEV): the product of uniforms below a mean of 20,
// rejection from a Lorentzian above it; G, SQ and ALXM are kept for a repeated mean
float nr_poidev(float xm, NoiseState& s) {
  const float PI = 3.141592654f;
  float em, t, y;
  if (xm < 20.0f) {
    if (xm != s.oldm) {
      s.oldm = xm;
      s.g = expf(-xm);
    }
    em = -1.0f;
    t = 1.0f;
    do {
      em = em + 1.0f;
      t = t * nr_ran(s);
    } while (!(t <= s.g));
    return em;
  }
  if (xm != s.oldm) {
    s.oldm = xm;
    s.sq = sqrtf(2.0f * xm);
    s.alxm = logf(xm);
    s.g = (float)((double)(xm * s.alxm) - lgamma_ph((double)(xm + 1.0f)));
  }
  for (;;) {
    do {
      y = tanf(PI * nr_ran(s));
      em = s.sq * y + xm;
    } while (!(em >= 0.0f));
    em = (float)(int32_t)em;
    t = (float)((double)(0.9f * (1.0f + y * y)) * exp((double)(em * s.alxm) - lgamma_ph((double)(em + 1.0f)) - (double)s.g));
    if (nr_ran(s) <= t) return em;
  }
}
}  // namespace

CTEM_API int32_t ctem_poisson_noise(int32_t numsx, int32_t numsy, int32_t npat, float* patterns, int64_t* state) {
  if (numsx <= 0 || numsy <= 0 || npat < 0 || !state || (npat > 0 && !patterns))
    return dfail(CTEM_E_ARG, "ctem_poisson_noise: bad argument");
  const int64_t npix = (int64_t)numsx * numsy;
  for (int64_t e = 0; e < npix * npat; e++)
    if (!(patterns[e] >= 0.0f) || patterns[e] == INFINITY)
      return dfail(CTEM_E_ARG, "ctem_poisson_noise: pattern value %g at %lld is not finite and non-negative",
                   (double)patterns[e], (long long)e);
  NoiseState s;
  if (state[1] == 0) {  // the saved variables' initial values
    memset(&s, 0, sizeof s);
    s.oldm = -1.0f;
  } else {
    memcpy(&s, &state[2], sizeof s);
  }
  s.idum = (int32_t)state[0];
  for (int32_t a = 0; a < npat; a++) {
    float* pat = patterns + npix * a;
    for (int i = 0; i < numsx; i++)
      for (int j = 0; j < numsy; j++) pat[(int64_t)j * numsx + i] = nr_poidev(pat[(int64_t)j * numsx + i], s);
  }
  state[0] = s.idum;
  state[1] = 1;
  memcpy(&state[2], &s, sizeof s);
  return CTEM_OK;
}

CTEM_API int32_t ctem_bin_pattern(int32_t numsx, int32_t numsy, int32_t binning, int32_t gamma, float gammavalue,
                                  const float* pattern, float* binned) {
  if (numsx <= 0 || numsy <= 0 || binning <= 0 || !pattern || !binned)
    return dfail(CTEM_E_ARG, "ctem_bin_pattern: bad argument");
  // the program's blocks past the pattern's edge read and write outside its arrays
  if (numsx % binning || numsy % binning)
    return dfail(CTEM_E_ARG, "ctem_bin_pattern: binning %d does not divide %d x %d", binning, numsx, numsy);
  const int binx = numsx / binning, biny = numsy / binning;
  const float bindx = 1.0f / ((float)binning * (float)binning);
  if (binning == 1) {
    memcpy(binned, pattern, sizeof(float) * (size_t)numsx * numsy);
  } else {
    // binned(i/binning+1, j/binning+1) = sum(EBSDpattern(i:i+binning-1, j:j+binning-1)), then
    // binned * bindx
    for (int jb = 0; jb < biny; jb++)
      for (int ib = 0; ib < binx; ib++) {
        KahanF s;
        for (int j = jb * binning; j < jb * binning + binning; j++)
          for (int i = ib * binning; i < ib * binning + binning; i++) s.add(pattern[(int64_t)j * numsx + i]);
        binned[(int64_t)jb * binx + ib] = s.s * bindx;
      }
  }
  if (gamma) {
    const int64_t nb = (int64_t)binx * biny;
    for (int64_t e = 0; e < nb; e++) binned[e] = powf(binned[e], gammavalue);
  }
  return CTEM_OK;
}

}  // extern "C"
