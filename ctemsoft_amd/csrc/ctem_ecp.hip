// Electron channeling patterns (include/ctem_ecp.h): the geometry steps of the reference
// program's set-up (host, float32/float64 as the reference computes them) and the kernels of the
// beam loop that the master-pattern path does not already have:
//   k_ecp_depth_dots   ScatMat: per-depth intensity d_z = Re psi_z^T Sgh conj(psi_z), as the
//                      complex product U = Sgh conj(Psi) on v_mfma_f64_16x16x4_f64 (one 32 x 32
//                      tile of rows x depths per workgroup) and the column dot with Psi
//   k_ecp_scatmat_out  the running sums of d_z over the depths, stored at each thickness's
//                      step count (CalcLghSM's Lghsum, CTEMECP.f90:600-616)
//   k_ecp_augment      Blochwv: the block matrices t [[A, 0], [eps Sgh^T, -A^H]] of the
//                      thickness integral, for t = startthick and t = thickinc
//   k_ecp_bloch_steps  y = E0 e_1, then y <- E y per thickness; Re psi^H phi / eps is the
//                      integral of psi^H Sgh^T psi over 0..t
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ctem_ecp.h"
#include "ctem_internal.h"
#include "ctem_sg.h"

#pragma clang fp contract(off)

namespace ctem {

typedef double double4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int64_t ecp_sidx(const DevLut& L, int h, int k, int l) {
  return (int64_t)(h - L.slo[0]) + (int64_t)L.sdim[0] * ((k - L.slo[1]) + (int64_t)L.sdim[1] * (l - L.slo[2]));
}

// One workgroup per (direction, 32-row tile ti, 32-depth tile tz); four waves, each a 16 x 16
// quadrant of the tile. U(r, z) = sum_c S(r, c) conj(psi_z(c)) with S(r, c) = Sgh(hkl_r - hkl_c)
// from the table; the epilogue forms Re psi_z(r) U(r, z) and sums the tile's rows in a fixed order
// into partial[(b * rtiles + ti) * nzp + z].
__global__ __launch_bounds__(256) void k_ecp_depth_dots(Bucket B, BeamView V, DevRefl R, DevLut L,
                                                        const double2* __restrict__ Psi, int64_t spsi, int nz,
                                                        int ztiles, double* __restrict__ partial) {
  const int npad = B.npad, rtiles = npad / 32, nzp = ztiles * 32;
  const int per_b = rtiles * ztiles;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int b = lb / per_b, t = lb % per_b;
  const int ti = t % rtiles, tz = t / rtiles;
  const int dir = B.dirs[b];
  const int n = V.nns[dir];
  const int32_t* sidx = V.sidx + V.soff[dir];
  const double2* P = Psi + (int64_t)b * spsi;
  __shared__ double2 Sg[32][33];  // S(row, c) of the chunk
  __shared__ double2 Pc[32][33];  // conj psi_z(c): [c][z]
  __shared__ int rh[32][3], chh[32][3];
  __shared__ double red[2][32];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, l16 = lane & 15, l4 = lane >> 4;
  const int wr = (wave >> 1) * 16, wc = (wave & 1) * 16;
  if (tid < 32) {
    const int a = ti * 32 + tid;
    for (int d = 0; d < 3; d++) rh[tid][d] = (a < n) ? R.hkl[3 * sidx[a] + d] : 0;
  }
  double4v ur = {0, 0, 0, 0}, ui = {0, 0, 0, 0};
  const bool live = ti * 32 < n;
  if (live) {
    for (int c0 = 0; c0 < n; c0 += 32) {
      __syncthreads();
      if (tid < 32) {
        const int c = c0 + tid;
        for (int d = 0; d < 3; d++) chh[tid][d] = (c < n) ? R.hkl[3 * sidx[c] + d] : 0;
      }
      __syncthreads();
      for (int e = tid; e < 1024; e += 256) {
        const int cc = e & 31, rr = e >> 5;
        double2 v = make_double2(0.0, 0.0);
        if (ti * 32 + rr < n && c0 + cc < n)
          v = L.slut[ecp_sidx(L, rh[rr][0] - chh[cc][0], rh[rr][1] - chh[cc][1], rh[rr][2] - chh[cc][2])];
        Sg[rr][cc] = v;
        const int zz = e >> 5, z = tz * 32 + zz;
        double2 pv = make_double2(0.0, 0.0);
        if (z < nz && c0 + cc < n) pv = P[(int64_t)z * npad + c0 + cc];
        Pc[cc][zz] = make_double2(pv.x, -pv.y);
      }
      __syncthreads();
      const int kn = min(32, n - c0);
      for (int k4 = 0; k4 < kn; k4 += 4) {
        const double2 a = Sg[wr + l16][k4 + l4];
        const double2 h = Pc[k4 + l4][wc + l16];
        ur = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, h.x, ur, 0, 0, 0);
        ur = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, h.y, ur, 0, 0, 0);
        ui = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, h.y, ui, 0, 0, 0);
        ui = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, h.x, ui, 0, 0, 0);
      }
    }
  }
  // lane l holds rows wr + l/16 + 4q, column (depth) wc + l%16
  const int z = tz * 32 + wc + l16;
  double v = 0.0;
  if (live && z < nz)
    for (int q = 0; q < 4; q++) {
      const int gi = ti * 32 + wr + l4 + 4 * q;
      if (gi < n) {
        const double2 ps = P[(int64_t)z * npad + gi];
        v += ps.x * ur[q] - ps.y * ui[q];
      }
    }
  v += xchg16(v);
  v += xchg32(v);
  __syncthreads();
  if (l4 == 0) red[wave >> 1][wc + l16] = v;
  __syncthreads();
  if (tid < 32) {
    const int zo = tz * 32 + tid;
    if (zo < nz) partial[((int64_t)b * rtiles + ti) * nzp + zo] = red[0][tid] + red[1][tid];
  }
}

// One thread per direction: the running sum over the depths (row tiles added in order), stored
// as fl32(sum / nat) at the thickness index store[z] (-1: none) of each depth.
__global__ void k_ecp_scatmat_out(int nb, const int32_t* __restrict__ dirs, int rtiles, int nz, int nzp,
                                  const double* __restrict__ partial, const int32_t* __restrict__ store,
                                  double nat, int numk, float* __restrict__ sr) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int dir = dirs[b];
  double run = 0.0;
  for (int z = 0; z < nz; z++) {
    double d = 0.0;
    for (int ti = 0; ti < rtiles; ti++) d += partial[((int64_t)b * rtiles + ti) * nzp + z];
    run += d;
    const int t = store[z];
    if (t >= 0) sr[(int64_t)t * numk + dir] = (float)(run / nat);
  }
}

// The block matrices of the Blochwv integral, N = 2 npad, column-major: for direction slot b,
// M[b] = t0 X and M[nb + b] = dt X with X = [[A, 0], [eps S^T, -A^H]], A = A1 (the assembled
// 2 pi i DynMat / fl32(2 kn), npad x npad) and S^T(r, c) = Sgh(hkl_c - hkl_r) on the strong beams.
// One workgroup per 32 x 32 tile of X.
__global__ __launch_bounds__(256) void k_ecp_augment(Bucket B, BeamView V, DevRefl R, DevLut L,
                                                     const double2* __restrict__ A1, double eps, double t0,
                                                     double dt, double2* __restrict__ M) {
  const int npad = B.npad, N = 2 * npad, tiles = N / 32;
  const int b = blockIdx.x / (tiles * tiles);
  const int t = blockIdx.x % (tiles * tiles);
  const int ti = t % tiles, tj = t / tiles;
  const int dir = B.dirs[b];
  const int n = V.nns[dir];
  const int32_t* sidx = V.sidx + V.soff[dir];
  const double2* Ab = A1 + (int64_t)b * npad * npad;
  double2* M0 = M + (int64_t)b * N * N;
  double2* MD = M + (int64_t)(B.nb + b) * N * N;
  for (int e = threadIdx.x; e < 1024; e += 256) {
    const int r = ti * 32 + (e & 31), c = tj * 32 + (e >> 5);
    double2 v = make_double2(0.0, 0.0);
    if (r < npad && c < npad) {
      v = Ab[(int64_t)c * npad + r];
    } else if (r >= npad && c < npad) {
      const int rr = r - npad;
      if (rr < n && c < n) {
        const int32_t* hc = R.hkl + 3 * sidx[c];
        const int32_t* hr = R.hkl + 3 * sidx[rr];
        const double2 sv = L.slut[ecp_sidx(L, hc[0] - hr[0], hc[1] - hr[1], hc[2] - hr[2])];
        v = make_double2(eps * sv.x, eps * sv.y);
      }
    } else if (r >= npad && c >= npad) {
      const double2 a = Ab[(int64_t)(r - npad) * npad + (c - npad)];
      v = make_double2(-a.x, a.y);
    }
    const int64_t o = (int64_t)c * N + r;
    M0[o] = make_double2(t0 * v.x, t0 * v.y);
    MD[o] = make_double2(dt * v.x, dt * v.y);
  }
}

// One workgroup per direction slot: y = E0 e_1, then for each further thickness y <- Ed y, with
// E0 = E[b], Ed = E[nb + b] (N x N, column-major, N = 2 npad); per thickness
// sr[it * numk + dir] = fl32(((Re psi^H phi) / eps / thick[it]) / nat) with psi = y[0..n-1],
// phi = y[npad..npad+n-1]. Threads take rows r = tid % N (and r + 256 ...), columns split in
// G = 256 / N groups for N < 256, partial sums added in group order.
__global__ __launch_bounds__(256) void k_ecp_bloch_steps(Bucket B, BeamView V, const double2* __restrict__ E,
                                                         int nt, const float* __restrict__ thick, double eps,
                                                         double nat, int numk, float* __restrict__ sr) {
  extern __shared__ double2 ys[];  // y [N], y' [N], group partials [G][N]
  const int npad = B.npad, N = 2 * npad;
  const int b = blockIdx.x;
  const int dir = B.dirs[b];
  const int n = V.nns[dir];
  const double2* E0 = E + (int64_t)b * N * N;
  const double2* Ed = E + (int64_t)(B.nb + b) * N * N;
  const int G = N < 256 ? 256 / N : 1;
  double2* y = ys;
  double2* y2 = ys + N;
  double2* gp = ys + 2 * N;
  __shared__ double red[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int r = tid; r < N; r += 256) y[r] = E0[r];
  __syncthreads();
  for (int it = 0; it < nt; it++) {
    if (it > 0) {
      if (G > 1) {
        const int r = tid % N, g = tid / N;
        double sx = 0.0, sy = 0.0;
        if (g < G)
          for (int c = g; c < N; c += G) {
            const double2 e = Ed[(int64_t)c * N + r];
            const double2 yv = y[c];
            sx += e.x * yv.x - e.y * yv.y;
            sy += e.x * yv.y + e.y * yv.x;
          }
        if (g < G) gp[g * N + r] = make_double2(sx, sy);
        __syncthreads();
        if (tid < N) {
          double ax = 0.0, ay = 0.0;
          for (int q = 0; q < G; q++) { ax += gp[q * N + tid].x; ay += gp[q * N + tid].y; }
          y2[tid] = make_double2(ax, ay);
        }
      } else {
        for (int r = tid; r < N; r += 256) {
          double sx = 0.0, sy = 0.0;
          for (int c = 0; c < N; c++) {
            const double2 e = Ed[(int64_t)c * N + r];
            const double2 yv = y[c];
            sx += e.x * yv.x - e.y * yv.y;
            sy += e.x * yv.y + e.y * yv.x;
          }
          y2[r] = make_double2(sx, sy);
        }
      }
      __syncthreads();
      double2* tmp = y; y = y2; y2 = tmp;
    }
    double v = 0.0;
    for (int r = tid; r < n; r += 256) v += y[r].x * y[npad + r].x + y[r].y * y[npad + r].y;
    v = wave_sum(v);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (tid == 0) {
      const double val = ((red[0] + red[1]) + red[2]) + red[3];
      sr[(int64_t)it * numk + dir] = (float)((val / eps / (double)thick[it]) / nat);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------
// MatrixExponential 'Pade' (math.f90:187-315), batched: the diagonal Pade approximant of degree q
// with 2^j scaling, (q, j) from ilev = nint(log10(max |a_ij|)) + 3 (ilev >= 1; qPade, jPade of
// :209-211), E = Dqq^-1 Nqq, then j squarings. The scattering matrix of the channeling-pattern
// program comes from this approximant, whose truncation error at small ||A|| (q = 2, 3 with
// j = 0) is far above double rounding; the same approximant here gives the program's matrix.

// per matrix: modA = max |a_ij| (A already holds A z0), ilev -> q, j; A <- A / 2^j (= B);
// N = D = I; C = B. qm1[b] = q - 1, jj[b] = j. err bit 4: modA > 1000 (the reference stops).
__global__ __launch_bounds__(256) void k_pade_prep(int npad, double2* __restrict__ A, double2* __restrict__ N,
                                                   double2* __restrict__ D, double2* __restrict__ Cb,
                                                   int32_t* __restrict__ qm1, int32_t* __restrict__ jj,
                                                   int32_t* __restrict__ err) {
  const int b = blockIdx.x;
  const int64_t per = (int64_t)npad * npad;
  double2* Ab = A + b * per;
  __shared__ double red[4];
  __shared__ int js;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double m = 0.0;
  for (int64_t e = tid; e < per; e += 256) {
    const double2 v = Ab[e];
    const double a = hypot(v.x, v.y);
    m = (a > m || a != a) ? a : m;
  }
  for (int o = 32; o >= 1; o >>= 1) {
    const double t = __shfl_xor(m, o, 64);
    m = (t > m || t != t) ? t : m;
  }
  if (lane == 0) red[wave] = m;
  __syncthreads();
  if (tid == 0) {
    double mm = red[0];
    for (int w = 1; w < 4; w++) mm = (red[w] > mm || red[w] != red[w]) ? red[w] : mm;
    const int qt[6] = {2, 3, 4, 4, 4, 4}, jt[6] = {0, 0, 1, 5, 8, 11};
    int ilev = 1;
    if (!(mm == mm) || mm > 1000.0) {
      atomicOr(err, 4);
    } else if (mm > 0.0) {
      ilev = (int)round(log10(mm)) + 3;
      if (ilev <= 0) ilev = 1;
      if (ilev > 6) ilev = 6;
    }
    qm1[b] = qt[ilev - 1] - 1;
    jj[b] = jt[ilev - 1];
    js = jt[ilev - 1];
  }
  __syncthreads();
  const double sc = ldexp(1.0, -js);
  double2* Nb = N + b * per;
  double2* Db = D + b * per;
  double2* Cbb = Cb + b * per;
  for (int64_t e = tid; e < per; e += 256) {
    const double2 v = Ab[e];
    const double2 bv = make_double2(v.x * sc, v.y * sc);
    Ab[e] = bv;
    Cbb[e] = bv;
    const bool dg = (e % npad) == (e / npad);
    Nb[e] = make_double2(dg ? 1.0 : 0.0, 0.0);
    Db[e] = Nb[e];
  }
}

// step icnt of the Pade sums: pref = prod_{i <= icnt} (q - i + 1) / i / (2q - i + 1), the
// reference's running product; N += pref C, D += (sgn pref) C with sgn = (-1)^icnt, for the
// matrices with icnt <= q
__global__ __launch_bounds__(256) void k_pade_accum(int npad, int icnt, const int32_t* __restrict__ qm1,
                                                    const double2* __restrict__ Cb, double2* __restrict__ N,
                                                    double2* __restrict__ D) {
  const int b = blockIdx.y;
  const int q = qm1[b] + 1;
  if (icnt > q) return;
  double pref = 1.0, sgn = -1.0;
  for (int i = 1; i <= icnt; i++) {
    pref = pref * (double)(q - i + 1) / (double)i / (double)(2 * q - i + 1);
    if (i < icnt) sgn = -sgn;
  }
  const double sp = sgn * pref;
  const int64_t per = (int64_t)npad * npad;
  const double2* C = Cb + b * per;
  double2* Nb = N + b * per;
  double2* Db = D + b * per;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < per; e += (int64_t)gridDim.x * 256) {
    const double2 c = C[e];
    double2 nv = Nb[e], dv = Db[e];
    nv.x = nv.x + pref * c.x; nv.y = nv.y + pref * c.y;
    dv.x = dv.x + sp * c.x; dv.y = dv.y + sp * c.y;
    Nb[e] = nv;
    Db[e] = dv;
  }
}

// X <- D^-1 X for each matrix (npad x npad, column-major): Gauss-Jordan elimination with partial
// pivoting (largest |re| + |im| in the column, the first of equals). With onchip (npad <= 64) the
// two matrices are held in LDS for the elimination. err bit 2: a zero pivot.
__global__ __launch_bounds__(256) void k_zsolve(int n, double2* __restrict__ Dg, double2* __restrict__ Xg, int onchip,
                                                int32_t* __restrict__ err) {
  extern __shared__ double2 sh[];
  const int b = blockIdx.x;
  const int64_t per = (int64_t)n * n;
  double2* rowD = sh;
  double2* rowX = sh + n;
  double2* mult = sh + 2 * n;
  double2* D = Dg + b * per;
  double2* X = Xg + b * per;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (onchip) {
    double2* Dl = sh + 3 * n;
    double2* Xl = Dl + per;
    for (int64_t e = tid; e < per; e += 256) { Dl[e] = D[e]; Xl[e] = X[e]; }
    D = Dl;
    X = Xl;
  }
  __shared__ double rv[4];
  __shared__ int ri[4];
  __shared__ int piv_s;
  for (int k = 0; k < n; k++) {
    __syncthreads();
    double best = -1.0;
    int bi = n;
    for (int r = k + tid; r < n; r += 256) {
      const double2 v = D[(int64_t)k * n + r];
      const double a = fabs(v.x) + fabs(v.y);
      if (a > best) { best = a; bi = r; }
    }
    for (int o = 32; o >= 1; o >>= 1) {
      const double tb = __shfl_xor(best, o, 64);
      const int ti = __shfl_xor(bi, o, 64);
      if (tb > best || (tb == best && ti < bi)) { best = tb; bi = ti; }
    }
    if (lane == 0) { rv[wave] = best; ri[wave] = bi; }
    __syncthreads();
    if (tid == 0) {
      double bb = rv[0];
      int ii = ri[0];
      for (int w = 1; w < 4; w++)
        if (rv[w] > bb || (rv[w] == bb && ri[w] < ii)) { bb = rv[w]; ii = ri[w]; }
      if (!(bb > 0.0)) { atomicOr(err, 2); ii = k; }
      piv_s = ii;
    }
    __syncthreads();
    const int p = piv_s;
    for (int c = tid; c < n; c += 256) {
      const int64_t ok = (int64_t)c * n + k, op = (int64_t)c * n + p;
      if (c >= k) {
        const double2 dk = D[ok], dp = D[op];
        D[ok] = dp;
        D[op] = dk;
        rowD[c] = dp;
      }
      const double2 xk = X[ok], xp = X[op];
      X[ok] = xp;
      X[op] = xk;
      rowX[c] = xp;
    }
    __syncthreads();
    // reciprocal of the pivot: conj(d) / |d|^2
    const double2 d = rowD[k];
    const double den = d.x * d.x + d.y * d.y;
    const double2 inv = (den > 0.0) ? make_double2(d.x / den, -d.y / den) : make_double2(1.0, 0.0);
    for (int r = tid; r < n; r += 256) mult[r] = (r == k) ? make_double2(0.0, 0.0) : D[(int64_t)k * n + r];
    for (int c = tid; c < n; c += 256) {
      if (c > k) {
        const double2 v = rowD[c];
        const double2 w = make_double2(v.x * inv.x - v.y * inv.y, v.x * inv.y + v.y * inv.x);
        rowD[c] = w;
        D[(int64_t)c * n + k] = w;
      }
      const double2 v = rowX[c];
      const double2 w = make_double2(v.x * inv.x - v.y * inv.y, v.x * inv.y + v.y * inv.x);
      rowX[c] = w;
      X[(int64_t)c * n + k] = w;
    }
    __syncthreads();
    if (tid == 0) D[(int64_t)k * n + k] = make_double2(1.0, 0.0);
    for (int64_t e = tid; e < per; e += 256) {
      const int r = (int)(e % n), c = (int)(e / n);
      if (r == k) continue;
      const double2 m = mult[r];
      if (c > k) {
        const double2 u = rowD[c];
        double2 v = D[e];
        v.x = v.x - (m.x * u.x - m.y * u.y);
        v.y = v.y - (m.x * u.y + m.y * u.x);
        D[e] = v;
      } else if (c == k) {
        D[e] = make_double2(0.0, 0.0);
      }
      const double2 u = rowX[c];
      double2 v = X[e];
      v.x = v.x - (m.x * u.x - m.y * u.y);
      v.y = v.y - (m.x * u.y + m.y * u.x);
      X[e] = v;
    }
  }
  __syncthreads();
  if (onchip)
    for (int64_t e = tid; e < per; e += 256) Xg[b * per + e] = X[e];
}

void launch_pade_prep(hipStream_t s, int32_t npad, int32_t nb, double2* A, double2* N, double2* D, double2* C,
                      int32_t* qm1, int32_t* jj, int32_t* err) {
  hipLaunchKernelGGL(k_pade_prep, dim3(nb), dim3(256), 0, s, npad, A, N, D, C, qm1, jj, err);
}

void launch_pade_accum(hipStream_t s, int32_t npad, int32_t nb, int32_t icnt, const int32_t* qm1, const double2* C,
                       double2* N, double2* D) {
  const int64_t per = (int64_t)npad * npad;
  const int gx = (int)std::min<int64_t>(64, (per + 255) / 256);
  hipLaunchKernelGGL(k_pade_accum, dim3(gx, nb), dim3(256), 0, s, npad, icnt, qm1, C, N, D);
}

void launch_zsolve(hipStream_t s, int32_t npad, int32_t nb, double2* D, double2* X, int32_t* err) {
  const int onchip = npad <= 64 ? 1 : 0;
  const size_t lds = (size_t)3 * npad * sizeof(double2) + (onchip ? (size_t)2 * npad * npad * sizeof(double2) : 0);
  hipLaunchKernelGGL(k_zsolve, dim3(nb), dim3(256), lds, s, npad, D, X, onchip, err);
}

void launch_ecp_depth_dots(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R, const DevLut& L,
                           const double2* Psi, int64_t spsi, int32_t nz, double* partial) {
  const int rtiles = B.npad / 32, ztiles = (nz + 31) / 32;
  hipLaunchKernelGGL(k_ecp_depth_dots, dim3(B.nb * rtiles * ztiles), dim3(256), 0, s, B, V, R, L, Psi, spsi, nz,
                     ztiles, partial);
}

void launch_ecp_scatmat_out(hipStream_t s, const Bucket& B, int32_t nz, const double* partial, const int32_t* store,
                            double nat, int32_t numk, float* sr) {
  const int rtiles = B.npad / 32, nzp = (nz + 31) / 32 * 32;
  hipLaunchKernelGGL(k_ecp_scatmat_out, dim3((B.nb + 127) / 128), dim3(128), 0, s, B.nb, B.dirs, rtiles, nz, nzp,
                     partial, store, nat, numk, sr);
}

void launch_ecp_augment(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R, const DevLut& L,
                        const double2* A1, double eps, double t0, double dt, double2* M) {
  const int tiles = 2 * B.npad / 32;
  hipLaunchKernelGGL(k_ecp_augment, dim3(tiles * tiles * B.nb), dim3(256), 0, s, B, V, R, L, A1, eps, t0, dt, M);
}

void launch_ecp_bloch_steps(hipStream_t s, const Bucket& B, const BeamView& V, const double2* E, int32_t nt,
                            const float* thick, double eps, double nat, int32_t numk, float* sr) {
  const int N = 2 * B.npad, G = N < 256 ? 256 / N : 1;
  const size_t lds = (size_t)(2 + G) * N * sizeof(double2);
  hipLaunchKernelGGL(k_ecp_bloch_steps, dim3(B.nb), dim3(256), lds, s, B, V, E, nt, thick, eps, nat, numk, sr);
}

}  // namespace ctem

// ------------------------------------------------------------------------------------------
// host geometry of ECpattern

namespace {

thread_local std::string g_ecp_err;

int32_t ecp_fail(int32_t code, const char* m) {
  g_ecp_err = m;
  return code;
}

// CalcDotDouble 'r' (crystal.f90:451-469): dot_product(p, matmul(rmt, q))
double dotd_r(const double p[3], const double q[3], const double* rmt) {
  double v[3];
  for (int i = 0; i < 3; i++) v[i] = ((0.0 + rmt[3 * i] * q[0]) + rmt[3 * i + 1] * q[1]) + rmt[3 * i + 2] * q[2];
  return ((0.0 + p[0] * v[0]) + p[1] * v[1]) + p[2] * v[2];
}

// matmul(t, dmt): d(j) = sum_i t(i) dmt(i, j)
void vec_mat(const double t[3], const double* m, double d[3]) {
  for (int j = 0; j < 3; j++) d[j] = ((0.0 + t[0] * m[j]) + t[1] * m[3 + j]) + t[2] * m[6 + j];
}

}  // namespace

extern "C" CTEM_API const char* ctem_ecp_last_error(void) { return g_ecp_err.c_str(); }

extern "C" int32_t ctem_ecp_foil_normal(const int32_t fn[3], const double dmt[9], const double rmt[9], float dynfn[3]) {
  if (!fn || !dmt || !rmt || !dynfn) return ecp_fail(CTEM_E_ARG, "ctem_ecp_foil_normal: null argument");
  const double t[3] = {(double)(float)fn[0], (double)(float)fn[1], (double)(float)fn[2]};
  double d[3];
  vec_mat(t, dmt, d);
  float p[3] = {(float)d[0], (float)d[1], (float)d[2]};
  const float x = std::sqrt(ctem::dot_r(p, p, rmt));
  for (int i = 0; i < 3; i++) dynfn[i] = (x != 0.0f) ? p[i] / x : 0.0f;
  return CTEM_OK;
}

extern "C" int32_t ctem_ecp_shortest_g(const int32_t k[3], const double dmt[9], const double rmt[9], int32_t ga_out[3]) {
  if (!k || !dmt || !rmt || !ga_out) return ecp_fail(CTEM_E_ARG, "ctem_ecp_shortest_g: null argument");
  const int u = k[0], v = k[1], w = k[2];
  int nzero[3], snz = 0;
  for (int i = 0; i < 3; i++) { nzero[i] = (k[i] == 0); snz += nzero[i]; }
  int ga[3], gb[3];
  if (snz == 3) return ecp_fail(CTEM_E_ARG, "ctem_ecp_shortest_g: beam direction [000]");
  if (snz == 0) {
    ga[0] = v; ga[1] = -u; ga[2] = 0;
    gb[0] = w; gb[1] = 0; gb[2] = -u;
  } else if (snz == 1) {
    for (int i = 0; i < 3; i++) { ga[i] = nzero[i]; gb[i] = 1 - nzero[i]; }
    if (nzero[0]) { gb[0] *= 0; gb[1] *= w; gb[2] *= -v; }
    if (nzero[1]) { gb[0] *= w; gb[1] *= 0; gb[2] *= -u; }
    if (nzero[2]) { gb[0] *= v; gb[1] *= -u; gb[2] *= 0; }
  } else {
    if (nzero[0] && nzero[1]) { ga[0] = 1; ga[1] = 0; ga[2] = 0; gb[0] = 0; gb[1] = 1; gb[2] = 0; }
    if (nzero[0] && nzero[2]) { ga[0] = 0; ga[1] = 0; ga[2] = 1; gb[0] = 1; gb[1] = 0; gb[2] = 0; }
    if (nzero[1] && nzero[2]) { ga[0] = 0; ga[1] = 1; ga[2] = 0; gb[0] = 0; gb[1] = 0; gb[2] = 1; }
  }
  // the search over integer combinations (ima, imb, ina, inb) in -5..5 (symmetry.f90:1180-1225):
  // the pair with the smallest squared index sum, the last one in loop order among equals
  constexpr int inm = 5;
  int minsum = 50, ml[4] = {0, 0, 0, 0};
  bool found = false;
  for (int ima = -inm; ima <= inm; ima++)
    for (int imb = -inm; imb <= inm; imb++)
      for (int ina = -inm; ina <= inm; ina++)
        for (int inb = -inm; inb <= inm; inb++) {
          int el[6];
          for (int i = 0; i < 3; i++) {
            el[i] = imb * ga[i] - ima * gb[i];
            el[3 + i] = ina * gb[i] - inb * ga[i];
          }
          const int denom = ina * imb - inb * ima;
          int fit = 100;
          if (denom != 0) {
            float fsum = 0.0f;
            int gi[6];
            for (int i = 0; i < 6; i++) {
              const float fel = (float)el[i] / (float)denom;
              gi[i] = (int)fel;
              fsum += std::fabs((float)gi[i] - fel);
            }
            if (fsum == 0.0f) {
              fit = 0;
              for (int i = 0; i < 6; i++) fit += gi[i] * gi[i];
            }
          }
          if (fit <= minsum) {
            minsum = fit;
            ml[0] = ima; ml[1] = imb; ml[2] = ina; ml[3] = inb;
            found = true;
          }
        }
  const int den = ml[2] * ml[1] - ml[3] * ml[0];
  if (!found || den == 0) return ecp_fail(CTEM_E_RANGE, "ctem_ecp_shortest_g: no lattice basis of the zone found");
  int gone[3], gtwo[3];
  for (int i = 0; i < 3; i++) {
    gone[i] = (ml[1] * ga[i] - ml[0] * gb[i]) / den;
    gtwo[i] = (ml[2] * gb[i] - ml[3] * ga[i]) / den;
  }
  // CalcCross(gone, gtwo, 'r', 'r') then CalcDot(., k, 'r'): swap when negative
  const double cr[3] = {(double)((float)gone[1] * (float)gtwo[2] - (float)gone[2] * (float)gtwo[1]),
                        (double)((float)gone[2] * (float)gtwo[0] - (float)gone[0] * (float)gtwo[2]),
                        (double)((float)gone[0] * (float)gtwo[1] - (float)gone[1] * (float)gtwo[0])};
  double xs[3];
  vec_mat(cr, dmt, xs);
  const float gsave[3] = {(float)xs[0], (float)xs[1], (float)xs[2]};
  const float kf[3] = {(float)k[0], (float)k[1], (float)k[2]};
  const float fit = ctem::dot_r(gsave, kf, rmt);
  for (int i = 0; i < 3; i++) ga_out[i] = (fit < 0.0f) ? gtwo[i] : gone[i];
  return CTEM_OK;
}

extern "C" int32_t ctem_ecp_convergence(const int32_t ga[3], const double rmt[9], double mlambda, float* thetac,
                                        float* ktmax, float* bragg, float* galen) {
  if (!ga || !rmt || !thetac || !ktmax) return ecp_fail(CTEM_E_ARG, "ctem_ecp_convergence: null argument");
  const float g[3] = {(float)ga[0], (float)ga[1], (float)ga[2]};
  const float len = std::sqrt(ctem::dot_r(g, g, rmt));
  const float tt = 2.0f * ::asinf((0.5f * (float)mlambda) * len);  // CalcDiffAngle (diffraction.f90:321-333)
  const float br = tt * 0.5f;
  if (!(br > 0.0f)) return ecp_fail(CTEM_E_ARG, "ctem_ecp_convergence: zero Bragg angle (ga = 0)");
  if (*ktmax != 0.0f) *thetac = ((*ktmax * 2.0f) * br) * 1000.0f;
  else *ktmax = *thetac / (2000.0f * br);
  if (bragg) *bragg = br;
  if (galen) *galen = len;
  return CTEM_OK;
}

extern "C" int32_t ctem_ecp_kvectors(const int32_t k[3], const int32_t ga[3], float ktmax, int32_t npx,
                                     const double dmt[9], const double rmt[9], double mlambda, int32_t cap,
                                     int32_t* ij, double* kvec, double* kt_out, double* kn, int32_t* numk) {
  if (!k || !ga || !dmt || !rmt || !numk) return ecp_fail(CTEM_E_ARG, "ctem_ecp_kvectors: null argument");
  if (npx < 0 || npx > 4096) return ecp_fail(CTEM_E_ARG, "ctem_ecp_kvectors: npx must be 0..4096");
  const double gad[3] = {(double)ga[0], (double)ga[1], (double)ga[2]};
  const double glen = std::sqrt(dotd_r(gad, gad, rmt));
  if (!(glen > 0.0)) return ecp_fail(CTEM_E_ARG, "ctem_ecp_kvectors: ga = 0");
  double gan[3];
  for (int i = 0; i < 3; i++) gan[i] = gad[i] / glen;
  const double delta = 2.0 * (double)ktmax * glen / (2.0 * (double)(float)npx + 1.0);
  const double kd[3] = {(double)k[0], (double)k[1], (double)k[2]};
  double kstar[3], gperp[3], cr[3];
  vec_mat(kd, dmt, kstar);  // TransSpace d -> r
  cr[0] = (gad[1] * kstar[2] - gad[2] * kstar[1]) / 1.0;  // CalcCross r -> r, iv = 0
  cr[1] = (gad[2] * kstar[0] - gad[0] * kstar[2]) / 1.0;
  cr[2] = (gad[0] * kstar[1] - gad[1] * kstar[0]) / 1.0;
  vec_mat(cr, dmt, gperp);
  {
    const double x = std::sqrt(dotd_r(gperp, gperp, rmt));
    if (x != 0.0) for (int i = 0; i < 3; i++) gperp[i] = gperp[i] / x;
    const double y = std::sqrt(dotd_r(kstar, kstar, rmt));
    if (y != 0.0) for (int i = 0; i < 3; i++) kstar[i] = kstar[i] / y;
  }
  const double il2 = 1.0 / (mlambda * mlambda);
  int32_t cnt = 0;
  const bool fill = ij || kvec || kt_out || kn;
  auto add = [&](int i, int j, const float kt[3]) -> bool {
    if (fill && cnt >= cap) return false;
    if (fill) {
      const float l = std::sqrt(ctem::dot_r(kt, kt, rmt));
      const float ktlen = l * l;
      const double sq = std::sqrt(il2 - (double)ktlen);
      double kr[3];
      for (int d = 0; d < 3; d++) kr[d] = (double)(float)((double)kt[d] + sq * kstar[d]);
      if (ij) { ij[2 * cnt] = i; ij[2 * cnt + 1] = j; }
      if (kvec) for (int d = 0; d < 3; d++) kvec[3 * cnt + d] = kr[d];
      if (kt_out) for (int d = 0; d < 3; d++) kt_out[3 * cnt + d] = (double)kt[d];
      if (kn) kn[cnt] = dotd_r(kr, kstar, rmt);
    }
    cnt++;
    return true;
  };
  // the centre (Lauexy = 0): kt = -0 gan - 0 gperp
  {
    float kt[3];
    for (int d = 0; d < 3; d++) kt[d] = (float)(-(0.0 * gan[d]) - 0.0 * gperp[d]);
    if (!add(0, 0, kt)) return ecp_fail(CTEM_E_RANGE, "ctem_ecp_kvectors: cap too small");
  }
  const int nx = 2 * npx;
  const int64_t ijmax = (int64_t)npx * npx;
  for (int i = -nx; i <= nx; i++)
    for (int j = -nx; j <= nx; j++) {
      if (i == 0 && j == 0) continue;
      if ((int64_t)i * i + (int64_t)j * j > ijmax) continue;
      float kt[3];
      const double xi = 0.0 + (double)(float)i * delta, yj = 0.0 + (double)(float)j * delta;
      for (int d = 0; d < 3; d++) kt[d] = (float)(-(xi * gan[d]) - yj * gperp[d]);
      if (!add(i, j, kt)) return ecp_fail(CTEM_E_RANGE, "ctem_ecp_kvectors: cap too small");
    }
  *numk = cnt;
  return CTEM_OK;
}
