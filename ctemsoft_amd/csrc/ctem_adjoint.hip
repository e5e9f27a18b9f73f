// Reverse-mode kernels of the master-pattern path (FP64 form, and the mixed form's own kernels
// below): the gradient of a real function
// f(dots) of the per-direction, per-site intensities dots_p = Re sum_gh Lgh(g,h) S_p(g,h)
// with respect to the complex LUT entries U(g) and the per-site Debye-Waller factors B_p.
//
// Forward (per direction, beam set fixed):
//   D(U)   Bethe-corrected DynMat (gvectors.f90:1123-1173)
//   A      = alpha D,  alpha = 2 pi i dz / fl32(2 kn)
//   E      = expm(A),  psi_1 = E e_1,  psi_{k+1} = E psi_k
//   Lgh    = c sum_k lam_k psi_k psi_k^H,  c = dz / thick
//   dots_p = Re sum_gh Lgh(g,h) S_p(g-h)
// Complex adjoints x_bar are taken as df/dRe(x) + i df/dIm(x). With g_p = df/ddots_p:
//   Lgh_bar = sum_p g_p conj(S_p),  M = c (Lgh_bar + Lgh_bar^H)
//   phi_Z = lam_Z M psi_Z,  phi_k = lam_k M psi_k + E^H phi_{k+1}
//   E_bar = sum_k phi_k psi_{k-1}^H  (psi_0 = e_1)
//   A_bar = L_exp(A^H, E_bar)  (the Frechet derivative of expm at A^H in the direction E_bar,
//           the upper-right block of expm([[A^H, E_bar], [0, A^H]]))
//   D_bar = conj(alpha) A_bar
//   U_bar(g-h) += D_bar(g,h)                                       (g != h)
//   U_bar(g-w) += -lam/2 sum_{h != g} D_bar(g,h) conj(U(w-h)) / s_w  (Bethe, g != h)
//   U_bar(w-h) += -lam/2 sum_{g != h} conj(U(g-w)) D_bar(g,h) / s_w
//   U_bar(g-w) += -lam/s_w Re(D_bar(g,g)) U(g-w)                    (diagonal |U|^2 term)
//   B_bar_p    = g_p Re sum_gh Lgh(g,h) dS_p/dB_p (g-h)
// LUT cells are addressed as the forward assembly reads them (column-major offsets of the
// Fortran bounds, checked against the allocation).
#include <hip/hip_runtime.h>

#include "ctem_internal.h"

namespace ctem {

__device__ __forceinline__ int lut_cell(const DevLut& L, int h, int k, int l) {
  const int idx = (h - L.lo[0]) + L.dim[0] * ((k - L.lo[1]) + L.dim[1] * (l - L.lo[2]));
  return ((unsigned)idx < (unsigned)L.ncell) ? idx : -1;
}

__device__ __forceinline__ double2 lut_val(const DevLut& L, int idx) {
  if (idx < 0) return make_double2(0.0, 0.0);
  const float2 u = L.lut[idx];
  return make_double2((double)u.x, (double)u.y);
}

__device__ __forceinline__ void lut_add(double* glut, int idx, double re, double im) {
  if (idx < 0) return;
  atomicAdd(glut + 2 * idx, re);
  atomicAdd(glut + 2 * idx + 1, im);
}

// The LUT gradient is accumulated in nrep copies (glut[replica][2 ncell]), the workgroup's
// linear index choosing the copy: every direction adds into the same few hundred cells of small
// g - h, and one copy for the whole GPU serialises those atomics (the direct term alone took
// three times the rest of its kernel). Workgroups are dispatched to the XCDs round-robin, so with
// a multiple of 8 copies each copy is written from one XCD. launch_adj_lut_reduce sums them.
__device__ __forceinline__ double* lut_replica(double* glut, int ncell, int nrep) {
  const unsigned lin = blockIdx.x + gridDim.x * blockIdx.y;
  return glut + (size_t)(lin % (unsigned)nrep) * 2 * (size_t)ncell;
}

__device__ __forceinline__ int64_t s_index(const DevLut& L, int h, int k, int l) {
  return (int64_t)(h - L.slo[0]) + (int64_t)L.sdim[0] * ((k - L.slo[1]) + (int64_t)L.sdim[1] * (l - L.slo[2]));
}

// M = c sum_p g_p (conj(S_p(g-h)) + S_p(h-g)) on the n x n block, zero in the padding
__global__ void k_adj_m(Bucket B, BeamView V, DevRefl R, DevLut L, const double* __restrict__ gdots,
                        double c, double2* __restrict__ M) {
  const int npad = B.npad, b = blockIdx.y;
  const int dir = B.dirs[b];
  const int n = V.nns[dir];
  const int32_t* sidx = V.sidx + V.soff[dir];
  const int64_t per = (int64_t)npad * npad;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < per; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % npad), j = (int)(e / npad);
    double2 v = make_double2(0.0, 0.0);
    if (i < n && j < n) {
      const int32_t* gi = R.hkl + 3 * sidx[i];
      const int32_t* gj = R.hkl + 3 * sidx[j];
      const int64_t cp = s_index(L, gi[0] - gj[0], gi[1] - gj[1], gi[2] - gj[2]);
      const int64_t cm = s_index(L, gj[0] - gi[0], gj[1] - gi[1], gj[2] - gi[2]);
      for (int p = 0; p < L.nsite; p++) {
        const double gp = gdots[(int64_t)dir * L.nsite + p];
        const double2 sp = L.slut[(int64_t)p * L.sncell + cp];
        const double2 sm = L.slut[(int64_t)p * L.sncell + cm];
        v.x += c * gp * (sp.x + sm.x);
        v.y += c * gp * (-sp.y + sm.y);
      }
    }
    M[(int64_t)b * per + e] = v;
  }
}

// Hm (nz x npad, column-major): row k holds psi_k^H, psi_0 = e_1
__global__ void k_adj_hm(int npad, int nz, const double2* __restrict__ Psi, int64_t spsi,
                         double2* __restrict__ Hm) {
  const int b = blockIdx.y;
  const int64_t tot = (int64_t)npad * nz;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(e % nz), j = (int)(e / nz);
    double2 v;
    if (k == 0) v = make_double2(j == 0 ? 1.0 : 0.0, 0.0);
    else {
      const double2 p = Psi[(int64_t)b * spsi + (int64_t)(k - 1) * npad + j];
      v = make_double2(p.x, -p.y);
    }
    Hm[(int64_t)b * tot + e] = v;
  }
}

// backward depth recursion: Phi[:,k-1] = phi_k, phi_Z = lam_Z Q_Z, phi_k = lam_k Q_k + E^H phi_{k+1}
// (Q = M Psi). One workgroup per matrix, phi in LDS; (E^H phi)_r = sum_c conj(E(c, r)) phi_c is the
// dot product of column r of E (contiguous) with phi: each wave takes rows r = wave, wave + 4, ..,
// its lanes read the column in 64-entry runs and the wave sums them in a fixed order.
__global__ __launch_bounds__(256) void k_adj_depth(int npad, int nz, const double2* __restrict__ E,
                                                   const double2* __restrict__ Q, const double* __restrict__ lam,
                                                   double2* __restrict__ Phi) {
  extern __shared__ double2 ph[];  // [2 npad]: phi_{k+1} and phi_k
  const int b = blockIdx.x;
  const double2* Eb = E + (int64_t)b * npad * npad;
  const double2* Qb = Q + (int64_t)b * npad * nz;
  double2* Pb = Phi + (int64_t)b * npad * nz;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double2* cur = ph;
  double2* nxt = ph + npad;
  for (int r = tid; r < npad; r += blockDim.x) {
    const double2 q = Qb[(int64_t)(nz - 1) * npad + r];
    const double2 v = make_double2(lam[nz - 1] * q.x, lam[nz - 1] * q.y);
    cur[r] = v;
    Pb[(int64_t)(nz - 1) * npad + r] = v;
  }
  __syncthreads();
  for (int k = nz - 2; k >= 0; k--) {
    const double lk = lam[k];
    for (int r = wave; r < npad; r += 4) {
      const double2* col = Eb + (int64_t)r * npad;
      double sr = 0.0, si = 0.0;
      for (int c = lane; c < npad; c += 64) {
        const double2 e = col[c], x = cur[c];
        // conj(e) * x
        sr += e.x * x.x + e.y * x.y;
        si += e.x * x.y - e.y * x.x;
      }
      sr = wave_sum(sr);
      si = wave_sum(si);
      if (lane == 0) {
        const double2 q = Qb[(int64_t)k * npad + r];
        const double2 v = make_double2(lk * q.x + sr, lk * q.y + si);
        nxt[r] = v;
        Pb[(int64_t)k * npad + r] = v;
      }
    }
    __syncthreads();
    double2* t = cur; cur = nxt; nxt = t;
  }
}

__global__ __launch_bounds__(256) void k_adj_pair(int npad, const double2* __restrict__ A,
                                                  const double2* __restrict__ Eb, double* __restrict__ sigma,
                                                  double2* __restrict__ PX, double2* __restrict__ PY) {
  const int b = blockIdx.x;
  const int64_t per = (int64_t)npad * npad;
  const double2* Ab = A + b * per;
  const double2* Gb = Eb + b * per;
  __shared__ double red[2][4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double na = 0.0, ng = 0.0;
  for (int j = wave; j < npad; j += 4) {
    double sa = 0.0, sg = 0.0;
    for (int i = lane; i < npad; i += 64) {
      const double2 a = Ab[(int64_t)j * npad + i], g = Gb[(int64_t)j * npad + i];
      sa += hypot(a.x, a.y);
      sg += hypot(g.x, g.y);
    }
    for (int o = 32; o > 0; o >>= 1) { sa += __shfl_xor(sa, o); sg += __shfl_xor(sg, o); }
    na = fmax(na, sa);
    ng = fmax(ng, sg);
  }
  if (lane == 0) { red[0][wave] = na; red[1][wave] = ng; }
  __syncthreads();
  na = fmax(fmax(red[0][0], red[0][1]), fmax(red[0][2], red[0][3]));
  ng = fmax(fmax(red[1][0], red[1][1]), fmax(red[1][2], red[1][3]));
  double sg = 0.0;
  if (ng > 0.0 && isfinite(ng)) {
    int ea, eg;
    frexp(na > 0.0 ? na : 1.0, &ea);
    frexp(ng, &eg);
    sg = ldexp(1.0, (ea - 1) - eg);
  }
  if (tid == 0) sigma[b] = sg;
  for (int64_t e = tid; e < per; e += 256) {
    const int i = (int)(e % npad), j = (int)(e / npad);
    const double2 a = Ab[(int64_t)i * npad + j];  // A^H(i,j) = conj(A(j,i))
    PX[b * per + e] = make_double2(a.x, -a.y);
    const double2 g = Gb[e];
    PY[b * per + e] = make_double2(sg * g.x, sg * g.y);
  }
}

// ||[[X, Y], [0, X]]||_1 = max_j (sum_i |X_ij| + |Y_ij|), the squaring count s with
// norm / 2^s <= theta, and X, Y scaled by 2^-s in place (error bits as k_norm_scale)
__global__ __launch_bounds__(256) void k_norm_scale_pair(int npad, double2* __restrict__ X, double2* __restrict__ Y,
                                                         double theta, int smax, int* __restrict__ sq,
                                                         int* __restrict__ err) {
  const int b = blockIdx.x;
  const int64_t per = (int64_t)npad * npad;
  double2* Xb = X + b * per;
  double2* Yb = Y + b * per;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ double wmax[4];
  __shared__ int s_sh;
  double cmax = 0.0;
  for (int j = wave; j < npad; j += 4) {
    double sum = 0.0;
    for (int i = lane; i < npad; i += 64) {
      const double2 u = Xb[(int64_t)j * npad + i], v = Yb[(int64_t)j * npad + i];
      sum += hypot(u.x, u.y) + hypot(v.x, v.y);
    }
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
    cmax = fmax(cmax, sum);
    if (!(sum == sum)) cmax = sum;
  }
  if (lane == 0) wmax[wave] = cmax;
  __syncthreads();
  if (tid == 0) {
    double nrm = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    for (int w = 0; w < 4; w++) if (!(wmax[w] == wmax[w])) nrm = wmax[w];
    int s = 0;
    if (!isfinite(nrm)) {
      atomicOr(err, 2);
    } else {
      while (nrm > theta && s < smax) { nrm *= 0.5; s++; }
      if (nrm > theta) atomicOr(err, 4);
    }
    s_sh = s;
    sq[b] = s;
  }
  __syncthreads();
  const double f = ldexp(1.0, -s_sh);
  if (s_sh > 0)
    for (int64_t e = tid; e < per; e += 256) {
      const double2 u = Xb[e], v = Yb[e];
      Xb[e] = make_double2(u.x * f, u.y * f);
      Yb[e] = make_double2(v.x * f, v.y * f);
    }
}

// D_bar = conj(alpha) A_bar / sigma from the derivative block of the exponential; conj(alpha) =
// -i pre with pre = 2 pi dz / fl32(2 kn); zero outside the n x n block
// (ldt = 2 npad: T2 is the dense block matrix, the block at column offset npad; ldt = npad: T2 is
// the Y part of the pair form)
__global__ void k_adj_extract(Bucket B, BeamView V, const double* __restrict__ kn, double dz,
                              const double2* __restrict__ T2, int ldt, const double* __restrict__ sigma,
                              double2* __restrict__ Dbar) {
  const int npad = B.npad, b = blockIdx.y;
  const int dir = B.dirs[b];
  const int n = V.nns[dir];
  const int64_t per = (int64_t)npad * npad;
  const int64_t tstride = (int64_t)ldt * ldt;
  const int coff = ldt == npad ? 0 : npad;
  const double pre = 2.0 * 3.141592653589793238 * dz / (double)(float)(2.0 * kn[dir]);
  const double sg = sigma[b];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < per; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % npad), j = (int)(e / npad);
    double2 v = make_double2(0.0, 0.0);
    if (i < n && j < n && sg > 0.0) {
      const double2 a = T2[(int64_t)b * tstride + (int64_t)(j + coff) * ldt + i];
      // (-i pre) (a.x + i a.y) / sigma
      v = make_double2(pre * a.y / sg, -pre * a.x / sg);
    }
    Dbar[b * per + e] = v;
  }
}

// LUT gradient, terms indexed by (g, w): one workgroup per (direction, 32-row tile of g,
// 32-chunk of weak beams). X(g,w) = sum_{h != g} D_bar(g,h) conj(U(w-h)) by chunks of h.
__global__ __launch_bounds__(256) void k_adj_lut_gw(Bucket B, BeamView V, DevRefl R, DevLut L,
                                                    double mlambda, const double2* __restrict__ Dbar,
                                                    double* __restrict__ glut, int wchunks, int nrep) {
  const int npad = B.npad, tiles = npad / 32;
  const int b = blockIdx.y;
  const int ti = blockIdx.x / wchunks, wc = blockIdx.x % wchunks;
  const int dir = B.dirs[b];
  const int n = V.nns[dir], nw = V.nnw[dir];
  const int w0 = wc * 32;
  if (ti >= tiles || ti * 32 >= n || w0 >= nw) return;
  const int32_t* sidx = V.sidx + V.soff[dir];
  const int32_t* widx = V.widx + V.woff[dir];
  const float* wsg = V.wsg + V.woff[dir];
  const double2* Db = Dbar + (int64_t)b * npad * npad;
  double* gl = lut_replica(glut, L.ncell, nrep);
  __shared__ int gh[32][3], wh[32][3], hh[32][3];
  __shared__ double2 Dt[32][33], Uc[32][33];
  const int tid = threadIdx.x, r = tid & 31, cq = tid >> 5;
  if (tid < 32) {
    const int g = ti * 32 + tid, w = w0 + tid;
    for (int d = 0; d < 3; d++) {
      gh[tid][d] = (g < n) ? R.hkl[3 * sidx[g] + d] : 0;
      wh[tid][d] = (w < nw) ? R.hkl[3 * widx[w] + d] : 0;
    }
  }
  double xr[4] = {0, 0, 0, 0}, xi[4] = {0, 0, 0, 0};
  for (int h0 = 0; h0 < n; h0 += 32) {
    __syncthreads();
    if (tid < 32) {
      const int h = h0 + tid;
      for (int d = 0; d < 3; d++) hh[tid][d] = (h < n) ? R.hkl[3 * sidx[h] + d] : 0;
    }
    __syncthreads();
    for (int e = tid; e < 1024; e += 256) {
      const int a = e & 31, c = e >> 5;
      // Dt[g][h] without the diagonal; Uc[h][w] = conj(U(w - h))
      const int g = ti * 32 + a, h = h0 + c;
      double2 dv = make_double2(0.0, 0.0);
      if (g < n && h < n && g != h) dv = Db[(int64_t)h * npad + g];
      Dt[a][c] = dv;
      const int w = w0 + a, hq = h0 + c;
      double2 uv = make_double2(0.0, 0.0);
      if (w < nw && hq < n) {
        const double2 u = lut_val(L, lut_cell(L, wh[a][0] - hh[c][0], wh[a][1] - hh[c][1], wh[a][2] - hh[c][2]));
        uv = make_double2(u.x, -u.y);
      }
      Uc[c][a] = uv;
    }
    __syncthreads();
    for (int c = 0; c < 32; c++) {
      const double2 d = Dt[r][c];
      for (int q = 0; q < 4; q++) {
        const double2 u = Uc[c][cq + 8 * q];
        xr[q] += d.x * u.x - d.y * u.y;
        xi[q] += d.x * u.y + d.y * u.x;
      }
    }
  }
  const int g = ti * 32 + r;
  if (g >= n) return;
  const double2 dgg = Db[(int64_t)g * npad + g];
  for (int q = 0; q < 4; q++) {
    const int wl = cq + 8 * q, w = w0 + wl;
    if (w >= nw) continue;
    const int idx = lut_cell(L, gh[r][0] - wh[wl][0], gh[r][1] - wh[wl][1], gh[r][2] - wh[wl][2]);
    const double sw = (double)wsg[w];
    const double f = -0.5 * mlambda / sw;
    const double2 u = lut_val(L, idx);
    const double fd = -mlambda / sw * dgg.x;  // diagonal |U(g-w)|^2 / s_w term
    lut_add(gl, idx, f * xr[q] + fd * u.x, f * xi[q] + fd * u.y);
  }
}

// LUT gradient, terms indexed by (w, h): Y(w,h) = sum_{g != h} conj(U(g-w)) D_bar(g,h), and
// the direct term D_bar(g,h) into U(g-h) (done by the first weak chunk).
__global__ __launch_bounds__(256) void k_adj_lut_wh(Bucket B, BeamView V, DevRefl R, DevLut L,
                                                    double mlambda, const double2* __restrict__ Dbar,
                                                    double* __restrict__ glut, int wchunks, int nrep) {
  const int npad = B.npad, tiles = npad / 32;
  const int b = blockIdx.y;
  const int tj = blockIdx.x / (wchunks + 1), wc = blockIdx.x % (wchunks + 1) - 1;
  const int dir = B.dirs[b];
  const int n = V.nns[dir], nw = V.nnw[dir];
  if (tj >= tiles || tj * 32 >= n) return;
  const int32_t* sidx = V.sidx + V.soff[dir];
  const int32_t* widx = V.widx + V.woff[dir];
  const float* wsg = V.wsg + V.woff[dir];
  const double2* Db = Dbar + (int64_t)b * npad * npad;
  double* gl = lut_replica(glut, L.ncell, nrep);
  const int tid = threadIdx.x, r = tid & 31, cq = tid >> 5;
  if (wc < 0) {
    // direct term: the columns h of tile tj, all rows g != h; the 32 lanes of a half wave run
    // over consecutive rows of one column
    for (int q = 0; q < 4; q++) {
      const int h = tj * 32 + cq + 8 * q;
      if (h >= n) continue;
      const int32_t* hv = R.hkl + 3 * sidx[h];
      for (int g = r; g < n; g += 32) {
        if (g == h) continue;
        const int32_t* gv = R.hkl + 3 * sidx[g];
        const double2 d = Db[(int64_t)h * npad + g];
        lut_add(gl, lut_cell(L, gv[0] - hv[0], gv[1] - hv[1], gv[2] - hv[2]), d.x, d.y);
      }
    }
    return;
  }
  const int w0 = wc * 32;
  if (w0 >= nw) return;
  __shared__ int hh[32][3], wh[32][3], gg[32][3];
  __shared__ double2 Dt[32][33], Uc[32][33];
  if (tid < 32) {
    const int h = tj * 32 + tid, w = w0 + tid;
    for (int d = 0; d < 3; d++) {
      hh[tid][d] = (h < n) ? R.hkl[3 * sidx[h] + d] : 0;
      wh[tid][d] = (w < nw) ? R.hkl[3 * widx[w] + d] : 0;
    }
  }
  double yr[4] = {0, 0, 0, 0}, yi[4] = {0, 0, 0, 0};
  for (int g0 = 0; g0 < n; g0 += 32) {
    __syncthreads();
    if (tid < 32) {
      const int g = g0 + tid;
      for (int d = 0; d < 3; d++) gg[tid][d] = (g < n) ? R.hkl[3 * sidx[g] + d] : 0;
    }
    __syncthreads();
    for (int e = tid; e < 1024; e += 256) {
      const int a = e & 31, c = e >> 5;
      // Dt[g][h] (rows g0.., columns of tile tj) without the diagonal; Uc[g][w] = conj(U(g - w))
      const int g = g0 + a, h = tj * 32 + c;
      double2 dv = make_double2(0.0, 0.0);
      if (g < n && h < n && g != h) dv = Db[(int64_t)h * npad + g];
      Dt[a][c] = dv;
      const int w = w0 + c;
      double2 uv = make_double2(0.0, 0.0);
      if (g < n && w < nw) {
        const double2 u = lut_val(L, lut_cell(L, gg[a][0] - wh[c][0], gg[a][1] - wh[c][1], gg[a][2] - wh[c][2]));
        uv = make_double2(u.x, -u.y);
      }
      Uc[a][c] = uv;
    }
    __syncthreads();
    // thread (r = h, cq): Y(w, h) for w = cq + 8q
    for (int a = 0; a < 32; a++) {
      const double2 d = Dt[a][r];
      for (int q = 0; q < 4; q++) {
        const double2 u = Uc[a][cq + 8 * q];
        yr[q] += u.x * d.x - u.y * d.y;
        yi[q] += u.x * d.y + u.y * d.x;
      }
    }
  }
  const int h = tj * 32 + r;
  if (h >= n) return;
  for (int q = 0; q < 4; q++) {
    const int wl = cq + 8 * q, w = w0 + wl;
    if (w >= nw) continue;
    const int idx = lut_cell(L, wh[wl][0] - hh[r][0], wh[wl][1] - hh[r][1], wh[wl][2] - hh[r][2]);
    const double f = -0.5 * mlambda / (double)wsg[w];
    lut_add(gl, idx, f * yr[q], f * yi[q]);
  }
}

// ------------------------------------------------------------------------------------------
// Mixed form: the forward exponential as E = I + F (FP32 F), the Frechet derivative from FP32
// matrices on split-FP16 products. It is taken as L_exp(A^H, Ebar) = L_exp(A, Ebar^H)^H, so that
// the pair exponential starts from A itself.

// E = I + F in FP64 for the backward depth recursion
__global__ void k_adj_e64(int npad, const float2* __restrict__ F, double2* __restrict__ E) {
  const int b = blockIdx.y;
  const int64_t per = (int64_t)npad * npad;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < per; e += (int64_t)gridDim.x * blockDim.x) {
    const float2 f = F[b * per + e];
    const bool diag = (e % npad) == (e / npad);
    E[b * per + e] = make_double2((double)f.x + (diag ? 1.0 : 0.0), (double)f.y);
  }
}

__device__ __forceinline__ double dnanmax(double a, double b) { return (a != a) ? a : ((b != b) ? b : fmax(a, b)); }

// max over the workgroup (256 threads), a NaN kept
__device__ __forceinline__ double block_nanmax(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v = dnanmax(v, __shfl_xor(v, o));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  return dnanmax(dnanmax(red[0], red[1]), dnanmax(red[2], red[3]));
}

// The pair (X, Y) = (A 2^-s, sigma 2^-s Ebar^H) from the FP32 A (read and scaled in place:
// X may alias A) and the FP64 Ebar: sigma the power of two that brings ||Ebar^H||_1 (the largest
// row sum of |Ebar|) just below ||A||_1, s from ||[[X, Y], [0, X]]||_1 = max_j (colsum_j |A| +
// sigma rowsum_j |Ebar|) / 2^s <= theta (error bits as k_norm_scale), and the max |re|, |im| of X
// and Y into xmax[b], ymax[b] (zeroed by the caller). One workgroup per matrix; the transpose
// goes through 32 x 32 LDS tiles so that reads and writes are both column runs.
__global__ __launch_bounds__(256) void k_adj_pair32(int npad, float2* __restrict__ A, const double2* __restrict__ Eb,
                                                    double theta, int smax, int* __restrict__ sq,
                                                    int* __restrict__ err, double* __restrict__ sigma,
                                                    float2* __restrict__ Y, float* __restrict__ xmax,
                                                    float* __restrict__ ymax) {
  extern __shared__ double cs[];  // [2 npad]: column sums of |A|, row sums of |Ebar|
  __shared__ double2 tile[32][33];
  __shared__ double red[4];
  __shared__ int s_sh;
  __shared__ double sg_sh;
  const int b = blockIdx.x;
  const int64_t per = (int64_t)npad * npad;
  float2* Ab = A + b * per;
  const double2* Gb = Eb + b * per;
  double* cA = cs;
  double* rE = cs + npad;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int j = wave; j < npad; j += 4) {
    double sa = 0.0;
    for (int i = lane; i < npad; i += 64) {
      const float2 a = Ab[(int64_t)j * npad + i];
      sa += (double)hypotf(a.x, a.y);
    }
    sa = wave_sum(sa);
    if (lane == 0) cA[j] = sa;
  }
  // row sums: consecutive threads take consecutive rows, so each column is read in runs
  for (int j = tid; j < npad; j += 256) {
    double se = 0.0;
    for (int i = 0; i < npad; i++) {
      const double2 g = Gb[(int64_t)i * npad + j];
      se += hypot(g.x, g.y);
    }
    rE[j] = se;
  }
  __syncthreads();
  double na = 0.0, ng = 0.0;
  for (int j = tid; j < npad; j += 256) {
    na = dnanmax(na, cA[j]);
    ng = dnanmax(ng, rE[j]);
  }
  na = block_nanmax(na, red);
  ng = block_nanmax(ng, red);
  if (tid == 0) {
    double sg = 0.0;
    if (ng > 0.0 && isfinite(ng)) {
      int ea, eg;
      frexp((na > 0.0 && isfinite(na)) ? na : 1.0, &ea);
      frexp(ng, &eg);
      sg = ldexp(1.0, (ea - 1) - eg);
    }
    sg_sh = sg;
    sigma[b] = sg;
  }
  __syncthreads();
  const double sg = sg_sh;
  double nrm = 0.0;
  for (int j = tid; j < npad; j += 256) nrm = dnanmax(nrm, cA[j] + sg * rE[j]);
  nrm = block_nanmax(nrm, red);
  if (tid == 0) {
    int sc = 0;
    if (!isfinite(nrm)) {
      atomicOr(err, 2);
    } else {
      while (nrm > theta && sc < smax) { nrm *= 0.5; sc++; }
      if (nrm > theta) atomicOr(err, 4);
    }
    s_sh = sc;
    sq[b] = sc;
  }
  __syncthreads();
  const float fx = ldexpf(1.0f, -s_sh);
  float vx = 0.0f;
  for (int64_t e = tid; e < per; e += 256) {
    const float2 a = Ab[e];
    const float2 w = make_float2(a.x * fx, a.y * fx);
    Ab[e] = w;
    vx = nmax(vx, nmax(fabsf(w.x), fabsf(w.y)));
  }
  wave_max_to(vx, xmax + b);
  // Y(r, c) = f conj(Ebar(c, r)); tile[rl][cl] holds Ebar(c0 + cl, r0 + rl)
  const double fy = sg * ldexp(1.0, -s_sh);
  float vy = 0.0f;
  const int tx = tid & 31, ty = tid >> 5;
  const int nt = npad / 32;
  for (int t = 0; t < nt * nt; t++) {
    const int r0 = (t % nt) * 32, c0 = (t / nt) * 32;
    __syncthreads();
    for (int q = 0; q < 4; q++) {
      const int rl = ty + 8 * q;
      tile[rl][tx] = Gb[(int64_t)(r0 + rl) * npad + c0 + tx];
    }
    __syncthreads();
    for (int q = 0; q < 4; q++) {
      const int cl = ty + 8 * q;
      const double2 g = tile[tx][cl];
      const float2 w = make_float2((float)(fy * g.x), (float)(-fy * g.y));
      Y[b * per + (int64_t)(c0 + cl) * npad + r0 + tx] = w;
      vy = nmax(vy, nmax(fabsf(w.x), fabsf(w.y)));
    }
  }
  wave_max_to(vy, ymax + b);
}

// D_bar = conj(alpha) A_bar from the mixed pair exponential's derivative part
// G = L_exp(A, sigma Ebar^H) = (sigma A_bar)^H: D_bar(i,j) = conj(alpha) conj(G(j,i)) / sigma, zero
// outside the n x n block. One workgroup per 32 x 32 tile of D_bar (G read through LDS).
__global__ __launch_bounds__(256) void k_adj_extract32(Bucket B, BeamView V, const double* __restrict__ kn,
                                                       double dz, const float2* __restrict__ G,
                                                       const double* __restrict__ sigma, double2* __restrict__ Dbar) {
  __shared__ float2 tile[32][33];
  const int npad = B.npad, b = blockIdx.y, nt = npad / 32;
  const int i0 = (blockIdx.x % nt) * 32, j0 = (blockIdx.x / nt) * 32;
  const int dir = B.dirs[b];
  const int n = V.nns[dir];
  const int64_t per = (int64_t)npad * npad;
  const double pre = 2.0 * 3.141592653589793238 * dz / (double)(float)(2.0 * kn[dir]);
  const double sg = sigma[b];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  // tile[il][jl] = G(j0 + jl, i0 + il)
  for (int q = 0; q < 4; q++) {
    const int il = ty + 8 * q;
    tile[il][tx] = G[b * per + (int64_t)(i0 + il) * npad + j0 + tx];
  }
  __syncthreads();
  for (int q = 0; q < 4; q++) {
    const int jl = ty + 8 * q;
    const int i = i0 + tx, j = j0 + jl;
    double2 v = make_double2(0.0, 0.0);
    if (i < n && j < n && sg > 0.0) {
      const float2 g = tile[tx][jl];
      // a = conj(g) = sigma A_bar(i, j);  (-i pre) a / sigma
      v = make_double2(-pre * (double)g.y / sg, -pre * (double)g.x / sg);
    }
    Dbar[b * per + (int64_t)j * npad + i] = v;
  }
}

// ------------------------------------------------------------------------------------------
static int grid_rows(int64_t per) { return (int)((per + 255) / 256 < 256 ? (per + 255) / 256 : 256); }

void launch_adj_m(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R, const DevLut& L,
                  const double* gdots, double c, double2* M) {
  const int64_t per = (int64_t)B.npad * B.npad;
  hipLaunchKernelGGL(k_adj_m, dim3(grid_rows(per), B.nb), dim3(256), 0, s, B, V, R, L, gdots, c, M);
}
void launch_adj_hm(hipStream_t s, int npad, int nb, int nz, const double2* Psi, int64_t spsi, double2* Hm) {
  hipLaunchKernelGGL(k_adj_hm, dim3(grid_rows((int64_t)npad * nz), nb), dim3(256), 0, s, npad, nz, Psi, spsi, Hm);
}
void launch_adj_depth(hipStream_t s, int npad, int nb, int nz, const double2* E, const double2* Q,
                      const double* lam, double2* Phi) {
  hipLaunchKernelGGL(k_adj_depth, dim3(nb), dim3(256), 2 * npad * sizeof(double2), s, npad, nz, E, Q, lam, Phi);
}
void launch_adj_pair(hipStream_t s, int npad, int nb, const double2* A, const double2* Ebar, double* sigma,
                     double2* PX, double2* PY) {
  hipLaunchKernelGGL(k_adj_pair, dim3(nb), dim3(256), 0, s, npad, A, Ebar, sigma, PX, PY);
}
void launch_norm_scale_pair(hipStream_t s, int npad, int nb, double2* X, double2* Y, double theta, int smax,
                            int32_t* sq, int32_t* err) {
  hipLaunchKernelGGL(k_norm_scale_pair, dim3(nb), dim3(256), 0, s, npad, X, Y, theta, smax, sq, err);
}
void launch_adj_extract(hipStream_t s, const Bucket& B, const BeamView& V, const double* kn, double dz,
                        const double2* T2, int ldt, const double* sigma, double2* Dbar) {
  const int64_t per = (int64_t)B.npad * B.npad;
  hipLaunchKernelGGL(k_adj_extract, dim3(grid_rows(per), B.nb), dim3(256), 0, s, B, V, kn, dz, T2, ldt, sigma,
                     Dbar);
}
void launch_adj_e64(hipStream_t s, int npad, int nb, const float2* F, double2* E) {
  hipLaunchKernelGGL(k_adj_e64, dim3(grid_rows((int64_t)npad * npad), nb), dim3(256), 0, s, npad, F, E);
}
void launch_adj_pair32(hipStream_t s, int npad, int nb, float2* A, const double2* Ebar, double theta, int smax,
                       int32_t* sq, int32_t* err, double* sigma, float2* Y, float* xmax, float* ymax) {
  hipLaunchKernelGGL(k_adj_pair32, dim3(nb), dim3(256), 2 * npad * sizeof(double), s, npad, A, Ebar, theta, smax, sq,
                     err, sigma, Y, xmax, ymax);
}
void launch_adj_extract32(hipStream_t s, const Bucket& B, const BeamView& V, const double* kn, double dz,
                          const float2* G, const double* sigma, double2* Dbar) {
  const int nt = B.npad / 32;
  hipLaunchKernelGGL(k_adj_extract32, dim3(nt * nt, B.nb), dim3(256), 0, s, B, V, kn, dz, G, sigma, Dbar);
}
// glut[0][e] = sum over the copies r of glut[r][e], in the order r = 0, 1, ..
__global__ void k_adj_lut_reduce(int64_t n2, int nrep, double* __restrict__ glut) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n2; e += (int64_t)gridDim.x * blockDim.x) {
    double v = glut[e];
    for (int r = 1; r < nrep; r++) v += glut[(int64_t)r * n2 + e];
    glut[e] = v;
  }
}
int adj_lut_replicas(int ncell) {
  const size_t one = 16 * (size_t)ncell;
  return (int)std::max<size_t>(1, std::min<size_t>(kLutReplicas, ((size_t)2 << 30) / std::max<size_t>(one, 1)));
}
void launch_adj_lut_reduce(hipStream_t s, int ncell, int nrep, double* glut) {
  const int64_t n2 = 2 * (int64_t)ncell;
  if (nrep > 1) hipLaunchKernelGGL(k_adj_lut_reduce, dim3(grid_rows(n2)), dim3(256), 0, s, n2, nrep, glut);
}
void launch_adj_lut(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R, const DevLut& L,
                    double mlambda, const double2* Dbar, double* glut, int max_nw, int nrep) {
  const int tiles = B.npad / 32;
  const int wchunks = std::max(1, (max_nw + 31) / 32);
  hipLaunchKernelGGL(k_adj_lut_gw, dim3(tiles * wchunks, B.nb), dim3(256), 0, s, B, V, R, L, mlambda, Dbar, glut, wchunks,
                     nrep);
  hipLaunchKernelGGL(k_adj_lut_wh, dim3(tiles * (wchunks + 1), B.nb), dim3(256), 0, s, B, V, R, L, mlambda, Dbar, glut,
                     wchunks, nrep);
}

}  // namespace ctem
