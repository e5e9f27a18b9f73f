// HIP kernels of the EBSD master-pattern path for gfx950 (CDNA4, wave64).
//
//   k_select       Bethe beam selection and ordering, one workgroup per direction
//                  (gvectors.f90:988-1120), bit-exact float32 excitation errors
//   k_assemble     Bethe-corrected DynMat (gvectors.f90:1123-1173): LUT gathers staged in
//                  LDS, weak-beam sum as a tiled complex product
//   k_norm_scale   1-norm, squaring count and power-of-two scaling per matrix
//   k_zgemm        batched complex GEMM on v_mfma_f64_16x16x4_f64 with a fused polynomial
//                  epilogue (Taylor / Paterson-Stockmeyer steps)
//   k_lgh_dots     Lgh = (dz/thick) Psi diag(lam) Psi^H (upper-triangle tiles, Lgh being
//                  Hermitian) and Re sum Lgh (.) S_p (CTEMEBSDmaster
//                  .f90:456, with Lgh as in :697-720) in fixed-order partial sums
//   k_scatter      float32 svals to the Laue-equivalent Lambert pixels (:462-479), the last
//                  direction in list order winning a pixel, as in the reference's serial loop
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "ctem_internal.h"
#include "ctem_laue.h"
#include "ctem_sg.h"

namespace ctem {

typedef double double4v __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------------------------------
// beam selection
// fn: the foil normal (float32, reciprocal components) shared by all directions, or null for
// FN = k of each direction
__global__ __launch_bounds__(256) void k_select(const double* __restrict__ kvec, int32_t numk,
                                                DevRefl R, const double* __restrict__ rmt_g,
                                                const float* __restrict__ fn_g, SelectOut out, int fill) {
  const int ik = blockIdx.x;
  if (ik >= numk) return;
  __shared__ double rmt[9];
  __shared__ int wave_s[4], wave_w[4];
  __shared__ int base_s, base_w;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid < 9) rmt[tid] = rmt_g[tid];
  if (tid == 0) { base_s = 0; base_w = 0; }
  __syncthreads();
  const float k[3] = {(float)kvec[3 * ik], (float)kvec[3 * ik + 1], (float)kvec[3 * ik + 2]};
  const float fn[3] = {fn_g ? fn_g[0] : k[0], fn_g ? fn_g[1] : k[1], fn_g ? fn_g[2] : k[2]};
  const int64_t soff = fill ? out.soff[ik] : 0;
  const int64_t woff = fill ? out.woff[ik] : 0;
  bool range_ok = true;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (int t0 = 0; t0 < R.nref; t0 += 256) {
    const int ig = t0 + tid;
    int cls = 0;
    float sg = 0.0f;
    if (ig == 0) {
      cls = 2;  // the transmitted beam is always strong, with s_g = 0
    } else if (ig < R.nref) {
      const float g[3] = {R.g[3 * ig], R.g[3 * ig + 1], R.g[3 * ig + 2]};
      sg = calcsg_fn(g, k, fn, rmt, &range_ok);
      const float sgp = fabsf(sg);
      if (R.dbd[ig]) {
        if (sgp <= R.sgcutoff) cls = 2;
      } else if (sgp <= R.cut1[ig]) {
        cls = (sgp <= R.cut2[ig]) ? 2 : 1;
      }
    }
    const uint64_t bs = __ballot(cls == 2);
    const uint64_t bw = __ballot(cls == 1);
    if (lane == 0) { wave_s[wave] = __popcll(bs); wave_w[wave] = __popcll(bw); }
    __syncthreads();
    if (fill) {
      int os = base_s, ow = base_w;
      for (int w = 0; w < wave; w++) { os += wave_s[w]; ow += wave_w[w]; }
      if (cls == 2) {
        const int64_t p = soff + os + __popcll(bs & lt);
        out.sidx[p] = ig;
        out.ssg[p] = sg;
      } else if (cls == 1) {
        const int64_t p = woff + ow + __popcll(bw & lt);
        out.widx[p] = ig;
        out.wsg[p] = sg;
      }
    }
    __syncthreads();
    if (tid == 0) {
      base_s += wave_s[0] + wave_s[1] + wave_s[2] + wave_s[3];
      base_w += wave_w[0] + wave_w[1] + wave_w[2] + wave_w[3];
    }
    __syncthreads();
  }
  if (!fill && tid == 0) {
    out.nns[ik] = base_s;
    out.nnw[ik] = base_w;
  }
  if (!range_ok) atomicOr(out.err, 1);
}

void launch_select(hipStream_t s, const double* k, int32_t numk, const DevRefl& R,
                   const double* rmt9, const float* fn, const SelectOut& out, int fill) {
  if (numk <= 0) return;
  hipLaunchKernelGGL(k_select, dim3(numk), dim3(256), 0, s, k, numk, R, rmt9, fn, out, fill);
}

// ------------------------------------------------------------------------------------------
// DynMat assembly
// S_p(g-h) table index; the table box holds every difference of two list entries
__device__ __forceinline__ int64_t slut_index(const DevLut& L, int h, int k, int l) {
  return (int64_t)(h - L.slo[0]) + (int64_t)L.sdim[0] * ((k - L.slo[1]) + (int64_t)L.sdim[1] * (l - L.slo[2]));
}

// One workgroup per 32x32 tile of one matrix. Per chunk of 32 weak beams the LUT values
// U(g_r - w) and U(w - h_c) are gathered (into registers one chunk ahead, then LDS); the
// weak-beam sum
// sum_w U(g-w) U(w-h) / s_w is a 32x32x32 complex product on v_mfma_f64_16x16x4_f64, each
// wave one 16x16 quadrant (four real MFMAs per complex step); the mixed form (F32) runs the
// same product on v_mfma_f32_16x16x4_f32 with FP32 accumulation. The diagonal sum
// sum_w |U(g-w)|^2 / s_w (gvectors.f90:1165) is kept per row, in FP64, by the first 32 threads.
template <bool F32>
__global__ __launch_bounds__(256) void k_assemble(Bucket B, BeamView V, DevRefl R, DevLut L,
                                                  const double* __restrict__ kn, double mlambda,
                                                  float dynupz, double dz, double2* __restrict__ A,
                                                  float2* __restrict__ A32, int scaled) {
#pragma clang fp contract(off)
  const int npad = B.npad, tiles = npad / 32;
  const int b = blockIdx.x / (tiles * tiles);
  const int t = blockIdx.x % (tiles * tiles);
  const int ti = t % tiles, tj = t / tiles;
  const int dir = B.dirs[b];
  const int n = V.nns[dir], nw = V.nnw[dir];
  const int32_t* sidx = V.sidx + V.soff[dir];
  const float* ssg = V.ssg + V.soff[dir];
  const int32_t* widx = V.widx + V.woff[dir];
  const float* wsg = V.wsg + V.woff[dir];
  double2* Ab = A + (int64_t)b * npad * npad;

  __shared__ int rh[32][3], ch[32][3];
  __shared__ double winv[32], dsum_s[32], dpart_s[8][32];
  typedef typename std::conditional<F32, float2, double2>::type CT;  // staged LUT values
  typedef typename std::conditional<F32, f32x4, double4v>::type AccT;
  __shared__ CT Ua[32][33], Ub[32][33];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, l4 = lane >> 4;
  const int wr = (wave >> 1) * 16, wc = (wave & 1) * 16;
  if (tid < 32) {
    const int a = ti * 32 + tid, c = tj * 32 + tid;
    for (int d = 0; d < 3; d++) {
      rh[tid][d] = (a < n) ? R.hkl[3 * sidx[a] + d] : 0;
      ch[tid][d] = (c < n) ? R.hkl[3 * sidx[c] + d] : 0;
    }
  }
  __syncthreads();
  AccT accr = {0, 0, 0, 0}, acci = {0, 0, 0, 0};
  double dpart = 0.0;  // diagonal tiles: this thread's share of sum_w |U(g_r - w)|^2 / s_w
  const bool live = (ti * 32 < n) && (tj * 32 < n);
  const bool diag_tile = (ti == tj);
  // register prefetch of the next chunk: thread t gathers rows/columns t & 31 for the weak
  // beams (t >> 5) + 8m, m = 0..3, while the MFMAs of the current chunk run
  const int prr = tid & 31, pww = tid >> 5;
  const bool row_ok = ti * 32 + prr < n, col_ok = tj * 32 + prr < n;
  float2 pa[4], pb[4];
  float psg[4];
  auto prefetch = [&](int w0) {
    for (int m = 0; m < 4; m++) {
      const int w = w0 + pww + 8 * m;
      pa[m] = make_float2(0.f, 0.f);
      pb[m] = make_float2(0.f, 0.f);
      psg[m] = 1.0f;
      if (w < nw) {
        const int32_t* hw = R.hkl + 3 * widx[w];
        const int h0 = hw[0], h1 = hw[1], h2 = hw[2];
        psg[m] = wsg[w];
        if (row_ok) pa[m] = lut_get(L, rh[prr][0] - h0, rh[prr][1] - h1, rh[prr][2] - h2);
        if (col_ok) pb[m] = lut_get(L, h0 - ch[prr][0], h1 - ch[prr][1], h2 - ch[prr][2]);
      }
    }
  };
  if (live) {
    prefetch(0);
    for (int w0 = 0; w0 < nw; w0 += 32) {
      __syncthreads();
      for (int m = 0; m < 4; m++) {
        const int ww = pww + 8 * m;
        Ua[prr][ww] = CT{pa[m].x, pa[m].y};  // U(g_r - w)
        Ub[ww][prr] = CT{pb[m].x, pb[m].y};  // U(w - h_c)
        if (prr == 0) winv[ww] = 1.0 / (double)psg[m];  // cmplx(1.D0/weaksg, 0) of gvectors.f90:1156
        if (diag_tile) {
          // |U(g_r - w)|^2 / s_w (gvectors.f90:1165); entries past nw are 0 with s = 1
          const double m2 = hypot((double)pa[m].x, (double)pa[m].y);
          dpart += m2 * m2 / (double)psg[m];
        }
      }
      __syncthreads();
      if (w0 + 32 < nw) prefetch(w0 + 32);
      const int wn = min(32, nw - w0);
      // entries past nw are zero, so the product may run over whole steps of 4
      for (int k4 = 0; k4 < wn; k4 += 4) {
        const CT a = Ua[wr + l16][k4 + l4];
        const CT u = Ub[k4 + l4][wc + l16];
        if constexpr (F32) {
          const float iv = (float)winv[k4 + l4];
          const float bx = u.x * iv, by = u.y * iv;  // U(w - h) / s_w
          accr = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, bx, accr, 0, 0, 0);
          accr = __builtin_amdgcn_mfma_f32_16x16x4f32(-a.y, by, accr, 0, 0, 0);
          acci = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, by, acci, 0, 0, 0);
          acci = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, bx, acci, 0, 0, 0);
        } else {
          const double iv = winv[k4 + l4];
          const double2 bq = make_double2(u.x * iv, u.y * iv);  // U(w - h) / s_w
          accr = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, bq.x, accr, 0, 0, 0);
          accr = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, bq.y, accr, 0, 0, 0);
          acci = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, bq.y, acci, 0, 0, 0);
          acci = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, bq.x, acci, 0, 0, 0);
        }
      }
    }
  }
  // diagonal sums: the eight partial sums of each row, added in a fixed order
  if (diag_tile) {
    dpart_s[pww][prr] = dpart;
    __syncthreads();
    if (tid < 32) {
      double d = 0.0;
      for (int q = 0; q < 8; q++) d += dpart_s[q][tid];
      dsum_s[tid] = d;
    }
  }
  __syncthreads();
  // write the tile: lane l holds rows wr + l/16 + 4e (f64 MFMA) or wr + 4(l/16) + e (f32),
  // column wc + l%16
  const double half = 0.5 * mlambda;
  // scaled 1: 2 pi dz / fl32(2 kn) (Bloch waves, W / cmplx(2 kn)); 2: pi mLambda dz (the
  // scattering matrix of CalcLghSM, DynMat * dcmplx(0, cPi * mLambda), CTEMECP.f90:586)
  const double pre = scaled == 1   ? 2.0 * 3.141592653589793238 * dz / (double)(float)(2.0 * kn[dir])
                     : scaled == 2 ? 3.141592653589793238 * mlambda * dz
                                   : 0.0;
  for (int e = 0; e < 4; e++) {
    const int r = F32 ? wr + 4 * l4 + e : wr + l4 + 4 * e, c = wc + l16;
    const int gi = ti * 32 + r, gj = tj * 32 + c;
    double vr = 0.0, vi = 0.0;
    if (gi < n && gj < n) {
      if (gi != gj) {
        const float2 u = lut_get(L, rh[r][0] - ch[c][0], rh[r][1] - ch[c][1], rh[r][2] - ch[c][2]);
        vr = (double)u.x - half * (double)accr[e];
        vi = (double)u.y - half * (double)acci[e];
      } else {
        const double sd = dsum_s[r] * mlambda / 2.0;
        vr = 2.0 * (double)ssg[gi] / mlambda - sd;
        vi = (double)dynupz;
      }
    }
    if (gi < npad && gj < npad) {
      // scaled: A = i * pre * DynMat, in FP64 or (A32) rounded to FP32
      const double2 v = scaled ? make_double2(-pre * vi, pre * vr) : make_double2(vr, vi);
      if (A32)
        A32[(int64_t)b * npad * npad + (int64_t)gj * npad + gi] = make_float2((float)v.x, (float)v.y);
      else
        Ab[(int64_t)gj * npad + gi] = v;
    }
  }
}

void launch_assemble(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R,
                     const DevLut& L, const double* kn, double mlambda, float dynupz, double dz,
                     double2* A) {
  const int tiles = B.npad / 32;
  hipLaunchKernelGGL(k_assemble<false>, dim3(tiles * tiles * B.nb), dim3(256), 0, s, B, V, R, L, kn,
                     mlambda, dynupz, dz, A, (float2*)nullptr, 1);
}

void launch_assemble_scatmat(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R,
                             const DevLut& L, double mlambda, float dynupz, double dz, double2* A) {
  const int tiles = B.npad / 32;
  hipLaunchKernelGGL(k_assemble<false>, dim3(tiles * tiles * B.nb), dim3(256), 0, s, B, V, R, L,
                     (const double*)nullptr, mlambda, dynupz, dz, A, (float2*)nullptr, 2);
}

void launch_assemble32(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R,
                       const DevLut& L, const double* kn, double mlambda, float dynupz, double dz,
                       float2* A32) {
  const int tiles = B.npad / 32;
  hipLaunchKernelGGL(k_assemble<true>, dim3(tiles * tiles * B.nb), dim3(256), 0, s, B, V, R, L, kn,
                     mlambda, dynupz, dz, (double2*)nullptr, A32, 1);
}

void launch_dynmat(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R,
                   const DevLut& L, double mlambda, float dynupz, double2* D) {
  const int tiles = B.npad / 32;
  hipLaunchKernelGGL(k_assemble<false>, dim3(tiles * tiles * B.nb), dim3(256), 0, s, B, V, R, L,
                     (const double*)nullptr, mlambda, dynupz, 0.0, D, (float2*)nullptr, 0);
}

// ------------------------------------------------------------------------------------------
// norm and scaling
// amod (may be null) receives max |a_ij| of the unscaled matrix (the quantity the reference's
// MatrixExponential picks its Pade degree and squaring count from, math.f90:217-219)
__global__ __launch_bounds__(256) void k_norm_scale(int npad, double2* __restrict__ A,
                                                    double theta, int smax, int* __restrict__ sq,
                                                    int* __restrict__ err, double* __restrict__ amod) {
  const int b = blockIdx.x;
  double2* Ab = A + (int64_t)b * npad * npad;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ double wmax[4], wemax[4];
  __shared__ int s_sh;
  double cmax = 0.0, emax = 0.0;
  for (int j = wave; j < npad; j += 4) {
    double sum = 0.0;
    for (int i = lane; i < npad; i += 64) {
      const double2 v = Ab[(int64_t)j * npad + i];
      const double h = hypot(v.x, v.y);
      sum += h;
      emax = fmax(emax, h);
    }
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
    cmax = fmax(cmax, sum);
    if (!(sum == sum)) cmax = sum;  // keep a NaN
  }
  for (int o = 32; o > 0; o >>= 1) emax = fmax(emax, __shfl_xor(emax, o));
  if (lane == 0) { wmax[wave] = cmax; wemax[wave] = emax; }
  __syncthreads();
  if (tid == 0) {
    if (amod) amod[b] = fmax(fmax(wemax[0], wemax[1]), fmax(wemax[2], wemax[3]));
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
  if (s_sh > 0) {
    for (int64_t e = tid; e < (int64_t)npad * npad; e += 256) {
      double2 v = Ab[e];
      Ab[e] = make_double2(v.x * f, v.y * f);
    }
  }
}

void launch_norm_scale(hipStream_t s, int32_t npad, int32_t nb, double2* A, double theta,
                       int32_t smax, int32_t* sq, int32_t* err, double* amod) {
  hipLaunchKernelGGL(k_norm_scale, dim3(nb), dim3(256), 0, s, npad, A, theta, smax, sq, err, amod);
}

// ------------------------------------------------------------------------------------------
// batched complex GEMM, FP64 MFMA. Workgroup tile 64x64, 4 waves of 32x32, K chunks of 16
// staged through LDS. MFMA 16x16x4 f64 operand layout (lane l): A[l%16][l/16],
// B[l/16][l%16]; result element e of lane l is C[l/16 + 4*e][l%16] (rows interleaved by 4,
// unlike the f32/bf16 16x16 variants whose lane holds four consecutive rows).
#define ZG_LDS 65
__global__ __launch_bounds__(256) void k_zgemm(ZGemm g) {
  const int tm = (g.m + 63) / 64, tn = (g.n + 63) / 64;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int b = lb / (tm * tn);
  const int t = lb % (tm * tn);
  const int bm = (t % tm) * 64, bn = (t / tm) * 64;
  if (g.active && g.step >= g.active[b]) return;
  const double2* X = g.X + (int64_t)b * g.sx;
  const double2* Y = g.Y + (int64_t)b * g.sy;
  double2* C = g.C + (int64_t)b * g.sc;

  __shared__ double As_r[16][ZG_LDS], As_i[16][ZG_LDS], Bs_r[16][ZG_LDS], Bs_i[16][ZG_LDS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int l16 = lane & 15, l4 = lane >> 4;
  // waves whose 32x32 sub-tile lies wholly in the padding only help to stage operands
  const bool wave_live = (bm + wm < g.m) && (bn + wn < g.n);

  double4v cr[2][2], ci[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) { cr[i][j] = (double4v){0, 0, 0, 0}; ci[i][j] = (double4v){0, 0, 0, 0}; }

  // register-prefetch pipeline: the global loads of chunk k0+16 are in flight while the
  // MFMAs of chunk k0 run from LDS
  double2 ra[4], rb[4];
  auto gload = [&](int k0) {
    for (int r = 0; r < 4; r++) {
      const int mm = tid & 63, kk = (tid >> 6) + 4 * r;
      const int row = bm + mm, col = k0 + kk;
      ra[r] = (row < g.m && col < g.k) ? X[(int64_t)col * g.ldx + row] : make_double2(0.0, 0.0);
    }
    for (int r = 0; r < 4; r++) {
      const int kk = tid & 15, nn = (tid >> 4) + 16 * r;
      const int row = k0 + kk, col = bn + nn;
      rb[r] = (row < g.k && col < g.n) ? Y[(int64_t)col * g.ldy + row] : make_double2(0.0, 0.0);
    }
  };
  gload(0);
  for (int k0 = 0; k0 < g.k; k0 += 16) {
    __syncthreads();
    for (int r = 0; r < 4; r++) {
      const int mm = tid & 63, kk = (tid >> 6) + 4 * r;
      As_r[kk][mm] = ra[r].x;
      As_i[kk][mm] = ra[r].y;
      const int kb = tid & 15, nn = (tid >> 4) + 16 * r;
      Bs_r[kb][nn] = rb[r].x;
      Bs_i[kb][nn] = rb[r].y;
    }
    __syncthreads();
    if (k0 + 16 < g.k) gload(k0 + 16);
    if (!wave_live) continue;
#pragma unroll
    for (int ks = 0; ks < 16; ks += 4) {
      double a_r[2], a_i[2], b_r[2], b_i[2];
      for (int i = 0; i < 2; i++) {
        a_r[i] = As_r[ks + l4][wm + 16 * i + l16];
        a_i[i] = As_i[ks + l4][wm + 16 * i + l16];
        b_r[i] = Bs_r[ks + l4][wn + 16 * i + l16];
        b_i[i] = Bs_i[ks + l4][wn + 16 * i + l16];
      }
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
          cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r[i], b_r[j], cr[i][j], 0, 0, 0);
          ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r[i], b_i[j], ci[i][j], 0, 0, 0);
          cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a_i[i], b_i[j], cr[i][j], 0, 0, 0);
          ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_i[i], b_r[j], ci[i][j], 0, 0, 0);
        }
    }
  }
  if (!wave_live) return;
  // epilogue
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int e = 0; e < 4; e++) {
        const int row = bm + wm + 16 * i + l4 + 4 * e;
        const int col = bn + wn + 16 * j + l16;
        if (row < g.m && col < g.n) {
          double vr = g.alpha * cr[i][j][e], vi = g.alpha * ci[i][j][e];
          if (row == col) vr += g.betaI;
          for (int q = 0; q < 4; q++)
            if (g.M[q]) {
              const double2 mv = g.M[q][(int64_t)b * g.sm + (int64_t)col * g.ldm + row];
              vr += g.beta[q] * mv.x;
              vi += g.beta[q] * mv.y;
            }
          C[(int64_t)col * g.ldc + row] = make_double2(vr, vi);
        }
      }
}

void launch_zgemm(hipStream_t s, const ZGemm& g) {
  const int tm = (g.m + 63) / 64, tn = (g.n + 63) / 64;
  hipLaunchKernelGGL(k_zgemm, dim3(tm * tn * g.nb), dim3(256), 0, s, g);
}

// ------------------------------------------------------------------------------------------
// elementwise helpers
struct ComboArgs { const double2* M[4]; double beta[4]; };

__global__ __launch_bounds__(256) void k_combo(int npad, int nb, double2* __restrict__ C,
                                               double betaI, ComboArgs a) {
  const int64_t per = (int64_t)npad * npad;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < per * nb; e += (int64_t)gridDim.x * 256) {
    const int64_t loc = e % per;
    const int row = (int)(loc % npad), col = (int)(loc / npad);
    double vr = (row == col) ? betaI : 0.0, vi = 0.0;
    for (int q = 0; q < 4; q++)
      if (a.M[q]) { const double2 v = a.M[q][e]; vr += a.beta[q] * v.x; vi += a.beta[q] * v.y; }
    C[e] = make_double2(vr, vi);
  }
}

void launch_combo(hipStream_t s, int32_t npad, int32_t nb, double2* C, double betaI,
                  const double2* const M[4], const double beta[4]) {
  ComboArgs a;
  for (int q = 0; q < 4; q++) { a.M[q] = M[q]; a.beta[q] = beta[q]; }
  const int64_t tot = (int64_t)npad * npad * nb;
  const int grid = (int)((tot + 255) / 256 < 65536 ? (tot + 255) / 256 : 65536);
  hipLaunchKernelGGL(k_combo, dim3(grid), dim3(256), 0, s, npad, nb, C, betaI, a);
}

__global__ __launch_bounds__(256) void k_parity_copy(int npad, const int* __restrict__ sq,
                                                     const double2* __restrict__ src,
                                                     double2* __restrict__ dst) {
  const int b = blockIdx.y;
  if ((sq[b] & 1) == 0) return;
  const int64_t per = (int64_t)npad * npad;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < per; e += (int64_t)gridDim.x * 256)
    dst[b * per + e] = src[b * per + e];
}

void launch_parity_copy(hipStream_t s, int32_t npad, int32_t nb, const int32_t* sq,
                        const double2* src, double2* dst) {
  const int64_t per = (int64_t)npad * npad;
  const int gx = (int)((per + 255) / 256 < 64 ? (per + 255) / 256 : 64);
  hipLaunchKernelGGL(k_parity_copy, dim3(gx, nb), dim3(256), 0, s, npad, sq, src, dst);
}

// ------------------------------------------------------------------------------------------
// depth recursion psi_{k+1} = E psi_k, k = 1..nz-1, psi_1 = E e_1 (CTEMEBSDmaster.f90:697-720
// restated as a propagation). Persistent: one 1024-thread workgroup per CU walks the
// directions of the chunk, so the live set of E matrices (one per CU) stays in the
// last-level caches while it is read nz-1 times. Rows are split over 256 threads and the
// columns in four quarters whose partial sums meet in LDS; accumulation is in FP64 for both
// storage types. With add_identity the stored matrix is F = E - I (mixed form).
template <typename CS>
__device__ __forceinline__ double2 ld2(const CS* p) { const CS v = *p; return make_double2((double)v.x, (double)v.y); }

template <typename CS>
__global__ __launch_bounds__(1024) void k_depth(int npad, int nb, int nz, const CS* __restrict__ E,
                                                int add_identity, double2* __restrict__ Psi, int64_t spsi) {
  extern __shared__ double2 sm[];
  double2* x = sm;              // [npad]
  double2* part = sm + npad;    // [4][npad]
  const int tid = threadIdx.x, q = tid >> 8, r0 = tid & 255;
  const int64_t per = (int64_t)npad * npad;
  const int cq = (npad + 3) / 4;
  const int c0 = q * cq, c1 = min(npad, c0 + cq);
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    const CS* Eb = E + (int64_t)b * per;
    double2* P = Psi + (int64_t)b * spsi;
    __syncthreads();
    for (int i = tid; i < npad; i += 1024) {
      double2 v = ld2(Eb + i);
      if (add_identity && i == 0) v.x += 1.0;
      x[i] = v;
      P[i] = v;
    }
    __syncthreads();
    for (int k = 1; k < nz; k++) {
      for (int r = r0; r < npad; r += 256) {
        double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
        int c = c0;
        // eight loads in flight before their products; even offsets into a, odd ones into b
        for (; c + 8 <= c1; c += 8) {
          double2 e[8];
#pragma unroll
          for (int j = 0; j < 8; j++) e[j] = ld2(Eb + (int64_t)(c + j) * npad + r);
#pragma unroll
          for (int j = 0; j < 8; j += 2) {
            const double2 x0 = x[c + j], x1 = x[c + j + 1];
            ar += e[j].x * x0.x - e[j].y * x0.y; ai += e[j].x * x0.y + e[j].y * x0.x;
            br += e[j + 1].x * x1.x - e[j + 1].y * x1.y; bi += e[j + 1].x * x1.y + e[j + 1].y * x1.x;
          }
        }
        for (; c + 1 < c1; c += 2) {
          const double2 e0 = ld2(Eb + (int64_t)c * npad + r), e1 = ld2(Eb + (int64_t)(c + 1) * npad + r);
          const double2 x0 = x[c], x1 = x[c + 1];
          ar += e0.x * x0.x - e0.y * x0.y; ai += e0.x * x0.y + e0.y * x0.x;
          br += e1.x * x1.x - e1.y * x1.y; bi += e1.x * x1.y + e1.y * x1.x;
        }
        if (c < c1) {
          const double2 e0 = ld2(Eb + (int64_t)c * npad + r), x0 = x[c];
          ar += e0.x * x0.x - e0.y * x0.y; ai += e0.x * x0.y + e0.y * x0.x;
        }
        part[q * npad + r] = make_double2(ar + br, ai + bi);
      }
      __syncthreads();
      for (int r = tid; r < npad; r += 1024) {
        double2 y = part[r];
        for (int qq = 1; qq < 4; qq++) { const double2 t = part[qq * npad + r]; y.x += t.x; y.y += t.y; }
        if (add_identity) { y.x += x[r].x; y.y += x[r].y; }
        P[(int64_t)k * npad + r] = y;
        part[r] = y;  // staged; copied into x after the barrier
      }
      __syncthreads();
      for (int r = tid; r < npad; r += 1024) x[r] = part[r];
      __syncthreads();
    }
  }
}

// Depth recursion with the whole matrix held on chip (FP32 storage, npad <= 224): E is read
// from memory once per direction instead of once per depth step. The 1024 threads form a
// 32 x 32 grid; thread (tr, tc) keeps the RT x RT block of rows tr*RT.. and columns tc*RT..
// (RT = npad/32) in registers, or for npad = 224 five of its seven rows in registers and two
// in LDS. Per step each thread forms its RT partial row sums in FP32 (RT terms each), the 32
// threads of a row group (half a wave) add them (a reduce-scatter of the 2 RT real and
// imaginary sums within each 16-lane DPP row, then one exchange between the two rows), and lane
// tc < RT of the group forms entry tr*RT + tc of the new vector psi_{k+1} = E psi_k, which is
// kept in FP64 in LDS (double-buffered, with an FP32 copy for the next step's products) and
// written out.
// Reduce-scatter of 16 values over the 16 lanes of a DPP row (lanes 16 m .. 16 m + 15): at each
// of four levels a lane keeps half of its values and adds its partner's copies of them (partners
// at lane distance 8 by row_ror 8, mirrored within 8 lanes by row_half_mirror, then quad_perm
// [2,3,0,1] and [1,0,3,2]); afterwards lane l holds the row's sum of value l % 16.
__device__ __forceinline__ float row_reduce_scatter16(const float (&v)[16], int lane) {
  const bool b3 = lane & 8, b2 = lane & 4, b1 = lane & 2, b0 = lane & 1;
  float a[8], c[4], d[2];
#pragma unroll
  for (int m = 0; m < 8; m++) a[m] = (b3 ? v[m + 8] : v[m]) + dpp_f32<0x128>(b3 ? v[m] : v[m + 8]);
#pragma unroll
  for (int m = 0; m < 4; m++) c[m] = (b2 ? a[m + 4] : a[m]) + dpp_f32<0x141>(b2 ? a[m] : a[m + 4]);
#pragma unroll
  for (int m = 0; m < 2; m++) d[m] = (b1 ? c[m + 2] : c[m]) + dpp_f32<0x4E>(b1 ? c[m] : c[m + 2]);
  return (b0 ? d[1] : d[0]) + dpp_f32<0xB1>(b0 ? d[0] : d[1]);
}

template <int NPAD>
__global__ __launch_bounds__(1024) void k_depth_onchip(int nb, int nz, const float2* __restrict__ E,
                                                       int add_identity, double2* __restrict__ Psi,
                                                       int64_t spsi) {
  constexpr int RT = NPAD / 32;
  constexpr int LR = (RT == 7) ? 2 : 0;  // rows of the block kept in LDS
  constexpr int RR = RT - LR;            // rows kept in registers
  __shared__ double2 xs[2][NPAD];
  __shared__ float2 xf32[2][NPAD];
  __shared__ float2 Elds[LR > 0 ? LR : 1][RT][1024];
  const int tid = threadIdx.x, tr = tid >> 5, tc = tid & 31;
  const int64_t per = (int64_t)NPAD * NPAD;
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    const float2* Eb = E + (int64_t)b * per;
    double2* P = Psi + (int64_t)b * spsi;
    float2 er[RR][RT];
    __syncthreads();  // the previous direction is done with xs and Elds
    for (int i = 0; i < RT; i++)
      for (int j = 0; j < RT; j++) {
        const float2 v = Eb[(int64_t)(tc * RT + j) * NPAD + tr * RT + i];
        if (i < RR) er[i < RR ? i : 0][j] = v;
        else if constexpr (LR > 0) Elds[i - RR][j][tid] = v;
      }
    for (int i = tid; i < NPAD; i += 1024) {
      double2 v = make_double2((double)Eb[i].x, (double)Eb[i].y);
      if (add_identity && i == 0) v.x += 1.0;
      xs[0][i] = v;
      xf32[0][i] = make_float2((float)v.x, (float)v.y);
      P[i] = v;
    }
    __syncthreads();
    for (int k = 1; k < nz; k++) {
      const int cur = (k - 1) & 1;
      float2 xf[RT];
#pragma unroll
      for (int j = 0; j < RT; j++) xf[j] = xf32[cur][tc * RT + j];
      float ar[RT], ai[RT];
#pragma unroll
      for (int i = 0; i < RT; i++) {
        ar[i] = 0.f;
        ai[i] = 0.f;
#pragma unroll
        for (int j = 0; j < RT; j++) {
          float2 e;
          if (i < RR) e = er[i < RR ? i : 0][j];
          else e = Elds[LR > 0 ? i - RR : 0][j][tid];
          ar[i] = fmaf(e.x, xf[j].x, ar[i]); ar[i] = fmaf(-e.y, xf[j].y, ar[i]);
          ai[i] = fmaf(e.x, xf[j].y, ai[i]); ai[i] = fmaf(e.y, xf[j].x, ai[i]);
        }
      }
      // the 32 partial sums of row r meet in FP32 (each holds RT terms); the recursion
      // itself is carried in FP64. Value i of the reduce-scatter is the real sum of row i,
      // value 8 + i its imaginary sum.
      float v[16];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        v[i] = (i < RT) ? ar[i < RT ? i : 0] : 0.f;
        v[8 + i] = (i < RT) ? ai[i < RT ? i : 0] : 0.f;
      }
      float sr = row_reduce_scatter16(v, tc);
      sr += xchg16(sr);  // the other DPP row of the half wave
      // lane tc < 8 takes the imaginary sum of row tc from lane tc + 8
      const float si = dpp_f32<0x128>(sr);
      if (tc < RT) {
        const int r = tr * RT + tc;
        double yr = sr, yi = si;
        if (add_identity) { yr += xs[cur][r].x; yi += xs[cur][r].y; }
        const double2 y = make_double2(yr, yi);
        xs[cur ^ 1][r] = y;
        xf32[cur ^ 1][r] = make_float2((float)yr, (float)yi);
        P[(int64_t)k * NPAD + r] = y;
      }
      __syncthreads();
    }
  }
}

void launch_depth(hipStream_t s, int32_t npad, int32_t nb, int32_t nz, const void* E, int fp32,
                  int add_identity, double2* Psi, int64_t spsi, int32_t grid) {
  const size_t lds = (size_t)5 * npad * sizeof(double2);
  const int g = std::min(nb, grid);
  if (fp32 && npad <= 224 && npad % 32 == 0) {
    const float2* Ef = (const float2*)E;
    switch (npad) {
#define DEPTH_CASE(N) case N: hipLaunchKernelGGL(k_depth_onchip<N>, dim3(g), dim3(1024), 0, s, nb, nz, Ef, add_identity, Psi, spsi); return;
      DEPTH_CASE(32) DEPTH_CASE(64) DEPTH_CASE(96) DEPTH_CASE(128) DEPTH_CASE(160) DEPTH_CASE(192) DEPTH_CASE(224)
#undef DEPTH_CASE
    }
  }
  if (fp32)
    hipLaunchKernelGGL(k_depth<float2>, dim3(g), dim3(1024), lds, s, npad, nb, nz, (const float2*)E,
                       add_identity, Psi, spsi);
  else
    hipLaunchKernelGGL(k_depth<double2>, dim3(g), dim3(1024), lds, s, npad, nb, nz, (const double2*)E,
                       add_identity, Psi, spsi);
}

// Depth recursion over groups of four depths (mixed form, matrices too large to hold on chip):
// with F = E - I, F2 = E^2 - I and F4 = E^4 - I in FP32, psi_0 = e_1 + F e_1 (the first column
// of E), psi_1 = psi_0 + F psi_0, [psi_2, psi_3] = [psi_0, psi_1] + F2 [psi_0, psi_1], then
// [psi_{k+4} .. psi_{k+7}] = [psi_k .. psi_{k+3}] + F4 [psi_k .. psi_{k+3}]: each pass over a
// matrix streamed from memory serves four vectors, a quarter of the passes of k_depth. Rows over
// 512 threads, columns in two halves: FP32 products and sums within a half (from an FP32 copy of
// the vectors), the halves' sums and the recursion in FP64. LDS: 160 npad bytes; 86 registers,
// one workgroup per CU (a 64-register build for two spills, and measured 7% slower on corundum).
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4))) void k_depth_quad(
    int npad, int nb, int nz, const float2* __restrict__ F, const float2* __restrict__ F2,
    const float2* __restrict__ F4, double2* __restrict__ Psi, int64_t spsi) {
  extern __shared__ double2 sm[];
  double2* x = sm;                          // [4][npad] the current vectors
  float2* part = (float2*)(sm + 4 * npad);  // [2 halves][4][npad] FP32 partial sums
  float2* xf = part + 8 * npad;             // [4][npad] FP32 copy of x
  const int tid = threadIdx.x, h = tid >> 9, r0 = tid & 511;
  const int64_t per = (int64_t)npad * npad;
  const int ch = (npad + 1) / 2;
  const int c0 = h * ch, c1 = min(npad, c0 + ch);
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    const float2* Fb = F + (int64_t)b * per;
    const float2* F2b = F2 + (int64_t)b * per;
    const float2* F4b = F4 + (int64_t)b * per;
    double2* P = Psi + (int64_t)b * spsi;
    // out[v] = in[v] + M in[v] for the vectors in slots 0 .. NV-1, into slots o .. o+NV-1,
    // stored as depths k .. k+NV-1 (those below nz)
    auto pass = [&](const float2* M, auto nvc, int o, int k) __attribute__((always_inline)) {
      constexpr int NV = decltype(nvc)::value;
      for (int r = r0; r < npad; r += 512) {
        float ar[NV], ai[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) { ar[v] = 0.f; ai[v] = 0.f; }
        auto mac = [&](const float2 e, int c) __attribute__((always_inline)) {
#pragma unroll
          for (int v = 0; v < NV; v++) {
            const float2 u = xf[v * npad + c];
            ar[v] = fmaf(e.x, u.x, ar[v]); ar[v] = fmaf(-e.y, u.y, ar[v]);
            ai[v] = fmaf(e.x, u.y, ai[v]); ai[v] = fmaf(e.y, u.x, ai[v]);
          }
        };
        // four loads in flight before their products
        int c = c0;
        for (; c + 4 <= c1; c += 4) {
          float2 e[4];
#pragma unroll
          for (int j = 0; j < 4; j++) e[j] = M[(int64_t)(c + j) * npad + r];
#pragma unroll
          for (int j = 0; j < 4; j++) mac(e[j], c + j);
        }
        for (; c < c1; c++) mac(M[(int64_t)c * npad + r], c);
#pragma unroll
        for (int v = 0; v < NV; v++) part[(4 * h + v) * npad + r] = make_float2(ar[v], ai[v]);
      }
      __syncthreads();
      // one thread per (vector, row) reads x there and writes the output slot, so the output
      // slots may be the input slots
      for (int i = tid; i < NV * npad; i += 1024) {
        const int v = i / npad, r = i - v * npad;
        const float2 p0 = part[v * npad + r], p1 = part[(4 + v) * npad + r];
        const double2 xi = x[v * npad + r];
        const double2 y = make_double2(xi.x + ((double)p0.x + (double)p1.x), xi.y + ((double)p0.y + (double)p1.y));
        if (k + v < nz) P[(int64_t)(k + v) * npad + r] = y;
        x[(o + v) * npad + r] = y;
        xf[(o + v) * npad + r] = make_float2((float)y.x, (float)y.y);
      }
      __syncthreads();
    };
    __syncthreads();  // the previous direction is done with x and xf
    for (int i = tid; i < npad; i += 1024) {
      const float2 f = Fb[i];
      const double2 v = make_double2((double)f.x + (i == 0 ? 1.0 : 0.0), (double)f.y);
      x[i] = v;
      xf[i] = make_float2((float)v.x, (float)v.y);
      P[i] = v;
    }
    __syncthreads();
    if (nz > 1) pass(Fb, std::integral_constant<int, 1>{}, 1, 1);
    if (nz > 2) pass(F2b, std::integral_constant<int, 2>{}, 2, 2);
    for (int k = 4; k < nz; k += 4) pass(F4b, std::integral_constant<int, 4>{}, 0, k);
  }
}

void launch_depth_quad(hipStream_t s, int32_t npad, int32_t nb, int32_t nz, const float2* F, const float2* F2,
                       const float2* F4, double2* Psi, int64_t spsi, int32_t grid) {
  const size_t lds = (size_t)4 * npad * sizeof(double2) + (size_t)12 * npad * sizeof(float2);
  const int g = std::min(nb, 2 * grid);  // two rounds of one workgroup per CU
  hipLaunchKernelGGL(k_depth_quad, dim3(g), dim3(1024), lds, s, npad, nb, nz, F, F2, F4, Psi, spsi);
}

// ------------------------------------------------------------------------------------------
// Lgh and the Sgh contraction
// NSP = nsite (1 or 2): the S_p gathers of the epilogue are issued before the depth loop and
// wait in registers (same arithmetic and order); NSP = 0: gathered in the epilogue (any nsite).
// F32 (mixed form): Psi staged in FP32 and the products on v_mfma_f32_16x16x4_f32 with FP32
// accumulation over the depths (the contraction with S_p stays FP64); measured against the FP64
// form it adds ~1e-6 relative to the intensities of the mixed form (largest error over the Fe,
// Ni and corundum tests 7.8e-6, from 4.8e-6 with FP64 Lgh)
template <int NSP, bool F32>
__global__ __launch_bounds__(256) void k_lgh_dots(Bucket B, BeamView V, DevRefl R, DevLut L,
                                                  const double2* __restrict__ Psi, int64_t spsi,
                                                  int nz, const double* __restrict__ lam,
                                                  double dzt, double* __restrict__ partial,
                                                  int keep_dir, double2* __restrict__ keep_lgh) {
  const int npad = B.npad, tiles = npad / 32, ntri = tiles * (tiles + 1) / 2;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int b = lb / ntri;
  const int t = lb % ntri;
  // Lgh is Hermitian: only tiles on or above the diagonal (ti <= tj), t = tj (tj + 1) / 2 + ti
  int tj = 0;
  while ((tj + 1) * (tj + 2) / 2 <= t) tj++;
  const int ti = t - tj * (tj + 1) / 2;
  const bool offd = (ti != tj);
  const int dir = B.dirs[b];
  const int n = V.nns[dir];
  const int32_t* sidx = V.sidx + V.soff[dir];
  const double2* P = Psi + (int64_t)b * spsi;
  // Psi tiles depth-major (Pg[k][g]): the staging writes and the operand reads of 16 lanes
  // are contiguous rows
  typedef typename std::conditional<F32, float2, double2>::type PT;
  typedef typename std::conditional<F32, float, double>::type WT;
  typedef typename std::conditional<F32, f32x4, double4v>::type AccT;
  __shared__ PT Pg[32][32], Ph[32][32];
  __shared__ WT lw[32];
  __shared__ int rh[32][3], ch[32][3];
  __shared__ double red[8][4];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6, l16 = lane & 15, l4 = lane >> 4;
  const int wr = (wave >> 1) * 16, wc = (wave & 1) * 16;
  const int nsite = L.nsite;
  AccT accr = {0, 0, 0, 0}, acci = {0, 0, 0, 0};
  // accumulator entry q of lane l: row 4 (l / 16) + q (FP32 MFMA) or l / 16 + 4 q (FP64 MFMA)
  // of the wave's quadrant, column l % 16
  auto qrow = [&](int q) __attribute__((always_inline)) { return F32 ? wr + 4 * l4 + q : wr + l4 + 4 * q; };
  const bool live = (ti * 32 < n) && (tj * 32 < n);
  if (tid < 32) {
    const int a = ti * 32 + tid, c = tj * 32 + tid;
    for (int d = 0; d < 3; d++) {
      rh[tid][d] = (a < n) ? R.hkl[3 * sidx[a] + d] : 0;
      ch[tid][d] = (c < n) ? R.hkl[3 * sidx[c] + d] : 0;
    }
  }
  // S_p(h - g) = conj S_p(g - h) bit for bit (the table's sine terms are odd in d, the rest
  // depends on |d|), so the mirrored entry conj Lgh(g, h) of an off-diagonal tile meets
  // conj S_p(g - h) and adds the same term Re(Lgh(g, h) S_p(g - h)) again: one gather per entry
  constexpr int NS1 = NSP > 0 ? NSP : 1;
  double2 sp[NS1][4];  // S_p(g - h) of the lane's entries
#pragma unroll
  for (int p = 0; p < NS1; p++)
#pragma unroll
    for (int q = 0; q < 4; q++) sp[p][q] = make_double2(0.0, 0.0);
  if constexpr (NSP > 0) {
    __syncthreads();  // rh, ch
    if (live) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int r = qrow(q), c = wc + l16;
        if (ti * 32 + r < n && tj * 32 + c < n) {
          const int dh = rh[r][0] - ch[c][0], dk = rh[r][1] - ch[c][1], dl = rh[r][2] - ch[c][2];
          const int64_t cell = slut_index(L, dh, dk, dl);
#pragma unroll
          for (int p = 0; p < NSP; p++) sp[p][q] = L.slut[(int64_t)p * L.sncell + cell];
        }
      }
    }
  }
  if (live) {
    for (int k0 = 0; k0 < nz; k0 += 32) {
      __syncthreads();
      for (int e = tid; e < 1024; e += 256) {
        const int rr = e & 31, kk = e >> 5;
        const int k = k0 + kk;
        double2 vg = make_double2(0, 0), vh = make_double2(0, 0);
        if (k < nz) {
          vg = P[(int64_t)k * npad + ti * 32 + rr];
          vh = P[(int64_t)k * npad + tj * 32 + rr];
        }
        if constexpr (F32) {
          Pg[kk][rr] = make_float2((float)vg.x, (float)vg.y);
          Ph[kk][rr] = make_float2((float)vh.x, (float)vh.y);
        } else {
          Pg[kk][rr] = vg;
          Ph[kk][rr] = vh;
        }
      }
      if (tid < 32) lw[tid] = (WT)((k0 + tid < nz) ? lam[k0 + tid] : 0.0);
      __syncthreads();
      // sum_k (lam_k psi_k(g)) conj(psi_k(h)) on v_mfma_f64_16x16x4_f64 (F32:
      // v_mfma_f32_16x16x4_f32), each wave one 16x16 quadrant; entries past nz are zero
      const int kn = min(32, nz - k0);
      for (int k4 = 0; k4 < kn; k4 += 4) {
        const PT a = Pg[k4 + l4][wr + l16];
        const WT w = lw[k4 + l4];
        const WT ar = w * a.x, ai = w * a.y;
        const PT h = Ph[k4 + l4][wc + l16];
        if constexpr (F32) {
          accr = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, h.x, accr, 0, 0, 0);
          accr = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, h.y, accr, 0, 0, 0);
          acci = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, h.x, acci, 0, 0, 0);
          acci = __builtin_amdgcn_mfma_f32_16x16x4f32(-ar, h.y, acci, 0, 0, 0);
        } else {
          accr = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, h.x, accr, 0, 0, 0);
          accr = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, h.y, accr, 0, 0, 0);
          acci = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, h.x, acci, 0, 0, 0);
          acci = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar, h.y, acci, 0, 0, 0);
        }
      }
    }
  }
  __syncthreads();
  double part[4] = {0, 0, 0, 0};
  for (int q = 0; q < 4; q++) {
    const int r = qrow(q), c = wc + l16;
    const int gi = ti * 32 + r, gj = tj * 32 + c;
    if (live && gi < n && gj < n) {
      const double lr = dzt * (double)accr[q], li = dzt * (double)acci[q];
      if constexpr (NSP > 0) {
#pragma unroll
        for (int p = 0; p < NSP; p++) {
          const double v = lr * sp[p][q].x - li * sp[p][q].y;
          part[p] += v;
          if (offd) part[p] += v;  // the mirrored entry
        }
      } else {
        const int dh = rh[r][0] - ch[c][0], dk = rh[r][1] - ch[c][1], dl = rh[r][2] - ch[c][2];
        const int64_t cell = nsite ? slut_index(L, dh, dk, dl) : 0;
        for (int p = 0; p < nsite && p < 4; p++) {
          const double2 sv = L.slut[(int64_t)p * L.sncell + cell];
          const double v = lr * sv.x - li * sv.y;
          part[p] += v;
          if (offd) part[p] += v;  // the mirrored entry
        }
      }
      if (dir == keep_dir) {
        keep_lgh[(int64_t)gi * n + gj] = make_double2(lr, li);
        if (offd) keep_lgh[(int64_t)gj * n + gi] = make_double2(lr, -li);
      }
    }
  }
  // block reduction in fixed order: within each wave (lane moves), then waves in order
  for (int p = 0; p < 4; p++) {
    const double v = wave_sum(part[p]);
    if (lane == 0) red[wave][p] = v;
  }
  __syncthreads();
  if (tid < nsite && tid < 4) {
    const double v = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
    partial[((int64_t)b * ntri + t) * nsite + tid] = v;
  }
}

// The mixed form's Lgh and contraction with whole directions per workgroup: Lgh = sum over the
// depths, and the contraction is linear in it, so each direction's depths are split into nsplit
// parts, one workgroup (8 waves) per (direction, part). The part's Psi (FP32) is staged in LDS
// once, as Ps[k][g] (the depth-major layout of k_lgh_dots, rows past the part's depths zero), and
// the waves walk the 16 x 16 quadrants of the 32 x 32 tiles on and above the diagonal (the same
// quadrants, MFMA sequence and per-entry terms as k_lgh_dots<.., true>), each lane adding its
// entries' Re(Lgh S_p) terms to its own partial sums; partial[(b nsplit + part) nsite + p] is the
// fixed-order sum of the lanes' and waves' sums, k_reduce_dots adds the parts in order. The tile
// kernel re-reads the Psi rows of every tile pair from L2 (about eight times the matrix per
// direction); here they are read once, and with two parts three workgroups share a CU, so that
// one stages its Psi while the others run their MFMAs. keep_dir needs nsplit = 1.
template <int NSP>
__global__ __launch_bounds__(512) void k_lgh_dots_dir(Bucket B, BeamView V, DevRefl R, DevLut L,
                                                      const double2* __restrict__ Psi, int64_t spsi,
                                                      int nz, int nsplit, const double* __restrict__ lam,
                                                      double dzt, double* __restrict__ partial, int keep_dir,
                                                      double2* __restrict__ keep_lgh) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  const int npad = B.npad, tiles = npad / 32, ntri = tiles * (tiles + 1) / 2;
  const int kper = (nz + nsplit - 1) / nsplit;
  const int nzp = (kper + 3) & ~3;
  float2* Ps = (float2*)lds_raw;                              // [nzp][npad]
  float* lw = (float*)(Ps + (size_t)nzp * npad);              // [nzp]
  int* hk = (int*)(lw + nzp);                                 // [npad][3]
  double* red = (double*)(((uintptr_t)(hk + 3 * npad) + 7) & ~(uintptr_t)7);  // [8][4]
  const int b = blockIdx.x / nsplit, part = blockIdx.x % nsplit;
  const int k0 = part * kper, kn = min(kper, nz - k0);  // this part's depths k0 .. k0 + kn - 1
  const int dir = B.dirs[b];
  const int n = V.nns[dir];
  const int32_t* sidx = V.sidx + V.soff[dir];
  const double2* P = Psi + (int64_t)b * spsi + (int64_t)k0 * npad;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, l16 = lane & 15, l4 = lane >> 4;
  const int nsite = L.nsite;
  for (int e = tid; e < nzp * npad; e += 512) {
    const int k = e / npad;
    const double2 v = (k < kn) ? P[e] : make_double2(0.0, 0.0);
    Ps[e] = make_float2((float)v.x, (float)v.y);
  }
  for (int k = tid; k < nzp; k += 512) lw[k] = (k < kn) ? (float)lam[k0 + k] : 0.0f;
  for (int g = tid; g < npad; g += 512)
    for (int d = 0; d < 3; d++) hk[3 * g + d] = (g < n) ? R.hkl[3 * sidx[g] + d] : 0;
  __syncthreads();
  double acc[4] = {0, 0, 0, 0};
  for (int task = wave; task < 4 * ntri; task += 8) {
    const int t = task >> 2, quad = task & 3;
    int tj = 0;
    while ((tj + 1) * (tj + 2) / 2 <= t) tj++;
    const int ti = t - tj * (tj + 1) / 2;
    const bool offd = (ti != tj);
    const int r0 = ti * 32 + (quad >> 1) * 16, c0 = tj * 32 + (quad & 1) * 16;
    if (r0 >= n || c0 >= n) continue;  // wave-uniform
    // NSP = nsite (1 or 2): the S_p gathers of the lane's entries are issued before the depth
    // loop and wait in registers; NSP = 0: gathered after it (any nsite)
    constexpr int NS1 = NSP > 0 ? NSP : 1;
    double2 sp[NS1][4];
    if constexpr (NSP > 0) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int gi = r0 + 4 * l4 + q, gj = c0 + l16;
#pragma unroll
        for (int p = 0; p < NSP; p++) sp[p][q] = make_double2(0.0, 0.0);
        if (gi < n && gj < n) {
          const int64_t cell = slut_index(L, hk[3 * gi] - hk[3 * gj], hk[3 * gi + 1] - hk[3 * gj + 1],
                                          hk[3 * gi + 2] - hk[3 * gj + 2]);
#pragma unroll
          for (int p = 0; p < NSP; p++) sp[p][q] = L.slut[(int64_t)p * L.sncell + cell];
        }
      }
    }
    f32x4 accr = {0, 0, 0, 0}, acci = {0, 0, 0, 0};
    for (int k4 = 0; k4 < nzp; k4 += 4) {
      const float2 a = Ps[(k4 + l4) * npad + r0 + l16];
      const float w = lw[k4 + l4];
      const float ar = w * a.x, ai = w * a.y;
      const float2 h = Ps[(k4 + l4) * npad + c0 + l16];
      accr = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, h.x, accr, 0, 0, 0);
      accr = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, h.y, accr, 0, 0, 0);
      acci = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, h.x, acci, 0, 0, 0);
      acci = __builtin_amdgcn_mfma_f32_16x16x4f32(-ar, h.y, acci, 0, 0, 0);
    }
    for (int q = 0; q < 4; q++) {
      const int gi = r0 + 4 * l4 + q, gj = c0 + l16;
      if (gi >= n || gj >= n) continue;
      const double lr = dzt * (double)accr[q], li = dzt * (double)acci[q];
      if constexpr (NSP > 0) {
#pragma unroll
        for (int p = 0; p < NSP; p++) {
          const double v = lr * sp[p][q].x - li * sp[p][q].y;
          acc[p] += v;
          if (offd) acc[p] += v;  // the mirrored entry conj Lgh(g, h) meets conj S_p(g - h)
        }
      } else {
        const int dh = hk[3 * gi] - hk[3 * gj], dk = hk[3 * gi + 1] - hk[3 * gj + 1], dl = hk[3 * gi + 2] - hk[3 * gj + 2];
        const int64_t cell = nsite ? slut_index(L, dh, dk, dl) : 0;
        for (int p = 0; p < nsite && p < 4; p++) {
          const double2 sv = L.slut[(int64_t)p * L.sncell + cell];
          const double v = lr * sv.x - li * sv.y;
          acc[p] += v;
          if (offd) acc[p] += v;  // the mirrored entry
        }
      }
      if (dir == keep_dir) {
        keep_lgh[(int64_t)gi * n + gj] = make_double2(lr, li);
        if (offd) keep_lgh[(int64_t)gj * n + gi] = make_double2(lr, -li);
      }
    }
  }
  for (int p = 0; p < 4; p++) {
    const double v = wave_sum(acc[p]);
    if (lane == 0) red[wave * 4 + p] = v;
  }
  __syncthreads();
  if (tid < nsite && tid < 4) {
    double v = 0.0;
    for (int w = 0; w < 8; w++) v += red[w * 4 + tid];
    partial[((int64_t)b * nsplit + part) * nsite + tid] = v;
  }
}

// LDS of k_lgh_dots_dir for one (direction, part): Psi in FP32, the depth weights, the beams'
// indices, the wave sums
static size_t lgh_dir_lds(int npad, int nz, int nsplit) {
  const int kper = (nz + nsplit - 1) / nsplit;
  const size_t nzp = (size_t)((kper + 3) & ~3);
  return nzp * npad * sizeof(float2) + nzp * sizeof(float) + 3 * (size_t)npad * sizeof(int) + 8 + 8 * 4 * sizeof(double);
}

__global__ void k_reduce_dots(int nb, int ntile, int nsite, const int* __restrict__ dirs,
                              const double* __restrict__ partial, double* __restrict__ dots) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  for (int p = 0; p < nsite; p++) {
    double s = 0.0;
    for (int t = 0; t < ntile; t++) s += partial[((int64_t)b * ntile + t) * nsite + p];
    dots[(int64_t)dirs[b] * nsite + p] = s;
  }
}

void launch_lgh_dots(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R,
                     const DevLut& L, const double2* Psi, int64_t spsi, int32_t nz,
                     const double* lam, double dzt, double* partial, double* dots,
                     int32_t keep_dir, double2* keep_lgh, int32_t f32) {
  const int tiles = B.npad / 32, ntri = tiles * (tiles + 1) / 2;
  // the mixed form, when a direction's Psi fits in LDS: whole directions per workgroup, in two
  // depth parts (one part when one Lgh is kept: each part only holds its depths' share)
  const int nsplit = keep_dir >= 0 ? 1 : 2;
  if (f32 && lgh_dir_lds(B.npad, nz, nsplit) <= (size_t)150 * 1024 && L.nsite <= 4) {
#define LGH_DIR(NSP)                                                                                             \
  hipLaunchKernelGGL(k_lgh_dots_dir<NSP>, dim3(B.nb * nsplit), dim3(512), lgh_dir_lds(B.npad, nz, nsplit), s, B, V, \
                     R, L, Psi, spsi, nz, nsplit, lam, dzt, partial, keep_dir, keep_lgh)
    if (L.nsite == 1) LGH_DIR(1);
    else if (L.nsite == 2) LGH_DIR(2);
    else LGH_DIR(0);
#undef LGH_DIR
    hipLaunchKernelGGL(k_reduce_dots, dim3((B.nb + 127) / 128), dim3(128), 0, s, B.nb, nsplit, L.nsite, B.dirs,
                       partial, dots);
    return;
  }
  const dim3 grid(ntri * B.nb);
#define LGH_LAUNCH(NSP, F)                                                                          \
  hipLaunchKernelGGL((k_lgh_dots<NSP, F>), grid, dim3(256), 0, s, B, V, R, L, Psi, spsi, nz, lam, dzt, \
                     partial, keep_dir, keep_lgh)
  if (f32) {
    if (L.nsite == 1) LGH_LAUNCH(1, true);
    else if (L.nsite == 2) LGH_LAUNCH(2, true);
    else LGH_LAUNCH(0, true);
  } else {
    if (L.nsite == 1) LGH_LAUNCH(1, false);
    else if (L.nsite == 2) LGH_LAUNCH(2, false);
    else LGH_LAUNCH(0, false);
  }
#undef LGH_LAUNCH
  hipLaunchKernelGGL(k_reduce_dots, dim3((B.nb + 127) / 128), dim3(128), 0, s, B.nb,
                     ntri, L.nsite, B.dirs, partial, dots);
}

// ------------------------------------------------------------------------------------------
// Lambert scatter
// Several directions can map to one pixel (hexagonal grids: the (i,0) row is listed twice and
// neighbouring hexagonal cells share square-array pixels). The reference's serial loop leaves
// the value of the last direction in list order (CTEMEBSDmaster.f90:410-479); the scatter keeps
// that order in two passes: the largest direction index that reaches each pixel, then the
// stores of that direction only.
__global__ void k_scatter_owner(int numk, const int32_t* __restrict__ ij, int npx, int npy, int isym,
                                int32_t* __restrict__ owner) {
  const int ik = blockIdx.x * blockDim.x + threadIdx.x;
  if (ik >= numk) return;
  int32_t q[24];
  const int ne = laue_equiv(ij[2 * ik], ij[2 * ik + 1], isym, q);
  for (int e = 0; e < ne; e++) {
    const int i = q[2 * e], j = q[2 * e + 1];
    if (i < -npx || i > npx || j < -npy || j > npy) continue;
    atomicMax(&owner[(int64_t)(j + npy) * (2 * npx + 1) + (i + npx)], ik);
  }
}

__global__ void k_scatter(int numk, const int32_t* __restrict__ ij, const double* __restrict__ dots,
                          int nsite, float nat, int npx, int npy, int isym,
                          const int32_t* __restrict__ owner, float* __restrict__ sr) {
  const int ik = blockIdx.x * blockDim.x + threadIdx.x;
  if (ik >= numk) return;
  int32_t q[24];
  const int ne = laue_equiv(ij[2 * ik], ij[2 * ik + 1], isym, q);
  const int64_t plane = (int64_t)(2 * npx + 1) * (2 * npy + 1);
  for (int e = 0; e < ne; e++) {
    const int i = q[2 * e], j = q[2 * e + 1];
    if (i < -npx || i > npx || j < -npy || j > npy) continue;
    const int64_t pix = (int64_t)(j + npy) * (2 * npx + 1) + (i + npx);
    if (owner[pix] != ik) continue;
    for (int p = 0; p < nsite; p++)
      sr[p * plane + pix] = __fdiv_rn((float)dots[(int64_t)ik * nsite + p], nat);
  }
}

__global__ void k_fill_zero32(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    p[e] = 0u;
}

// flag 32 when a matrix needs more squarings than the schedule a run was launched with
__global__ void k_sq_check(int nb, const int32_t* __restrict__ sq, int smax, int32_t* __restrict__ err) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb && sq[b] > smax) atomicOr(err, 32);
}

void launch_sq_check(hipStream_t s, int32_t nb, const int32_t* sq, int32_t smax, int32_t* err) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_sq_check, dim3((nb + 255) / 256), dim3(256), 0, s, nb, sq, smax, err);
}

void launch_fill_zero(hipStream_t s, void* p, size_t bytes) {
  const int64_t n = (int64_t)(bytes / 4);
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_zero32, dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t*)p, n);
}

void launch_scatter(hipStream_t s, int32_t numk, const int32_t* ij, const double* dots,
                    int32_t nsite, float nat, int32_t npx, int32_t npy, int32_t isym,
                    int32_t* owner, float* sr) {
  if (numk <= 0) return;
  hipLaunchKernelGGL(k_scatter_owner, dim3((numk + 127) / 128), dim3(128), 0, s, numk, ij, npx, npy,
                     isym, owner);
  hipLaunchKernelGGL(k_scatter, dim3((numk + 127) / 128), dim3(128), 0, s, numk, ij, dots, nsite,
                     nat, npx, npy, isym, (const int32_t*)owner, sr);
}

}  // namespace ctem
