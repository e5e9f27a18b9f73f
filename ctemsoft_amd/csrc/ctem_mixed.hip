// Mixed-precision kernels: complex GEMM on split-BF16 MFMA (v_mfma_f32_16x16x32_bf16,
// gfx950) with FP32 accumulation, for FP32 complex matrices.
//
// Each FP32 operand x is split into bfloat16 terms x = x0 + x1 + x2 + O(2^-24 |x|)
// (x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1)). A real product is then
//   terms = 3:  x0*y0 + x0*y1 + x1*y0                       (~2^-16 relative)
//   terms = 6:  the above + x0*y2 + x1*y1 + x2*y0           (~FP32 accuracy)
// each term one MFMA with FP32 accumulation. Complex products use four real products
// (re*re, im*im, re*im, im*re) into three accumulators (re+, re-, im).
//
// Tile: 64x64 per workgroup of 4 waves (32x32 per wave = 2x2 MFMA blocks), K chunks of 32
// staged through LDS as bfloat16 planes with K contiguous, so that each lane reads its
// 8-element fragment with one 16-byte LDS load. Operand layout of the 16x16x32 MFMA
// (lane l): A[l%16][8*(l/16)+t], B[8*(l/16)+t][l%16], t = 0..7; result element e of lane l is
// C[4*(l/16)+e][l%16].
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include "ctem_internal.h"

namespace ctem {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

#define CG_KS 40  // LDS row stride in bf16 elements (32 used, padded against bank conflicts)

__device__ __forceinline__ void split3(float x, __bf16& a, __bf16& b, __bf16& c) {
  a = (__bf16)x;
  const float r1 = x - (float)a;
  b = (__bf16)r1;
  const float r2 = r1 - (float)b;
  c = (__bf16)r2;
}

template <int TERMS>
__global__ __launch_bounds__(256) void k_cgemm(CGemm g) {
  constexpr int NP = (TERMS == 6) ? 3 : 2;  // bf16 planes per real component
  const int tm = (g.m + 63) / 64, tn = (g.n + 63) / 64;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int b = lb / (tm * tn);
  const int t = lb % (tm * tn);
  const int bm = (t % tm) * 64, bn = (t / tm) * 64;
  if (g.active && g.step >= g.active[b]) return;
  const float2* X = g.X + (int64_t)b * g.sx;
  const float2* Y = g.Y + (int64_t)b * g.sy;
  float2* C = g.C + (int64_t)b * g.sc;

  // planes: [component re/im][term][row or col][k]
  __shared__ __attribute__((aligned(16))) __bf16 As[2][NP][64][CG_KS];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[2][NP][64][CG_KS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int l16 = lane & 15, l4 = lane >> 4;
  const bool wave_live = (bm + wm < g.m) && (bn + wn < g.n);

  f4v crp[2][2], crn[2][2], ci[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      crp[i][j] = (f4v){0, 0, 0, 0};
      crn[i][j] = (f4v){0, 0, 0, 0};
      ci[i][j] = (f4v){0, 0, 0, 0};
    }

  // register-prefetch pipeline: thread t stages row/col (t & 63) and k = 8*(t >> 6) .. +7 of
  // both tiles (16 lanes hit 16 distinct 16-byte bank groups in LDS); the global loads of the
  // next chunk are in flight during the MFMAs of the current one
  const int srow = tid & 63, skq = (tid >> 6) * 8;
  float2 ra[8], rb[8];
  auto gload = [&](int k0) {
    for (int j = 0; j < 8; j++) {
      const int kk = k0 + skq + j;
      ra[j] = (bm + srow < g.m && kk < g.k) ? X[(int64_t)kk * g.ldx + bm + srow] : make_float2(0.f, 0.f);
      rb[j] = (bn + srow < g.n && kk < g.k) ? Y[(int64_t)(bn + srow) * g.ldy + kk] : make_float2(0.f, 0.f);
    }
  };
  auto stage = [&](const float2* v, __bf16 (*P)[NP][64][CG_KS]) {
    __attribute__((aligned(16))) __bf16 pr[NP][8], pi[NP][8];
    for (int j = 0; j < 8; j++) {
      __bf16 a, bb, c;
      split3(v[j].x, a, bb, c);
      pr[0][j] = a; pr[1][j] = bb; if constexpr (NP == 3) pr[NP - 1][j] = c;
      split3(v[j].y, a, bb, c);
      pi[0][j] = a; pi[1][j] = bb; if constexpr (NP == 3) pi[NP - 1][j] = c;
    }
    for (int p = 0; p < NP; p++) {
      *(bf8v*)&P[0][p][srow][skq] = *(bf8v*)pr[p];
      *(bf8v*)&P[1][p][srow][skq] = *(bf8v*)pi[p];
    }
  };
  gload(0);
  for (int k0 = 0; k0 < g.k; k0 += 32) {
    __syncthreads();
    stage(ra, As);
    stage(rb, Bs);
    __syncthreads();
    if (k0 + 32 < g.k) gload(k0 + 32);
    if (wave_live) {
      bf8v a[2][2][NP], bq[2][2][NP];  // [block][re/im][term]
      for (int i = 0; i < 2; i++)
        for (int c = 0; c < 2; c++)
          for (int p = 0; p < NP; p++) {
            a[i][c][p] = *(const bf8v*)&As[c][p][wm + 16 * i + l16][8 * l4];
            bq[i][c][p] = *(const bf8v*)&Bs[c][p][wn + 16 * i + l16][8 * l4];
          }
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
#define MM(acc, ca, pa, cb, pb) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][ca][pa], bq[j][cb][pb], acc, 0, 0, 0)
          // smallest terms first
          if constexpr (TERMS == 6) {
            MM(crp[i][j], 0, 2, 0, 0); MM(crp[i][j], 0, 1, 0, 1); MM(crp[i][j], 0, 0, 0, 2);
            MM(crn[i][j], 1, 2, 1, 0); MM(crn[i][j], 1, 1, 1, 1); MM(crn[i][j], 1, 0, 1, 2);
            MM(ci[i][j], 0, 2, 1, 0); MM(ci[i][j], 0, 1, 1, 1); MM(ci[i][j], 0, 0, 1, 2);
            MM(ci[i][j], 1, 2, 0, 0); MM(ci[i][j], 1, 1, 0, 1); MM(ci[i][j], 1, 0, 0, 2);
          }
          MM(crp[i][j], 0, 1, 0, 0); MM(crp[i][j], 0, 0, 0, 1); MM(crp[i][j], 0, 0, 0, 0);
          MM(crn[i][j], 1, 1, 1, 0); MM(crn[i][j], 1, 0, 1, 1); MM(crn[i][j], 1, 0, 1, 0);
          MM(ci[i][j], 0, 1, 1, 0); MM(ci[i][j], 0, 0, 1, 1); MM(ci[i][j], 0, 0, 1, 0);
          MM(ci[i][j], 1, 1, 0, 0); MM(ci[i][j], 1, 0, 0, 1); MM(ci[i][j], 1, 0, 0, 0);
#undef MM
        }
    }
  }
  if (!wave_live) return;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int e = 0; e < 4; e++) {
        const int row = bm + wm + 16 * i + 4 * l4 + e;
        const int col = bn + wn + 16 * j + l16;
        if (row < g.m && col < g.n) {
          float vr = g.alpha * (crp[i][j][e] - crn[i][j][e]);
          float vi = g.alpha * ci[i][j][e];
          if (row == col) vr += g.betaI;
          for (int q = 0; q < 4; q++)
            if (g.M[q]) {
              const float2 mv = g.M[q][(int64_t)b * g.sm + (int64_t)col * g.ldm + row];
              vr += g.beta[q] * mv.x;
              vi += g.beta[q] * mv.y;
            }
          C[(int64_t)col * g.ldc + row] = make_float2(vr, vi);
        }
      }
}

void launch_cgemm(hipStream_t s, const CGemm& g) {
  const int tm = (g.m + 63) / 64, tn = (g.n + 63) / 64;
  if (g.terms == 6)
    hipLaunchKernelGGL(k_cgemm<6>, dim3(tm * tn * g.nb), dim3(256), 0, s, g);
  else
    hipLaunchKernelGGL(k_cgemm<3>, dim3(tm * tn * g.nb), dim3(256), 0, s, g);
}

// ------------------------------------------------------------------------------------------
__global__ void k_z2c(int64_t n, const double2* __restrict__ a, float2* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = make_float2((float)a[i].x, (float)a[i].y);
}
__global__ void k_c2z(int64_t n, const float2* __restrict__ a, double2* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = make_double2(a[i].x, a[i].y);
}
static int grid_for(int64_t n) { return (int)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536); }
void launch_z2c(hipStream_t s, int64_t count, const double2* src, float2* dst) {
  if (count > 0) hipLaunchKernelGGL(k_z2c, dim3(grid_for(count)), dim3(256), 0, s, count, src, dst);
}
void launch_c2z(hipStream_t s, int64_t count, const float2* src, double2* dst) {
  if (count > 0) hipLaunchKernelGGL(k_c2z, dim3(grid_for(count)), dim3(256), 0, s, count, src, dst);
}

struct CComboArgs { const float2* M[4]; float beta[4]; };
__global__ void k_ccombo(int npad, int nb, float2* __restrict__ C, float betaI, CComboArgs a) {
  const int64_t per = (int64_t)npad * npad;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < per * nb; e += (int64_t)gridDim.x * 256) {
    const int64_t loc = e % per;
    const int row = (int)(loc % npad), col = (int)(loc / npad);
    float vr = (row == col) ? betaI : 0.f, vi = 0.f;
    for (int q = 0; q < 4; q++)
      if (a.M[q]) { const float2 v = a.M[q][e]; vr += a.beta[q] * v.x; vi += a.beta[q] * v.y; }
    C[e] = make_float2(vr, vi);
  }
}
void launch_ccombo(hipStream_t s, int32_t npad, int32_t nb, float2* C, float betaI,
                   const float2* const M[4], const float beta[4]) {
  CComboArgs a;
  for (int q = 0; q < 4; q++) { a.M[q] = M[q]; a.beta[q] = beta[q]; }
  hipLaunchKernelGGL(k_ccombo, dim3(grid_for((int64_t)npad * npad * nb)), dim3(256), 0, s, npad, nb, C, betaI, a);
}

__global__ void k_cparity_copy(int npad, const int* __restrict__ sq, const float2* __restrict__ src,
                               float2* __restrict__ dst) {
  const int b = blockIdx.y;
  if ((sq[b] & 1) == 0) return;
  const int64_t per = (int64_t)npad * npad;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < per; e += (int64_t)gridDim.x * 256)
    dst[b * per + e] = src[b * per + e];
}
void launch_cparity_copy(hipStream_t s, int32_t npad, int32_t nb, const int32_t* sq,
                         const float2* src, float2* dst) {
  const int64_t per = (int64_t)npad * npad;
  const int gx = (int)((per + 255) / 256 < 64 ? (per + 255) / 256 : 64);
  hipLaunchKernelGGL(k_cparity_copy, dim3(gx, nb), dim3(256), 0, s, npad, sq, src, dst);
}

// Psi[:,0] = (I + F)[:,0]: first column of F plus e_1
__global__ void k_cfirst_col(int npad, const float2* __restrict__ F, float2* __restrict__ Psi, int64_t spsi) {
  const int b = blockIdx.x;
  for (int i = threadIdx.x; i < npad; i += blockDim.x) {
    float2 v = F[(int64_t)b * npad * npad + i];
    if (i == 0) v.x += 1.0f;
    Psi[(int64_t)b * spsi + i] = v;
  }
}
void launch_cfirst_col(hipStream_t s, int32_t npad, int32_t nb, const float2* E, float2* Psi,
                       int64_t spsi) {
  hipLaunchKernelGGL(k_cfirst_col, dim3(nb), dim3(256), 0, s, npad, E, Psi, spsi);
}

}  // namespace ctem
