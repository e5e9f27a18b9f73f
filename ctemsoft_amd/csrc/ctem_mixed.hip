// Mixed-precision kernels: complex GEMM on split 16-bit MFMA (v_mfma_f32_16x16x32_bf16 and
// v_mfma_f32_16x16x32_f16, gfx950) with FP32 accumulation, for FP32 complex matrices.
//
// Split-BF16: each FP32 operand x is split into bfloat16 terms x = x0 + x1 + x2 + O(2^-24 |x|)
// (x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1)). A real product is then
//   scheme 3:  x0*y0 + x0*y1 + x1*y0                       (~2^-16 relative)
//   scheme 6:  the above + x0*y2 + x1*y1 + x2*y0           (~FP32 accuracy)
// Split-FP16 (scheme 16): each operand matrix is scaled by 2^e with e = 14 - ceil(log2 max|x|)
// (max over the matrix, tracked by the kernel that wrote it), so that every scaled entry lies
// below 2^14 < 65504; then x 2^e = x0 + x1 + O(2^-22 |x| 2^e) with x0 = fp16(x 2^e),
// x1 = fp16(x 2^e - x0). The product x0*y0 + x0*y1 + x1*y0 has ~2^-21 error relative to
// max|x| max|y|, 32 times better than split-BF16 scheme 3 for the same three MFMAs, because
// FP16 carries 11 significand bits to BF16's 8; the scaling keeps residuals of entries down
// to ~2^-24 max|x| in the normal FP16 range. Non-finite entries pass through as in FP32.
// All launches have k a multiple of 8 (matrices padded to multiples of 32).
// Each term is one MFMA with FP32 accumulation. Complex products use four real products
// (re*re, im*im, re*im, im*re) into three accumulators (re+, re-, im).
//
// Tile: 64x64 per workgroup of 4 waves (32x32 per wave = 2x2 MFMA blocks), K chunks of 32
// staged through LDS as bfloat16 planes with K contiguous, so that each lane reads its
// 8-element fragment with one 16-byte LDS load. Operand layout of the 16x16x32 MFMA
// (lane l): A[l%16][8*(l/16)+t], B[8*(l/16)+t][l%16], t = 0..7; result element e of lane l is
// C[4*(l/16)+e][l%16].
#include <hip/hip_bf16.h>
#include <algorithm>
#include <type_traits>
#include <hip/hip_runtime.h>

#include "ctem_ebsd.h"
#include "ctem_internal.h"

namespace ctem {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef _Float16 h8v __attribute__((ext_vector_type(8)));

template <bool F16> struct Split;
template <> struct Split<false> {
  typedef __bf16 E;
  typedef bf8v V;
  __device__ static f4v mma(V a, V b, f4v c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};
template <> struct Split<true> {
  typedef _Float16 E;
  typedef h8v V;
  __device__ static f4v mma(V a, V b, f4v c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};

// LDS tiles hold 32 k values per row (or column) as four 16-byte chunks of 8, with the chunk
// index swizzled by the row: the fragment reads (8 lanes x 16 B of 8 different rows per LDS
// cycle) and the 4-byte staging writes of the Y tile then fall on distinct banks without
// padding
__device__ __forceinline__ int cg_swz(int row, int chunk) { return chunk ^ ((row >> 1) & 3); }

// two 16-bit values as one 32-bit word, the first in the low half
template <typename E>
__device__ __forceinline__ uint32_t pack2(E lo, E hi) {
  return (uint32_t)__builtin_bit_cast(uint16_t, lo) | ((uint32_t)__builtin_bit_cast(uint16_t, hi) << 16);
}

template <typename E>
__device__ __forceinline__ void split3(float x, E& a, E& b, E& c) {
  a = (E)x;
  const float r1 = x - (float)a;
  b = (E)r1;
  const float r2 = r1 - (float)b;
  c = (E)r2;
}

template <int TERMS, bool F16>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TERMS == 6 ? 2 : 3))) void k_cgemm(CGemm g) {
  typedef typename Split<F16>::E E;
  typedef typename Split<F16>::V V;
  constexpr int NP = (TERMS == 6) ? 3 : 2;  // 16-bit planes per real component
  const int tm = (g.m + 63) / 64, tn = (g.n + 63) / 64;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int b = lb / (tm * tn);
  const int t = lb % (tm * tn);
  const int bm = (t % tm) * 64, bn = (t / tm) * 64;
  if (g.active && g.step >= g.active[b]) return;
  const float2* X = g.X + (int64_t)b * g.sx;
  const float2* Y = g.Y + (int64_t)b * g.sy;
  float2* C = g.C + (int64_t)b * g.sc;
  int ex = 0, ey = 0;
  if constexpr (F16) {
    ex = f16_scale_exp(g.xmax[b]);
    ey = f16_scale_exp(g.ymax[b]);
  }

  // planes: [component re/im][term][row or col][k, swizzled chunks]
  __shared__ __attribute__((aligned(16))) E As[2][NP][64][32];
  __shared__ __attribute__((aligned(16))) E Bs[2][NP][64][32];
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

  // register-prefetch pipeline; the global loads of the next chunk are in flight during the
  // MFMAs of the current one. X (column-major, rows contiguous): thread t loads row t & 63 at
  // k = 8*(t >> 6) .. +7, each load coalesced over the 64 rows of a wave. Y (column-major, k
  // contiguous): in round j thread t loads column (t >> 4) + 16 j at k = 2*(t & 15), +1 as one
  // 16-byte load, 16 lanes covering the column's 256 contiguous bytes.
  const int arow = tid & 63, akq = tid >> 6;
  const int bkp = tid & 15, bcol = tid >> 4;
  // loads are branch-free: indices are clamped into the matrix and out-of-range values zeroed
  const bool arow_ok = bm + arow < g.m;
  const float2* xr0 = X + min(bm + arow, g.m - 1);
  float2 ra[8];
  float4 rb[4];
  auto gload = [&](int k0) {
    for (int j = 0; j < 8; j++) {
      const int kk = k0 + 8 * akq + j;
      const float2 v = xr0[(int64_t)min(kk, g.k - 1) * g.ldx];
      ra[j] = (arow_ok && kk < g.k) ? v : make_float2(0.f, 0.f);
    }
    for (int j = 0; j < 4; j++) {
      const int col = bn + bcol + 16 * j;
      const int kk = k0 + 2 * bkp;
      const float2* yc = Y + (int64_t)min(col, g.n - 1) * g.ldy;
      if (g.vec) {
        // k is even and ldy is even: the pair is 16-byte aligned; with k a multiple of 8 the
        // pair lies wholly inside or wholly outside the matrix
        const float4 v = *(const float4*)(yc + min(kk, g.k - 2));
        rb[j] = (col < g.n && kk < g.k) ? v : make_float4(0.f, 0.f, 0.f, 0.f);
      } else {
        const float2 v0 = yc[min(kk, g.k - 1)], v1 = yc[min(kk + 1, g.k - 1)];
        const bool ok = col < g.n;
        rb[j] = make_float4((ok && kk < g.k) ? v0.x : 0.f, (ok && kk < g.k) ? v0.y : 0.f,
                            (ok && kk + 1 < g.k) ? v1.x : 0.f, (ok && kk + 1 < g.k) ? v1.y : 0.f);
      }
    }
  };
  const float sx = ldexpf(1.0f, ex), sy = ldexpf(1.0f, ey);  // exact powers of two
  auto split_scaled = [&](float x, float sc, E (&o)[NP]) {
    if constexpr (F16) x *= sc;
    E a, bb, c;
    split3(x, a, bb, c);
    o[0] = a; o[1] = bb;
    if constexpr (NP == 3) o[2] = c;
  };
  auto stage_x = [&]() {
    __attribute__((aligned(16))) E pr[NP][8], pi[NP][8];
    for (int j = 0; j < 8; j++) {
      E o[NP];
      split_scaled(ra[j].x, sx, o);
      for (int p = 0; p < NP; p++) pr[p][j] = o[p];
      split_scaled(ra[j].y, sx, o);
      for (int p = 0; p < NP; p++) pi[p][j] = o[p];
    }
    const int ch = 8 * cg_swz(arow, akq);
    for (int p = 0; p < NP; p++) {
      *(V*)&As[0][p][arow][ch] = *(V*)pr[p];
      *(V*)&As[1][p][arow][ch] = *(V*)pi[p];
    }
  };
  auto stage_y = [&]() {
    for (int j = 0; j < 4; j++) {
      const int col = bcol + 16 * j;
      const int off = 8 * cg_swz(col, bkp >> 2) + 2 * (bkp & 3);
      E o0[NP], o1[NP];
      split_scaled(rb[j].x, sy, o0);
      split_scaled(rb[j].z, sy, o1);
      for (int p = 0; p < NP; p++) *(uint32_t*)&Bs[0][p][col][off] = pack2(o0[p], o1[p]);
      split_scaled(rb[j].y, sy, o0);
      split_scaled(rb[j].w, sy, o1);
      for (int p = 0; p < NP; p++) *(uint32_t*)&Bs[1][p][col][off] = pack2(o0[p], o1[p]);
    }
  };
  gload(0);
  // A wave whose 32x32 quadrant lies past the matrix edge only stages its share of the
  // tiles; the choice is made once, outside the loop, so that the live copy of the loop keeps
  // its accumulators in place across iterations.
  auto kloop = [&](auto live) {
  for (int k0 = 0; k0 < g.k; k0 += 32) {
    __syncthreads();
    stage_x();
    stage_y();
    __syncthreads();
    if (k0 + 32 < g.k) gload(k0 + 32);
    if constexpr (!decltype(live)::value) continue;
    V a[2][2][NP], bq[2][2][NP];  // [block][re/im][term]
    for (int i = 0; i < 2; i++) {
      const int ra_ = wm + 16 * i + l16, rb_ = wn + 16 * i + l16;
      for (int c = 0; c < 2; c++)
        for (int p = 0; p < NP; p++) {
          a[i][c][p] = *(const V*)&As[c][p][ra_][8 * cg_swz(ra_, l4)];
          bq[i][c][p] = *(const V*)&Bs[c][p][rb_][8 * cg_swz(rb_, l4)];
        }
    }
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) {
#define MM(acc, ca, pa, cb, pb) acc = Split<F16>::mma(a[i][ca][pa], bq[j][cb][pb], acc)
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
  };
  if (wave_live)
    kloop(std::true_type{});
  else
    kloop(std::false_type{});
  if (!wave_live) return;
  // undo the operand scaling (a power of two)
  const float unscale = ldexpf(1.0f, -(ex + ey));
  const float alpha = g.alpha * unscale, alpha2 = g.alpha2 * unscale;
  float2* C2 = g.C2 ? g.C2 + (int64_t)b * g.sc : nullptr;
  float vmax = 0.0f, vmax2 = 0.0f;
  // each lane holds rows row0 .. row0+3 of one column: contiguous in the column-major outputs
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int row0 = bm + wm + 16 * i + 4 * l4;
      const int col = bn + wn + 16 * j + l16;
      if (col >= g.n || row0 >= g.m) continue;
      float vr[4], vi[4], wr[4], wi[4];
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const float pr = crp[i][j][e] - crn[i][j][e], pi = ci[i][j][e];
        vr[e] = alpha * pr; vi[e] = alpha * pi;
        wr[e] = alpha2 * pr; wi[e] = alpha2 * pi;
        if (row0 + e == col) { vr[e] += g.betaI; wr[e] += g.betaI2; }
      }
      const int64_t mo = (int64_t)b * g.sm + (int64_t)col * g.ldm + row0;
      const int64_t co = (int64_t)col * g.ldc + row0;
      if (g.vec) {
        // m is a multiple of 4: the four rows are all inside
        for (int q = 0; q < 4; q++)
          if (g.M[q]) {
            const float4 m0 = *(const float4*)(g.M[q] + mo), m1 = *(const float4*)(g.M[q] + mo + 2);
            const float mr[4] = {m0.x, m0.z, m1.x, m1.z}, mi[4] = {m0.y, m0.w, m1.y, m1.w};
            for (int e = 0; e < 4; e++) {
              vr[e] += g.beta[q] * mr[e]; vi[e] += g.beta[q] * mi[e];
              wr[e] += g.beta2[q] * mr[e]; wi[e] += g.beta2[q] * mi[e];
            }
          }
        *(float4*)(C + co) = make_float4(vr[0], vi[0], vr[1], vi[1]);
        *(float4*)(C + co + 2) = make_float4(vr[2], vi[2], vr[3], vi[3]);
        if (C2) {
          *(float4*)(C2 + co) = make_float4(wr[0], wi[0], wr[1], wi[1]);
          *(float4*)(C2 + co + 2) = make_float4(wr[2], wi[2], wr[3], wi[3]);
        }
        for (int e = 0; e < 4; e++) {
          vmax = nmax(vmax, nmax(fabsf(vr[e]), fabsf(vi[e])));
          vmax2 = nmax(vmax2, nmax(fabsf(wr[e]), fabsf(wi[e])));
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; e++) {
          if (row0 + e >= g.m) continue;
          for (int q = 0; q < 4; q++)
            if (g.M[q]) {
              const float2 mv = g.M[q][mo + e];
              vr[e] += g.beta[q] * mv.x; vi[e] += g.beta[q] * mv.y;
              wr[e] += g.beta2[q] * mv.x; wi[e] += g.beta2[q] * mv.y;
            }
          C[co + e] = make_float2(vr[e], vi[e]);
          if (C2) C2[co + e] = make_float2(wr[e], wi[e]);
          vmax = nmax(vmax, nmax(fabsf(vr[e]), fabsf(vi[e])));
          vmax2 = nmax(vmax2, nmax(fabsf(wr[e]), fabsf(wi[e])));
        }
      }
    }
  if (g.cmax) wave_max_to(vmax, g.cmax + b);
  if (C2 && g.cmax2) wave_max_to(vmax2, g.cmax2 + b);
}

void launch_cgemm(hipStream_t s, const CGemm& gin) {
  CGemm g = gin;
  auto al16 = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  bool vec = g.m % 4 == 0 && g.k % 8 == 0 && g.ldy % 2 == 0 && g.ldc % 2 == 0 && g.ldm % 2 == 0 &&
             g.sy % 2 == 0 && g.sc % 2 == 0 && g.sm % 2 == 0 && al16(g.Y) && al16(g.C) && al16(g.C2);
  for (int q = 0; q < 4; q++) vec = vec && al16(g.M[q]);
  g.vec = vec ? 1 : 0;
  const int tm = (g.m + 63) / 64, tn = (g.n + 63) / 64;
  if (g.terms == CTEM_SPLIT_FP16X3)
    hipLaunchKernelGGL((k_cgemm<3, true>), dim3(tm * tn * g.nb), dim3(256), 0, s, g);
  else if (g.terms == CTEM_SPLIT_BF16X6)
    hipLaunchKernelGGL((k_cgemm<6, false>), dim3(tm * tn * g.nb), dim3(256), 0, s, g);
  else
    hipLaunchKernelGGL((k_cgemm<3, false>), dim3(tm * tn * g.nb), dim3(256), 0, s, g);
}

// ------------------------------------------------------------------------------------------
// 1-norm (max column sum of |a|, accumulated in FP64), squaring count s with
// ||A||_1 / 2^s <= theta, in-place scaling by 2^-s and the max |entry| afterwards; one
// workgroup per matrix
__global__ __launch_bounds__(256) void k_norm_scale32(int npad, float2* __restrict__ A, double theta,
                                                      int smax, int* __restrict__ sq, int* __restrict__ err,
                                                      float* __restrict__ amax) {
  const int b = blockIdx.x;
  float2* Ab = A + (int64_t)b * npad * npad;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ double wmax[4];
  __shared__ int s_sh;
  double cmax = 0.0;
  for (int j = wave; j < npad; j += 4) {
    double sum = 0.0;
    for (int i = lane; i < npad; i += 64) {
      const float2 v = Ab[(int64_t)j * npad + i];
      sum += (double)hypotf(v.x, v.y);
    }
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
    cmax = fmax(cmax, sum);
    if (!(sum == sum)) cmax = sum;  // keep a NaN
  }
  if (lane == 0) wmax[wave] = cmax;
  __syncthreads();
  if (tid == 0) {
    double nrm = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    for (int w = 0; w < 4; w++) if (!(wmax[w] == wmax[w])) nrm = wmax[w];
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
  const float f = ldexpf(1.0f, -s_sh);
  float vmax = 0.0f;
  for (int64_t e = tid; e < (int64_t)npad * npad; e += 256) {
    const float2 v = Ab[e];
    const float2 w = make_float2(v.x * f, v.y * f);
    Ab[e] = w;
    vmax = nmax(vmax, nmax(fabsf(w.x), fabsf(w.y)));
  }
  wave_max_to(vmax, amax + b);
}
void launch_norm_scale32(hipStream_t s, int32_t npad, int32_t nb, float2* A, double theta,
                         int32_t smax, int32_t* sq, int32_t* err, float* amax) {
  hipLaunchKernelGGL(k_norm_scale32, dim3(nb), dim3(256), 0, s, npad, A, theta, smax, sq, err, amax);
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
  const int gx = (int)std::min<int64_t>((per + 255) / 256, 64);
  hipLaunchKernelGGL(k_cparity_copy, dim3(gx, nb), dim3(256), 0, s, npad, sq, src, dst);
}

}  // namespace ctem
