// Mixed form on split-FP16 matrices (the default CTEM_SPLIT_FP16X3 scheme): every matrix of the
// Taylor stage and of the squarings is stored once, by the kernel that produces it, as four
// FP16 planes (real high/low, imaginary high/low) with one exponent per matrix (layout:
// split_index in ctem_internal.h). The products
// then copy their tiles into LDS unchanged, and the three FP16 MFMA products per real product
// (high*high + high*low + low*high, FP32 accumulation) see the same operands as the FP32 form
// split on the fly in ctem_mixed.hip.
//
// Tile layout of k_sgemm (64 x 64 output tile per 256-thread workgroup, 32 x 32 per wave):
// the X tile is staged as it lies in memory, k-major with the 64 rows contiguous, and the MFMA
// A operand (8 consecutive k of one row per lane) is read with the LDS transpose read
// ds_read_b64_tr_b16. Within a group of 16 lanes that read 4 halfs each, lane i receives
// element i % 4 of lane 4e + i / 4 as its element e; lane 4e + j reads rows r0 + 4j .. +3 at
// column k0 + e, so that lane i ends up with rows r0 + i at columns k0 .. k0 + 3.
#include <hip/hip_runtime.h>
#include <algorithm>
#include "ctem_ebsd.h"
#include "ctem_internal.h"

namespace ctem {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef _Float16 h4v __attribute__((ext_vector_type(4)));
typedef short s4v __attribute__((ext_vector_type(4)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
typedef unsigned int u2v __attribute__((ext_vector_type(2)));
#define LDS_PTR(T, p) ((__attribute__((address_space(3))) T*)(p))

// s -> high + low, both FP16 (round to nearest), |s| < 2^14
__device__ __forceinline__ void split2(float s, _Float16& hi, _Float16& lo) {
  hi = (_Float16)s;
  lo = (_Float16)(s - (float)hi);
}

// Stores four consecutive entries (rows r .. r+3 of one column, values already scaled by 2^e)
// in split form. dst points at this lane's rows in plane 0 of their 8-row group. The lane pair
// at lanes l and l ^ 16 holds the whole group (the even lane of the pair rows 0-3, the odd lane
// rows 4-7) and trades halves, so that the even lane stores the two real planes and the odd
// lane the two imaginary planes as 16-byte pieces. All 64 lanes must take part.
__device__ __forceinline__ void store_split4(_Float16* dst, const float (&re)[4], const float (&im)[4], bool odd) {
  h4v o0, o1, o2, o3;
#pragma unroll
  for (int e = 0; e < 4; e++) {
    _Float16 h, l;
    split2(re[e], h, l); o0[e] = h; o1[e] = l;
    split2(im[e], h, l); o2[e] = h; o3[e] = l;
  }
  // the even lane sends its imaginary rows 0-3, the odd lane its real rows 4-7:
  // v_permlane16_swap(x, y) trades the odd 16-lane rows of x with the even rows of y (returned
  // as the new x, the new y), so the odd lane finds the even lane's y in the new x and the even
  // lane the odd lane's x in the new y
  const u2v a0 = __builtin_bit_cast(u2v, o0), a1 = __builtin_bit_cast(u2v, o1);
  const u2v a2 = __builtin_bit_cast(u2v, o2), a3 = __builtin_bit_cast(u2v, o3);
  u2v r0, r1;
#pragma unroll
  for (int w = 0; w < 2; w++) {
    const auto t0 = __builtin_amdgcn_permlane16_swap(a0[w], a2[w], false, false);
    const auto t1 = __builtin_amdgcn_permlane16_swap(a1[w], a3[w], false, false);
    r0[w] = odd ? t0[0] : t0[1];
    r1[w] = odd ? t1[0] : t1[1];
  }
  const h4v q0 = __builtin_bit_cast(h4v, r0), q1 = __builtin_bit_cast(h4v, r1);
  const h8v w0 = odd ? __builtin_shufflevector(q0, o2, 0, 1, 2, 3, 4, 5, 6, 7)
                     : __builtin_shufflevector(o0, q0, 0, 1, 2, 3, 4, 5, 6, 7);
  const h8v w1 = odd ? __builtin_shufflevector(q1, o3, 0, 1, 2, 3, 4, 5, 6, 7)
                     : __builtin_shufflevector(o1, q1, 0, 1, 2, 3, 4, 5, 6, 7);
  // the group starts 4 rows earlier on odd lanes; their planes are the imaginary ones
  _Float16* gb = dst - (odd ? 4 : 0) + (odd ? 16 : 0);
  *(h8v*)gb = w0;
  *(h8v*)(gb + 8) = w1;
}

// One workgroup of 4 waves per 64 x (32 JN) tile of C; wave w computes rows 32 (w >> 1) .. +31
// and columns 16 JN (w & 1) .. +16 JN - 1 as 2 x JN MFMA subtiles of 16 x 16.
template <int JN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(JN == 4 ? 2 : 3))) void k_sgemm(SGemm g) {
  constexpr int BN = 32 * JN;  // tile columns
  constexpr int YG = JN / 2;   // Y group blocks staged per thread
  const int n = g.n;
  const int tm = (n + 63) >> 6, tn = (n + BN - 1) / BN;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  // symmetric mode (BN = 64): tiles (ti, tj) with ti <= tj, t = tj (tj + 1) / 2 + ti
  const int ntile = g.sym ? tm * (tm + 1) / 2 : tm * tn;
  const int t = lb % ntile;
  const int b = g.order ? g.order[lb / ntile] : lb / ntile;
  int bm, bn;
  if (g.sym) {
    int tj = 0;
    while ((tj + 1) * (tj + 2) / 2 <= t) tj++;
    bm = (t - tj * (tj + 1) / 2) * 64;
    bn = tj * 64;
  } else {
    bm = (t % tm) * 64;
    bn = (t / tm) * BN;
  }
  const bool offd = g.sym && bm != bn;
  if (!g.order && g.active && g.step >= g.active[b] + g.active_add) return;
  const int64_t per = (int64_t)n * n;
  const int64_t mb = (int64_t)b * 4 * per;
  const _Float16* X = (const _Float16*)g.X + mb;
  const _Float16* Y = (const _Float16*)g.Y + mb;

  // X tile: [plane][k 32][row 64], 32-byte row segments swizzled by k (segment ^ (k & 3) ^
  // ((k >> 3) & 3); the k >> 3 term gives the transpose reads of neighbouring 16-lane groups,
  // at k and k + 8, different segments, and halved the LDS bank-conflict cycles measured with
  // k & 3 alone);
  // Y tile: [plane][col BN][k 32], 16-byte k chunks swizzled by the column (chunk ^ ((col >> 1) & 3))
  // (one array: after the k loop it holds the transposed FP32 copy of an off-diagonal tile)
  __shared__ __attribute__((aligned(16))) _Float16 Ls[4 * 32 * 64 + 4 * BN * 32];
  _Float16* const As = Ls;
  _Float16* const Bs = Ls + 4 * 32 * 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * (16 * JN);
  const int l16 = lane & 15, l4 = lane >> 4;
  // column subtiles of this wave inside the matrix (n is a multiple of 32): 0, 2 or JN
  const int jlive = (bm + wm < n) ? max(0, min(JN, (n - bn - wn) / 16)) : 0;

  // staging, the 64-byte group block (8 rows, 4 planes) per thread: X rows bm + 8 (t & 7) .. +7
  // of column k0 + (t >> 3); Y columns bn + (t >> 2) + 64 m at rows k0 + 8 (t & 3) .. +7. Rows
  // and columns past the matrix are clamped onto its last ones: they feed only output rows and
  // columns that are not written.
  const int xk = tid >> 3, xr = 8 * (tid & 7);
  const _Float16* xp = X + split_index(n, min(bm + xr, n - 8), xk, 0);
  const int xlds = xk * 64 + 16 * ((xr >> 4) ^ (xk & 3) ^ ((xk >> 3) & 3)) + (xr & 8);
  const int yq = tid & 3;
  const _Float16* yp[YG];
  int ylds[YG];
  #pragma unroll
  for (int m = 0; m < YG; m++) {
    const int ycol = (tid >> 2) + 64 * m;
    yp[m] = Y + split_index(n, 8 * yq, min(bn + ycol, n - 1), 0);
    ylds[m] = ycol * 32 + 8 * (yq ^ ((ycol >> 1) & 3));
  }
  // Symmetric mode (JN = 2): only the tiles on and above the diagonal are stored, so the
  // chunks of X left of the diagonal tile (k0 < bm) are staged from X^T (the Y-style columns
  // bm + (t >> 2), rows k0 + 8 (t & 3)) into the Y-style layout, and the chunks of Y below it
  // (k0 >= bn + 64) from Y^T (column k0 + (t >> 3), rows bn + 8 (t & 7)) into the X-style
  // layout; the fragment reads of such a chunk swap between the transpose read and the
  // 16-byte read.
  auto xtrans = [&](int k0) __attribute__((always_inline)) { return g.sym && k0 < bm; };
  auto ytrans = [&](int k0) __attribute__((always_inline)) { return g.sym && k0 >= bn + 64; };
  const _Float16* xtp = X + split_index(n, 8 * yq, min(bm + (tid >> 2), n - 1), 0);
  const _Float16* ytp = Y + split_index(n, min(bn + xr, n - 8), xk, 0);
  u4v rx[4], ry[YG][4];
  auto gload = [&](int k0) __attribute__((always_inline)) {
    const _Float16* xs = xtrans(k0) ? xtp + 4 * k0 : xp + (int64_t)k0 * 4 * n;
    #pragma unroll
    for (int p = 0; p < 4; p++) rx[p] = *(const u4v*)(xs + 8 * p);
    if constexpr (YG == 1) {
      const _Float16* ys = ytrans(k0) ? ytp + (int64_t)k0 * 4 * n : yp[0] + 4 * k0;
      #pragma unroll
      for (int p = 0; p < 4; p++) ry[0][p] = *(const u4v*)(ys + 8 * p);
    } else {
      #pragma unroll
      for (int m = 0; m < YG; m++)
        #pragma unroll
        for (int p = 0; p < 4; p++) ry[m][p] = *(const u4v*)(yp[m] + 4 * k0 + 8 * p);
    }
  };

  f4v crp[2][JN], crn[2][JN], ci[2][JN];
  #pragma unroll
  for (int i = 0; i < 2; i++)
    #pragma unroll
    for (int j = 0; j < JN; j++) {
      crp[i][j] = (f4v){0, 0, 0, 0};
      crn[i][j] = (f4v){0, 0, 0, 0};
      ci[i][j] = (f4v){0, 0, 0, 0};
    }
  // fragment offsets: A (transpose read, lane reads rows r0 + 4 (l16 & 3) .. +3 at column
  // 8 l4 + (l16 >> 2), and 4 columns further for the upper half), B (16 bytes at chunk l4)
  const int kl = 8 * l4 + (l16 >> 2);
  int aoff[2], boff[JN], aofft[2], bofft[JN];  // (and the same reads on the swapped layouts)
  #pragma unroll
  for (int i = 0; i < 2; i++) {
    const int r0 = wm + 16 * i, row = r0 + l16;
    aoff[i] = kl * 64 + 16 * ((r0 >> 4) ^ (kl & 3) ^ ((kl >> 3) & 3)) + 4 * (l16 & 3);
    aofft[i] = row * 32 + 8 * (l4 ^ ((row >> 1) & 3));
  }
  #pragma unroll
  for (int j = 0; j < JN; j++) {
    const int c0 = wn + 16 * j;
    bofft[j] = kl * 64 + 16 * ((c0 >> 4) ^ (kl & 3) ^ ((kl >> 3) & 3)) + 4 * (l16 & 3);
  }
  #pragma unroll
  for (int j = 0; j < JN; j++) {
    const int col = wn + 16 * j + l16;
    boff[j] = col * 32 + 8 * (l4 ^ ((col >> 1) & 3));
  }

  // exponents and maxima of the operands (read before the loop: its barriers would hold
  // these loads back to the epilogue)
  const float xmx = g.xmax[b], ymx = g.ymax[b];
  const int ey = g.ye[b];
  const int exy = g.xe[b] + ey;
  // with M[0] = X, beta_0 X enters the accumulators (units 2^-(ex+ey)) from the X tiles staged
  // at k0 = each 32-column block of the output: X[rows][cols] = As[k = col - k0][rows]
  const bool mx0 = g.m0_is_x != 0;
  const float dm0 = mx0 ? (g.beta[0] / g.alpha) * ldexpf(1.0f, ey) : 0.0f;
  float mmx[2] = {0.f, 0.f};
  int mex[2] = {0, 0};
  #pragma unroll
  for (int q = 0; q < 2; q++)
    if (g.M[q]) { mmx[q] = g.mmax[q][b]; mex[q] = g.me[q][b]; }
  // raw X = Y: the 64-byte block of 8 FP32 complex rows sits where the split group block of the
  // same rows would (both 8 bytes per entry), and is split here into the four plane rows
  const float fraw = g.xraw ? ldexpf(1.0f, g.xe[b] - g.rs[b]) : 0.0f;
  auto split_block = [&](u4v (&r)[4]) __attribute__((always_inline)) {
    h8v p0, p1, p2, p3;
    #pragma unroll
    for (int q = 0; q < 4; q++) {
      const f4v v = __builtin_bit_cast(f4v, r[q]);  // rows 2q, 2q + 1: re, im, re, im
      #pragma unroll
      for (int t = 0; t < 2; t++) {
        _Float16 h, l;
        split2(v[2 * t] * fraw, h, l); p0[2 * q + t] = h; p1[2 * q + t] = l;
        split2(v[2 * t + 1] * fraw, h, l); p2[2 * q + t] = h; p3[2 * q + t] = l;
      }
    }
    r[0] = __builtin_bit_cast(u4v, p0); r[1] = __builtin_bit_cast(u4v, p1);
    r[2] = __builtin_bit_cast(u4v, p2); r[3] = __builtin_bit_cast(u4v, p3);
  };

  gload(0);
  // The number of live column subtiles of the wave (0, 2 or JN) is chosen once, outside the
  // loop, so that each copy of the loop keeps its accumulators in place across iterations.
  auto kloop = [&](auto live) __attribute__((always_inline)) {
    constexpr int JL = decltype(live)::value;
    for (int k0 = 0; k0 < n; k0 += 32) {
      const bool xt = xtrans(k0), yt = ytrans(k0);
      __syncthreads();
      if constexpr (YG == 1) {
        const int xo = xt ? ylds[0] : xlds, yo = yt ? xlds : ylds[0];
        if (g.xraw) {
          split_block(rx);
          split_block(ry[0]);
        }
        #pragma unroll
        for (int p = 0; p < 4; p++) *(u4v*)&As[p * 2048 + xo] = rx[p];
        #pragma unroll
        for (int p = 0; p < 4; p++) *(u4v*)&Bs[p * 2048 + yo] = ry[0][p];
      } else {
        #pragma unroll
        for (int p = 0; p < 4; p++) *(u4v*)&As[p * 2048 + xlds] = rx[p];
        #pragma unroll
        for (int m = 0; m < YG; m++)
          #pragma unroll
          for (int p = 0; p < 4; p++) *(u4v*)&Bs[p * 32 * BN + ylds[m]] = ry[m][p];
      }
      __syncthreads();
      if (k0 + 32 < n) gload(k0 + 32);
      if constexpr (JL == 0) continue;
      // 8 consecutive k of one row (or column) per lane: transpose reads on the X-style layout,
      // 16-byte reads on the Y-style one
      auto trread = [&](const _Float16* base) __attribute__((always_inline)) {
        const s4v lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16(LDS_PTR(s4v, base));
        const s4v hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16(LDS_PTR(s4v, base + 256));
        return __builtin_bit_cast(h8v, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
      };
      h8v a[2][2][2];  // [block][re/im][high/low]
      if (!xt) {
        #pragma unroll
        for (int i = 0; i < 2; i++)
          #pragma unroll
          for (int P = 0; P < 4; P++) a[i][P >> 1][P & 1] = trread(&As[P * 2048 + aoff[i]]);
      } else {
        #pragma unroll
        for (int i = 0; i < 2; i++)
          #pragma unroll
          for (int P = 0; P < 4; P++) a[i][P >> 1][P & 1] = *(const h8v*)&As[P * 2048 + aofft[i]];
      }
      #pragma unroll
      for (int j = 0; j < JL; j++) {
        h8v bq[2][2];
        if (!yt) {
          #pragma unroll
          for (int P = 0; P < 4; P++) bq[P >> 1][P & 1] = *(const h8v*)&Bs[P * 32 * BN + boff[j]];
        } else {
          #pragma unroll
          for (int P = 0; P < 4; P++) bq[P >> 1][P & 1] = trread(&Bs[P * 32 * BN + bofft[j]]);
        }
        #pragma unroll
        for (int i = 0; i < 2; i++) {
#define MM(acc, ca, ha, cb, hb) acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i][ca][ha], bq[cb][hb], acc, 0, 0, 0)
          // smallest terms first
          MM(crp[i][j], 0, 1, 0, 0); MM(crp[i][j], 0, 0, 0, 1); MM(crp[i][j], 0, 0, 0, 0);
          MM(crn[i][j], 1, 1, 1, 0); MM(crn[i][j], 1, 0, 1, 1); MM(crn[i][j], 1, 0, 1, 0);
          MM(ci[i][j], 0, 1, 1, 0); MM(ci[i][j], 0, 0, 1, 1); MM(ci[i][j], 0, 0, 1, 0);
          MM(ci[i][j], 1, 1, 0, 0); MM(ci[i][j], 1, 0, 0, 1); MM(ci[i][j], 1, 0, 0, 0);
#undef MM
        }
      }
      if (mx0) {
        #pragma unroll
        for (int hb = 0; hb < JL / 2; hb++)
          if (k0 == bn + wn + 32 * hb) {
            #pragma unroll
            for (int i = 0; i < 2; i++)
              #pragma unroll
              for (int jj = 0; jj < 2; jj++) {
                const int j = 2 * hb + jj;
                const int kk = 16 * jj + l16, r = wm + 16 * i + 4 * l4;
                const int o = kk * 64 + 16 * ((r >> 4) ^ (kk & 3) ^ ((kk >> 3) & 3)) + (r & 15);
                const h4v p0 = *(const h4v*)&As[o], p1 = *(const h4v*)&As[2048 + o];
                const h4v p2 = *(const h4v*)&As[4096 + o], p3 = *(const h4v*)&As[6144 + o];
                #pragma unroll
                for (int e = 0; e < 4; e++) {
                  crp[i][j][e] += dm0 * ((float)p0[e] + (float)p1[e]);
                  ci[i][j][e] += dm0 * ((float)p2[e] + (float)p3[e]);
                }
              }
          }
      }
    }
  };
  if (jlive == JN)
    kloop(std::integral_constant<int, JN>{});
  else if (jlive == 2 && JN > 2)
    kloop(std::integral_constant<int, 2>{});
  else
    kloop(std::integral_constant<int, 0>{});

  // output exponents from bounds on the entries (the same in every workgroup of the matrix)
  const float xy = (float)n * xmx * ymx;
  float bnd = fabsf(g.alpha) * xy + fabsf(g.betaI), bnd2 = fabsf(g.alpha2) * xy + fabsf(g.betaI2);
  #pragma unroll
  for (int q = 0; q < 2; q++)
    if (g.M[q]) {
      bnd += fabsf(g.beta[q]) * mmx[q];
      bnd2 += fabsf(g.beta2[q]) * mmx[q];
    }
  const int ec = f16_scale_exp(bnd), ec2 = f16_scale_exp(bnd2);
  if (t == 0 && tid == 0) {
    g.ce[b] = ec;
    if (g.C2) g.ce2[b] = ec2;
  }
  // the FP32 copy for the depth kernel (F, and E^2 - I, E^4 - I after it) is read in full: off-diagonal
  // tiles also store their transpose there; after the matrix's last product no C is stored
  float2* Cf = nullptr;
  bool last = false;
  if (g.Cf && g.f_sq[b] == g.f_at) {
    Cf = g.Cf + (int64_t)b * per;
    last = g.Cf2 == nullptr;
  } else if (g.Cf2 && g.f_sq[b] + 1 == g.f_at) {
    Cf = g.Cf2 + (int64_t)b * per;
    last = g.Cf4 == nullptr;
  } else if (g.Cf4 && g.f_sq[b] + 2 == g.f_at) {
    Cf = g.Cf4 + (int64_t)b * per;
    last = true;
  }
  const bool mirror = offd && Cf != nullptr;
  if (jlive == 0 && !mirror) return;
  // the accumulators carry 2^(ex + ey); the outputs are stored times 2^ec
  const float fa = g.alpha * ldexpf(1.0f, ec - exy), fa2 = g.alpha2 * ldexpf(1.0f, ec2 - exy);
  const float fi = g.betaI * ldexpf(1.0f, ec), fi2 = g.betaI2 * ldexpf(1.0f, ec2);
  float fm[2] = {0.f, 0.f}, fm2[2] = {0.f, 0.f};
  #pragma unroll
  for (int q = 0; q < 2; q++)
    if (g.M[q]) {
      fm[q] = g.beta[q] * ldexpf(1.0f, ec - mex[q]);
      fm2[q] = g.beta2[q] * ldexpf(1.0f, ec2 - mex[q]);
    }
  _Float16* C = (_Float16*)g.C + mb;
  _Float16* C2 = g.C2 ? (_Float16*)g.C2 + mb : nullptr;
  const float cunscale = ldexpf(1.0f, -ec);
  float vmax = 0.0f, vmax2 = 0.0f;
  // Transposed FP32 copy of an off-diagonal tile (rows bm.., columns bn..) into columns bm..,
  // rows bn..: the lanes put their values (scaled like C) into LDS as float2, LDS column R (tile row)
  // holding the 64 tile columns in 32 chunks of 16 bytes, chunk ch at
  // ch ^ 2 ((ch >> 4) & 1) ^ (R & 1) ^ 8 ((R >> 2) & 1) (no bank conflicts for the puts, 32
  // lanes at a time, nor for the 16-byte reads of eight consecutive rows per thread).
  auto mchunk = [](int R, int ch) __attribute__((always_inline)) {
    return ch ^ (2 * ((ch >> 4) & 1)) ^ (R & 1) ^ (8 * ((R >> 2) & 1));
  };
  auto mirror_put = [&](int i, int j, const float (&re)[4], const float (&im)[4]) __attribute__((always_inline)) {
    float2* Lf = (float2*)Ls;
    const int cc = wn + 16 * j + l16;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int R = wm + 16 * i + 4 * l4 + e;
      Lf[R * 64 + 2 * mchunk(R, cc >> 1) + (cc & 1)] = make_float2(re[e], im[e]);
    }
  };
  // each lane holds rows row0 .. row0+3 of one column: 8 contiguous bytes per plane, the four
  // planes 16 bytes apart in one 64-byte group block (lane pairs l, l ^ 16 share a group)
  const bool odd = (l4 & 1) != 0;
  auto store_split = [&](_Float16* dst, const float (&re)[4], const float (&im)[4], float& mx) __attribute__((always_inline)) {
    store_split4(dst, re, im, odd);
#pragma unroll
    for (int e = 0; e < 4; e++) mx = nmax(mx, nmax(fabsf(re[e]), fabsf(im[e])));
  };

  // n is a multiple of 32: every lane of a live subtile holds valid rows and columns
  auto epi = [&](auto live) __attribute__((always_inline)) {
    constexpr int JL = decltype(live)::value;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < JL; j++) {
        const int row0 = bm + wm + 16 * i + 4 * l4;
        const int col = bn + wn + 16 * j + l16;
        const int64_t o = split_index(n, row0, col, 0);
        float vr[4], vi[4], wr[4], wi[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const float pr = crp[i][j][e] - crn[i][j][e], pi = ci[i][j][e];
          vr[e] = fa * pr; vi[e] = fa * pi;
          wr[e] = fa2 * pr; wi[e] = fa2 * pi;
          if (row0 + e == col) { vr[e] += fi; wr[e] += fi2; }
        }
#pragma unroll
        for (int q = 0; q < 2; q++)
          if (g.M[q] && !(q == 0 && mx0)) {
            float mr[4], mi[4];
            if (g.mraw[q]) {
              const float4* Mf = (const float4*)((const float2*)g.M[q] + (int64_t)b * per + (int64_t)col * n + row0);
              const float4 u0 = Mf[0], u1 = Mf[1];
              mr[0] = u0.x; mi[0] = u0.y; mr[1] = u0.z; mi[1] = u0.w;
              mr[2] = u1.x; mi[2] = u1.y; mr[3] = u1.z; mi[3] = u1.w;
            } else {
              const _Float16* Mq = (const _Float16*)g.M[q] + mb + o;
              const h4v m0 = *(const h4v*)Mq, m1 = *(const h4v*)(Mq + 8);
              const h4v m2 = *(const h4v*)(Mq + 16), m3 = *(const h4v*)(Mq + 24);
#pragma unroll
              for (int e = 0; e < 4; e++) {
                mr[e] = (float)m0[e] + (float)m1[e];
                mi[e] = (float)m2[e] + (float)m3[e];
              }
            }
#pragma unroll
            for (int e = 0; e < 4; e++) {
              vr[e] += fm[q] * mr[e]; vi[e] += fm[q] * mi[e];
              wr[e] += fm2[q] * mr[e]; wi[e] += fm2[q] * mi[e];
            }
          }
        // a matrix's last product hands over its FP32 copy only: no later product reads C
        if (!last) store_split(C + o, vr, vi, vmax);
        if (C2) store_split(C2 + o, wr, wi, vmax2);
        if (Cf) {
          float4* qf = (float4*)(Cf + (int64_t)col * n + row0);
          qf[0] = make_float4(vr[0] * cunscale, vi[0] * cunscale, vr[1] * cunscale, vi[1] * cunscale);
          qf[1] = make_float4(vr[2] * cunscale, vi[2] * cunscale, vr[3] * cunscale, vi[3] * cunscale);
        }
        if (mirror) mirror_put(i, j, vr, vi);
      }
  };
  if (mirror) __syncthreads();  // every wave is done with the k loop's tiles
  if (jlive == JN)
    epi(std::integral_constant<int, JN>{});
  else if (jlive == 2)
    epi(std::integral_constant<int, 2>{});
  // true maxima (the scale factors are powers of two)
  wave_max_to(vmax * ldexpf(1.0f, -ec), g.cmax + b);
  if (C2) wave_max_to(vmax2 * ldexpf(1.0f, -ec2), g.cmax2 + b);
  if (!mirror) return;

  // the transposed tile from LDS: thread item (column R, rows 8 q .. 8 q + 7) reads 64 bytes
  __syncthreads();
  const float2* Lf = (const float2*)Ls;
#pragma unroll
  for (int p = 0; p < 2; p++) {
    const int it = tid + 256 * p, R = it >> 3, q = it & 7;
    if (bm + R >= n || bn + 8 * q >= n) continue;
    f4v* df = (f4v*)(Cf + (int64_t)(bm + R) * n + bn + 8 * q);
#pragma unroll
    for (int k = 0; k < 4; k++) df[k] = *(const f4v*)&Lf[R * 64 + 2 * mchunk(R, 4 * q + k)] * cunscale;
  }
}

void launch_sgemm(hipStream_t s, const SGemm& gin) {
  SGemm g = gin;
  // the staged X tiles give beta_0 M_0 only where M_0 is X and no second output needs it
  g.m0_is_x = (g.M[0] == g.X && g.me[0] == g.xe && g.C2 == nullptr && g.alpha != 0.0f) ? 1 : 0;
  const int nmat = g.order ? g.nord : g.nb;
  if (nmat <= 0) return;
  // 32 x 32 per wave: the wider 32 x 64 wave tile (JN = 4) needs two fewer waves per SIMD, which
  // costs more than its 25% saving in LDS reads per MFMA gains (measured at n = 192, 224, 256)
  constexpr int JN = 2;
  const int tm = (g.n + 63) / 64, tn = (g.n + 32 * JN - 1) / (32 * JN);
  const int ntile = g.sym ? tm * (tm + 1) / 2 : tm * tn;  // symmetric mode needs square tiles (JN = 2)
  hipLaunchKernelGGL(k_sgemm<JN>, dim3(ntile * nmat), dim3(256), 0, s, g);
}

// Mixed-form assembly for the split path: A = i pre DynMat (gvectors.f90:1123-1173) in FP32,
// with the weak-beam sum W = sum_w U(g - w) U(w - h) / s_w as a split-FP16 MFMA product. One
// 64 x 64 tile per 256-thread workgroup, 32 x 32 per wave. The direction's weak beams (hkl and
// 1/s_w) are staged in LDS once; per chunk of 32 weak beams thread t gathers, one chunk ahead
// in registers, U(g_r - w) for its tile row r = t & 63 and U(w - h_c) for its tile column
// c = t & 63 at the weak beams 8 (t >> 6) .. +7 of the chunk, and writes them split into FP16
// planes [plane][row or column][32 weak beams] (16-byte chunks swizzled as in k_sgemm). Scales:
// U by 2^ea from the LUT's max entry, U / s_w by 2^eb from that max over the direction's
// smallest |s_w|. The diagonal sums sum_w |U(g_r - w)|^2 / s_w (gvectors.f90:1165) are taken
// in FP64 during the gathers and added in a fixed order. With sym set (a complex-symmetric
// DynMat) only the tiles on and above the diagonal are assembled; an off-diagonal tile then
// also gives the column partials of its transpose (its row sums).
__device__ __forceinline__ int as_swz(int row, int chunk) { return chunk ^ ((row >> 1) & 3); }

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
void k_assemble_split(Bucket B, BeamView V, DevRefl R, DevLut L, const double* __restrict__ kn,
                      double mlambda, float dynupz, double dz, float lutmax, float2* __restrict__ A32,
                      double* __restrict__ colpart, float* __restrict__ amax, int sym) {
  const int npad = B.npad, tiles = (npad + 63) >> 6;
  const int ntile = sym ? tiles * (tiles + 1) / 2 : tiles * tiles;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int b = lb / ntile, t = lb % ntile;
  int ti, tj;
  if (sym) {  // t = tj (tj + 1) / 2 + ti, ti <= tj
    tj = 0;
    while ((tj + 1) * (tj + 2) / 2 <= t) tj++;
    ti = t - tj * (tj + 1) / 2;
  } else {
    ti = t % tiles;
    tj = t / tiles;
  }
  const int bm = ti * 64, bn = tj * 64;
  const int dir = B.dirs[b];
  const int n = V.nns[dir], nw = V.nnw[dir];
  const int nwpad = (nw + 31) & ~31;
  const int32_t* sidx = V.sidx + V.soff[dir];
  const float* ssg = V.ssg + V.soff[dir];
  const int32_t* widx = V.widx + V.woff[dir];
  const float* wsg = V.wsg + V.woff[dir];

  // LUT offsets are linear in (h, k, l): idx(g - h) = lin(g) - lin(h) - lin(lo), lin(h, k, l) =
  // h + d0 (k + d1 l), the same integer as the column-major offset lut_get computes
  extern __shared__ int4 wk[];  // [nwpad] weak beams: lin(w), bits of fl32(1/s_w), of 1/s_w (FP64)
  __shared__ int rl[64], cl[64];  // lin of the tile's rows (minus lin(lo)) and columns
  __shared__ __attribute__((aligned(16))) _Float16 As[4 * 64 * 32];
  __shared__ __attribute__((aligned(16))) _Float16 Bs[4 * 64 * 32];
  __shared__ float red[4];
  __shared__ double dps[4][64], cs[8][64], rs[2][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int l16 = lane & 15, l4 = lane >> 4;
  const bool wave_live = (bm + wm < n) && (bn + wn < n);
  const bool diag_tile = (ti == tj);

  const float2* lut = L.lut;
  const int d0 = L.dim[0], d1 = L.dim[1], ncell = L.ncell;
  const int linlo = L.lo[0] + d0 * (L.lo[1] + d1 * L.lo[2]);
  int32_t* lerr = L.err;
  auto lin = [&](const int32_t* h) __attribute__((always_inline)) { return h[0] + d0 * (h[1] + d1 * h[2]); };
  if (tid < 64) {
    const int a = bm + tid;
    rl[tid] = ((a < n) ? lin(R.hkl + 3 * sidx[a]) : 0) - linlo;
  } else if (tid < 128) {
    const int c = bn + tid - 64;
    cl[tid - 64] = (c < n) ? lin(R.hkl + 3 * sidx[c]) : 0;
  }
  float mi = 0.0f;
  for (int w = tid; w < nwpad; w += 256) {
    if (w < nw) {
      const int id = widx[w];
      const float sw = wsg[w];
      const double rv = 1.0 / (double)sw;
      const float iv = (float)rv;  // cmplx(1.D0/weaksg, 0) of gvectors.f90:1156
      wk[w] = make_int4(lin(R.hkl + 3 * id), __float_as_int(iv), __double2loint(rv), __double2hiint(rv));
      mi = nmax(mi, fabsf(iv));
    } else {
      wk[w] = make_int4(0, 0, 0, 0);
    }
  }
  mi = nmax(mi, dpp_f32<0xB1>(mi));
  mi = nmax(mi, dpp_f32<0x4E>(mi));
  mi = nmax(mi, dpp_f32<0x124>(mi));
  mi = nmax(mi, dpp_f32<0x128>(mi));
  mi = nmax(mi, xchg16(mi));
  mi = nmax(mi, xchg32(mi));
  if (lane == 0) red[wave] = mi;
  __syncthreads();
  mi = nmax(nmax(red[0], red[1]), nmax(red[2], red[3]));
  const int ea = f16_scale_exp(lutmax), eb = f16_scale_exp(lutmax * mi);
  const float sa = ldexpf(1.0f, ea), sb = ldexpf(1.0f, eb);
  // lut_get on a linear offset: outside the allocation -> 0 and error bit 8
  auto U = [&](int idx) __attribute__((always_inline)) -> float2 {
    if ((unsigned)idx >= (unsigned)ncell) {
      atomicOr(lerr, 8);
      return make_float2(0.0f, 0.0f);
    }
    return lut[idx];
  };

  // gathers: row (or column) p = t & 63, weak beams w0 + pk .. +7
  const int pr = tid & 63, pk = 8 * (tid >> 6);
  const bool prow_ok = bm + pr < n, pcol_ok = bn + pr < n;
  const int rlp = rl[pr], clp = cl[pr] + linlo;
  float2 ga[8], gb[8];
  double dpart = 0.0;
  auto gather = [&](int w0) __attribute__((always_inline)) {
    #pragma unroll
    for (int j = 0; j < 8; j++) {
      const int w = w0 + pk + j;
      const int q = wk[w].x;
      ga[j] = (prow_ok && w < nw) ? U(rlp - q) : make_float2(0.f, 0.f);        // U(g_r - w)
      gb[j] = (pcol_ok && w < nw) ? U(q - clp) : make_float2(0.f, 0.f);        // U(w - h_c)
    }
  };
  const int alds = pr * 32 + 8 * as_swz(pr, tid >> 6);
  auto stage = [&](int w0) __attribute__((always_inline)) {
    h8v oa[4], ob[4];
    #pragma unroll
    for (int j = 0; j < 8; j++) {
      const float iv = __int_as_float(wk[w0 + pk + j].y);
      _Float16 h, l;
      split2(ga[j].x * sa, h, l); oa[0][j] = h; oa[1][j] = l;
      split2(ga[j].y * sa, h, l); oa[2][j] = h; oa[3][j] = l;
      split2(gb[j].x * iv * sb, h, l); ob[0][j] = h; ob[1][j] = l;  // U(w - h) / s_w
      split2(gb[j].y * iv * sb, h, l); ob[2][j] = h; ob[3][j] = l;
      if (diag_tile) {
        // |U(g_r - w)|^2 / s_w in FP64 (entries past nw are 0)
        const int w = w0 + pk + j;
        const double x = ga[j].x, y = ga[j].y;
        dpart += (x * x + y * y) * __hiloint2double(wk[w].w, wk[w].z);
      }
    }
    #pragma unroll
    for (int p = 0; p < 4; p++) {
      *(h8v*)&As[p * 2048 + alds] = oa[p];
      *(h8v*)&Bs[p * 2048 + alds] = ob[p];
    }
  };

  f4v crp[2][2], crn[2][2], ci[2][2];
  #pragma unroll
  for (int i = 0; i < 2; i++)
    #pragma unroll
    for (int j = 0; j < 2; j++) {
      crp[i][j] = (f4v){0, 0, 0, 0};
      crn[i][j] = (f4v){0, 0, 0, 0};
      ci[i][j] = (f4v){0, 0, 0, 0};
    }
  int aoff[2], boff[2];
  #pragma unroll
  for (int i = 0; i < 2; i++) {
    const int ra = wm + 16 * i + l16, cb = wn + 16 * i + l16;
    aoff[i] = ra * 32 + 8 * as_swz(ra, l4);
    boff[i] = cb * 32 + 8 * as_swz(cb, l4);
  }
  gather(0);
  auto kloop = [&](auto live) __attribute__((always_inline)) {
    for (int w0 = 0; w0 < nwpad; w0 += 32) {
      __syncthreads();
      stage(w0);
      __syncthreads();
      if (w0 + 32 < nwpad) gather(w0 + 32);
      if constexpr (!decltype(live)::value) continue;
      h8v a[2][2][2], bq[2][2];  // [block][re/im][high/low]; B one column block at a time
      #pragma unroll
      for (int i = 0; i < 2; i++)
        #pragma unroll
        for (int c = 0; c < 2; c++)
          #pragma unroll
          for (int h = 0; h < 2; h++) a[i][c][h] = *(const h8v*)&As[(2 * c + h) * 2048 + aoff[i]];
      #pragma unroll
      for (int j = 0; j < 2; j++) {
        #pragma unroll
        for (int c = 0; c < 2; c++)
          #pragma unroll
          for (int h = 0; h < 2; h++) bq[c][h] = *(const h8v*)&Bs[(2 * c + h) * 2048 + boff[j]];
        #pragma unroll
        for (int i = 0; i < 2; i++) {
#define MM(acc, ca, ha, cb, hb) acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i][ca][ha], bq[cb][hb], acc, 0, 0, 0)
          MM(crp[i][j], 0, 1, 0, 0); MM(crp[i][j], 0, 0, 0, 1); MM(crp[i][j], 0, 0, 0, 0);
          MM(crn[i][j], 1, 1, 1, 0); MM(crn[i][j], 1, 0, 1, 1); MM(crn[i][j], 1, 0, 1, 0);
          MM(ci[i][j], 0, 1, 1, 0); MM(ci[i][j], 0, 0, 1, 1); MM(ci[i][j], 0, 0, 1, 0);
          MM(ci[i][j], 1, 1, 0, 0); MM(ci[i][j], 1, 0, 0, 1); MM(ci[i][j], 1, 0, 0, 0);
#undef MM
        }
      }
    }
  };
  if (wave_live)
    kloop(std::true_type{});
  else
    kloop(std::false_type{});

  // diagonal sums of the tile rows: four partial sums per row, added in a fixed order
  if (diag_tile) {
    dps[tid >> 6][pr] = dpart;
    __syncthreads();
    if (tid < 64) dps[0][tid] = ((dps[0][tid] + dps[1][tid]) + dps[2][tid]) + dps[3][tid];
    __syncthreads();
  }
  // write the tile: lane holds rows bm + wm + 16 i + 4 l4 + e of column bn + wn + 16 j + l16
  const double half = 0.5 * mlambda;
  const double pre = 2.0 * 3.141592653589793238 * dz / (double)(float)(2.0 * kn[dir]);
  const double wsc = ldexp(1.0, -(ea + eb));
  float2* Ab = A32 + (int64_t)b * npad * npad;
  float fmx = 0.0f;
  double csum[2] = {0.0, 0.0};
  const bool rows = sym && ti != tj;  // row sums wanted: column partials of the transpose
  #pragma unroll
  for (int i = 0; i < 2; i++) {
    double rsum[4] = {0.0, 0.0, 0.0, 0.0};
    #pragma unroll
    for (int j = 0; j < 2; j++) {
      const int cc = wn + 16 * j + l16, gj = bn + cc;
      const int row0 = bm + wm + 16 * i + 4 * l4;
      float2 f[4];
      #pragma unroll
      for (int e = 0; e < 4; e++) {
        const int rr = wm + 16 * i + 4 * l4 + e;
        const int gi = bm + rr;
        double vr = 0.0, vi = 0.0;
        if (gi < n && gj < n) {
          if (gi != gj) {
            const float2 u = U(rl[rr] - cl[cc]);
            const float wr = crp[i][j][e] - crn[i][j][e], wi = ci[i][j][e];
            vr = (double)u.x - half * ((double)wr * wsc);
            vi = (double)u.y - half * ((double)wi * wsc);
          } else {
            const double sd = dps[0][rr] * mlambda / 2.0;
            vr = 2.0 * (double)ssg[gi] / mlambda - sd;
            vi = (double)dynupz;
          }
        }
        f[e] = make_float2((float)(-pre * vi), (float)(pre * vr));
      }
      // npad is a multiple of 32: the four rows are all inside the matrix or all outside
      if (row0 < npad && gj < npad) {
        float4* q = (float4*)(Ab + (int64_t)gj * npad + row0);
        q[0] = make_float4(f[0].x, f[0].y, f[1].x, f[1].y);
        q[1] = make_float4(f[2].x, f[2].y, f[3].x, f[3].y);
        #pragma unroll
        for (int e = 0; e < 4; e++) {
          const double m = (double)hypotf(f[e].x, f[e].y);
          csum[j] += m;
          rsum[e] += m;
          fmx = nmax(fmx, nmax(fabsf(f[e].x), fabsf(f[e].y)));
        }
      }
    }
    if (rows) {
      // the row sums over the wave's 32 columns: the 16 lanes of a row, in a fixed order
      #pragma unroll
      for (int e = 0; e < 4; e++) {
        double v = rsum[e];
        v += dpp_f64<0xB1>(v);
        v += dpp_f64<0x4E>(v);
        v += dpp_f64<0x124>(v);
        v += dpp_f64<0x128>(v);
        if (l16 == 0) rs[wn >> 5][wm + 16 * i + 4 * l4 + e] = v;
      }
    }
  }
  // column 1-norm partials of the tile: eight per column (two waves x four lane groups),
  // added in a fixed order
  #pragma unroll
  for (int j = 0; j < 2; j++) cs[(wm >> 5) * 4 + l4][wn + 16 * j + l16] = csum[j];
  __syncthreads();
  if (tid < 64 && bn + tid < npad) {
    double d = 0.0;
    for (int q = 0; q < 8; q++) d += cs[q][tid];
    colpart[((int64_t)b * tiles + ti) * npad + bn + tid] = d;
  }
  // the row sums of the two waves of a row block: column partials of tile-row tj at columns bm..
  if (rows && tid < 64 && bm + tid < npad) colpart[((int64_t)b * tiles + tj) * npad + bm + tid] = rs[0][tid] + rs[1][tid];
  wave_max_to(fmx, amax + b);
}

void launch_assemble_split(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R,
                           const DevLut& L, const double* kn, double mlambda, float dynupz, double dz,
                           float lutmax, int32_t nwmax, float2* A32, double* colpart, float* amax, int32_t sym) {
  const int tiles = (B.npad + 63) / 64;
  const int ntile = sym ? tiles * (tiles + 1) / 2 : tiles * tiles;
  const size_t lds = sizeof(int4) * (size_t)std::max(32, (nwmax + 31) & ~31);
  hipLaunchKernelGGL(k_assemble_split, dim3(ntile * B.nb), dim3(256), lds, s, B, V, R, L, kn, mlambda,
                     dynupz, dz, lutmax, A32, colpart, amax, sym);
}

// 1-norm (max column sum of |a|: the column partials of the assembly added in a fixed order),
// squaring count s with ||A||_1 / 2^s <= theta, and the split exponent and max |entry| of
// B = A 2^-s (the max of A times 2^-s, exactly); one workgroup per matrix. B itself is not
// stored: the Taylor products split A while staging it.
__global__ __launch_bounds__(256) void k_norm_split(int npad, const double* __restrict__ colpart,
                                                    const float* __restrict__ amax, double theta, int smax,
                                                    int* __restrict__ sq, int* __restrict__ err,
                                                    int* __restrict__ bexp, float* __restrict__ bmax) {
  const int b = blockIdx.x;
  const int tiles = (npad + 63) / 64;  // row partials per column (64-row assembly tiles)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ double wnorm[4];
  double cnorm = 0.0;
  for (int j = tid; j < npad; j += 256) {
    double sum = 0.0;
    for (int t = 0; t < tiles; t++) sum += colpart[((int64_t)b * tiles + t) * npad + j];
    cnorm = fmax(cnorm, sum);
    if (!(sum == sum)) cnorm = sum;  // keep a NaN
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double v = __shfl_xor(cnorm, o);
    cnorm = (cnorm != cnorm) ? cnorm : ((v != v) ? v : fmax(cnorm, v));  // a NaN wins
  }
  if (lane == 0) wnorm[wave] = cnorm;
  __syncthreads();
  if (tid == 0) {
    double nrm = fmax(fmax(wnorm[0], wnorm[1]), fmax(wnorm[2], wnorm[3]));
    for (int w = 0; w < 4; w++)
      if (!(wnorm[w] == wnorm[w])) nrm = wnorm[w];
    const float m = amax[b];
    int sc = 0;
    if (!isfinite(nrm) || !(m == m)) {
      atomicOr(err, 2);
    } else {
      while (nrm > theta && sc < smax) { nrm *= 0.5; sc++; }
      if (nrm > theta) atomicOr(err, 4);
    }
    const float mb = ldexpf(m, -sc);
    sq[b] = sc;
    bexp[b] = f16_scale_exp(mb);
    bmax[b] = mb;
  }
}

void launch_norm_split(hipStream_t s, int32_t npad, int32_t nb, const double* colpart, const float* amax,
                       double theta, int32_t smax, int32_t* sq, int32_t* err, int32_t* bexp, float* bmax) {
  hipLaunchKernelGGL(k_norm_split, dim3(nb), dim3(256), 0, s, npad, colpart, amax, theta, smax, sq, err, bexp,
                     bmax);
}

}  // namespace ctem
