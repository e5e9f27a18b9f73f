// Internal interfaces between the C-ABI layer (ctem_api.hip) and the kernels
// (ctem_kernels.hip, ctem_mixed.hip). Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ctem {

// Workgroups are dispatched round-robin over the 8 XCDs of the MI355X (XCD = blockIdx % 8),
// each with its own L2. This maps the hardware block index to a logical index such that
// consecutive logical indices (the tiles of one matrix) run on the same XCD and share its L2.
__device__ __forceinline__ int xcd_logical_block(int bid, int nblk) {
  constexpr int X = 8;
  const int per = nblk / X, rem = nblk % X;
  const int xcd = bid % X, idx = bid / X;
  return (xcd < rem) ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// NaN-propagating max of non-negative values
__device__ __forceinline__ float nmax(float a, float b) { return (a != a) ? a : ((b != b) ? b : fmaxf(a, b)); }

// Cross-lane moves without LDS (the whole wave active). DPP within 16-lane rows: CTRL 0xB1 =
// quad_perm [1,0,3,2] (lane ^ 1), 0x4E = [2,3,0,1] (lane ^ 2), 0x124 / 0x128 = row_ror 4 / 8,
// 0x141 = row_half_mirror.
template <int CTRL>
__device__ __forceinline__ unsigned dpp_u32(unsigned x) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float x) { return __uint_as_float(dpp_u32<CTRL>(__float_as_uint(x))); }
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  return __hiloint2double((int)dpp_u32<CTRL>((unsigned)__double2hiint(x)), (int)dpp_u32<CTRL>((unsigned)__double2loint(x)));
}
// The value of lane l ^ 16 (l ^ 32): v_permlane16_swap (v_permlane32_swap) trades the odd 16-lane
// (32-lane) rows of its first operand with the even rows of its second and returns both; with
// the same value in both, the odd rows find their partner's value in the first result and the
// even rows in the second.
__device__ __forceinline__ unsigned xchg16(unsigned v) {
  const auto t = __builtin_amdgcn_permlane16_swap(v, v, false, false);
  return (threadIdx.x & 16) ? t[0] : t[1];
}
__device__ __forceinline__ unsigned xchg32(unsigned v) {
  const auto t = __builtin_amdgcn_permlane32_swap(v, v, false, false);
  return (threadIdx.x & 32) ? t[0] : t[1];
}
__device__ __forceinline__ float xchg16(float v) { return __uint_as_float(xchg16(__float_as_uint(v))); }
__device__ __forceinline__ float xchg32(float v) { return __uint_as_float(xchg32(__float_as_uint(v))); }
__device__ __forceinline__ double xchg16(double v) {
  return __hiloint2double((int)xchg16((unsigned)__double2hiint(v)), (int)xchg16((unsigned)__double2loint(v)));
}
__device__ __forceinline__ double xchg32(double v) {
  return __hiloint2double((int)xchg32((unsigned)__double2hiint(v)), (int)xchg32((unsigned)__double2loint(v)));
}
// sum over the 64 lanes, the same in every lane and in a fixed order: the quads (lane ^ 1,
// lane ^ 2), the four quads of a row (row_ror 4, 8), then the rows (lane ^ 16, lane ^ 32)
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x124>(v);
  v += dpp_f64<0x128>(v);
  v += xchg16(v);
  v += xchg32(v);
  return v;
}

// per-matrix max |entry|: wave reduction, one atomic per wave (non-negative floats order as
// their bit patterns, a NaN above every finite value); the whole wave must be active
__device__ __forceinline__ void wave_max_to(float v, float* slot) {
  v = nmax(v, dpp_f32<0xB1>(v));
  v = nmax(v, dpp_f32<0x4E>(v));
  v = nmax(v, dpp_f32<0x124>(v));
  v = nmax(v, dpp_f32<0x128>(v));
  v = nmax(v, xchg16(v));
  v = nmax(v, xchg32(v));
  if ((threadIdx.x & 63) == 0 && v > 0.0f) atomicMax((unsigned int*)slot, __float_as_uint(v));
  if ((threadIdx.x & 63) == 0 && v != v) atomicMax((unsigned int*)slot, 0x7fc00000u);
}

// scaling exponent of the split-FP16 operands: max |x| 2^e lies in [2^13, 2^14); e is held to
// [-60, 60] so that products of two scale factors stay normal FP32 numbers (a matrix with
// max |x| < 2^-46 is then scaled below 2^14, still far above the FP16 subnormal range)
__device__ __forceinline__ int f16_scale_exp(float m) {
  if (!(m > 0.0f) || !(m <= 3.4e38f)) return 0;
  int q;
  frexpf(m, &q);
  return min(max(14 - q, -60), 60);
}

// Reflection table on the device, structure of arrays in master-list order.
struct DevRefl {
  int32_t nref;
  const float* g;        // [3*nref] float(hkl)
  const int32_t* hkl;    // [3*nref]
  const float* cut1;     // [nref] cutoff * fl32(|Ucg| mLambda)
  const float* cut2;     // [nref] weakcutoff * fl32(|Ucg| mLambda)
  const uint8_t* dbd;    // [nref] double-diffraction flags
  float sgcutoff;
};

struct DevLut {
  const float2* lut;     // complex LUT values (float32, exactly the reference's values)
  int32_t lo[3], dim[3]; // Fortran bounds of the LUT box
  int32_t ncell;
  int32_t* err;          // bit 8: a lookup fell outside the LUT allocation
  const double2* slut;   // [nsite * sncell] per-site structure-matrix table S_p(g-h)
  int32_t slo[3], sdim[3];  // its box, which holds every difference of two list entries
  int64_t sncell;
  int32_t nsite;
};

// U(g-h) as Compute_DynMat reads it: LUT(h,k,l) at the column-major offset of the Fortran
// bounds, without a per-axis bounds check (gvectors.f90:1133-1160). When a beam set spans
// more than the LUT box along an axis (hexagonal axes, where list entries reach |h+k| up to
// twice the box), an offset that still lies inside the array reads the aliased entry, as the
// reference does; an offset outside the array (which the reference reads from neighbouring
// heap memory) returns 0 and raises error bit 8.
__device__ __forceinline__ float2 lut_get(const DevLut& L, int h, int k, int l) {
  const int idx = (h - L.lo[0]) + L.dim[0] * ((k - L.lo[1]) + L.dim[1] * (l - L.lo[2]));
  if ((unsigned)idx >= (unsigned)L.ncell) {
    atomicOr(L.err, 8);
    return make_float2(0.0f, 0.0f);
  }
  return L.lut[idx];
}

struct SelectOut {       // per-direction beam sets (CSR over directions)
  int32_t* nns;          // [numk]
  int32_t* nnw;          // [numk]
  const int64_t* soff;   // [numk] offsets into sidx/ssg (fill pass)
  const int64_t* woff;
  int32_t* sidx; float* ssg;
  int32_t* widx; float* wsg;
  int32_t* err;          // set to 1 when an angle leaves the reproduced cosf branch
};

// Beam selection. fill=0 counts, fill=1 writes the lists. fn (float[3] on the device, or null):
// the foil normal of the excitation errors, null for FN = k per direction.
void launch_select(hipStream_t s, const double* k, int32_t numk, const DevRefl& R,
                   const double* rmt9, const float* fn, const SelectOut& out, int fill);

// Per-bucket descriptors for the dense stages. All matrices of a bucket are n_pad x n_pad,
// column-major, with batch stride n_pad*n_pad; dirs[b] names the plan direction of slot b.
struct Bucket {
  int32_t npad;
  int32_t nb;            // directions in this chunk
  const int32_t* dirs;   // [nb] plan direction indices (device)
};

struct BeamView {        // read-only view of the selection lists
  const int32_t* nns; const int32_t* nnw;
  const int64_t* soff; const int64_t* woff;
  const int32_t* sidx; const float* ssg;
  const int32_t* widx; const float* wsg;
};

// A = (2 pi i dz / fl32(2 kn)) * DynMat, Bethe-corrected (gvectors.f90:1123-1173), FP64.
void launch_assemble(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R,
                     const DevLut& L, const double* kn, double mlambda, float dynupz, double dz,
                     double2* A);
// A = i pi mLambda dz DynMat, FP64 (the scattering matrix of the channeling-pattern program)
void launch_assemble_scatmat(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R,
                             const DevLut& L, double mlambda, float dynupz, double dz, double2* A);
// the same scaled matrix rounded to FP32 (mixed form)
void launch_assemble32(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R,
                       const DevLut& L, const double* kn, double mlambda, float dynupz, double dz,
                       float2* A32);
// the same for the split path, with the weak-beam sum on split-FP16 MFMA (64 x 64 tiles);
// also the column sums of |a| per 64-row tile (colpart[b][tile row][column], FP64) and the
// max |re|, |im| per matrix into amax[b] (zeroed by the caller). lutmax = max |re|, |im| over
// the LUT; nwmax = the largest weak-beam count of the bucket. (An exponent for writing A in
// split form straight from here would have to come from a bound on its entries; on corundum
// such bounds run many bits above the true maximum and cost the low planes their precision.)
void launch_assemble_split(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R,
                           const DevLut& L, const double* kn, double mlambda, float dynupz, double dz,
                           float lutmax, int32_t nwmax, float2* A32, double* colpart, float* amax,
                           int32_t sym);
// same matrix, unscaled (DynMat itself), for parity checks
void launch_dynmat(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R,
                   const DevLut& L, double mlambda, float dynupz, double2* D);

// ||A||_1 per matrix, squaring count s with ||A||_1 / 2^s <= theta, in-place scaling by 2^-s.
// sq[b] receives s; err is set on a non-finite norm or s > smax. amod (may be null) receives
// max |a_ij| of each unscaled matrix.
void launch_norm_scale(hipStream_t s, int32_t npad, int32_t nb, double2* A, double theta,
                       int32_t smax, int32_t* sq, int32_t* err, double* amod);

// Batched complex GEMM C = alpha*X*Y + betaI*I + sum_j beta_j*M_j on FP64 MFMA.
struct ZGemm {
  int32_t m, n, k;       // C is m x n, X is m x k, Y is k x n (column-major)
  int32_t ldx, ldy, ldc, ldm;
  const double2* X; int64_t sx;
  const double2* Y; int64_t sy;
  double2* C; int64_t sc;
  double alpha;
  double betaI;
  const double2* M[4]; double beta[4]; int64_t sm;
  const int32_t* active; int32_t step;   // skip matrix b when step >= active[b]
  int32_t nb;
};
void launch_zgemm(hipStream_t s, const ZGemm& g);

// elementwise C = betaI*I + sum_j beta_j*M_j (all npad x npad)
void launch_combo(hipStream_t s, int32_t npad, int32_t nb, double2* C, double betaI,
                  const double2* const M[4], const double beta[4]);
// copy matrices whose squaring count is odd from src to dst (so the result sits in dst)
void launch_parity_copy(hipStream_t s, int32_t npad, int32_t nb, const int32_t* sq,
                        const double2* src, double2* dst);

// depth recursion on the GPU's last-level caches: Psi[b][:,k] = E_b^(k+1) e_1, FP64
// accumulation; E is double2 (fp32 = 0) or float2 (fp32 = 1) holding E, or E - I when
// add_identity is set. grid = number of persistent workgroups (normally the CU count).
void launch_depth(hipStream_t s, int32_t npad, int32_t nb, int32_t nz, const void* E, int fp32,
                  int add_identity, double2* Psi, int64_t spsi, int32_t grid);
// the depth recursion from F = E - I and F2 = E^2 - I (FP32, column-major): psi_1 = E e_1,
// psi_2 = E psi_1, then psi_{k+2} = E^2 psi_k for two depths per pass over F2
void launch_depth_quad(hipStream_t s, int32_t npad, int32_t nb, int32_t nz, const float2* F, const float2* F2,
                       const float2* F4, double2* Psi, int64_t spsi, int32_t grid);

// Lgh = (dz/thick) Psi diag(lam) Psi^H and the per-site contraction with S_p(g-h):
// partial sums per 32x32 tile, reduced in fixed order into dots[dir*nsite + p]. f32: the Lgh
// products in FP32 (mixed form)
void launch_lgh_dots(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R,
                     const DevLut& L, const double2* Psi, int64_t spsi, int32_t nz,
                     const double* lam, double dzt, double* partial, double* dots,
                     int32_t keep_dir, double2* keep_lgh, int32_t f32);

// scatter fl32 svals to the Laue-equivalent pixels of sr[nsite][2npy+1][2npx+1]; owner
// [(2npy+1)(2npx+1)] int32 must hold -1 on entry (the last direction in list order wins a pixel)
void launch_scatter(hipStream_t s, int32_t numk, const int32_t* ij, const double* dots,
                    int32_t nsite, float nat, int32_t npx, int32_t npy, int32_t isym,
                    int32_t* owner, float* sr);

// ---- mixed precision (ctem_mixed.hip) ----
// Complex GEMM on split 16-bit MFMA (BF16 or FP16 planes) with FP32 accumulation for FP32
// complex matrices
// (float2, column-major); same argument meaning as ZGemm.
struct CGemm {
  int32_t m, n, k;
  int32_t ldx, ldy, ldc, ldm;
  const float2* X; int64_t sx;
  const float2* Y; int64_t sy;
  float2* C; int64_t sc;
  float alpha;
  float betaI;
  const float2* M[4]; float beta[4]; int64_t sm;
  const int32_t* active; int32_t step;
  int32_t nb;
  int32_t terms;         // split scheme: CTEM_SPLIT_BF16X3, CTEM_SPLIT_BF16X6 or CTEM_SPLIT_FP16X3
  int32_t* err;          // reserved for kernel-side error bits
  const float* xmax;     // [nb] max |entry| of X and Y per matrix (split-FP16 scaling)
  const float* ymax;
  float* cmax;           // [nb] receives max |entry| of C per matrix (zeroed by the caller), or null
  // optional second output from the same product:
  // C2 = alpha2*X*Y + betaI2*I + sum_j beta2_j*M_j (same M, strides and leading dimension as C)
  float2* C2;
  float alpha2, betaI2, beta2[4];
  float* cmax2;
  int32_t vec;           // set by launch_cgemm: 16-byte vector access to Y, C, C2 and M is aligned
};
void launch_cgemm(hipStream_t s, const CGemm& g);

// Split-FP16 complex matrices (mixed form with CTEM_SPLIT_FP16X3): four FP16 values per entry
// (real high, real low, imaginary high, imaginary low) and one exponent e per matrix: entry =
// (high + low) * 2^-e. Storage is column-major in groups of 8 rows: column c of matrix b
// starts at b * 4 n^2 + c * 4 n halfs, and holds for each group of rows 8g .. 8g+7 the four
// planes one after the other (64 bytes), so that 8 consecutive rows of one plane are 16
// contiguous bytes and all planes of an entry lie in one 64-byte block. A matrix takes 8 bytes
// per entry, as float2 does. Alongside each buffer the true max |re|, |im| over the matrix is
// kept in FP32 (bounds for the next exponents).
__host__ __device__ __forceinline__ int64_t split_index(int n, int row, int col, int plane) {
  return (int64_t)col * 4 * n + (row >> 3) * 32 + plane * 8 + (row & 7);
}
// C = alpha X Y + betaI I + sum_j beta_j M_j (and optionally C2 with alpha2, betaI2, beta2
// from the same product), n a multiple of 32. The output exponents come from the bound
// |alpha| n max|X| max|Y| + |betaI| + sum |beta_j| max|M_j|.
struct SGemm {
  int32_t n;
  const uint16_t* X; const int32_t* xe; const float* xmax;
  const uint16_t* Y; const int32_t* ye; const float* ymax;
  uint16_t* C; int32_t* ce; float* cmax;
  float alpha, betaI;
  const uint16_t* M[2]; const int32_t* me[2]; const float* mmax[2]; float beta[2];
  uint16_t* C2; int32_t* ce2; float* cmax2;
  float alpha2, betaI2, beta2[2];
  const int32_t* active; int32_t step;  // matrices with step >= active[b] + active_add are skipped
  int32_t active_add;
  int32_t nb;
  // optional: process only the matrices order[0 .. nord-1] (then active is not consulted)
  const int32_t* order; int32_t nord;
  int32_t m0_is_x;       // set by launch_sgemm: M[0] is X (its term is taken from the staged X tiles)
  // optional FP32 copy of C (column-major float2, n x n per matrix) for the matrices with
  // f_sq[b] == f_at: the last product of each matrix hands its result to the depth kernel
  float2* Cf; const int32_t* f_sq; int32_t f_at;
  // optional FP32 copy of the product after it (f_sq[b] + 1 == f_at): E^2 - I for the depth
  // recursion over groups of depths; with Cf2 set, a matrix's product at f_sq still stores C
  float2* Cf2;
  // optional FP32 copy of the product two after it (f_sq[b] + 2 == f_at): E^4 - I (needs Cf2;
  // with Cf4 set, the product at f_sq + 1 still stores C)
  float2* Cf4;
  // symmetric operands and result (X, Y, M and C polynomials in one complex-symmetric matrix):
  // only the 64 x 64 tiles on or above the diagonal of C are computed and stored (C2 alike),
  // and only those of X and Y are read (the rest through the symmetry); Cf is written in full
  int32_t sym;
  // raw operands: X = Y (xraw) and M[q] (mraw[q]) point at the assembled FP32 matrix A (float2,
  // column-major); the operand is B = A 2^-rs[b] (rs the squaring counts), X and Y split on the
  // fly into the staged tiles (scaled by 2^xe as split operands are), M[q] with me[q] = rs
  int32_t xraw; const int32_t* rs; int32_t mraw[2];
};
void launch_sgemm(hipStream_t s, const SGemm& g);
// zero `bytes` bytes (a multiple of 4) of device memory with a kernel (no memset node, so that
// the sequence can be captured into a graph)
void launch_fill_zero(hipStream_t s, void* p, size_t bytes);
// sets bit 32 of *err when some sq[b] > smax (a run launched with a stale squaring schedule)
void launch_sq_check(hipStream_t s, int32_t nb, const int32_t* sq, int32_t smax, int32_t* err);
// The column-sum partials and max |entry| of launch_assemble_split -> ||A||_1 per matrix, the
// squaring count s, and the split exponent (from the exact max) and max |entry| of B = A 2^-s
void launch_norm_split(hipStream_t s, int32_t npad, int32_t nb, const double* colpart, const float* amax,
                       double theta, int32_t smax, int32_t* sq, int32_t* err, int32_t* bexp, float* bmax);
// FP32 form of launch_norm_scale: ||A||_1 per matrix, squaring count, in-place scaling by
// 2^-s, and the max |entry| of the scaled matrix into amax[b] (zeroed by the caller)
void launch_norm_scale32(hipStream_t s, int32_t npad, int32_t nb, float2* A, double theta,
                         int32_t smax, int32_t* sq, int32_t* err, float* amax);
void launch_cparity_copy(hipStream_t s, int32_t npad, int32_t nb, const int32_t* sq,
                         const float2* src, float2* dst);

// ---- electron channeling patterns (ctem_ecp.hip) ----
// L.slut: the channeling structure matrix table Sgh(g-h) (one plane)
// ScatMat: per-depth Re psi_z^T Sgh conj(psi_z), partial sums per 32-row tile:
// partial[(b * npad/32 + tile) * nzp + z], nzp = nz rounded up to 32
void launch_ecp_depth_dots(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R, const DevLut& L,
                           const double2* Psi, int64_t spsi, int32_t nz, double* partial);
// running sums over the depths, sr[t * numk + dir] = fl32(sum / nat) where store[z] = t >= 0
void launch_ecp_scatmat_out(hipStream_t s, const Bucket& B, int32_t nz, const double* partial, const int32_t* store,
                            double nat, int32_t numk, float* sr);
// Blochwv: M[b] = t0 X_b, M[nb + b] = dt X_b, X = [[A1, 0], [eps Sgh^T, -A1^H]] (2 npad square)
void launch_ecp_augment(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R, const DevLut& L,
                        const double2* A1, double eps, double t0, double dt, double2* M);
// y = E[b] e_1, y <- E[nb + b] y per further thickness; sr[it * numk + dir] = fl32(Re psi^H phi / eps
// / thick[it] / nat)
void launch_ecp_bloch_steps(hipStream_t s, const Bucket& B, const BeamView& V, const double2* E, int32_t nt,
                            const float* thick, double eps, double nat, int32_t numk, float* sr);

// MatrixExponential 'Pade' (math.f90:187-315), batched over nb npad x npad matrices A (holding
// A z0): prep scales A to B = A / 2^j in place and sets N = D = I, C = B, qm1[b] = q - 1, jj[b] = j;
// accum adds step icnt (1..4) of the Pade sums for the matrices with icnt <= q (C = B^icnt
// supplied by the caller's products); zsolve overwrites X with D^-1 X (D is destroyed).
void launch_pade_prep(hipStream_t s, int32_t npad, int32_t nb, double2* A, double2* N, double2* D, double2* C,
                      int32_t* qm1, int32_t* jj, int32_t* err);
void launch_pade_accum(hipStream_t s, int32_t npad, int32_t nb, int32_t icnt, const int32_t* qm1, const double2* C,
                       double2* N, double2* D);
void launch_zsolve(hipStream_t s, int32_t npad, int32_t nb, double2* D, double2* X, int32_t* err);

// ---- reverse mode (ctem_adjoint.hip) ----
void launch_adj_m(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R, const DevLut& L,
                  const double* gdots, double c, double2* M);
void launch_adj_hm(hipStream_t s, int npad, int nb, int nz, const double2* Psi, int64_t spsi, double2* Hm);
void launch_adj_depth(hipStream_t s, int npad, int nb, int nz, const double2* E, const double2* Q,
                      const double* lam, double2* Phi);
// X = A^H, Y = sigma Ebar: the pair form of [[A^H, sigma Ebar], [0, A^H]]
void launch_adj_pair(hipStream_t s, int npad, int nb, const double2* A, const double2* Ebar, double* sigma,
                     double2* PX, double2* PY);
// ||[[X, Y], [0, X]]||_1 per matrix, its squaring count and the scaling of X and Y in place
void launch_norm_scale_pair(hipStream_t s, int npad, int nb, double2* X, double2* Y, double theta, int smax,
                            int32_t* sq, int32_t* err);
// D_bar from the derivative block: T2 with leading dimension ldt (npad: the pair's Y part)
void launch_adj_extract(hipStream_t s, const Bucket& B, const BeamView& V, const double* kn, double dz,
                        const double2* T2, int ldt, const double* sigma, double2* Dbar);
// glut: nrep copies of the 2 ncell gradient entries (zeroed by the caller); the copy a workgroup
// adds into follows its index. launch_adj_lut_reduce leaves their sum in copy 0. nrep =
// adj_lut_replicas(ncell): kLutReplicas, fewer when the copies would pass 2 GiB.
constexpr int kLutReplicas = 256;
int adj_lut_replicas(int ncell);
void launch_adj_lut(hipStream_t s, const Bucket& B, const BeamView& V, const DevRefl& R, const DevLut& L,
                    double mlambda, const double2* Dbar, double* glut, int max_nw, int nrep);
void launch_adj_lut_reduce(hipStream_t s, int ncell, int nrep, double* glut);
// mixed form: E = I + F in FP64 from the FP32 F
void launch_adj_e64(hipStream_t s, int npad, int nb, const float2* F, double2* E);
// the pair (A 2^-s, sigma 2^-s Ebar^H) in FP32 (A scaled in place), its squaring count and the
// max |re|, |im| of both parts (xmax, ymax zeroed by the caller)
void launch_adj_pair32(hipStream_t s, int npad, int nb, float2* A, const double2* Ebar, double theta, int smax,
                       int32_t* sq, int32_t* err, double* sigma, float2* Y, float* xmax, float* ymax);
// D_bar from the derivative part G = L_exp(A, sigma Ebar^H) (FP32)
void launch_adj_extract32(hipStream_t s, const Bucket& B, const BeamView& V, const double* kn, double dz,
                          const float2* G, const double* sigma, double2* Dbar);

}  // namespace ctem
