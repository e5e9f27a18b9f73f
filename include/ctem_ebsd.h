/* ctem_ebsd.h -- C-ABI of the MI355X-native EBSD master-pattern Bloch-wave path.
 *
 * This library replaces the per-beam-direction loop of CTEMsoft's CTEMEBSDmaster
 * (src_programs/CTEMEBSDmaster.f90:406-484): Bethe beam selection and ordering, assembly
 * of the dynamical matrix from the Ucg look-up table, the depth-integrated Lgh intensity
 * matrix, the Sgh contraction and the Lambert-grid scatter. The caller (a Fortran host
 * through ISO_C_BINDING, C, or Python through ctypes) keeps namelist input, the per-energy
 * setup and the master-file output.
 *
 * Conventions
 *   - Every function returns int32_t status: CTEM_OK (0) or a negative CTEM_E* code; it never
 *     aborts. ctem_last_error() gives a message for the calling thread.
 *   - Plain pointers and sizes only. Host buffers are caller-owned; device buffers are owned
 *     by the context or plan that allocated them. Work is ordered on the context's HIP
 *     stream and every call that returns host data synchronises it.
 *   - Complex numbers are interleaved (re, im) doubles. 3x3 matrices are row-major
 *     (m[3*i+j] = M(i,j)). The potential look-up table is in the reference's Fortran order
 *     LUT(h,k,l), h fastest, with lower bounds lut_lo[3] and extents lut_dim[3]
 *     (gvectors.f90:1316-1317).
 *   - One host thread per context.
 */
#ifndef CTEM_EBSD_H
#define CTEM_EBSD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define CTEM_API __attribute__((visibility("default")))
#else
#define CTEM_API
#endif

#define CTEM_OK 0
#define CTEM_E_ARG -1       /* invalid argument */
#define CTEM_E_HIP -2       /* HIP runtime error (no device, out of memory, launch failure) */
#define CTEM_E_STATE -3     /* call out of order (e.g. no crystal uploaded) */
#define CTEM_E_RANGE -4     /* input outside what the kernels support */
#define CTEM_E_NUMERIC -5   /* non-finite values or norm overflow in the matrix exponential */

/* precision of the Lgh stage */
#define CTEM_PREC_FP64 0    /* complex FP64 throughout (FP64 MFMA) */
#define CTEM_PREC_MIXED 1   /* split 16-bit MFMA products with FP32 accumulation; FP64 norm,
                               squaring-count choice and depth accumulation; Lgh products in
                               FP32 and the final contraction in FP64 (intensities within 1e-4,
                               measured <= 1.2e-5, of CTEM_PREC_FP64) */

/* split schemes of the mixed form's complex products (ctem_plan_set_mixed_terms) */
#define CTEM_SPLIT_BF16X3 3   /* two BF16 planes, 3 MFMA products: ~2^-16 relative */
#define CTEM_SPLIT_BF16X6 6   /* three BF16 planes, 6 MFMA products: ~FP32 accuracy */
#define CTEM_SPLIT_FP16X3 16  /* two FP16 planes of x*2^12, 3 MFMA products: ~2^-21 relative;
                                 entries must stay below 16 in magnitude (CTEM_E_NUMERIC) */

typedef struct ctem_ctx ctem_ctx;
typedef struct ctem_plan ctem_plan;

/* Per-energy crystal setup, as the reference host computes it (CTEMEBSDmaster.f90:358-396):
 * CalcWaveLength (diffraction.f90:229-302), Compute_ReflectionList 'ALL'
 * (gvectors.f90:1262-1444), Set_Bethe_Parameters (gvectors.f90:1196-1234), CalcOrbit
 * (symmetry.f90:614-676). */
typedef struct ctem_crystal {
  int32_t nref;               /* master reflection list length (DynNbeamsLinked); entry 0 is 000 */
  const int32_t* hkl;         /* [3*nref] Miller indices in list order */
  const double* ucg;          /* [2*nref] Ucg per list entry (float32 values widened) */
  const uint8_t* dbdiff;      /* [nref] double-diffraction flag of each list entry */
  int32_t lut_lo[3];          /* lower bounds of LUT (= -2*imh, -2*imk, -2*iml) */
  int32_t lut_dim[3];         /* extents of LUT */
  const double* lut;          /* [2*prod(lut_dim)] complex LUT, Fortran order */
  double mlambda;             /* electron wavelength [nm] (mLambda) */
  float dynupz;               /* DynUpz = Vpmod of 000 [V] */
  float weakcutoff, cutoff, sgcutoff; /* &BetheList */
  double rmt[9];              /* reciprocal metric tensor */
  int32_t nsite;              /* atoms in the asymmetric unit (numset) */
  const int32_t* site_z;      /* [nsite] atomic numbers */
  const float* site_b;        /* [nsite] Debye-Waller B [nm^2] */
  const int32_t* site_norbit; /* [nsite] orbit sizes */
  const double* orbit;        /* [3*sum(norbit)] orbit positions, site by site */
} ctem_crystal;

/* ----- context ----------------------------------------------------------------------- */
CTEM_API int32_t ctem_ctx_create(int32_t device, ctem_ctx** out);
CTEM_API int32_t ctem_ctx_destroy(ctem_ctx* ctx);
CTEM_API const char* ctem_last_error(void);
/* device name and compute-unit count, for logs */
CTEM_API int32_t ctem_device_info(ctem_ctx* ctx, char* name, int32_t name_len, int32_t* cu_count);

/* Upload the per-energy crystal setup (the state Compute_ReflectionList and CalcOrbit leave
 * for the beam loop, CTEMEBSDmaster.f90:381-401 and :427). It becomes the context's current
 * crystal: plans created afterwards use it, and each plan keeps its crystal until it is
 * destroyed, so the bins of an energy series can be resident together (one upload and one
 * plan per bin). Precomputes the per-site structure-matrix table S_p(g-h) of
 * CTEMEBSDmaster.f90:418-449 and its derivative in B over the box of all g-h. */
CTEM_API int32_t ctem_upload_crystal(ctem_ctx* ctx, const ctem_crystal* c);

/* ----- incident directions ------------------------------------------------------------ */
/* RoscaLambert incident directions of Calckvectors (kvectors.f90:421-542, AddkVector
 * :839-947) for Laue group isym (1..12) on the square (usehex=0) or hexagonal grid.
 * hemisphere = +1 samples the northern hemisphere as the reference does; -1 returns the
 * same Lambert pixels mirrored to the southern hemisphere (k_z -> -k_z before the
 * transformation to crystal coordinates). Returns the count in *numk; pass ij/k/kn = NULL to
 * query the count only. dsm is the direct structure matrix. */
CTEM_API int32_t ctem_kvectors(int32_t npx, int32_t npy, int32_t isym, int32_t usehex, int32_t hemisphere,
                      const double dsm[9], double mlambda, int32_t cap, int32_t* ij, double* k,
                      double* kn, int32_t* numk);

/* Laue-symmetry equivalent Lambert pixels (Apply2DLaueSymmetry, Lambert.f90:1942-2046);
 * iequiv holds up to 12 (i,j) pairs. */
CTEM_API int32_t ctem_laue_equiv(int32_t ipx, int32_t ipy, int32_t isym, int32_t* iequiv, int32_t* nequiv);

/* ----- batched master-pattern plan ---------------------------------------------------- */
/* A plan holds numk incident directions resident in device memory plus the depth profile:
 * lam[nz] float32 weights for depths z_k = k*dz, k = 1..nz, the reference's integer
 * thickness `thick` (CTEMEBSDmaster.f90:142,306) and kn = 1/lambda per direction. */
CTEM_API int32_t ctem_plan_create(ctem_ctx* ctx, int32_t numk, const double* k /*3*numk*/,
                         const double* kn /*numk*/, int32_t nz, const float* lam, double dz,
                         double thick, ctem_plan** out);
CTEM_API int32_t ctem_plan_destroy(ctem_plan* plan);

/* Beam selection on the GPU (gvectors.f90:988-1120): per direction the strong and weak
 * reflections in list order with their float32 excitation errors, bit-exact with the
 * reference. Optionally copies the counts to the host. */
CTEM_API int32_t ctem_plan_select(ctem_plan* plan, int32_t* nns /*numk or NULL*/, int32_t* nnw /*numk or NULL*/);
/* Copy one direction's beam set to the host (after ctem_plan_select). */
CTEM_API int32_t ctem_plan_get_beams(ctem_plan* plan, int32_t ik, int32_t* strong, float* strongsg,
                            int32_t* weak, float* weaksg);

/* Run the whole path for every direction of the plan: selection (if not done), DynMat
 * assembly, Lgh and the Sgh contraction, in the given precision. Results stay on the
 * device until fetched. keep_lgh_ik (or -1) names one direction whose Lgh is kept for
 * ctem_plan_get_lgh. */
CTEM_API int32_t ctem_plan_run(ctem_plan* plan, int32_t precision, int32_t keep_lgh_ik);

/* No counterpart in the reference (the launch pattern of the batched loop). on = 1: once a run
 * of the plan has read back the squaring counts of its matrices (the first run in each
 * precision; the same directions and crystal give the same counts on every run), later runs
 * need no host synchronisation inside, and the plan captures its whole run -- assembly,
 * matrix exponentials, depth recursion and contraction of every bucket -- into a HIP graph
 * once and replays it as one launch. The graph is captured again when the options, the
 * selection or the context's work space change. A run that needs more squarings than
 * recorded (not possible for unchanged inputs) is detected on the device and run again
 * without the graph. Per-kernel timing (ctem_plan_timing) is not recorded in a graph replay;
 * the total is. Runs that keep an Lgh (keep_lgh_ik >= 0) do not use the graph. Default 0. */
CTEM_API int32_t ctem_plan_set_graph(ctem_plan* plan, int32_t on);
/* No counterpart in the reference (launch tuning; results do not depend on it). Directions are
 * grouped into buckets of one padded matrix size, a multiple of bucket_multiple (32, default,
 * or 64), and each bucket runs in chunks of as many directions as fit in chunk_gib GiB of work
 * space (0.25..256, default 24). ctemsoft_amd/autotune.py picks both per GPU and workload. */
CTEM_API int32_t ctem_plan_set_tuning(ctem_plan* plan, int32_t bucket_multiple, double chunk_gib);
/* No counterpart in the reference. Run several plans of one context (the energy bins of a
 * series) as one captured HIP graph: every plan's bucket layout and squaring counts are
 * established first (eager runs where missing), then all their runs are captured into one
 * graph, kept on the context and replayed while the same plans, options and work space are
 * used. Results are those of ctem_plan_run on each plan, bit for bit. */
CTEM_API int32_t ctem_plans_run(ctem_plan* const* plans, int32_t nplans, int32_t precision);

/* Split schemes (CTEM_SPLIT_*) of the mixed form for the Taylor products (on the scaled
 * matrix, ||B||_1 <= 3/4 with CTEM_SPLIT_FP16X3 for both, 1/2 otherwise) and for the squarings;
 * default CTEM_SPLIT_FP16X3 for both. */
CTEM_API int32_t ctem_plan_set_mixed_terms(ctem_plan* plan, int32_t taylor_terms, int32_t square_terms);
/* No counterpart in the reference (a switch of the mixed form). mode 1 (default): when the
 * crystal's structure-factor LUT is inversion-symmetric, U(-g) = U(g) to within 2^-20 of its max
 * (centrosymmetric structure, origin on an inversion centre), the split mixed form treats the
 * structure matrix as complex symmetric and computes each product's tiles on and above the
 * diagonal only; mode 0: full products. *in_effect (may be NULL): 1 if the symmetric products
 * will be used. The FP64 form always uses full products. */
CTEM_API int32_t ctem_plan_set_symmetry(ctem_plan* plan, int32_t mode, int32_t* in_effect);

/* Re sum_gh Lgh(g,h) Sgh(g,h,p) per direction and site, in double, before the reference's
 * float32 store (CTEMEBSDmaster.f90:456): dots[ik*nsite + p]. */
CTEM_API int32_t ctem_plan_get_dots(ctem_plan* plan, double* dots);
/* The reference's float32 svals: fl32(fl32(dot)/fl32(sum of orbit sizes)) (:456-459). */
CTEM_API int32_t ctem_plan_get_svals(ctem_plan* plan, float* svals);
/* Lgh (n x n complex, row-major) of the direction named in ctem_plan_run; *n on return. */
CTEM_API int32_t ctem_plan_get_lgh(ctem_plan* plan, double* lgh, int32_t cap, int32_t* n);
/* DynMat of one direction assembled on the GPU (row-major n x n), for parity tests. */
CTEM_API int32_t ctem_plan_dynmat(ctem_plan* plan, int32_t ik, double* dynmat, int32_t cap, int32_t* n);

/* Reverse mode, FP64: given gdots[ik*nsite + p] = df/ddots for a real function f of the plan's
 * intensities (ctem_plan_get_dots), returns df/dU for every LUT entry (glut[2*cell] real part,
 * glut[2*cell+1] imaginary part, cells in the Fortran order of the uploaded LUT) and df/dB_p
 * per site (gb[p], B in nm^2). The beam sets are held fixed; the forward matrices are
 * recomputed, not stored. The reference has no derivatives; this entry serves the fitting of
 * Ucg and Debye-Waller factors to a measured master pattern. Up to 512 strong beams per
 * direction. */
CTEM_API int32_t ctem_plan_adjoint(ctem_plan* plan, const double* gdots, double* glut, double* gb);
/* The same gradient with the mixed forward: psi (and so the intensities the function of gdots was
 * evaluated on) from the plan's mixed form of ctem_plan_run, the backward depth recursion in FP64,
 * and the Frechet derivative of the exponential on split-FP16 products with FP32 accumulation,
 * its squaring stages recomputed. No counterpart in the reference, as ctem_plan_adjoint. Same
 * arguments and limits. */
CTEM_API int32_t ctem_plan_adjoint_mixed(ctem_plan* plan, const double* gdots, double* glut, double* gb);

/* Scatter svals into a Lambert master array sr[nsite][2*npy+1][2*npx+1] (float32), applying
 * the Laue-symmetry equivalents per direction (CTEMEBSDmaster.f90:462-479). ij holds the
 * (i,j) of every direction of the plan. sr is a host buffer. */
CTEM_API int32_t ctem_plan_scatter(ctem_plan* plan, const int32_t* ij, int32_t npx, int32_t npy,
                          int32_t isym, float* sr);

/* Timing of the last ctem_plan_run, measured with HIP events on the context's stream:
 * total milliseconds and milliseconds in the dominant (matrix-exponential GEMM) kernels,
 * with the algorithmic flop count of those kernels. */
CTEM_API int32_t ctem_plan_timing(ctem_plan* plan, double* total_ms, double* gemm_ms, double* gemm_flops,
                         int32_t* gemm_launches);

/* No counterpart in the reference (measurement). The algorithmic flop count of the last
 * FP64 run of the plan, summed over its directions: total = 8n^2w + 8n^3(q+j) + (32/3)n^3 +
 * 16n^2Z + 4n^2P per direction (n strong and w weak beams, Z depths, P sites), products =
 * 8n^3(q+j), solve = (32/3)n^3, where q and j are the Pade degree and squaring count that
 * the reference's MatrixExponential (math.f90:208-219) picks from max |a_ij| of
 * A = 2 pi i dz DynMat / fl32(2kn). The count does not depend on the algorithm the library
 * runs; it prices the work of the path. */
CTEM_API int32_t ctem_plan_alg_flops(ctem_plan* plan, double* total, double* products, double* solve);

/* ----- single-direction drop-in for CalcLgh3 ------------------------------------------ */
/* Same arguments and meaning as the reference's external subroutine CalcLgh3
 * (CTEMEBSDmaster.f90:610-630): DMat and Lgh are nn x nn complex, column-major (Fortran),
 * lambdaE[izz] float32. Computed on the GPU of ctx in FP64. The reference's unset `kin` and
 * unused `debug` are accepted and ignored. */
CTEM_API int32_t ctem_calclgh3(ctem_ctx* ctx, const double* dmat, double* lgh, double thick, double kn,
                      int32_t nn, int32_t gzero, double depthstep, const float* lambdae,
                      int32_t izz);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): ctem_calclgh3 with
 * the same arguments and results, whose work space is a fixed array in device memory that the
 * library holds per GPU (8 MiB + 64 KiB, kept until the process ends) instead of buffers
 * allocated on the context. It calls no hipMalloc or hipFree. Limits: nn <= 256 strong beams
 * and izz <= 256 depth steps (CTEM_E_RANGE beyond). Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (several host threads, or contexts on the same device
 * used concurrently) race and must be serialised by the caller. */
CTEM_API int32_t ctem_calclgh3_static(ctem_ctx* ctx, const double* dmat, double* lgh, double thick, double kn,
                             int32_t nn, int32_t gzero, double depthstep, const float* lambdae,
                             int32_t izz);

#ifdef __cplusplus
}
#endif
#endif
