/* CPU restatement of CTEMsoft's EBSD master-pattern beam loop (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: this library is the checker that tests/, the package's
 * smoke check and bench.py's CPU-baseline leg compare the HIP kernels against. It is never
 * linked into, loaded by or called from the product library.
 *
 * Every function cites the reference routine it restates. Conventions: 3x3 metric and
 * structure matrices are row-major (m[i*3+j] = M(i,j)); complex matrices are row-major
 * interleaved (re,im) doubles, D[2*(i*n+j)] = Re M(i,j); the potential look-up table is in
 * the reference's Fortran order LUT(h,k,l) with lower bounds lo[3] and extents dim[3].
 */
#ifndef CTEM_ORACLE_H
#define CTEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* excitation error s_g in float32: CalcsgSingle, diffraction.f90:649-674 */
float orc_calcsg(const float g[3], const float k[3], const float fn[3], const double rmt[9]);

/* beam selection for one incident direction: Compute_DynMat 'BLOCHBETHE' selection and
 * ordering, gvectors.f90:988-1120. Returns 0, fills strong/weak reflection-list indices
 * (0-based, list order; strong[0] = 0 is the transmitted beam) and their float32 s_g. */
int orc_select_beams(int nref, const int32_t* hkl, const double* ucg /* 2*nref */,
                     const uint8_t* dbdiff /* per list entry */, double mlambda,
                     float weakcutoff, float cutoff, float sgcutoff, const double rmt[9],
                     const double kvec[3], int32_t* strong, float* strongsg, int32_t* nns,
                     int32_t* weak, float* weaksg, int32_t* nnw);

/* the same with the foil normal fn (float32 reciprocal components; NULL: FN = k), as the
 * channeling-pattern program selects with its DynFN (CTEMECP.f90:204-205,299) */
int orc_select_beams_fn(int nref, const int32_t* hkl, const double* ucg, const uint8_t* dbdiff,
                        double mlambda, float weakcutoff, float cutoff, float sgcutoff,
                        const double rmt[9], const double kvec[3], const float* fn, int32_t* strong,
                        float* strongsg, int32_t* nns, int32_t* weak, float* weaksg, int32_t* nnw);

/* Bethe-corrected dynamical matrix (n x n complex): gvectors.f90:1123-1173 */
int orc_dynmat(int nns, const int32_t* strong, const float* strongsg, int nnw,
               const int32_t* weak, const float* weaksg, const int32_t* hkl,
               const double* lut, const int32_t lo[3], const int32_t dim[3], double mlambda,
               float dynupz, double* dynmat);

/* structure matrix Sgh for one site: CTEMEBSDmaster.f90:418-449 (float32 terms) */
int orc_sgh(int nns, const int32_t* strong, const int32_t* hkl, const double rmt[9],
            int32_t z, float b_nm2, int norbit, const double* orbit /* norbit x 3 */,
            double* sgh);

/* the channeling structure matrix, all sites (occupancy included): CTEMECP.f90:318-350 */
int orc_sgh_ecp(int nns, const int32_t* strong, const int32_t* hkl, const double rmt[9], int nsite,
                const int32_t* z, const float* b_nm2, const float* occ, const int32_t* norbit,
                const double* orbit, double* sgh);

/* depth-integrated Lgh = (dz/thick) sum_k lam_k psi_k psi_k^H with
 * psi_k = expm(2 pi i k dz DynMat / fl32(2 kn)) e_1, evaluated in extended precision by
 * scaling and squaring; equal to CalcLgh3 (CTEMEBSDmaster.f90:610-722) up to rounding. */
int orc_lgh(int n, const double* dynmat, double kn, double dz, double thick, int nz,
            const float* lam, double* lgh);

/* Re sum_gh Lgh(g,h) Sgh(g,h): the double sum before the float32 store, :456-459 */
double orc_lgh_dot_sgh(int n, const double* lgh, const double* sgh);

/* RoscaLambert incident directions: kvectors.f90:421-542 (+AddkVector :839-947). Writes up
 * to cap directions; returns numk (or -1 if cap is too small). */
int orc_kvectors(int npx, int npy, int isym, int usehex, const double dsm[9], double mlambda,
                 int cap, int32_t* ij /* 2*cap */, double* k /* 3*cap */, double* kn);

/* Laue-equivalent Lambert pixels: Apply2DLaueSymmetry, Lambert.f90:1942-2046 */
int orc_laue_equiv(int ipx, int ipy, int isym, int32_t* iequiv /* 2*12 */);

/* ---- EBSD detector patterns (CTEMEBSD.f90) ---------------------------------------------- */

/* Float32 Lambert square inverse, Lambert2DSquareInverseSingle (Lambert.f90:405-433), times
 * the grid scale scl: xy[0], xy[1]; returns ierr (1 when |1 - |xyz|^2| > 1e-6: xy = 0) */
int orc_lambert_inverse_single(const float xyz[3], float scl, float xy[2]);

/* Detector set-up, CTEMEBSD.f90:414-518: pixel direction cosines rg[3][numsy][numsx]
 * (x fastest; the reference's rgx/rgy/rgz(numsx,numsy)), the MC energy histograms
 * accum_e(numEbins, -nsx:nsx, -nsy:nsy) (int32, Fortran order) interpolated to the pixels and
 * scaled by 0.25 for bins Emin..Emax (1-based, inclusive; the other bins 0), as
 * wdet[numsy][numsx][numEbins]; num_el = nint(sum(wdet)). sig is the MC file's sample tilt
 * (float32 as the program keeps it), thetac the detector tilt [deg], L and delta [micron],
 * xpc/ypc the pattern centre [pixels]. */
int orc_ebsd_detector(float L, float thetac, float delta, int numsx, int numsy, float xpc, float ypc,
                      float sig, int numEbins, int nsx, int nsy, const int32_t* accum_e, int Emin,
                      int Emax, float* rg, float* wdet, int64_t* num_el);

/* One noise-free pattern, CTEMEBSD.f90:525-556: the rotated pixel directions (quaternion q,
 * float32 quat_mult, quaternions.f90:255-265), the Lambert gather of the site-summed master
 * sr(-npx:npx, -npy:npy, nE) over bins Emin..Emax weighted by wdet, times prefactor (double,
 * stored float32). out[numsy][numsx] (x fastest). */
int orc_ebsd_pattern(int numsx, int numsy, const float* rg, const float* wdet, int numEbins, int Emin,
                     int Emax, int npx, int npy, const float* sr, const float q[4], double prefactor,
                     float* out);

/* Poisson deviates, POIDEV with the generator ran (noise.f90:136-228, Numerical Recipes), with
 * the routines' saved variables in st[] (zero-initialise, then st[0] = idum, e.g. -1). Applied
 * in the program's order: i (x) outer, j (y) inner. */
void orc_poisson_noise(float* pattern, int numsx, int numsy, int64_t* st);

/* the program's extra rotation: q = quat_mult(ax2qu(axisangle), q), axisangle[3] in degrees
 * (CTEMEBSD.f90:264-270; rotations.f90 ax2om/om2qu) */
void orc_apply_axisangle(int n, const float axisangle[4], float* quats);

/* sr = sum(srtmp, 4): per element over nset sites (CTEMEBSD.f90:404), the runtime's SUM */
void orc_site_sum(int64_t per, int nset, const float* srtmp, float* sr);

/* binning (binning divides numsx, numsy) and gamma, CTEMEBSD.f90:575-590 */
void orc_bin_pattern(int numsx, int numsy, int binning, int gamma, float gammavalue, const float* pat,
                     float* binned);

/* Euler angles (radians, Bunge) to a quaternion, eu2qu (rotations.f90:1322-1351), float32 */
void orc_eu2qu(const float e[3], float q[4]);

/* ---- Monte Carlo (CTEMMC.f90) ---- */
/* the generator of rng.f95 (seed: rng_seed; uniform: rng_uniform, float32) and its jump by n draws */
typedef struct { int32_t s[4]; } orc_mcrng;
void orc_mcrng_seed(orc_mcrng* r, int32_t seed);
float orc_mcrng_uniform(orc_mcrng* r);
void orc_mcrng_jump(orc_mcrng* r, uint64_t n);
/* single_run (CTEMMC.f90:304-526): mode 0 the program's one stream, mode 1 one block of 16384
 * draws of that stream per electron; kahan: compensated SUM in the direction normalisation */
int64_t orc_mc_run(double sig, double omega, int32_t numsx, int32_t primeseed, int64_t num_el, double EkeV,
                   double Ehistmin, double Ebinsize, double depthmax, double depthstep, double Ze, double density,
                   double at_wt, int mode, int kahan, int32_t* accum_e, int32_t* accum_z);

#ifdef __cplusplus
}
#endif
#endif
