/* ctem_ecp.h -- electron channeling patterns (ECP) on the GPU.
 *
 * Replaces the beam loop of the reference's channeling-pattern program, ECpattern
 * (src_programs/CTEMECP.f90:292-380): for every incident direction of the illumination disk,
 * Bethe beam selection with the program's foil normal, the Bethe-corrected dynamical matrix, the
 * thickness-integrated intensity matrix Lgh at each requested thickness, and the contraction
 * sr = Re sum_gh Lgh(g,h) Sgh(g,h) / sum(nat) with the channeling structure matrix
 * Sgh = sum over sites and orbits of Z^2 occ exp(-B s^2) exp(2 pi i (g-h).r) (:312-350). Both of
 * the program's solution modes are provided:
 *   CTEM_ECP_BLOCHWV (CalcLgh, :472-534): Lgh(t) = (1/t) int_0^t psi psi^H dz with
 *     psi(z) = exp(2 pi i z DynMat / fl32(2 kn)) e_1. The reference diagonalises DynMat; here the
 *     integral comes from one matrix exponential of the block matrix
 *     [[A, 0], [eps Sgh^T, -A^H]] (C. Van Loan, IEEE Trans. Autom. Control 23 (1978) 395) per
 *     direction and step, so that no eigensolver runs on the GPU.
 *   CTEM_ECP_SCATMAT (CalcLghSM, :559-618): the scattering matrix Azz = exp(i pi lambda
 *     zintstep DynMat) applied step by step to e_1 and Lgh the running sum of psi psi^H, stored
 *     when the step count reaches nint(thick/zintstep).
 * The host keeps the program's set-up (namelist, crystal, wavelength, reflection list with the
 * convergence-angle window: ctem_setup_create_ecp in ctem_setup.h) and output; this header also
 * declares the set-up's geometry steps (foil normal, ga, convergence, incident directions) with
 * the program's float32/float64 semantics.
 *
 * Conventions as in ctem_ebsd.h (int32_t status, never aborts, ctem_last_error()). An ECP plan is a
 * ctem_plan: ctem_plan_select, ctem_plan_get_beams, ctem_plan_dynmat, ctem_plan_timing and
 * ctem_plan_destroy apply to it; ctem_plan_run does not (ctem_ecp_run does).
 */
#ifndef CTEM_ECP_H
#define CTEM_ECP_H

#include <stdint.h>

#include "ctem_ebsd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CTEM_ECP_BLOCHWV 0
#define CTEM_ECP_SCATMAT 1

/* ---- host geometry of ECpattern (no GPU) ---- */

/* message of the last failed geometry call (ctem_ecp_foil_normal .. ctem_ecp_kvectors) of this
 * thread; the plan calls below report through ctem_last_error() */
CTEM_API const char* ctem_ecp_last_error(void);

/* DynFN: TransSpace(float(fn), 'd', 'r') then NormVec(.., 'r') (CTEMECP.f90:204-205), float32 */
CTEM_API int32_t ctem_ecp_foil_normal(const int32_t fn[3], const double dmt[9], const double rmt[9], float dynfn[3]);
/* ga of ShortestG(k, ga, gb, isym) (CTEMECP.f90:218; symmetry.f90:1133-1309): the shortest pair of
 * reciprocal lattice vectors spanning the zone of k, ordered so that (ga x gb).k >= 0. (The
 * program's further changes for 3- and 6-fold zones touch gb only.) */
CTEM_API int32_t ctem_ecp_shortest_g(const int32_t k[3], const double dmt[9], const double rmt[9], int32_t ga[3]);
/* The convergence relation of CTEMECP.f90:227-233 with bragg = CalcDiffAngle(ga)/2 and
 * galen = CalcLength(float(ga), 'r') (:239): with *ktmax != 0 on entry *thetac is computed
 * from it, otherwise *ktmax from *thetac [mrad]. bragg, galen may be NULL. */
CTEM_API int32_t ctem_ecp_convergence(const int32_t ga[3], const double rmt[9], double mlambda, float* thetac,
                                      float* ktmax, float* bragg, float* galen);
/* The incident directions of CalckvectorsSymmetry(k, ga, ktmax, npx, npx, numk, 1, npx^2, 0)
 * (CTEMECP.f90:245-253; kvectors.f90:573-690, Add_knode :839-947): the disk i^2 + j^2 <= npx^2
 * of the grid with step delta = 2 ktmax |ga| / (2 npx + 1) along ga and ga x k, the centre first,
 * then i ascending, j ascending. Per direction: ij[2], k[3] (float32 values in double, as the
 * program stores them), kt[3] and kn. Pass NULL arrays to query *numk only. */
CTEM_API int32_t ctem_ecp_kvectors(const int32_t k[3], const int32_t ga[3], float ktmax, int32_t npx,
                                   const double dmt[9], const double rmt[9], double mlambda, int32_t cap,
                                   int32_t* ij, double* kvec, double* kt, double* kn, int32_t* numk);

/* ---- the beam loop on the GPU ---- */

typedef struct ctem_ecp_params {
  int32_t mode;            /* CTEM_ECP_BLOCHWV or CTEM_ECP_SCATMAT (compmode) */
  int32_t numthick;        /* thicknesses startthick + thickinc * (i - 1), i = 1..numthick [nm] */
  float startthick, thickinc;
  float zintstep;          /* ScatMat integration step [nm] */
  float fn[3];             /* DynFN (ctem_ecp_foil_normal) */
  const float* site_occ;   /* [nsite] site occupancies (the .xtal file's ATOM_pos(:,4)) */
} ctem_ecp_params;

/* A plan over numk directions (k[3*numk], kn[numk] as ctem_ecp_kvectors returns them) on the
 * context's current crystal (uploaded from ctem_setup_create_ecp). Precomputes the channeling
 * structure matrix table Sgh(g-h) of CTEMECP.f90:318-350 over the box of all g-h. Limits:
 * 1..4096 thicknesses; BLOCHWV needs startthick > 0 and thickinc > 0 (the program divides by
 * thick); SCATMAT needs zintstep > 0 and at most 65536 steps nint(max thick / zintstep). */
CTEM_API int32_t ctem_ecp_plan_create(ctem_ctx* ctx, int32_t numk, const double* k, const double* kn,
                                      const ctem_ecp_params* prm, ctem_plan** out);
/* The beam loop for every direction (selection if not yet done), FP64 throughout. */
CTEM_API int32_t ctem_ecp_run(ctem_plan* plan);
/* sr[it*numk + ik] = fl32(Re sum Lgh Sgh / float(sum(nat))) (CTEMECP.f90:366-369); directions
 * whose step count never meets a ScatMat thickness (a thickness below zintstep/2) hold 0 there,
 * as the program's zeroed Lgh gives. */
CTEM_API int32_t ctem_ecp_get_sr(ctem_plan* plan, float* sr);
/* The program's output array sr(2npx+1, 2npx+1, numthick) (:290,366) as [it][j+npx][i+npx]
 * from the directions' ij; pixels outside the disk are 0 (the program leaves them unassigned). */
CTEM_API int32_t ctem_ecp_image(ctem_plan* plan, const int32_t* ij, int32_t npx, float* img);

#ifdef __cplusplus
}
#endif
#endif
