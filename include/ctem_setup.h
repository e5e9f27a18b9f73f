/* ctem_setup.h -- native per-energy setup of the master-pattern program.
 *
 * What the reference's host computes before the beam loop (CTEMEBSDmaster.f90:238-401), from
 * the crystal file's contents and the energy: metric and structure matrices (CalcMatrices,
 * crystal.f90:102-184), the space group's operators and the point group's (GenerateSymmetry,
 * symmetry.f90:368-450, with the point-group pass of ReadDataFile, files.f90:354-367), atom
 * positions and orbits (CalcPositions/CalcOrbit, symmetry.f90:614-676,773-863), the Laue group
 * of the Lambert sampling (CTEMEBSDmaster.f90:249-282), the relativistic wavelength with the
 * mean inner potential (CalcWaveLength, diffraction.f90:229-302), and the master reflection list
 * with its potential look-up table (Compute_ReflectionList 'ALL', gvectors.f90:1262-1444;
 * CalcUcg with the Weickenmeier-Kohl scattering factors and absorption, diffraction.f90:515-626,
 * others.f90:68-778). The float32 results (Ucg, the LUT, DynUpz) equal the reference's bit for
 * bit on the same C library; the result feeds ctem_upload_crystal directly.
 *
 * Conventions as in ctem_ebsd.h: int32_t status, never aborts, caller-owned host buffers.
 * Host code only; no GPU is needed.
 */
#ifndef CTEM_SETUP_H
#define CTEM_SETUP_H

#include <stdint.h>

#include "ctem_ebsd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The .xtal file's contents (files.f90:271-277): crystal system (1 cubic .. 7), lattice
 * parameters [nm, degrees], space group 1..237 (231..237: rhombohedral settings) and origin
 * setting (1 or 2; needed for the groups with two origins), and per site of the asymmetric
 * unit the atomic number (1..98) and x, y, z, occupancy, B [nm^2] (float32, as stored). */
typedef struct ctem_xtal {
  int32_t xtal_system;
  double a, b, c, alpha, beta, gamma;
  int32_t sgnum, sgset;
  int32_t ntype;
  const int32_t* atom_type; /* [ntype] */
  const float* atom_pos;    /* [5*ntype] */
} ctem_xtal;

/* Scalars of a setup (the reference's module variables of the same names). Matrices are
 * row-major 3x3. */
typedef struct ctem_setup_values {
  double mlambda, mrelcor, mpsihat, msigma; /* wavelength [nm], gamma, potential [V], sigma */
  double vol;                               /* cell volume [nm^3] */
  float dynupz;                             /* Vpmod of 000 [V] */
  int32_t isym, usehex;                     /* Laue group of the sampling and its grid */
  int32_t nsym, npt;                        /* space-group and point-group operator counts */
  double dmt[9], rmt[9], dsm[9], rsm[9];
} ctem_setup_values;

typedef struct ctem_setup ctem_setup;

/* Build the setup for one energy (ekev in keV, dmin in nm, Bethe cutoffs as in
 * BetheParameters.nml). */
CTEM_API int32_t ctem_setup_create(const ctem_xtal* x, float ekev, float dmin, float weakcutoff,
                                   float cutoff, float sgcutoff, ctem_setup** out);
/* The same setup with the look-up-table box -- CalcUcg of every lattice-allowed g, the bulk of
 * the work (143,745 cells for corundum at dmin 0.05 nm) -- computed on GPU `device` with the
 * C library's float32 expf, logf, sincosf and cabsf restated for the device (one thread per
 * cell). The result is bit-equal to ctem_setup_create's; the list order, 000 (and with it DynUpz
 * and the wavelength) and everything else are the host's. *lut_ms (may be NULL) receives the
 * kernel time. */
CTEM_API int32_t ctem_setup_create_gpu(int32_t device, const ctem_xtal* x, float ekev, float dmin,
                                       float weakcutoff, float cutoff, float sgcutoff, double* lut_ms,
                                       ctem_setup** out);
/* The set-up of the electron channeling pattern program (CTEMECP.f90:158-238): the same crystal
 * data with the wavelength from `voltage` in V (CalcWaveLength(dble(voltage)), :197-198) and the
 * master reflection list of Compute_ReflectionList 'ALL' for zone axis k with the convergence-angle
 * window convang = thetac/1000 [rad] (:238; gvectors.f90:1295-1301,1349-1410): of each family
 * only the members at angles pi/2 -+ 3 convang to k, then the zone's own members (g.k = 0), are
 * listed; LUT cells of the inner box that no listed entry took stay 0, as in the reference. */
CTEM_API int32_t ctem_setup_create_ecp(const ctem_xtal* x, float voltage, float dmin, const int32_t k[3],
                                       float convang, float weakcutoff, float cutoff, float sgcutoff,
                                       ctem_setup** out);
/* A ctem_crystal that points into the setup's arrays (valid until ctem_setup_destroy), for
 * ctem_upload_crystal. */
CTEM_API int32_t ctem_setup_crystal(const ctem_setup* s, ctem_crystal* out);
CTEM_API int32_t ctem_setup_info(const ctem_setup* s, ctem_setup_values* out);
/* CalcDensity (crystal.f90:1388-1417) of the setup's crystal, float32 as the reference returns them:
 * density [g/cm^3], average atomic number and average atomic weight [g/mol] over the atoms of the
 * cell (occupancy-weighted weights), as the Monte Carlo program takes them (CTEMMC.f90:191-194). */
CTEM_API int32_t ctem_setup_density(const ctem_setup* s, float* dens, float* avZ, float* avA);
CTEM_API int32_t ctem_setup_destroy(ctem_setup* s);
/* Message of the last failed ctem_setup_* call of this thread. */
CTEM_API const char* ctem_setup_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
