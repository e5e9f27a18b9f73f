/* ctem_mc.h -- Monte Carlo electron trajectories (backscattered-electron energy and depth
 * histograms) on the GPU.
 *
 * Replaces the trajectory loop of the reference's Monte Carlo program, single_run
 * (src_programs/CTEMMC.f90:304-526): electrons enter a sample tilted by sig (and omega) at EkeV,
 * take continuous-slowing-down steps with screened-Rutherford deflections (the program's
 * expressions in double precision), and those that leave the surface are counted by exit energy
 * and exit direction on the square Lambert grid (accum_e) and, coarsened tenfold, by exit energy and
 * the depth of their last scattering event (accum_z). The host keeps the program's set-up (namelist,
 * crystal, CalcDensity: ctem_setup_density in ctem_setup.h) and its output file.
 *
 * Random numbers are those of the program's generator (rng.f95: Marsaglia's lag-3 subtraction
 * modulo 2147483579 combined with the 69069 congruential generator, rng_seed(primeseed + 1)):
 *   CTEM_MC_SEQUENTIAL  one stream for all electrons in order, as the program draws them (one GPU
 *                       lane; the program's output, for checking and for small runs);
 *   CTEM_MC_BLOCKED     electron e draws from the same stream starting at draw 16384 e (no
 *                       trajectory draws more than 1 + 3 x 5000), reached by jumping the
 *                       generator; electrons run in parallel and the histograms do not depend on
 *                       how they are spread over lanes. Statistically the program's estimate; not
 *                       the program's particular sample.
 *
 * Histogram layout (Fortran order of the program's arrays, energy fastest): accum_e
 * [numsx][numsx][numEbins] (index [j+nx][i+nx][iE-1], the program's accum_e(iE, i, j)), accum_z
 * [nzw][nzw][numzbins][numEbins] with nzw = 2 (nx / 10) + 1, nx = (numsx - 1) / 2. Hits the program
 * would write outside its arrays (an energy window that is not a whole number of bins; numsx
 * with nx mod 10 >= 5, where the depth map's pixel index reaches nx / 10 + 1) are counted in
 * *dropped instead.
 *
 * Conventions as in ctem_ebsd.h: int32_t status, never aborts, ctem_mc_last_error() gives a
 * message for the calling thread; one host thread per object.
 */
#ifndef CTEM_MC_H
#define CTEM_MC_H

#include <stdint.h>

#include "ctem_ebsd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CTEM_MC_BLOCKED 0
#define CTEM_MC_SEQUENTIAL 1

/* &MCdata (CTEMMC.f90:150-171; defaults :153-166) and CalcDensity's values widened to double as the
 * program passes them (Ze = dble(avZ), density = dble(dens), at_wt = dble(avA), :191-194) */
typedef struct ctem_mc_params {
  double sig, omega;                  /* sample tilt angles [degrees] */
  int32_t numsx;                      /* Lambert grid points along x and y (odd) */
  int32_t primeseed;
  int64_t num_el;                     /* electrons */
  double EkeV, Ehistmin, Ebinsize;    /* [keV] */
  double depthmax, depthstep;         /* [nm] */
  double Ze, density, at_wt;
} ctem_mc_params;

typedef struct ctem_mc ctem_mc;

/* Derived constants and zeroed histograms on `device`. numEbins = int((EkeV - Ehistmin)/Ebinsize)
 * + 1, numzbins = int(depthmax/depthstep) + 1 (:200-201). */
CTEM_API int32_t ctem_mc_create(int32_t device, const ctem_mc_params* p, ctem_mc** out);
CTEM_API int32_t ctem_mc_dims(const ctem_mc* mc, int32_t* numEbins, int32_t* numzbins, int32_t* nx, int32_t* nzw);
/* Add the hits of electrons first .. first + count - 1 to the histograms (CTEM_MC_BLOCKED; ranges
 * may be run in any order or split, e.g. over GPUs), or of electrons 0 .. count - 1 of the program's
 * single stream (CTEM_MC_SEQUENTIAL, first = 0). Synchronous. */
CTEM_API int32_t ctem_mc_run(ctem_mc* mc, int32_t mode, int64_t first, int64_t count);
CTEM_API int32_t ctem_mc_reset(ctem_mc* mc);
/* copy the histograms (either pointer may be NULL) and the dropped-hit count to the host */
CTEM_API int32_t ctem_mc_get(ctem_mc* mc, int32_t* accum_e, int32_t* accum_z, int64_t* dropped);
/* device addresses of the histograms, valid until ctem_mc_destroy (for a consumer on the GPU) */
CTEM_API int32_t ctem_mc_device(ctem_mc* mc, int32_t** accum_e, int32_t** accum_z);
/* No counterpart in the reference (measurement): trajectory steps taken since the last reset,
 * the scattering events of :434-521 (the entry of each electron not counted) */
CTEM_API int32_t ctem_mc_steps(ctem_mc* mc, int64_t* steps);
/* milliseconds of the last ctem_mc_run's kernel (HIP events) */
CTEM_API int32_t ctem_mc_last_ms(const ctem_mc* mc, double* ms);
CTEM_API int32_t ctem_mc_destroy(ctem_mc* mc);
CTEM_API const char* ctem_mc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
