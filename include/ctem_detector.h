/* ctem_detector.h -- EBSD detector patterns from a master pattern, on the GPU.
 *
 * Replaces the computation of the reference's detector-pattern program, ComputeEBSDPatterns
 * (src_programs/CTEMEBSD.f90:117-673): for every orientation, every scintillator pixel's
 * direction is rotated into the crystal frame, mapped to the square Lambert projection and the
 * energy-resolved master pattern is interpolated bilinearly there, weighted per energy bin with
 * the Monte Carlo exit-energy histogram interpolated to that pixel, and summed over the
 * energy window; the sum times the intensity prefactor is the pattern. The host keeps what the
 * program reads from its files (the &EBSDdata namelist, the angle file, the Monte Carlo energy
 * histograms, the master pattern) and what it does after the sum (Poisson noise, binning,
 * gamma), for which this header also declares host routines with the program's semantics.
 *
 * Lambert positions on the edge of the square (directions on the equator) read one row or column
 * past the master's or the histograms' last; the program reads whatever memory follows its
 * array there, here such values are 0.
 *
 * Results: the noise-free patterns equal the reference program's float32 values bit for bit
 * (the float32 steps in the program's order, the C library's atanf restated for the device,
 * the Fortran runtime's compensated SUM where the program calls SUM); ctem_poisson_noise
 * reproduces POIDEV's sequence, so the noisy patterns equal the program's output file.
 *
 * Conventions as in ctem_ebsd.h: int32_t status (CTEM_OK or a negative CTEM_E_* code), never
 * aborts, ctem_detector_last_error() gives a message for the calling thread; caller-owned host
 * buffers; one host thread per detector. Patterns are float32 [n][numsy][numsx] (x fastest: the
 * program's EBSDpattern(numsx, numsy) per orientation); quaternions are float32 [n][4] (scalar
 * first, the program's quatang(1:4, i)).
 */
#ifndef CTEM_DETECTOR_H
#define CTEM_DETECTOR_H

#include <stdint.h>

#include "ctem_ebsd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ctem_detector ctem_detector;

/* The detector and beam of &EBSDdata (CTEMEBSD.f90:167-200; defaults :180-198), float32 and
 * double as the program declares them: scintillator distance L and pixel size delta [micron],
 * detector tilt thetac [degrees], numsx x numsy pixels, pattern centre xpc, ypc [pixels], the
 * energy window energymin..energymax [keV], beam current [nA] and dwell time [micro s]. */
typedef struct ctem_ebsd_geometry {
  float L, thetac, delta;
  int32_t numsx, numsy;
  float xpc, ypc;
  float energymin, energymax;
  double beamcurrent, dwelltime;
} ctem_ebsd_geometry;

/* The Monte Carlo file's energy histograms as the program reads them (CTEMEBSD.f90:299-330):
 * numEbins bins from Ehistmin in steps of Ebinsize [keV], the half sizes nsx, nsy of the
 * Lambert grid (the file's numsx = 2 nsx + 1), the sample tilt MCsig [degrees] and
 * accum_e(numEbins, -nsx:nsx, -nsy:nsy) in Fortran order (energy fastest). */
typedef struct ctem_mc_energy {
  int32_t numEbins, nsx, nsy;
  double Ehistmin, Ebinsize, MCsig;
  const int32_t* accum_e;
} ctem_mc_energy;

/* <- the program's set-up before its orientation loop (CTEMEBSD.f90:299-335,411-518): the energy
 * window Emin..Emax, the pixel direction cosines, the exit-energy histograms interpolated to
 * the pixels, num_el and the intensity prefactor; uploads what the patterns need to `device`.
 * The program leaves the detector histograms of the bins outside Emin..Emax unassigned and
 * still sums them into num_el (:470-508); here they are zero. */
CTEM_API int32_t ctem_detector_create(int32_t device, const ctem_ebsd_geometry* geom, const ctem_mc_energy* mc,
                                      ctem_detector** out);
/* Emin, Emax (1-based bins), num_el and the prefactor (CTEMEBSD.f90:336-343,508,525); any
 * pointer may be NULL. num_el is 64-bit: the program's is a default integer, and for detector
 * totals past 2^31 its NINT (and so the prefactor's sign and size) is undefined. */
CTEM_API int32_t ctem_detector_info(const ctem_detector* d, int32_t* emin, int32_t* emax, int64_t* num_el,
                                    double* prefactor);
/* <- the master file's srtmp(-npx:npx, -npy:npy, nE, numset) (Fortran order) and its site sum
 * sr = sum(srtmp, 4) (CTEMEBSD.f90:374-404); nE must equal the histograms' numEbins
 * (CTEM_E_ARG otherwise, where the program stops). Square Lambert grids only (the program's
 * interpolation is the square one). */
CTEM_API int32_t ctem_detector_set_master(ctem_detector* d, int32_t npx, int32_t npy, int32_t nE, int32_t numset,
                                          const float* srtmp);
/* <- the orientation loop without the noise (CTEMEBSD.f90:525-556): n patterns for the
 * quaternions quats[n][4] into patterns[n][numsy][numsx] (host buffers; synchronous) */
CTEM_API int32_t ctem_detector_patterns(ctem_detector* d, int32_t n, const float* quats, float* patterns);
/* The same on device buffers (d_quats, d_patterns allocated by the caller on the detector's
 * device), enqueued on the detector's stream; ctem_detector_sync waits for it. */
CTEM_API int32_t ctem_detector_patterns_device(ctem_detector* d, int32_t n, const float* d_quats, float* d_patterns);
CTEM_API int32_t ctem_detector_sync(ctem_detector* d);
/* milliseconds of the last ctem_detector_patterns(_device) kernel (HIP events on its stream) */
CTEM_API int32_t ctem_detector_last_ms(const ctem_detector* d, double* ms);
CTEM_API int32_t ctem_detector_destroy(ctem_detector* d);
CTEM_API const char* ctem_detector_last_error(void);

/* ---- host routines with the program's semantics (no GPU) ---- */

/* quats[i] = eu2qu(eul[i] * dtor) with dtor = 0.0174533 (CTEMEBSD.f90:148,244;
 * rotations.f90:1322-1351); eul in degrees, float32 (Bunge phi1, Phi, phi2) */
CTEM_API int32_t ctem_eu2qu_deg(int32_t n, const float* eul_deg, float* quats);
/* The program's optional extra rotation (CTEMEBSD.f90:264-270): axisangle[4] = (axis, angle in
 * degrees); when the angle is not 0, every quaternion becomes quat_mult(ax2qu(axisangle), q). */
CTEM_API int32_t ctem_apply_axisangle(int32_t n, const float* axisangle, float* quats);
/* POIDEV applied to the pixels in the program's order (CTEMEBSD.f90:558-563: x outer, y inner)
 * with the generator state of the program's noise routines (noise.f90:136-228) in state[16]:
 * zero it and set state[0] = idum (the program uses -1) before the first pattern, then pass it
 * on unchanged from pattern to pattern. Values must be finite and non-negative (CTEM_E_ARG:
 * the program's POIDEV never returns for NaN or infinity). */
CTEM_API int32_t ctem_poisson_noise(int32_t numsx, int32_t numsy, int32_t npat, float* patterns, int64_t* state);
/* The program's scaling step (CTEMEBSD.f90:575-590): binned(numsx/binning, numsy/binning) =
 * sum of each binning x binning block (the Fortran runtime's SUM) times 1/binning^2, then
 * raised to gammavalue when gamma != 0 */
CTEM_API int32_t ctem_bin_pattern(int32_t numsx, int32_t numsy, int32_t binning, int32_t gamma, float gammavalue,
                                  const float* pattern, float* binned);

#ifdef __cplusplus
}
#endif
#endif
