! Instrumentation driver for the compiled reference's Monte Carlo set-up (test infrastructure only).
!
! Written for this repository; it links against the reference's own modules, compiled from their
! sources by oracle/Makefile, and performs the set-up of CTEMMC.f90's DoMCsimulation
! (src_programs/CTEMMC.f90:174-234): CrystalData, CalcWaveLength (which places the atoms and so
! sets numat), CalcDensity, InitLambertParameters, rng_seed(primeseed + 1); it dumps the density,
! average Z and A (float32), the Lambert constants the trajectories use and the first ndraw
! numbers of rng_uniform, as one little-endian stream file for tests/golden/make_golden.py.
program ctemref_mc

use local
use files
use io
use crystalvars
use crystal
use symmetryvars
use symmetry
use constants
use diffraction, only: CalcWaveLength
use rng
use Lambert

IMPLICIT NONE

character(fnlen)  :: xtalname, outname
real(kind=dbl)    :: EkeV
integer(kind=irg) :: primeseed, ndraw, i
real(kind=sgl)    :: dens, avZ, avA
type(rng_t)       :: r
real(kind=sgl), allocatable :: u(:)
integer, parameter :: ou = 77, iu = 78

namelist /refmc/ xtalname, EkeV, primeseed, ndraw, outname

xtalname = 'undefined'; EkeV = 30.D0; primeseed = 932117; ndraw = 1000; outname = 'mcdump.bin'
open(unit=iu, file='refmc.nml', status='old', DELIM='APOSTROPHE')
read(unit=iu, nml=refmc)
close(unit=iu)

SG%SYM_reduce = .TRUE.
call CrystalData(xtalname)
call CalcWaveLength(EkeV*1000.D0, 3)
call CalcDensity(dens, avZ, avA)
call InitLambertParameters
call rng_seed(r, primeseed + 1)
allocate(u(ndraw))
do i = 1, ndraw
  u(i) = rng_uniform(r)
end do
open(unit=ou, file=trim(outname), access='stream', form='unformatted', status='replace')
write(ou) int(20262, 4), ndraw, cell%ATOM_ntype
write(ou) dens, avZ, avA
write(ou) LPs%sPio2, LPs%sPi2, cell%vol
write(ou) numat(1:cell%ATOM_ntype)
write(ou) u
close(ou)

end program ctemref_mc
