! Instrumentation driver for the compiled reference's electron channeling pattern set-up
! (test infrastructure only).
!
! Written for this repository; it links against the reference's own modules, compiled from
! their sources by oracle/Makefile, and performs the set-up of CTEMECP.f90's ECpattern
! (src_programs/CTEMECP.f90:158-265) step for step, then dumps what the beam loop receives:
!   - CalcWaveLength with the voltage in V (:197-198), CalcPositions, the foil normal DynFN
!     (TransSpace + NormVec, :204-205);
!   - GetPatternSymmetry / ShortestG -> ga, gb (:208-221), the Bragg angle of ga and the
!     ktmax / thetac relation (:227-233);
!   - Compute_ReflectionList 'ALL' with the convergence-angle window (:238;
!     gvectors.f90:1262-1444) -> LUT, dbdiff, master reflection list;
!   - CalckvectorsSymmetry (:245-253; kvectors.f90:573-690) -> every incident direction;
!   - Compute_DynMat 'BLOCHBETHE' with the foil normal (:299) for the first ndyn directions ->
!     beam sets and DynMat.
! The output is one little-endian stream file parsed by tests/golden/make_golden.py.
program ctemref_ecp

use local
use files
use io
use crystalvars
use crystal
use symmetryvars
use symmetry
use constants
use diffraction
use dynamical
use gvectors
use kvectors

IMPLICIT NONE

character(fnlen)  :: xtalname, outname
real(kind=sgl)    :: voltage, dmin, thetac, ktmax, bragg, galen, klaue(2)
integer(kind=irg) :: k(3), fn(3), ga(3), gb(3), npix, ndyn, i, j, dgn, isym, numk, ijmax, nref, ip, n, &
                     maxHOLZ, ik, npx, npy
real(kind=dbl)    :: ctmp(192,3)
integer, parameter :: ou = 77, iu = 78
type(kvectorlist),pointer :: kp

namelist /refecp/ xtalname, voltage, k, fn, dmin, thetac, ktmax, npix, outname, ndyn

xtalname = 'undefined'; voltage = 30000.0; k = (/0,0,1/); fn = (/0,0,1/); dmin = 0.025
thetac = 0.0; ktmax = 0.0; npix = 8; outname = 'ecpdump.bin'; ndyn = 1
maxHOLZ = 0

open(unit=iu, file='refecp.nml', status='old', DELIM='APOSTROPHE')
read(unit=iu, nml=refecp)
close(unit=iu)

SG%SYM_reduce = .TRUE.
call CrystalData(xtalname)
call CalcWaveLength(dble(voltage), 3)
call CalcPositions('v')
call TransSpace(float(fn), DynFN, 'd', 'r')
call NormVec(DynFN, 'r')
j = 0
do i = 1, 32
  if (SGPG(i).le.cell%SYM_SGnum) j = i
end do
dgn = GetPatternSymmetry(k, j, .TRUE.)
isym = WPPG(dgn)
call ShortestG(k, ga, gb, isym)
call GetHOLZGeometry(float(ga), float(gb), k, fn)
bragg = CalcDiffAngle(ga(1), ga(2), ga(3))*0.5
if (ktmax.ne.0.0) then
  thetac = (ktmax * 2.0 * bragg)*1000.0
else
  ktmax = thetac / (2000.0 * bragg)
end if
call Compute_ReflectionList(dmin, k, ga, gb, 'ALL', .FALSE., maxHOLZ, thetac/1000.0)
nref = DynNbeamsLinked
galen = CalcLength(float(ga), 'r')
npx = npix
npy = npx
klaue = (/ 0.0, 0.0 /)
ijmax = float(npx)**2
isym = 1
call CalckvectorsSymmetry(dble(k), dble(ga), dble(ktmax), npx, npy, numk, isym, ijmax, klaue)
call Set_Bethe_Parameters(.TRUE.)

open(unit=ou, file=trim(outname), access='stream', form='unformatted', status='replace')
write(ou) int(20261, 4), numk, nref, cell%ATOM_ntype, cell%SYM_SGnum, ga, gb
write(ou) size(LUT,1), size(LUT,2), size(LUT,3), lbound(LUT,1), lbound(LUT,2), lbound(LUT,3)
write(ou) ktmax, thetac, bragg, galen, DynFN
write(ou) mLambda, dble(DynUpz), cell%vol
write(ou) cell%dmt, cell%rmt
write(ou) LUT
write(ou) merge(1_4, 0_4, dbdiff)
rltmpa => reflist%next
do i = 1, nref
  write(ou) rltmpa%hkl, rltmpa%Ucg
  rltmpa => rltmpa%next
end do
do ip = 1, cell%ATOM_ntype
  call CalcOrbit(ip, n, ctmp)
  write(ou) cell%ATOM_type(ip), cell%ATOM_pos(ip,1:5), n
  write(ou) ctmp(1:n,1:3)
end do
kp => khead
do ik = 1, numk
  write(ou) kp%i, kp%j, kp%k, kp%kt, kp%kn
  kp => kp%next
end do
kp => khead
do ik = 1, min(ndyn, numk)
  call Compute_DynMat('BLOCHBETHE', kp%k, kp%kt, .FALSE.)
  write(ou) BetheParameter%nns, BetheParameter%nnw
  write(ou) BetheParameter%stronghkl(1:3,1:BetheParameter%nns), BetheParameter%strongsg(1:BetheParameter%nns)
  write(ou) DynMat
  kp => kp%next
end do
close(ou)

end program ctemref_ecp
