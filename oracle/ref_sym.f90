! Symmetry-table driver for the compiled reference (test infrastructure only).
!
! Written for this repository; it links against the reference's own modules (compiled from
! its sources by oracle/Makefile) and dumps two integer/geometry tables that the master
! pattern's scatter and direction list depend on, so that the oracle's restatement and the
! library can be pinned against the reference for every Laue group:
!   - Apply2DLaueSymmetry (Lambert.f90:1942-2046) for isym = 1..13 over every pixel (ipx, ipy)
!     of the box -nbox..nbox squared: nequiv and iequiv(2,12);
!   - Calckvectors 'RoscaLambert' (kvectors.f90:421-542, AddkVector :839-947) for isym = 1..12
!     on the square and the hexagonal grid: numk, then (i, j, k(3), kn) per direction;
!   - the Laue group and grid of every space group 1..230 by the program's rule
!     (CTEMEBSDmaster.f90:249-282) on the reference's SGPG/PGLaueinv tables.
! The structure matrix and wavelength come from the namelist (they only enter through the
! final transformation of AddkVector). The output is one little-endian stream file parsed by
! tests/golden/make_golden.py.
program ctemref_sym

use local
use constants
use crystalvars
use symmetryvars
use diffraction
use kvectors
use Lambert

IMPLICIT NONE

integer(kind=irg) :: nbox, npx, npy, npyhex, isym, ipx, ipy, iequiv(2,12), nequiv, ihex, numk, ijmax, ik
integer(kind=irg) :: isg, i, j
integer(kind=irg),parameter :: LaueTest(11) = (/ 149, 151, 153, 156, 158, 160, 161, 164, 165, 166, 167 /)
real(kind=dbl)    :: dsm(3,3), wavelength
character(fnlen)  :: outname
logical           :: usehex
type(kvectorlist),pointer :: kp
integer, parameter :: ou = 77, iu = 78

namelist /refsym/ nbox, npx, dsm, wavelength, outname

nbox = 16; npx = 11; dsm = 0.D0; wavelength = 0.01D0; outname = 'refsym.bin'
open(unit=iu, file='refsym.nml', status='old', DELIM='APOSTROPHE')
read(unit=iu, nml=refsym)
close(unit=iu)

cell%dsm = dsm
mLambda = wavelength

open(unit=ou, file=trim(outname), access='stream', form='unformatted', status='replace')
write(ou) int(20261, 4), nbox, npx

! Laue-group equivalents of every pixel of the box
do isym = 1, 13
  do ipy = -nbox, nbox
    do ipx = -nbox, nbox
      iequiv = 0
      call Apply2DLaueSymmetry(ipx, ipy, isym, iequiv, nequiv)
      write(ou) nequiv, iequiv
    end do
  end do
end do

! incident directions for every Laue group, square and hexagonal grid; the program passes
! npyhex = nint(2 npy / sqrt(3)) as npy on the hexagonal grid (CTEMEBSDmaster.f90:390-401)
npy = npx
npyhex = nint(2.0*float(npy)/sqrt(3.0))
do ihex = 0, 1
  usehex = (ihex.eq.1)
  do isym = 1, 12
    ijmax = 0
    if (usehex) then
      call Calckvectors((/0.D0,0.D0,1.D0/), (/0.D0,0.D0,0.D0/), 0.D0, npx, npyhex, numk, isym, ijmax, &
                        'RoscaLambert', usehex)
    else
      call Calckvectors((/0.D0,0.D0,1.D0/), (/0.D0,0.D0,0.D0/), 0.D0, npx, npy, numk, isym, ijmax, &
                        'RoscaLambert', usehex)
    end if
    write(ou) isym, ihex, numk
    kp => khead
    do ik = 1, numk
      write(ou) kp%i, kp%j, kp%k, kp%kn
      kp => kp%next
    end do
  end do
end do
! the Laue group and grid the master-pattern program picks for every space group
! (CTEMEBSDmaster.f90:249-282, with the reference's SGPG and PGLaueinv tables)
do isg = 1, 230
  j = 0
  do i = 1, 32
    if (SGPG(i).le.isg) j = i
  end do
  isym = PGLaueinv(j)
  if (isym.eq.7) then
    do i = 1, 11
      if (isg.eq.LaueTest(i)) isym = 12
    end do
  end if
  write(ou) isg, isym, merge(1, 0, ((isym.ge.6).and.(isym.le.9)).or.(isym.eq.12))
end do
close(ou)

end program ctemref_sym
