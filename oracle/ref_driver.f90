! Instrumentation driver for the compiled reference (test infrastructure only).
!
! Written for this repository; it links against the reference's own modules, compiled
! from its sources by oracle/Makefile, and dumps the intermediate quantities of the EBSD
! master-pattern beam loop (CTEMEBSDmaster.f90:358-484) for a chosen set of incident
! directions, so that the CPU restatement in oracle/ctem_oracle.c and the HIP kernels can
! be compared with the reference's own numbers:
!   - per-energy setup: CalcWaveLength (diffraction.f90:229), Compute_ReflectionList 'ALL'
!     (gvectors.f90:1262-1444) -> LUT, dbdiff, master reflection list;
!   - Calckvectors 'RoscaLambert' (kvectors.f90:421-542) -> all incident directions;
!   - per direction: Compute_DynMat 'BLOCHBETHE' (gvectors.f90:960-1176) -> beam sets and
!     DynMat; the program's own CalcLgh3 (CTEMEBSDmaster.f90:610-722) -> Lgh.
! The output is one little-endian stream file parsed by tests/golden/make_golden.py.
program ctemref_driver

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

character(fnlen)  :: xtalname, lamfile, kselfile, outname, kfile
real(kind=sgl)    :: ekev, dmin
integer(kind=irg) :: npx, izz, dolgh, nsel, i, j, ik, isel, npy, npyhex, isym, numk, ijmax, nref, ip, n
real(kind=dbl)    :: depthstep, thick, kin
real(kind=dbl)    :: kcur(3), ktcur(3), kncur
real(kind=dbl),allocatable :: kf(:,:), knf(:)
logical           :: usehex, switchmirror
integer(kind=irg),parameter :: LaueTest(11) = (/ 149, 151, 153, 156, 158, 160, 161, 164, 165, 166, 167 /)
real(kind=sgl),allocatable :: lam(:)
integer(kind=irg),allocatable :: ksel(:), kdone(:)
complex(kind=dbl),allocatable :: Lgh(:,:)
real(kind=dbl) :: ctmp(192,3)
integer(kind=irg) :: gbox(3), hm(3), gg(3), ix, iy, iz, lutfull, nkf
complex(kind=dbl),allocatable :: lutx(:,:,:)
real(kind=sgl) :: gtmp(3)
integer, parameter :: ou = 77, iu = 78
type(kvectorlist),pointer :: kp

! dolgh: 0 = beam sets only, 1 = beam sets + DynMat, 2 = beam sets + DynMat + Lgh
! lutfull: 1 = also dump CalcUcg over the box that holds every difference of two list entries
! kfile: a stream file (int32 count, then per direction k(3) and kn as doubles) whose wave vectors
!        replace the selected ones of the list (e.g. the library's southern hemisphere, which the
!        program itself does not compute); the output numbers them 1..count
namelist /refdrv/ xtalname, ekev, dmin, npx, izz, depthstep, thick, lamfile, kselfile, outname, dolgh, &
                  lutfull, kfile

xtalname = 'undefined'; ekev = 20.0; dmin = 0.1; npx = 50; izz = 1; depthstep = 1.D0; thick = 1.D0
lamfile = 'undefined'; kselfile = 'undefined'; outname = 'refdump.bin'; dolgh = 0; lutfull = 0; kfile = 'undefined'

open(unit=iu, file='refdrv.nml', status='old', DELIM='APOSTROPHE')
read(unit=iu, nml=refdrv)
close(unit=iu)

allocate(lam(izz))
open(unit=iu, file=trim(lamfile), access='stream', form='unformatted', status='old')
read(iu) lam
close(iu)

! crystal, symmetry and atom positions (CTEMEBSDmaster.f90:238-247)
SG%SYM_reduce = .TRUE.
call CrystalData(xtalname)
call CalcPositions('v')

! Laue group selection (CTEMEBSDmaster.f90:249-282)
j = 0
do i = 1, 32
  if (SGPG(i).le.cell%SYM_SGnum) j = i
end do
isym = PGLaueinv(j)
switchmirror = .FALSE.
if (isym.eq.7) then
  do i = 1, 11
    if (cell%SYM_SGnum.eq.LaueTest(i)) switchmirror = .TRUE.
  end do
end if
if (switchmirror) isym = 12
usehex = .FALSE.
if (((isym.ge.6).and.(isym.le.9)).or.(isym.eq.12)) usehex = .TRUE.
npy = npx
npyhex = nint(2.0*float(npy)/sqrt(3.0))

call Set_Bethe_Parameters()
call CalcWaveLength(dble(ekev*1000.0), 3)
if (associated(reflist)) call Delete_gvectorlist()
nullify(reflist)
! Compute_ReflectionList (gvectors.f90:1333-1337) flags visited family members in the module
! array refdone, sized to the box -imh:imh x -imk:imk x -iml:iml. With hexagonal axes a family
! member of an in-box g can have |h+k| up to 2*imh, so those flag writes land outside the
! array: they alias other entries or, at the box ends, corrupt the heap (corundum, SG 167,
! crashes). The routine allocates refdone with stat= and keeps an existing allocation, so the
! driver allocates it first with the bounds of the LUT box (-2*imh:2*imh, ...). For cubic
! crystals every family member of an in-box g stays in the box and nothing changes.
gbox(1:3) = 1
do i = 1, 3
  do
    gbox(i) = gbox(i) + 1
    gtmp = 0.0
    gtmp(i) = float(gbox(i))
    if (1.0/CalcLength(gtmp, 'r') .lt. dmin) exit
  end do
end do
if (allocated(refdone)) deallocate(refdone)
allocate(refdone(-2*gbox(1):2*gbox(1), -2*gbox(2):2*gbox(2), -2*gbox(3):2*gbox(3)))
call Compute_ReflectionList(dmin, (/0,0,1/), (/1,0,0/), (/0,1,0/), 'ALL', .FALSE., 0)
nref = DynNbeamsLinked

if (usehex) then
  call Calckvectors((/0.D0,0.D0,1.D0/), (/0.D0,0.D0,0.D0/), 0.D0, npx, npyhex, numk, isym, ijmax, 'RoscaLambert', usehex)
else
  call Calckvectors((/0.D0,0.D0,1.D0/), (/0.D0,0.D0,0.D0/), 0.D0, npx, npy, numk, isym, ijmax, 'RoscaLambert', usehex)
end if

if (trim(kselfile).eq.'undefined') then
  nsel = numk
  allocate(ksel(nsel))
  do i = 1, nsel
    ksel(i) = i
  end do
else
  open(unit=iu, file=trim(kselfile), access='stream', form='unformatted', status='old')
  read(iu) nsel
  allocate(ksel(nsel))
  read(iu) ksel
  close(iu)
end if

if (trim(kfile).ne.'undefined') then
  open(unit=iu, file=trim(kfile), access='stream', form='unformatted', status='old')
  read(iu) nkf
  allocate(kf(3,nkf), knf(nkf))
  do i = 1, nkf
    read(iu) kf(1:3,i), knf(i)
  end do
  close(iu)
  deallocate(ksel)
  nsel = nkf
  allocate(ksel(nsel))
  do i = 1, nsel
    ksel(i) = i
  end do
end if

open(unit=ou, file=trim(outname), access='stream', form='unformatted', status='replace')
! header
write(ou) int(20260, 4), isym, merge(1,0,usehex), numk, nref, nsel, cell%ATOM_ntype, cell%SYM_SGnum
write(ou) size(LUT,1), size(LUT,2), size(LUT,3), lbound(LUT,1), lbound(LUT,2), lbound(LUT,3)
write(ou) mLambda, mRelcor, mPsihat, mSigma, dble(DynUpz), dble(BetheParameter%weakcutoff), &
          dble(BetheParameter%cutoff), dble(BetheParameter%sgcutoff), cell%vol
write(ou) cell%dmt, cell%rmt, cell%dsm, cell%rsm
! LUT and double-diffraction flags (Fortran order)
write(ou) LUT
write(ou) merge(1_4, 0_4, dbdiff)
! master reflection list in list order
rltmpa => reflist%next
do i = 1, nref
  write(ou) rltmpa%hkl, rltmpa%Ucg
  rltmpa => rltmpa%next
end do
! sites: atomic number, position, occupancy, B, and the orbit (CalcOrbit, symmetry.f90:614-676)
do ip = 1, cell%ATOM_ntype
  call CalcOrbit(ip, n, ctmp)
  write(ou) cell%ATOM_type(ip), cell%ATOM_pos(ip,1:5), n
  write(ou) ctmp(1:n,1:3)
end do
! all incident directions in list order
kp => khead
do ik = 1, numk
  write(ou) kp%i, kp%j, kp%k, kp%kn
  kp => kp%next
end do

! optional: the potential coefficients over the box -2*hm:2*hm (hm = largest |index| of the list
! along each axis), CalcUcg (diffraction.f90:515-626) for every allowed g, 0 elsewhere, as the
! reference fills its own LUT (gvectors.f90:1342-1427) over the smaller box
if (lutfull.eq.1) then
  hm = 0
  rltmpa => reflist%next
  do i = 1, nref
    hm = max(hm, abs(rltmpa%hkl))
    rltmpa => rltmpa%next
  end do
  allocate(lutx(-2*hm(1):2*hm(1), -2*hm(2):2*hm(2), -2*hm(3):2*hm(3)))
  lutx = dcmplx(0.D0, 0.D0)
  do iz = -2*hm(3), 2*hm(3)
    do iy = -2*hm(2), 2*hm(2)
      do ix = -2*hm(1), 2*hm(1)
        gg = (/ ix, iy, iz /)
        if (IsGAllowed(gg)) then
          call CalcUcg(gg)
          lutx(ix,iy,iz) = rlp%Ucg
        end if
      end do
    end do
  end do
  write(ou) hm
  write(ou) lutx
  deallocate(lutx)
end if

! per selected direction (beam loop body, CTEMEBSDmaster.f90:410-453)
allocate(kdone(numk))
kp => khead
ik = 1
do isel = 1, nsel
  if (allocated(kf)) then
    kcur = kf(1:3,isel)
    kncur = knf(isel)
    ktcur = 0.D0
  else
    do while (ik.lt.ksel(isel))
      kp => kp%next
      ik = ik + 1
    end do
    kcur = kp%k
    kncur = kp%kn
    ktcur = kp%kt
  end if
  DynFN = kcur
  call Compute_DynMat('BLOCHBETHE', kcur, ktcur, .FALSE.)
  write(ou) ksel(isel), BetheParameter%nns, BetheParameter%nnw
  write(ou) BetheParameter%strongID(1:BetheParameter%nns)
  write(ou) BetheParameter%strongsg(1:BetheParameter%nns)
  write(ou) BetheParameter%weakhkl(1:3,1:BetheParameter%nnw)
  write(ou) BetheParameter%weaksg(1:BetheParameter%nnw)
  if (dolgh.ge.1) write(ou) DynMat
  if (dolgh.ge.2) then
    allocate(Lgh(BetheParameter%nns, BetheParameter%nns))
    call CalcLgh3(DynMat, Lgh, thick, kncur, BetheParameter%nns, 1, kin, 0, depthstep, lam, izz)
    write(ou) Lgh
    deallocate(Lgh)
  end if
end do
close(ou)

end program ctemref_driver
