! Fortran host of the batched path: the EBSD master pattern of one energy with the per-direction
! beam loop of CTEMEBSDmaster (src_programs/CTEMEBSDmaster.f90:406-484) replaced by one plan of
! the MI355X library, called through ISO_C_BINDING (ctem_ebsd_iface).
!
! The host keeps what the reference's host does: the crystal and its symmetry (CrystalData,
! CalcPositions), the Laue-group choice, the Bethe parameters (BetheParameters.nml), the
! wavelength, the master reflection list with its Ucg look-up table (Compute_ReflectionList), the
! incident directions (Calckvectors) and the site orbits (CalcOrbit) -- all computed by the
! reference's own modules, which this program links -- and the master-file output in the
! reference's record layout (:540-568). The beam loop is ctem_upload_crystal, ctem_plan_create,
! ctem_plan_run and ctem_plan_scatter.
!
! Input: namelist file `masterhost.nml` (&masterhost: xtalname, ekev, dmin, npx, izz, depthstep,
! thick, lamfile, outname, energyfile, precision, device) and the depth weights lambdaE(1:izz) as
! a float32 stream file. Exit status 0 on success, 1 on a library error, 2 on an input error.
program ctem_master_host

use, intrinsic :: iso_c_binding
use local
use files
use crystalvars
use crystal
use symmetryvars
use symmetry
use constants
use diffraction
use dynamical
use gvectors
use kvectors
use ctem_ebsd_iface

IMPLICIT NONE

character(fnlen)  :: xtalname, lamfile, outname, energyfile
real(kind=sgl)    :: ekev, dmin
real(kind=dbl)    :: depthstep, thick
integer(kind=irg) :: npx, npy, npyhex, izz, precision, device, isym, i, j, ip, n, numk, ijmax, nref, numset
integer(kind=irg) :: gbox(3), ntot, io
logical           :: usehex, switchmirror
! space groups of Laue class -3m whose Lambert sampling uses the second setting (isym 12)
integer(kind=irg),parameter :: LaueTest(11) = (/ 149, 151, 153, 156, 158, 160, 161, 164, 165, 166, 167 /)
real(kind=sgl)    :: gtmp(3)
real(kind=dbl)    :: ctmp(192,3)
type(kvectorlist),pointer :: kp

! arrays handed to the library (contiguous, target)
type(ctem_crystal) :: cr
integer(c_int32_t), allocatable, target :: hkl(:,:), ij(:,:), site_z(:), site_no(:)
complex(c_double_complex), allocatable, target :: ucgl(:), lutc(:,:,:)
integer(c_int8_t), allocatable, target :: dbd(:)
real(c_double), allocatable, target :: kk(:,:), kn(:), orb(:,:)
real(c_float), allocatable, target :: site_b(:), lam(:)
real(c_float), allocatable, target :: sr(:,:,:)
type(c_ptr) :: ctx, plan
integer(c_int32_t) :: rc

integer, parameter :: iu = 78, ou = 79
namelist /masterhost/ xtalname, ekev, dmin, npx, izz, depthstep, thick, lamfile, outname, energyfile, &
                      precision, device

xtalname = 'undefined'; ekev = 20.0; dmin = 0.1; npx = 50; izz = 1; depthstep = 1.D0; thick = 1.D0
lamfile = 'undefined'; outname = 'master.data'; energyfile = 'undefined'; precision = 0; device = 0

open(unit=iu, file='masterhost.nml', status='old', DELIM='APOSTROPHE', iostat=io)
if (io.ne.0) then
  write (*,*) 'masterhost.nml not found'
  stop 2
end if
read(unit=iu, nml=masterhost)
close(unit=iu)
allocate(lam(izz))
open(unit=iu, file=trim(lamfile), access='stream', form='unformatted', status='old', iostat=io)
if (io.ne.0) then
  write (*,*) 'depth-weight file not found: ', trim(lamfile)
  stop 2
end if
read(iu) lam
close(iu)

! ---- per-energy setup with the reference's modules (CTEMEBSDmaster.f90:238-401) ----------
SG%SYM_reduce = .TRUE.
call CrystalData(xtalname)
call CalcPositions('v')
numset = cell%ATOM_ntype

! Laue group of the space group: the last point group whose first space group does not exceed
! SGnum, its Laue class, and the second -3m setting for the listed trigonal groups
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
usehex = (((isym.ge.6).and.(isym.le.9)).or.(isym.eq.12))
npy = npx
npyhex = nint(2.0*float(npy)/sqrt(3.0))

call Set_Bethe_Parameters()
call CalcWaveLength(dble(ekev*1000.0), 3)
if (associated(reflist)) call Delete_gvectorlist()
nullify(reflist)
! Compute_ReflectionList marks visited family members in `refdone`, sized to the box
! -imh:imh (...); with hexagonal axes a member of an in-box g can reach |h+k| = 2 imh, so the
! array is allocated here over the LUT box first (the routine keeps an existing allocation).
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
  call Calckvectors((/0.D0,0.D0,1.D0/), (/0.D0,0.D0,0.D0/), 0.D0, npx, npyhex, numk, isym, ijmax, &
                    'RoscaLambert', usehex)
else
  call Calckvectors((/0.D0,0.D0,1.D0/), (/0.D0,0.D0,0.D0/), 0.D0, npx, npy, numk, isym, ijmax, &
                    'RoscaLambert', usehex)
end if

! ---- the setup in the library's layout -------------------------------------------------------
allocate(hkl(3,nref), ucgl(nref), dbd(nref))
rltmpa => reflist%next
do i = 1, nref
  hkl(:,i) = rltmpa%hkl
  ucgl(i) = rltmpa%Ucg
  dbd(i) = merge(1_c_int8_t, 0_c_int8_t, dbdiff(rltmpa%hkl(1), rltmpa%hkl(2), rltmpa%hkl(3)))
  rltmpa => rltmpa%next
end do
allocate(lutc(lbound(LUT,1):ubound(LUT,1), lbound(LUT,2):ubound(LUT,2), lbound(LUT,3):ubound(LUT,3)))
lutc = LUT
allocate(site_z(numset), site_b(numset), site_no(numset))
ntot = 0
do ip = 1, numset
  call CalcOrbit(ip, n, ctmp)
  site_no(ip) = n
  ntot = ntot + n
end do
allocate(orb(3, ntot))
ntot = 0
do ip = 1, numset
  call CalcOrbit(ip, n, ctmp)
  site_z(ip) = cell%ATOM_type(ip)
  site_b(ip) = cell%ATOM_pos(ip,5)
  orb(1:3, ntot+1:ntot+n) = transpose(ctmp(1:n,1:3))
  ntot = ntot + n
end do
cr%nref = nref
cr%hkl = c_loc(hkl)
cr%ucg = c_loc(ucgl)
cr%dbdiff = c_loc(dbd)
cr%lut_lo = lbound(LUT)
cr%lut_dim = shape(LUT)
cr%lut = c_loc(lutc)
cr%mlambda = mLambda
cr%dynupz = DynUpz
cr%weakcutoff = BetheParameter%weakcutoff
cr%cutoff = BetheParameter%cutoff
cr%sgcutoff = BetheParameter%sgcutoff
cr%rmt = reshape(transpose(cell%rmt), (/9/))
cr%nsite = numset
cr%site_z = c_loc(site_z)
cr%site_b = c_loc(site_b)
cr%site_norbit = c_loc(site_no)
cr%orbit = c_loc(orb)

allocate(kk(3,numk), kn(numk), ij(2,numk))
kp => khead
do i = 1, numk
  kk(:,i) = kp%k
  kn(i) = kp%kn
  ij(:,i) = (/ kp%i, kp%j /)
  kp => kp%next
end do

! ---- the beam loop on the GPU ---------------------------------------------------------------
rc = ctem_ctx_create(int(device, c_int32_t), ctx)
if (rc.ne.CTEM_OK) call die('ctem_ctx_create')
rc = ctem_upload_crystal(ctx, cr)
if (rc.ne.CTEM_OK) call die('ctem_upload_crystal')
rc = ctem_plan_create(ctx, int(numk, c_int32_t), kk, kn, int(izz, c_int32_t), lam, depthstep, thick, plan)
if (rc.ne.CTEM_OK) call die('ctem_plan_create')
rc = ctem_plan_run(plan, int(precision, c_int32_t), -1_c_int32_t)
if (rc.ne.CTEM_OK) call die('ctem_plan_run')
! the square-array result (Apply2DLaueSymmetry + store, :462-479); on the hexagonal grid the
! library fills the hexagonal sampling array the reference calls srhex
if (usehex) then
  allocate(sr(-npx:npx, -npyhex:npyhex, numset))
  sr = 0.0
  rc = ctem_plan_scatter(plan, ij, int(npx, c_int32_t), int(npyhex, c_int32_t), int(isym, c_int32_t), sr)
else
  allocate(sr(-npx:npx, -npy:npy, numset))
  sr = 0.0
  rc = ctem_plan_scatter(plan, ij, int(npx, c_int32_t), int(npy, c_int32_t), int(isym, c_int32_t), sr)
end if
if (rc.ne.CTEM_OK) call die('ctem_plan_scatter')
rc = ctem_plan_destroy(plan)
rc = ctem_ctx_destroy(ctx)

! ---- master file: the reference's records (:540-568), one energy ----------------------------
open(unit=ou, file=trim(outname), status='unknown', action='write', form='unformatted')
write (ou) progname_host()
write (ou) scversion
write (ou) xtalname
write (ou) energyfile
write (ou) npx, merge(npyhex, npy, usehex), 1, numset
write (ou) (/ ekev /)
write (ou) cell%ATOM_type(1:numset)
if (usehex) then
  write (ou) 'hexago'
else
  write (ou) 'square'
end if
write (ou) sr
close(unit=ou, status='keep')
write (*,"(A,I0,A,I0,A)") 'master pattern: ', numk, ' directions, ', numset, ' site(s) -> '//trim(outname)

contains

function progname_host() result(p)
  character(fnlen) :: p
  p = 'CTEMEBSDmaster.f90'
end function progname_host

subroutine die(what)
  character(*), intent(in) :: what
  write (*,*) trim(what)//' failed: '//ctem_error_message()
  stop 1
end subroutine die

end program ctem_master_host
