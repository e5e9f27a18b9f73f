! Fortran host check of the ISO_C_BINDING boundary: reads one direction's DynMat and depth
! profile from `lgh_in.bin` (stream file written by tests/test_fortran_boundary.py), calls
! ctem_calclgh3 -- the GPU stand-in for CalcLgh3 (CTEMEBSDmaster.f90:610-722) with the same
! arguments -- and writes Lgh to `lgh_out.bin`. Exit status 0 on success, 1 on a library error.
program ctem_lgh_check

use, intrinsic :: iso_c_binding
use ctem_ebsd_iface

implicit none

integer(c_int32_t) :: nn, izz, rc
real(c_double) :: thick, kn, depthstep
real(c_float), allocatable :: lambdae(:)
complex(c_double_complex), allocatable :: dmat(:,:), lgh(:,:)
type(c_ptr) :: ctx

open(unit=10, file='lgh_in.bin', access='stream', form='unformatted', status='old')
read(10) nn, izz
read(10) thick, kn, depthstep
allocate(lambdae(izz), dmat(nn,nn), lgh(nn,nn))
read(10) lambdae
read(10) dmat
close(10)

rc = ctem_ctx_create(0, ctx)
if (rc /= CTEM_OK) then
  write (*,*) 'ctem_ctx_create failed: ', ctem_error_message()
  stop 1
end if
rc = ctem_calclgh3(ctx, dmat, lgh, thick, kn, nn, 1, depthstep, lambdae, izz)
if (rc /= CTEM_OK) then
  write (*,*) 'ctem_calclgh3 failed: ', ctem_error_message()
  stop 1
end if
rc = ctem_ctx_destroy(ctx)

open(unit=11, file='lgh_out.bin', access='stream', form='unformatted', status='replace')
write(11) lgh
close(11)
write (*,'(A,I0)') 'ctem_lgh_check: Lgh computed on the GPU for nn = ', nn

end program ctem_lgh_check
