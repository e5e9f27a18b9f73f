! ISO_C_BINDING interface to the MI355X EBSD master-pattern library (include/ctem_ebsd.h).
!
! A Fortran host (CTEMsoft's CTEMEBSDmaster or any other) `use`s this module and calls the
! C-ABI directly; every function returns an int32 status (CTEM_OK = 0) and never stops the
! program, so the host maps non-zero codes to its own FatalError. Arrays are passed by
! address (c_loc of contiguous, target arrays); complex(c_double_complex) arrays have the
! interleaved (re, im) layout the C side expects.
module ctem_ebsd_iface

use, intrinsic :: iso_c_binding

implicit none

integer(c_int32_t), parameter :: CTEM_OK = 0, CTEM_E_ARG = -1, CTEM_E_HIP = -2, &
                                 CTEM_E_STATE = -3, CTEM_E_RANGE = -4, CTEM_E_NUMERIC = -5
integer(c_int32_t), parameter :: CTEM_PREC_FP64 = 0, CTEM_PREC_MIXED = 1
integer(c_int32_t), parameter :: CTEM_SPLIT_BF16X3 = 3, CTEM_SPLIT_BF16X6 = 6, CTEM_SPLIT_FP16X3 = 16

! per-energy crystal setup (struct ctem_crystal)
type, bind(C) :: ctem_crystal
  integer(c_int32_t) :: nref
  type(c_ptr)        :: hkl, ucg, dbdiff
  integer(c_int32_t) :: lut_lo(3), lut_dim(3)
  type(c_ptr)        :: lut
  real(c_double)     :: mlambda
  real(c_float)      :: dynupz, weakcutoff, cutoff, sgcutoff
  real(c_double)     :: rmt(9)
  integer(c_int32_t) :: nsite
  type(c_ptr)        :: site_z, site_b, site_norbit, orbit
end type ctem_crystal

interface

  function ctem_ctx_create(device, ctx) bind(C, name='ctem_ctx_create') result(rc)
    import :: c_int32_t, c_ptr
    integer(c_int32_t), value :: device
    type(c_ptr), intent(out)  :: ctx
    integer(c_int32_t)        :: rc
  end function

  function ctem_ctx_destroy(ctx) bind(C, name='ctem_ctx_destroy') result(rc)
    import :: c_int32_t, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int32_t) :: rc
  end function

  function ctem_last_error() bind(C, name='ctem_last_error') result(msg)
    import :: c_ptr
    type(c_ptr) :: msg
  end function

  function ctem_upload_crystal(ctx, crystal) bind(C, name='ctem_upload_crystal') result(rc)
    import :: c_int32_t, c_ptr, ctem_crystal
    type(c_ptr), value             :: ctx
    type(ctem_crystal), intent(in) :: crystal
    integer(c_int32_t)             :: rc
  end function

  function ctem_kvectors(npx, npy, isym, usehex, hemisphere, dsm, mlambda, cap, ij, k, kn, numk) &
           bind(C, name='ctem_kvectors') result(rc)
    import :: c_int32_t, c_double, c_ptr
    integer(c_int32_t), value :: npx, npy, isym, usehex, hemisphere, cap
    real(c_double), intent(in) :: dsm(9)
    real(c_double), value     :: mlambda
    type(c_ptr), value        :: ij, k, kn
    integer(c_int32_t), intent(out) :: numk
    integer(c_int32_t)        :: rc
  end function

  function ctem_laue_equiv(ipx, ipy, isym, iequiv, nequiv) bind(C, name='ctem_laue_equiv') result(rc)
    import :: c_int32_t
    integer(c_int32_t), value :: ipx, ipy, isym
    integer(c_int32_t), intent(out) :: iequiv(2,12), nequiv
    integer(c_int32_t) :: rc
  end function

  function ctem_plan_create(ctx, numk, k, kn, nz, lam, dz, thick, plan) &
           bind(C, name='ctem_plan_create') result(rc)
    import :: c_int32_t, c_double, c_float, c_ptr
    type(c_ptr), value        :: ctx
    integer(c_int32_t), value :: numk, nz
    real(c_double), intent(in) :: k(3,*), kn(*)
    real(c_float), intent(in)  :: lam(*)
    real(c_double), value     :: dz, thick
    type(c_ptr), intent(out)  :: plan
    integer(c_int32_t)        :: rc
  end function

  function ctem_plan_destroy(plan) bind(C, name='ctem_plan_destroy') result(rc)
    import :: c_int32_t, c_ptr
    type(c_ptr), value :: plan
    integer(c_int32_t) :: rc
  end function

  function ctem_plan_select(plan, nns, nnw) bind(C, name='ctem_plan_select') result(rc)
    import :: c_int32_t, c_ptr
    type(c_ptr), value :: plan, nns, nnw
    integer(c_int32_t) :: rc
  end function

  function ctem_plan_run(plan, precision, keep_lgh_ik) bind(C, name='ctem_plan_run') result(rc)
    import :: c_int32_t, c_ptr
    type(c_ptr), value        :: plan
    integer(c_int32_t), value :: precision, keep_lgh_ik
    integer(c_int32_t)        :: rc
  end function

  function ctem_plan_get_svals(plan, svals) bind(C, name='ctem_plan_get_svals') result(rc)
    import :: c_int32_t, c_float, c_ptr
    type(c_ptr), value :: plan
    real(c_float), intent(out) :: svals(*)
    integer(c_int32_t) :: rc
  end function

  function ctem_plan_get_dots(plan, dots) bind(C, name='ctem_plan_get_dots') result(rc)
    import :: c_int32_t, c_double, c_ptr
    type(c_ptr), value :: plan
    real(c_double), intent(out) :: dots(*)
    integer(c_int32_t) :: rc
  end function

  function ctem_plan_scatter(plan, ij, npx, npy, isym, sr) bind(C, name='ctem_plan_scatter') result(rc)
    import :: c_int32_t, c_float, c_ptr
    type(c_ptr), value        :: plan
    integer(c_int32_t), intent(in) :: ij(2,*)
    integer(c_int32_t), value :: npx, npy, isym
    real(c_float), intent(inout) :: sr(*)
    integer(c_int32_t)        :: rc
  end function

  function ctem_plan_set_mixed_terms(plan, taylor_terms, square_terms) &
           bind(C, name='ctem_plan_set_mixed_terms') result(rc)
    import :: c_int32_t, c_ptr
    type(c_ptr), value        :: plan
    integer(c_int32_t), value :: taylor_terms, square_terms
    integer(c_int32_t)        :: rc
  end function

  ! mode 1: symmetric products of the split mixed form for an inversion-symmetric LUT, 0: off;
  ! in_effect (may be c_null_ptr) receives whether they are used
  function ctem_plan_set_symmetry(plan, mode, in_effect) &
           bind(C, name='ctem_plan_set_symmetry') result(rc)
    import :: c_int32_t, c_ptr
    type(c_ptr), value        :: plan
    integer(c_int32_t), value :: mode
    type(c_ptr), value        :: in_effect
    integer(c_int32_t)        :: rc
  end function

  ! reverse mode: gdots(p, ik) = df/ddots -> glut = df/dLUT (complex, LUT order), gb(p) = df/dB_p
  function ctem_plan_adjoint(plan, gdots, glut, gb) bind(C, name='ctem_plan_adjoint') result(rc)
    import :: c_int32_t, c_double, c_double_complex, c_ptr
    type(c_ptr), value :: plan
    real(c_double), intent(in) :: gdots(*)
    complex(c_double_complex), intent(out) :: glut(*)
    real(c_double), intent(out) :: gb(*)
    integer(c_int32_t) :: rc
  end function

  ! the same gradient with the mixed forward and split-FP16 products in the exponential's derivative
  function ctem_plan_adjoint_mixed(plan, gdots, glut, gb) bind(C, name='ctem_plan_adjoint_mixed') result(rc)
    import :: c_int32_t, c_double, c_double_complex, c_ptr
    type(c_ptr), value :: plan
    real(c_double), intent(in) :: gdots(*)
    complex(c_double_complex), intent(out) :: glut(*)
    real(c_double), intent(out) :: gb(*)
    integer(c_int32_t) :: rc
  end function

  ! same argument list and meaning as the reference's external CalcLgh3
  ! (CTEMEBSDmaster.f90:610-630), with the context first and no kin/debug
  function ctem_calclgh3(ctx, dmat, lgh, thick, kn, nn, gzero, depthstep, lambdae, izz) &
           bind(C, name='ctem_calclgh3') result(rc)
    import :: c_int32_t, c_double, c_double_complex, c_float, c_ptr
    type(c_ptr), value        :: ctx
    complex(c_double_complex), intent(in)  :: dmat(*)
    complex(c_double_complex), intent(out) :: lgh(*)
    real(c_double), value     :: thick, kn, depthstep
    integer(c_int32_t), value :: nn, gzero, izz
    real(c_float), intent(in) :: lambdae(*)
    integer(c_int32_t)        :: rc
  end function

  ! ctem_calclgh3 without device allocations: work space in a fixed per-GPU array (nn <= 256,
  ! izz <= 256); concurrent calls on one GPU race (no counterpart in the reference)
  function ctem_calclgh3_static(ctx, dmat, lgh, thick, kn, nn, gzero, depthstep, lambdae, izz) &
           bind(C, name='ctem_calclgh3_static') result(rc)
    import :: c_int32_t, c_double, c_double_complex, c_float, c_ptr
    type(c_ptr), value        :: ctx
    complex(c_double_complex), intent(in)  :: dmat(*)
    complex(c_double_complex), intent(out) :: lgh(*)
    real(c_double), value     :: thick, kn, depthstep
    integer(c_int32_t), value :: nn, gzero, izz
    real(c_float), intent(in) :: lambdae(*)
    integer(c_int32_t)        :: rc
  end function

end interface

contains

! the library's last error message as a Fortran string
function ctem_error_message() result(msg)
  character(len=:), allocatable :: msg
  type(c_ptr) :: p
  character(kind=c_char), pointer :: s(:)
  integer :: n
  interface
    function c_strlen(str) bind(C, name='strlen') result(l)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: str
      integer(c_size_t) :: l
    end function
  end interface
  p = ctem_last_error()
  n = int(c_strlen(p))
  call c_f_pointer(p, s, [n])
  allocate(character(len=n) :: msg)
  msg = transfer(s(1:n), msg)
end function ctem_error_message

end module ctem_ebsd_iface
