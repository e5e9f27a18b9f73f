! TEST INFRASTRUCTURE ONLY. A test double of the reference's `noise` module for an
! instrumented build of CTEMEBSD.f90 (oracle/Makefile, target _ref/CTEMEBSD_clean): POIDEV
! returns its argument, so the program writes its patterns before the Poisson step
! (CTEMEBSD.f90:560-565) and the noise-free values can be compared directly.
module noise
use local
IMPLICIT NONE
contains
real(kind=sgl) function POIDEV(XM, IDUM)
real(kind=sgl), INTENT(IN) :: XM
integer, INTENT(INOUT) :: IDUM
POIDEV = XM
end function POIDEV
end module noise
