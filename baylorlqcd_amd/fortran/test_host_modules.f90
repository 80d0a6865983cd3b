! test_host_modules.f90 -- the three host modules the qqcd_hip shim reads, as a
! minimal test host defines them: latdims (lattice extents, compile-time as
! in the reference's qqcd-test/user/latdims.f90), gmresrhs (the isignal
! first/later right-hand-side flag of qqcd-test/cfgsprops/quark/gmresrhs.f90)
! and lagfib (the random-number state of qqcd-test/cfgsprops/glue/lagfib.f90;
! this host seeds it through the library's initlagfib).
module latdims
   implicit none
   integer, parameter :: nx = 8, ny = 8, nz = 8, nt = 16
end module latdims

module gmresrhs
   implicit none
   integer :: isignal = 1
end module gmresrhs

module lagfib
   use iso_c_binding
   implicit none
   integer(c_int), dimension(0:520), save :: iwlagfib
   integer(c_int), save :: jrlagfib, krlagfib
   interface
      integer(c_int) function c_lagfib_init(myid, iseed, iw, jr, kr) bind(C, name="qqcd_lagfib_init")
         import :: c_int
         integer(c_int), value :: myid, iseed
         integer(c_int), intent(out) :: iw(521), jr, kr
      end function
   end interface
contains
   subroutine initlagfib(myid, iseed)
      integer, intent(in) :: myid, iseed
      if (c_lagfib_init(myid, iseed, iwlagfib, jrlagfib, krlagfib) /= 0) stop "initlagfib"
   end subroutine initlagfib
end module lagfib
