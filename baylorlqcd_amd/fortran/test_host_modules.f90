! test_host_modules.f90 -- the two host modules the qqcd_hip shim reads, as a
! minimal test host defines them: latdims (lattice extents, compile-time as
! in the reference's qqcd-test/user/latdims.f90) and gmresrhs (the isignal
! first/later right-hand-side flag of qqcd-test/cfgsprops/quark/gmresrhs.f90).
module latdims
   implicit none
   integer, parameter :: nx = 8, ny = 8, nz = 8, nt = 16
end module latdims

module gmresrhs
   implicit none
   integer :: isignal = 1
end module gmresrhs
