! qqcd_hip.f90 -- Fortran drop-in for the Dirac-operator and inverter layer.
!
! The procedures below carry the names and dummy-argument lists of the
! reference's module procedures (BaylorLQCD "qqcd"):
!   Hsingle, Hdouble, Hdbletm, gammamult, smear, twMshift, gamma5mult
!                                         qqcd-test/cfgsprops/quark/diracops.f90
!   gam5x                                 qqcd-test/cfgsprops/basics.f90
!   bicgstab, gmresdr, psc, gmresproject, vecdot
!                                         qqcd-test/cfgsprops/quark/inverters.f90
!   pointsource, fermprop                 qqcd-test/cfgsprops/quark/quark.f90
!   z4source, z2source, loopops, nsaverage
!                                         qqcd-test/cfgsprops/quark/disconloops.f90
!   GFwrite, GFread                       qqcd-test/cfgsprops/glue/gaugetools.f90
!   sweep                                 qqcd-test/cfgsprops/cfgsprops.f90
!   FPwrite, FPread                       qqcd-test/cfgsprops/quark/quark.f90
! and forward to the HIP library through its C ABI (include/qqcd_hip.h).  A
! host program replaces `use diracops` / `use inverters` (and the two quark
! procedures) with `use qqcd_hip`; everything else in the host is unchanged.
!
! The lattice extents come from the host's own latdims module (nx, ny, nz,
! nt) and the random stream from its lagfib module (iwlagfib, jrlagfib,
! krlagfib); the library starts on the first call.  Only one process is
! supported: the geometry arguments (vecbl, vecblinv, myid, nn, ldiv, nms,
! lvbc, ib, lbd, iblv) are checked against the single-process tables and
! otherwise unused.  Errors are reported as the reference reports them: a
! <MODULE>.ERROR file with the message, then stop
! (qqcd-test/cfgsprops/quark/diracops.f90:1077-1082).
!
! The link array is uploaded when an operator call first passes it and kept
! on the device; the drop-ins that change it in place (sweep, GFread,
! GFwrite's fixsu3) make the library forget that copy, so the next call
! uploads it again.  A host that changes u in place by its own code calls
! qqcd_gauge_changed(u, coact, bc); an array at a new address is picked up
! without it.

module qqcd_hip
   use iso_c_binding
   use latdims, only: nx, ny, nz, nt
   use lagfib, only: iwlagfib, jrlagfib, krlagfib
   implicit none
   private

   public :: Hsingle, Hdouble, Hdbletm, gammamult, gam5x, vecdot, bicgstab, gmresdr, psc, &
             gmresproject, pointsource, fermprop, fermpropR, qqcd_gauge_changed, qqcd_shutdown, &
             qqcd_solve_mixed_f, z4source, z2source, loopops, loopops_all, nsaverage, &
             GFwrite, GFread, FPwrite, FPread, sweep, smear, twMshift, gamma5mult

   logical, save :: started = .false.

   interface
      integer(c_int) function c_init(n, bc, device) bind(C, name="qqcd_init")
         import :: c_int
         integer(c_int), intent(in) :: n(4), bc(4)
         integer(c_int), value :: device
      end function
      integer(c_int) function c_finalize() bind(C, name="qqcd_finalize")
         import :: c_int
      end function
      type(c_ptr) function c_last_error() bind(C, name="qqcd_last_error")
         import :: c_ptr
      end function
      integer(c_int) function c_gauge_load(u, coact) bind(C, name="qqcd_gauge_load")
         import :: c_int, c_ptr
         type(c_ptr), value :: u, coact
      end function
      integer(c_int) function c_hsingle(gout, u, gin, idag, coact, bc, gblclr) bind(C, name="qqcd_hsingle")
         import :: c_int, c_ptr
         type(c_ptr), value :: gout, u, gin, coact, bc
         integer(c_int), value :: idag, gblclr
      end function
      integer(c_int) function c_hdouble(he, u, ge, idag, coact, kappa, bc) bind(C, name="qqcd_hdouble")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: he, u, ge, coact, bc
         integer(c_int), value :: idag
         real(c_double), value :: kappa
      end function
      integer(c_int) function c_hdbletm(he, u, ge, idag, coact, kappa, iflag, bc) bind(C, name="qqcd_hdbletm")
         import :: c_int, c_ptr
         type(c_ptr), value :: he, u, ge, coact, kappa, bc
         integer(c_int), value :: idag, iflag
      end function
      integer(c_int) function c_gammamult(ei, oi, eo, oo, mu, ieo, ibleo) bind(C, name="qqcd_gammamult")
         import :: c_int, c_ptr
         type(c_ptr), value :: ei, oi, eo, oo
         integer(c_int), value :: mu, ieo, ibleo
      end function
      integer(c_int) function c_gam5x(a) bind(C, name="qqcd_gam5x")
         import :: c_int, c_ptr
         type(c_ptr), value :: a
      end function
      integer(c_int) function c_vecdot(a, b, adb) bind(C, name="qqcd_vecdot")
         import :: c_int, c_ptr
         type(c_ptr), value :: a, b, adb
      end function
      integer(c_int) function c_pointsource(be, bo, src, icsrc, idsrc) bind(C, name="qqcd_pointsource")
         import :: c_int, c_ptr
         type(c_ptr), value :: be, bo, src
         integer(c_int), value :: icsrc, idsrc
      end function
      integer(c_int) function c_gmresdr(phi, x, gm, resmax, itermin, itercount, u, coact, kappa, iflag, bc) &
            bind(C, name="qqcd_gmresdr")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: phi, x, gm, itercount, u, coact, kappa, bc
         real(c_double), value :: resmax
         integer(c_int), value :: itermin, iflag
      end function
      integer(c_int) function c_gmresproject(b, xshift, nshift, gm, resmax, itermin, itercount, u, coact, kappa, &
                                             iflag, bc, isignal, mvp, gdr, ldgdr) bind(C, name="qqcd_gmresproject")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: b, xshift, gm, itercount, u, coact, kappa, bc, mvp, gdr
         real(c_double), value :: resmax
         integer(c_int), value :: nshift, itermin, iflag, isignal, ldgdr
      end function
      integer(c_int) function c_psc(b, x, gm, resmax, itermin, itercount, u, coact, kappa, iflag, bc, &
                                    isignal, resproj) bind(C, name="qqcd_psc")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: b, x, gm, itercount, u, coact, kappa, bc
         real(c_double), value :: resmax, resproj
         integer(c_int), value :: itermin, iflag, isignal
      end function
      integer(c_int) function c_bicgstab(phi, x, resmax, itermin, itercount, u, coact, kappa, iflag, bc) &
            bind(C, name="qqcd_bicgstab")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: phi, x, itercount, u, coact, kappa, bc
         real(c_double), value :: resmax
         integer(c_int), value :: itermin, iflag
      end function
      integer(c_int) function c_fermprop(be, bo, kappa, inverter, coact, bc, resmax, itermin, ngm, vectype, u, &
                                         xe, xo, isignal, itercount) bind(C, name="qqcd_fermprop")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: be, bo, kappa, coact, bc, ngm, u, xe, xo, itercount
         real(c_double), value :: resmax
         integer(c_int), value :: inverter, itermin, vectype, isignal
      end function
      integer(c_int) function c_solve_mixed(phi, x, gm, tol, basis, opp, recon, itol, keep, u, coact, kappa, &
                                            bc, itercount) bind(C, name="qqcd_solve_mixed")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: phi, x, gm, u, coact, kappa, bc, itercount
         real(c_double), value :: tol, itol
         integer(c_int), value :: basis, opp, recon, keep
      end function
      integer(c_int) function c_z4source_lagfib(be, bo, iw, jr, kr) bind(C, name="qqcd_z4source_lagfib")
         import :: c_int, c_ptr
         type(c_ptr), value :: be, bo
         integer(c_int), intent(inout) :: iw(521), jr, kr
      end function
      integer(c_int) function c_z2source_lagfib(be, bo, iw, jr, kr) bind(C, name="qqcd_z2source_lagfib")
         import :: c_int, c_ptr
         type(c_ptr), value :: be, bo
         integer(c_int), intent(inout) :: iw(521), jr, kr
      end function
      integer(c_int) function c_loopops(ksub, xe, xo, u, be, bo, bc, Je, Jo) bind(C, name="qqcd_loopops")
         import :: c_int, c_ptr
         integer(c_int), value :: ksub
         type(c_ptr), value :: xe, xo, u, be, bo, bc, Je, Jo
      end function
      integer(c_int) function c_loopsum(ksub, xe, xo, u, be, bo, bc, io, cur, Jtime) bind(C, name="qqcd_loopsum")
         import :: c_int, c_ptr, c_double
         integer(c_int), value :: ksub
         type(c_ptr), value :: xe, xo, u, be, bo, bc, io, Jtime
         real(c_double), value :: cur
      end function
      integer(c_int) function c_smear(he, ho, u, ge, go, asmear, nsmear, bc) bind(C, name="qqcd_smear")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: he, ho, u, ge, go, bc
         real(c_double), value :: asmear
         integer(c_int), value :: nsmear
      end function
      integer(c_int) function c_twmshift(hwhole, u, gwhole, idag, coact, kappa, iflag, bc) bind(C, name="qqcd_twmshift")
         import :: c_int, c_ptr
         type(c_ptr), value :: hwhole, u, gwhole, coact, kappa, bc
         integer(c_int), value :: idag, iflag
      end function
      integer(c_int) function c_gamma5mult(he, u, ge, idag, coact, kappa, iflag, bc) bind(C, name="qqcd_gamma5mult")
         import :: c_int, c_ptr
         type(c_ptr), value :: he, u, ge, coact, kappa, bc
         integer(c_int), value :: idag, iflag
      end function
      integer(c_int) function c_solve_log(cap, n, labels, counts, values) bind(C, name="qqcd_solve_log")
         import :: c_int, c_double, c_char
         integer(c_int), value :: cap
         integer(c_int), intent(out) :: n, counts(*)
         character(kind=c_char), intent(out) :: labels(*)
         real(c_double), intent(out) :: values(*)
      end function
      integer(c_int) function c_sweep(gaction, coact, nhit, nsweep, u, iw, jr, kr) bind(C, name="qqcd_sweep")
         import :: c_int, c_ptr
         integer(c_int), intent(in) :: gaction(4)
         type(c_ptr), value :: coact, u
         integer(c_int), value :: nhit, nsweep
         integer(c_int), intent(inout) :: iw(521), jr, kr
      end function
      integer(c_int) function c_gfwrite(n, u, filename, gaction, beta, alat, tad, icfgsave, myid) &
            bind(C, name="qqcd_gfwrite")
         import :: c_int, c_ptr, c_double, c_char
         integer(c_int), intent(in) :: n(4), gaction(4)
         type(c_ptr), value :: u
         character(kind=c_char), intent(in) :: filename(*)
         real(c_double), value :: beta
         real(c_double), intent(in) :: alat(4), tad(4)
         integer(c_int), value :: icfgsave, myid
      end function
      integer(c_int) function c_gfread(n, u, filename, gaction, beta, alat, tad, icfgsave, myidin) &
            bind(C, name="qqcd_gfread")
         import :: c_int, c_ptr, c_double, c_char
         integer(c_int), intent(in) :: n(4)
         type(c_ptr), value :: u
         character(kind=c_char), intent(in) :: filename(*)
         integer(c_int), intent(out) :: gaction(4), icfgsave, myidin
         real(c_double), intent(out) :: beta, alat(4), tad(4)
      end function
      integer(c_int) function c_fpwrite(n, xe, xo, filename, kappa, tmqcd, cSW, icsrc, idsrc, icfg, myid) &
            bind(C, name="qqcd_fpwrite")
         import :: c_int, c_ptr, c_double, c_char
         integer(c_int), intent(in) :: n(4)
         type(c_ptr), value :: xe, xo
         character(kind=c_char), intent(in) :: filename(*)
         real(c_double), value :: kappa, tmqcd, cSW
         integer(c_int), value :: icsrc, idsrc, icfg, myid
      end function
      integer(c_int) function c_fpread(n, xe, xo, filename, kappa, tmqcd, cSW, icsrc, idsrc, icfgin, myidin) &
            bind(C, name="qqcd_fpread")
         import :: c_int, c_ptr, c_double, c_char
         integer(c_int), intent(in) :: n(4)
         type(c_ptr), value :: xe, xo
         character(kind=c_char), intent(in) :: filename(*)
         real(c_double), intent(out) :: kappa, tmqcd, cSW
         integer(c_int), intent(out) :: icsrc, idsrc, icfgin, myidin
      end function
   end interface

contains

   subroutine fail(where, rc)
      character(len=*), intent(in) :: where
      integer(c_int), intent(in) :: rc
      character(kind=c_char), pointer :: msg(:)
      integer :: k, unit
      character(len=1024) :: text
      if (rc == 0) return
      text = ' '
      call c_f_pointer(c_last_error(), msg, [1024])
      do k = 1, 1024
         if (msg(k) == c_null_char) exit
         text(k:k) = msg(k)
      end do
      if (len_trim(text) == 0) write(text, "(a,i0)") "qqcd_hip: call failed with status ", rc
      open(newunit=unit, file=where//".ERROR", action="write", status="replace", form="formatted")
      write(unit, fmt=*) trim(text)
      close(unit, status="keep")
      stop
   end subroutine fail

   subroutine ensure_started(bc)
      integer, intent(in) :: bc(:)
      integer(c_int), target :: n(4), b(4)
      if (started) return
      n = [nx, ny, nz, nt]
      b = bc(1:4)
      call fail("QQCD_HIP", c_init(n, b, 0_c_int))
      started = .true.
   end subroutine ensure_started

   subroutine check_geometry(where, myid, ldiv, vecbl, vecblinv, iblv, lbd)
      character(len=*), intent(in) :: where
      integer, intent(in) :: myid
      logical, intent(in) :: ldiv(:)
      integer, intent(in), optional :: vecbl(:,:), vecblinv(:,:), iblv(:,:)
      logical, intent(in), optional :: lbd(:,:)
      integer :: ibl, mu, unit
      logical :: ok
      ok = myid == 0 .and. .not. any(ldiv)
      if (present(vecbl)) ok = ok .and. all(vecbl(1,1:8) == [1, 4, 6, 7, 10, 11, 13, 16]) &
                                  .and. all(vecbl(2,1:8) == [2, 3, 5, 8, 9, 12, 14, 15])
      if (present(vecblinv)) then
         do ibl = 1, 16
            ok = ok .and. vecblinv(2,ibl) == (ibl + 1)/2
         end do
      end if
      if (present(iblv) .and. present(lbd)) then
         do ibl = 1, 16
            do mu = 1, 4
               ok = ok .and. iblv(ibl,mu) == 1 + ieor(ibl - 1, ishft(1, mu - 1)) &
                       .and. (lbd(ibl,mu) .eqv. iand(ishft(ibl - 1, -(mu - 1)), 1) == 0)
            end do
         end do
      end if
      if (.not. ok) then
         open(newunit=unit, file=where//".ERROR", action="write", status="replace", form="formatted")
         write(unit, fmt=*) "qqcd_hip supports one process with the canonical block tables only"
         close(unit, status="keep")
         stop
      end if
   end subroutine check_geometry

   ! Append the last solve's log lines to <rwdir>CFGSPROPS.LOG as the reference's solvers
   ! write them (format (a12,i9,es17.10); qqcd-test/cfgsprops/quark/inverters.f90:3285-3291,
   ! :476-480), from process 0.
   subroutine append_log(rwdir, myid)
      character(len=*), intent(in) :: rwdir(:)
      integer, intent(in) :: myid
      integer, parameter :: cap = 100000
      integer(c_int) :: n, k
      integer(c_int), allocatable :: counts(:)
      real(c_double), allocatable :: values(:)
      character(kind=c_char), allocatable :: labels(:)
      character(len=12) :: lab
      integer :: unit, j
      if (myid /= 0) return
      allocate(counts(cap), values(cap), labels(12*cap))
      call fail("INVERTERS", c_solve_log(int(cap, c_int), n, labels, counts, values))
      if (n == 0) return
      open(newunit=unit, file=trim(rwdir(myid+1))//"CFGSPROPS.LOG", action="write", form="formatted", &
           status="unknown", position="append")
      do k = 1, min(n, cap)
         do j = 1, 12
            lab(j:j) = labels(12*(k-1)+j)
         end do
         write(unit=unit, fmt="(a12,i9,es17.10)") trim(lab), counts(k), values(k)
      end do
      close(unit, status="keep")
   end subroutine append_log

   ! Refresh the resident links after the host changed u in place.
   subroutine qqcd_gauge_changed(u, coact, bc)
      real(kind=8), intent(in), target, contiguous :: u(:,:,:,:,:)
      real(kind=8), intent(in), target, contiguous :: coact(:,:,:)
      integer, intent(in) :: bc(:)
      call ensure_started(bc)
      call fail("QQCD_HIP", c_gauge_load(c_loc(u), c_loc(coact)))
   end subroutine qqcd_gauge_changed

   subroutine qqcd_shutdown()
      if (started) call fail("QQCD_HIP", c_finalize())
      started = .false.
   end subroutine qqcd_shutdown

   ! ---- diracops ------------------------------------------------------

   subroutine Hsingle(gout,u,gin,idag,coact,bc,gblclr,vecbl,vecblinv,myid,nn,ldiv,nms,lvbc,ib,lbd,iblv,MRT)
      real(kind=8), intent(out), target, contiguous :: gout(:,:,:,:,:)
      real(kind=8), intent(in), target, contiguous :: u(:,:,:,:,:)
      real(kind=8), intent(inout), target, contiguous :: gin(:,:,:,:,:)
      integer, intent(in) :: idag, gblclr, myid, MRT
      real(kind=8), intent(in), target, contiguous :: coact(:,:,:)
      integer, intent(in), target :: bc(:)
      integer, intent(in) :: nms(:), vecbl(:,:), vecblinv(:,:), nn(:,:), iblv(:,:), lvbc(:,:,:), ib(:,:,:,:)
      logical, intent(in) :: ldiv(:), lbd(:,:)
      integer(c_int), target :: b(4)
      call check_geometry("DIRACOPS", myid, ldiv, vecbl, vecblinv, iblv, lbd)
      call ensure_started(bc)
      b = bc(1:4)
      call fail("DIRACOPS", c_hsingle(c_loc(gout), c_loc(u), c_loc(gin), idag, c_loc(coact), c_loc(b), gblclr))
   end subroutine Hsingle

   subroutine Hdouble(he,u,ge,idag,coact,kappa,bc,vecbl,vecblinv,myid,nn,ldiv,nms,lvbc,ib,lbd,iblv,MRT)
      real(kind=8), intent(out), target, contiguous :: he(:,:,:,:,:)
      real(kind=8), intent(in), target, contiguous :: u(:,:,:,:,:)
      real(kind=8), intent(inout), target, contiguous :: ge(:,:,:,:,:)
      integer, intent(in) :: idag, myid, MRT
      real(kind=8), intent(in), target, contiguous :: coact(:,:,:)
      real(kind=8), intent(in) :: kappa
      integer, intent(in) :: bc(:), nms(:), vecbl(:,:), vecblinv(:,:), nn(:,:), iblv(:,:), lvbc(:,:,:), ib(:,:,:,:)
      logical, intent(in) :: ldiv(:), lbd(:,:)
      integer(c_int), target :: b(4)
      call check_geometry("DIRACOPS", myid, ldiv, vecbl, vecblinv, iblv, lbd)
      call ensure_started(bc)
      b = bc(1:4)
      call fail("DIRACOPS", c_hdouble(c_loc(he), c_loc(u), c_loc(ge), idag, c_loc(coact), kappa, c_loc(b)))
   end subroutine Hdouble

   subroutine Hdbletm(he,u,GeeGooinv,ge,idag,coact,kappa,iflag,bc,vecbl,vecblinv,myid,nn,ldiv,nms,lvbc,ib,lbd, &
                      iblv,MRT)
      real(kind=8), intent(out), target, contiguous :: he(:,:,:,:,:)
      real(kind=8), intent(in), target, contiguous :: u(:,:,:,:,:)
      real(kind=8), intent(in) :: GeeGooinv(:,:,:,:,:)
      real(kind=8), intent(inout), target, contiguous :: ge(:,:,:,:,:)
      integer, intent(in) :: idag, iflag, myid, MRT
      real(kind=8), intent(in), target, contiguous :: coact(:,:,:)
      real(kind=8), intent(in), target, contiguous :: kappa(:)
      integer, intent(in) :: bc(:), nms(:), vecbl(:,:), vecblinv(:,:), nn(:,:), iblv(:,:), lvbc(:,:,:), ib(:,:,:,:)
      logical, intent(in) :: ldiv(:), lbd(:,:)
      integer(c_int), target :: b(4)
      call check_geometry("DIRACOPS", myid, ldiv, vecbl, vecblinv, iblv, lbd)
      call ensure_started(bc)
      b = bc(1:4)
      call fail("DIRACOPS", c_hdbletm(c_loc(he), c_loc(u), c_loc(ge), idag, c_loc(coact), c_loc(kappa), &
                                      iflag, c_loc(b)))
   end subroutine Hdbletm

   subroutine gammamult(evenIN, oddIN, evenOUT, oddOUT, mu, ieo, ibleo)
      real(kind=8), intent(in), target, contiguous :: evenIN(:,:,:,:,:), oddIN(:,:,:,:,:)
      real(kind=8), intent(inout), target, contiguous :: evenOUT(:,:,:,:,:), oddOUT(:,:,:,:,:)
      integer, intent(in) :: mu, ieo, ibleo
      if (.not. started) call ensure_started([1, 1, 1, 1])
      call fail("DIRACOPS", c_gammamult(c_loc(evenIN), c_loc(oddIN), c_loc(evenOUT), c_loc(oddOUT), mu, ieo, ibleo))
   end subroutine gammamult

   ! smear(he,ho,u,ge,go,asmear,nsmear,bc,...), diracops.f90:712: Gaussian smearing of (ge, go)
   subroutine smear(he,ho,u,ge,go,asmear,nsmear,bc,vecbl,vecblinv,myid,nn,ldiv,nms,lvbc,ib,lbd,iblv,MRT)
      real(kind=8), intent(out), target, contiguous :: he(:,:,:,:,:), ho(:,:,:,:,:)
      real(kind=8), intent(in), target, contiguous :: u(:,:,:,:,:), ge(:,:,:,:,:), go(:,:,:,:,:)
      real(kind=8), intent(in) :: asmear
      integer, intent(in) :: nsmear, myid, MRT
      integer, intent(in) :: bc(:), nms(:), vecbl(:,:), vecblinv(:,:), nn(:,:), iblv(:,:), lvbc(:,:,:), ib(:,:,:,:)
      logical, intent(in) :: ldiv(:), lbd(:,:)
      integer(c_int), target :: b(4)
      call check_geometry("DIRACOPS", myid, ldiv, vecbl, vecblinv, iblv, lbd)
      call ensure_started(bc)
      b = bc(1:4)
      call fail("DIRACOPS", c_smear(c_loc(he), c_loc(ho), c_loc(u), c_loc(ge), c_loc(go), asmear, &
                                    int(nsmear, c_int), c_loc(b)))
   end subroutine smear

   ! twMshift(hwhole,u,GeeGooinv,gwhole,idag,coact,kappa,iflag,bc,...), diracops.f90:1863:
   ! hwhole, gwhole(6,ntotal,4,2,8,2); iflag = -1 only, as the reference
   subroutine twMshift(hwhole,u,GeeGooinv,gwhole,idag,coact,kappa,iflag,bc,vecbl,vecblinv,myid,nn,ldiv,nms, &
                       lvbc,ib,lbd,iblv,MRT)
      real(kind=8), intent(out), target, contiguous :: hwhole(:,:,:,:,:,:)
      real(kind=8), intent(inout), target, contiguous :: gwhole(:,:,:,:,:,:)
      real(kind=8), intent(in), target, contiguous :: u(:,:,:,:,:), coact(:,:,:), kappa(:)
      real(kind=8), intent(in) :: GeeGooinv(:,:,:,:,:)
      integer, intent(in) :: idag, iflag, myid, MRT
      integer, intent(in) :: bc(:), nms(:), vecbl(:,:), vecblinv(:,:), nn(:,:), iblv(:,:), lvbc(:,:,:), ib(:,:,:,:)
      logical, intent(in) :: ldiv(:), lbd(:,:)
      integer(c_int), target :: b(4)
      call check_geometry("DIRACOPS", myid, ldiv, vecbl, vecblinv, iblv, lbd)
      if (size(hwhole, 6) /= 2 .or. size(gwhole, 6) /= 2) call fail("DIRACOPS", -1_c_int)
      call ensure_started(bc)
      b = bc(1:4)
      call fail("DIRACOPS", c_twmshift(c_loc(hwhole), c_loc(u), c_loc(gwhole), int(idag, c_int), c_loc(coact), &
                                       c_loc(kappa), int(iflag, c_int), c_loc(b)))
   end subroutine twMshift

   ! gamma5mult(he,u,GeeGooinv,ge,idag,coact,kappa,iflag,bc,...), diracops.f90:1750
   subroutine gamma5mult(he,u,GeeGooinv,ge,idag,coact,kappa,iflag,bc,vecbl,vecblinv,myid,nn,ldiv,nms,lvbc,ib, &
                         lbd,iblv,MRT)
      real(kind=8), intent(out), target, contiguous :: he(:,:,:,:,:)
      real(kind=8), intent(in), target, contiguous :: u(:,:,:,:,:), ge(:,:,:,:,:), coact(:,:,:), kappa(:)
      real(kind=8), intent(in) :: GeeGooinv(:,:,:,:,:)
      integer, intent(in) :: idag, iflag, myid, MRT
      integer, intent(in) :: bc(:), nms(:), vecbl(:,:), vecblinv(:,:), nn(:,:), iblv(:,:), lvbc(:,:,:), ib(:,:,:,:)
      logical, intent(in) :: ldiv(:), lbd(:,:)
      integer(c_int), target :: b(4)
      call check_geometry("DIRACOPS", myid, ldiv, vecbl, vecblinv, iblv, lbd)
      call ensure_started(bc)
      b = bc(1:4)
      call fail("DIRACOPS", c_gamma5mult(c_loc(he), c_loc(u), c_loc(ge), int(idag, c_int), c_loc(coact), &
                                         c_loc(kappa), int(iflag, c_int), c_loc(b)))
   end subroutine gamma5mult

   subroutine gam5x(a)
      real(kind=8), intent(inout), target, contiguous :: a(:,:,:,:,:)
      if (.not. started) call ensure_started([1, 1, 1, 1])
      call fail("BASICS", c_gam5x(c_loc(a)))
   end subroutine gam5x

   ! ---- inverters ------------------------------------------------------

   subroutine vecdot(a, b, adb, MRT2)
      real(kind=8), intent(in), target, contiguous :: a(:,:,:,:,:), b(:,:,:,:,:)
      real(kind=8), intent(out), target :: adb(2)
      integer, intent(in) :: MRT2
      if (.not. started) call ensure_started([1, 1, 1, 1])
      call fail("INVERTERS", c_vecdot(c_loc(a), c_loc(b), c_loc(adb)))
   end subroutine vecdot

   subroutine bicgstab(rwdir,phi,x,resmax,itermin,itercount,u,GeeGooinv,iflag,kappa,coact,bc,vecbl,vecblinv, &
                       myid,nn,ldiv,nms,lvbc,ib,lbd,iblv,MRT,MRT2)
      character(len=*), intent(in) :: rwdir(:)
      real(kind=8), intent(in), target, contiguous :: phi(:,:,:,:,:)
      real(kind=8), intent(inout), target, contiguous :: x(:,:,:,:,:)
      real(kind=8), intent(in) :: resmax
      integer, intent(in) :: itermin, iflag, myid, MRT, MRT2
      integer, intent(out), target :: itercount
      real(kind=8), intent(in), target, contiguous :: u(:,:,:,:,:), kappa(:), coact(:,:,:)
      real(kind=8), intent(in) :: GeeGooinv(:,:,:,:,:)
      integer, intent(in) :: bc(:), nms(:), vecbl(:,:), vecblinv(:,:), nn(:,:), iblv(:,:), lvbc(:,:,:), ib(:,:,:,:)
      logical, intent(in) :: ldiv(:), lbd(:,:)
      integer(c_int), target :: b(4)
      call check_geometry("INVERTERS", myid, ldiv, vecbl, vecblinv, iblv, lbd)
      call ensure_started(bc)
      b = bc(1:4)
      call fail("INVERTERS", c_bicgstab(c_loc(phi), c_loc(x), resmax, itermin, c_loc(itercount), c_loc(u), &
                                        c_loc(coact), c_loc(kappa), iflag, c_loc(b)))
      call append_log(rwdir, myid)
   end subroutine bicgstab

   subroutine gmresdr(rwdir,phi,x,GMRES,resmax,itermin,itercount,u,GeeGooinv,iflag,kappa,coact,bc,vecbl, &
                      vecblinv,myid,nn,ldiv,nms,lvbc,ib,lbd,iblv,MRT,MRT2)
      character(len=*), intent(in) :: rwdir(:)
      real(kind=8), intent(in), target, contiguous :: phi(:,:,:,:,:)
      real(kind=8), intent(inout), target, contiguous :: x(:,:,:,:,:)
      integer, intent(in), target :: GMRES(:)
      real(kind=8), intent(in) :: resmax
      integer, intent(in) :: itermin, iflag, myid, MRT, MRT2
      integer, intent(out), target :: itercount
      real(kind=8), intent(in), target, contiguous :: u(:,:,:,:,:), kappa(:), coact(:,:,:)
      real(kind=8), intent(in) :: GeeGooinv(:,:,:,:,:)
      integer, intent(in) :: bc(:), nms(:), vecbl(:,:), vecblinv(:,:), nn(:,:), iblv(:,:), lvbc(:,:,:), ib(:,:,:,:)
      logical, intent(in) :: ldiv(:), lbd(:,:)
      integer(c_int), target :: b(4), gm(2)
      call check_geometry("INVERTERS", myid, ldiv, vecbl, vecblinv, iblv, lbd)
      call ensure_started(bc)
      b = bc(1:4)
      gm = GMRES(1:2)
      call fail("INVERTERS", c_gmresdr(c_loc(phi), c_loc(x), c_loc(gm), resmax, itermin, c_loc(itercount), &
                                       c_loc(u), c_loc(coact), c_loc(kappa), iflag, c_loc(b)))
      call append_log(rwdir, myid)
   end subroutine gmresdr

   subroutine psc(rwdir,b,x,GMRES,resmax,itermin,itercount,u,GeeGooinv,iflag,kappa,coact,bc,vecbl, &
                  vecblinv,myid,nn,ldiv,nms,lvbc,ib,lbd,iblv,MRT,MRT2)
      use gmresrhs, only: isignal
      character(len=*), intent(in) :: rwdir(:)
      real(kind=8), intent(in), target, contiguous :: b(:,:,:,:,:)
      real(kind=8), intent(inout), target, contiguous :: x(:,:,:,:,:)
      integer, intent(in), target :: GMRES(:)
      real(kind=8), intent(in) :: resmax
      integer, intent(in) :: itermin, iflag, myid, MRT, MRT2
      integer, intent(out), target :: itercount
      real(kind=8), intent(in), target, contiguous :: u(:,:,:,:,:), kappa(:), coact(:,:,:)
      real(kind=8), intent(in) :: GeeGooinv(:,:,:,:,:)
      integer, intent(in) :: bc(:), nms(:), vecbl(:,:), vecblinv(:,:), nn(:,:), iblv(:,:), lvbc(:,:,:), ib(:,:,:,:)
      logical, intent(in) :: ldiv(:), lbd(:,:)
      integer(c_int), target :: bb(4), gm(2)
      call check_geometry("INVERTERS", myid, ldiv, vecbl, vecblinv, iblv, lbd)
      call ensure_started(bc)
      bb = bc(1:4)
      gm = GMRES(1:2)
      call fail("INVERTERS", c_psc(c_loc(b), c_loc(x), c_loc(gm), resmax, itermin, c_loc(itercount), c_loc(u), &
                                   c_loc(coact), c_loc(kappa), iflag, c_loc(bb), isignal, 0.0_c_double))
      call append_log(rwdir, myid)
   end subroutine psc

   ! gmresproject(rwdir,b,xshift,GMRES,resmax,itermin,itercount,u,GeeGooinv,iflag,kappa,coact,bc,
   ! vecbl,vecblinv,myid,nn,ldiv,nms,lvbc,ib,lbd,iblv,MRT,MRT2,isignal,mvp,gdr), inverters.f90:
   ! 7183-7271: GMRES-Proj over the resident deflation space.  Every shift of xshift is zeroed
   ! and the solution is xshift(:,:,:,:,:,1); isignal = 1 solves for the resident beg instead
   ! of b (the last deflation vector after psc's GMRES-DR); mvp = operator applications;
   ! gdr(mvp_j,1) = the least-squares residual norm after the Arnoldi step that was operator
   ! application mvp_j, other entries of gdr as they were.
   subroutine gmresproject(rwdir,b,xshift,GMRES,resmax,itermin,itercount,u,GeeGooinv,iflag,kappa,coact,bc, &
                           vecbl,vecblinv,myid,nn,ldiv,nms,lvbc,ib,lbd,iblv,MRT,MRT2,isignal,mvp,gdr)
      character(len=*), intent(in) :: rwdir(:)
      real(kind=8), intent(in), target, contiguous :: b(:,:,:,:,:)
      real(kind=8), intent(inout), target, contiguous :: xshift(:,:,:,:,:,:)
      integer, intent(in), target :: GMRES(:)
      real(kind=8), intent(in) :: resmax
      integer, intent(in) :: itermin, iflag, myid, MRT, MRT2, isignal
      integer, intent(out), target :: itercount
      integer, intent(inout), target :: mvp
      real(kind=8), intent(inout), target, contiguous :: gdr(:,:)
      real(kind=8), intent(in), target, contiguous :: u(:,:,:,:,:), kappa(:), coact(:,:,:)
      real(kind=8), intent(in) :: GeeGooinv(:,:,:,:,:)
      integer, intent(in) :: bc(:), nms(:), vecbl(:,:), vecblinv(:,:), nn(:,:), iblv(:,:), lvbc(:,:,:), ib(:,:,:,:)
      logical, intent(in) :: ldiv(:), lbd(:,:)
      integer(c_int), target :: bb(4), gm(2)
      call check_geometry("INVERTERS", myid, ldiv, vecbl, vecblinv, iblv, lbd)
      call ensure_started(bc)
      call file_dims("INVERTERS", size(xshift, 2), size(xshift, 1) == 6 .and. size(xshift, 3) == 4 .and. &
                     size(xshift, 4) == 2 .and. size(xshift, 5) == 8 .and. spinor_shape(b) .and. &
                     size(b, 2) == size(xshift, 2))
      bb = bc(1:4)
      gm = GMRES(1:2)
      call fail("INVERTERS", c_gmresproject(c_loc(b), c_loc(xshift), int(size(xshift, 6), c_int), c_loc(gm), &
                                            resmax, itermin, c_loc(itercount), c_loc(u), c_loc(coact), c_loc(kappa), &
                                            iflag, c_loc(bb), isignal, c_loc(mvp), c_loc(gdr), &
                                            int(size(gdr, 1), c_int)))
   end subroutine gmresproject

   ! Mixed-precision solve of the even-odd system (no counterpart in the
   ! reference): tol on ||phi - M x||/||phi||, GMRES = (m, k), operator
   ! storage opprec (1 FP32, 2 FP16) with recon 18 or 12 link reals.
   subroutine qqcd_solve_mixed_f(phi, x, GMRES, tol, opprec, recon, keep, u, kappa, coact, bc, itercount)
      real(kind=8), intent(in), target, contiguous :: phi(:,:,:,:,:)
      real(kind=8), intent(inout), target, contiguous :: x(:,:,:,:,:)
      integer, intent(in), target :: GMRES(:)
      real(kind=8), intent(in) :: tol
      integer, intent(in) :: opprec, recon, keep
      real(kind=8), intent(in), target, contiguous :: u(:,:,:,:,:), kappa(:), coact(:,:,:)
      integer, intent(in) :: bc(:)
      integer, intent(out), target :: itercount
      integer(c_int), target :: bb(4), gm(2)
      call ensure_started(bc)
      bb = bc(1:4)
      gm = GMRES(1:2)
      call fail("INVERTERS", c_solve_mixed(c_loc(phi), c_loc(x), c_loc(gm), tol, 1_c_int, opprec, recon, &
                                           0.0_c_double, keep, c_loc(u), c_loc(coact), c_loc(kappa), c_loc(bb), &
                                           c_loc(itercount)))
   end subroutine qqcd_solve_mixed_f

   ! ---- quark ----------------------------------------------------------

   subroutine pointsource(be,bo,src,icsrc,idsrc,myid,iblv,vecblinv)
      real(kind=8), intent(out), target, contiguous :: be(:,:,:,:,:), bo(:,:,:,:,:)
      integer, intent(in), target :: src(:)
      integer, intent(in) :: icsrc, idsrc, myid
      integer, intent(in) :: iblv(:,:), vecblinv(:,:)
      integer(c_int), target :: s(4)
      if (.not. started) call ensure_started([1, 1, 1, 1])
      s = src(1:4)
      call fail("QUARK", c_pointsource(c_loc(be), c_loc(bo), c_loc(s), icsrc, idsrc))
   end subroutine pointsource

   ! z4source(be,bo,nvhalf,myid), qqcd-test/cfgsprops/quark/disconloops.f90:4359: Z4 noise
   ! {1, -1, i, -i} on every (site, spin, colour) of both checkerboards from the host's own
   ! lagged-Fibonacci stream (module lagfib, qqcd-test/cfgsprops/glue/lagfib.f90): the draws are
   ! made on the device and iwlagfib, jrlagfib, krlagfib advanced past them, so the source and
   ! the stream afterwards are the reference's, bit for bit.
   subroutine z4source(be,bo,nvhalf,myid)
      real(kind=8), intent(out), target, contiguous :: be(:,:,:,:,:), bo(:,:,:,:,:)
      integer, intent(in) :: nvhalf, myid
      if (.not. started) call ensure_started([1, 1, 1, 1])
      if (nvhalf /= nx*ny*nz*nt/32) call fail("DISCONLOOPS", -1_c_int)
      call file_dims("DISCONLOOPS", size(be, 2), spinor_shape(be) .and. spinor_shape(bo) .and. &
                     size(bo, 2) == size(be, 2))
      call fail("DISCONLOOPS", c_z4source_lagfib(c_loc(be), c_loc(bo), iwlagfib, jrlagfib, krlagfib))
   end subroutine z4source

   ! z2source(be,bo,nvhalf), qqcd-test/cfgsprops/quark/disconloops.f90:4583: Z2 noise {1, -1}
   ! in the real part, from the same stream
   subroutine z2source(be,bo,nvhalf)
      real(kind=8), intent(out), target, contiguous :: be(:,:,:,:,:), bo(:,:,:,:,:)
      integer, intent(in) :: nvhalf
      if (.not. started) call ensure_started([1, 1, 1, 1])
      if (nvhalf /= nx*ny*nz*nt/32) call fail("DISCONLOOPS", -1_c_int)
      call file_dims("DISCONLOOPS", size(be, 2), spinor_shape(be) .and. spinor_shape(bo) .and. &
                     size(bo, 2) == size(be, 2))
      call fail("DISCONLOOPS", c_z2source_lagfib(c_loc(be), c_loc(bo), iwlagfib, jrlagfib, krlagfib))
   end subroutine z2source

   ! loopops(ksub,xe,xo,u,be,bo,Je,Jo,ieo,ibleo,bc,...), disconloops.f90:7769: the nine loop
   ! operators (point-split currents J_1..J_4, local scalar, local vector currents) of sink x and
   ! source b on the sites of slice (ieo, ibleo): Je for the gblclr=1 sites, Jo for gblclr=2.
   ! Every call computes all slices on the device; loopops_all hands them over in one call.
   subroutine loopops(ksub,xe,xo,u,be,bo,Je,Jo,ieo,ibleo,bc,vecbl,vecblinv,myid,nn,ldiv,nms,lvbc,ib,lbd,iblv,MRT)
      integer, intent(in) :: ksub, ieo, ibleo, myid, MRT
      real(kind=8), intent(inout), target, contiguous :: xe(:,:,:,:,:), xo(:,:,:,:,:), be(:,:,:,:,:), bo(:,:,:,:,:)
      real(kind=8), intent(in), target, contiguous :: u(:,:,:,:,:)
      real(kind=8), intent(out) :: Je(:,:,:), Jo(:,:,:)
      integer, intent(in) :: bc(:), nms(:), vecbl(:,:), vecblinv(:,:), nn(:,:), iblv(:,:), lvbc(:,:,:), ib(:,:,:,:)
      logical, intent(in) :: ldiv(:), lbd(:,:)
      real(kind=8), allocatable :: ae(:,:,:,:,:), ao(:,:,:,:,:)
      integer :: nvh
      call check_geometry("DISCONLOOPS", myid, ldiv, vecbl, vecblinv, iblv, lbd)
      nvh = nx*ny*nz*nt/32
      if (size(Je, 1) /= 2 .or. size(Je, 2) < nvh .or. size(Je, 3) < 9) call fail("DISCONLOOPS", -1_c_int)
      allocate(ae(2,nvh,9,2,8), ao(2,nvh,9,2,8))
      call loopops_all(ksub, xe, xo, u, be, bo, ae, ao, bc)
      Je(:,1:nvh,1:9) = ae(:,:,:,ieo,ibleo)
      Jo(:,1:nvh,1:9) = ao(:,:,:,ieo,ibleo)
   end subroutine loopops

   ! All (ieo, ibleo) slices of loopops at once: Je, Jo(2,nvhalf,9,2,8).
   subroutine loopops_all(ksub, xe, xo, u, be, bo, Je, Jo, bc)
      integer, intent(in) :: ksub
      real(kind=8), intent(in), target, contiguous :: xe(:,:,:,:,:), xo(:,:,:,:,:), be(:,:,:,:,:), bo(:,:,:,:,:)
      real(kind=8), intent(in), target, contiguous :: u(:,:,:,:,:)
      real(kind=8), intent(out), target, contiguous :: Je(:,:,:,:,:), Jo(:,:,:,:,:)
      integer, intent(in) :: bc(:)
      integer(c_int), target :: b(4)
      call ensure_started(bc)
      b = bc(1:4)
      call fail("DISCONLOOPS", c_loopops(int(ksub, c_int), c_loc(xe), c_loc(xo), c_loc(u), c_loc(be), c_loc(bo), &
                                         c_loc(b), c_loc(Je), c_loc(Jo)))
   end subroutine loopops_all

   ! nsaverage(Jtime,u,z2e,z2o,xe,xo,kappa,coact,bc,...,MRT2,ESHFES), disconloops.f90:3079:
   ! loopops and spacesum over every slice, momentum factors zero, the currents times kappa;
   ! Jtime(2,nt,5,9).  Only ESHFES = 1 (no eigenmode subtraction).
   subroutine nsaverage(Jtime,u,z2e,z2o,xe,xo,kappa,coact,bc,vecbl,vecblinv,myid,nn,ldiv,nms,lvbc,ib,lbd, &
                        iblv,MRT2,ESHFES)
      real(kind=8), intent(out), target, contiguous :: Jtime(:,:,:,:)
      real(kind=8), intent(inout), target, contiguous :: u(:,:,:,:,:)
      real(kind=8), intent(in) :: kappa
      real(kind=8), intent(inout), target, contiguous :: z2e(:,:,:,:,:), z2o(:,:,:,:,:), xe(:,:,:,:,:), xo(:,:,:,:,:)
      real(kind=8), intent(in) :: coact(:,:,:)
      integer, intent(in) :: bc(:), nms(:), vecbl(:,:), vecblinv(:,:), nn(:,:), iblv(:,:), lvbc(:,:,:), ib(:,:,:,:)
      logical, intent(in) :: ldiv(:), lbd(:,:)
      integer, intent(in) :: myid, MRT2, ESHFES
      integer(c_int), target :: b(4)
      if (ESHFES /= 1) call fail("DISCONLOOPS", -1_c_int)
      if (size(Jtime) /= 2*nt*5*9) call fail("DISCONLOOPS", -1_c_int)
      call check_geometry("DISCONLOOPS", myid, ldiv, vecbl, vecblinv, iblv, lbd)
      call ensure_started(bc)
      b = bc(1:4)
      call fail("DISCONLOOPS", c_loopsum(0_c_int, c_loc(xe), c_loc(xo), c_loc(u), c_loc(z2e), c_loc(z2o), &
                                         c_loc(b), c_null_ptr, kappa, c_loc(Jtime)))
   end subroutine nsaverage

   subroutine fermprop(rwdir,be,bo,nkappa,kappa,cSW,inverter,coact,bc,resmax,itermin,omegaM3R,u,vectype, &
                       kGMRES,xe,xo,myid,nn,ldiv,nms,lvbc,lvcl,ib,lbd,iblv,vecbl,vecblinv,ntmqcd,hoption,MRT,MRT2)
      use gmresrhs, only: isignal
      character(len=*), intent(in) :: rwdir(:)
      real(kind=8), intent(in), target, contiguous :: be(:,:,:,:,:)
      real(kind=8), intent(inout), target, contiguous :: bo(:,:,:,:,:)
      integer, intent(in) :: nkappa, hoption, inverter, kGMRES, ntmqcd, itermin, myid, MRT, MRT2
      real(kind=8), intent(in), target, contiguous :: kappa(:)
      real(kind=8), intent(in) :: cSW, resmax, omegaM3R
      real(kind=8), intent(in), target, contiguous :: coact(:,:,:)
      integer, intent(in) :: bc(:), nms(:)
      real(kind=8), intent(inout), target, contiguous :: u(:,:,:,:,:)
      integer, intent(inout) :: vectype
      real(kind=8), intent(inout), target, contiguous :: xe(:,:,:,:,:,:), xo(:,:,:,:,:,:)
      integer, intent(in) :: nn(:,:), iblv(:,:), vecbl(:,:), vecblinv(:,:), lvbc(:,:,:), lvcl(:,:,:), ib(:,:,:,:)
      logical, intent(in) :: ldiv(:), lbd(:,:)
      integer(c_int), target :: b(4), ngm(2), it
      if (nkappa /= -1 .or. ntmqcd /= 0) call fail("QUARK", -1_c_int)
      call check_geometry("QUARK", myid, ldiv, vecbl, vecblinv, iblv, lbd)
      call ensure_started(bc)
      b = bc(1:4)
      ngm = [vectype, kGMRES]
      call fail("QUARK", c_fermprop(c_loc(be), c_loc(bo), c_loc(kappa), inverter, c_loc(coact), c_loc(b), resmax, &
                                    itermin, c_loc(ngm), vectype, c_loc(u), c_loc(xe), c_loc(xo), isignal, c_loc(it)))
      if (inverter == 0 .and. vectype == 1) vectype = 0
      call append_log(rwdir, myid)
   end subroutine fermprop

   ! fermpropR(rwdir,be,bo,nkappa,...,MRT,MRT2,ir), quark.f90:846-931: fermprop with the Wilson
   ! r coefficient ir; ir = -1 negates coact(3,1,:) (:921-926), any other value leaves it.
   subroutine fermpropR(rwdir,be,bo,nkappa,kappa,cSW,inverter,coact,bc,resmax,itermin,omegaM3R,u,vectype, &
                        kGMRES,xe,xo,myid,nn,ldiv,nms,lvbc,lvcl,ib,lbd,iblv,vecbl,vecblinv,ntmqcd,hoption,MRT, &
                        MRT2,ir)
      character(len=*), intent(in) :: rwdir(:)
      real(kind=8), intent(in), target, contiguous :: be(:,:,:,:,:)
      real(kind=8), intent(inout), target, contiguous :: bo(:,:,:,:,:)
      integer, intent(in) :: nkappa, hoption, inverter, kGMRES, ntmqcd, itermin, myid, MRT, MRT2, ir
      real(kind=8), intent(in), target, contiguous :: kappa(:)
      real(kind=8), intent(in) :: cSW, resmax, omegaM3R
      real(kind=8), intent(in) :: coact(:,:,:)
      integer, intent(in) :: bc(:), nms(:)
      real(kind=8), intent(inout), target, contiguous :: u(:,:,:,:,:)
      integer, intent(inout) :: vectype
      real(kind=8), intent(inout), target, contiguous :: xe(:,:,:,:,:,:), xo(:,:,:,:,:,:)
      integer, intent(in) :: nn(:,:), iblv(:,:), vecbl(:,:), vecblinv(:,:), lvbc(:,:,:), lvcl(:,:,:), ib(:,:,:,:)
      logical, intent(in) :: ldiv(:), lbd(:,:)
      real(kind=8), target :: coactr(4,4,4)
      coactr = coact(1:4,1:4,1:4)
      if (ir == -1) coactr(3,1,:) = -coact(3,1,1:4)
      call fermprop(rwdir,be,bo,nkappa,kappa,cSW,inverter,coactr,bc,resmax,itermin,omegaM3R,u,vectype, &
                    kGMRES,xe,xo,myid,nn,ldiv,nms,lvbc,lvcl,ib,lbd,iblv,vecbl,vecblinv,ntmqcd,hoption,MRT,MRT2)
   end subroutine fermpropR

   ! ---- gauge fields ----------------------------------------------------

   ! sweep(gaction,coact,nhit,u,myid,nn,ldiv,nms,lv,ib,lbd,iblv,MRT), cfgsprops.f90:1973: one
   ! pseudo-heatbath sweep of u (in place) on the device, drawing from the host's lagfib stream
   ! exactly as the reference's sweep does (staples of the plaquette action: gaction = 1).  The
   ! operators pick up the new links on their next call.
   subroutine sweep(gaction,coact,nhit,u,myid,nn,ldiv,nms,lv,ib,lbd,iblv,MRT)
      integer, intent(in) :: gaction(:), nhit, myid, MRT
      real(kind=8), intent(in), target, contiguous :: coact(:,:,:)
      real(kind=8), intent(inout), target, contiguous :: u(:,:,:,:,:)
      integer, intent(in) :: nn(:,:), iblv(:,:), nms(:), lv(:,:,:), ib(:,:,:,:)
      logical, intent(in) :: ldiv(:), lbd(:,:)
      integer(c_int) :: ga(4)
      if (.not. started) call ensure_started([1, 1, 1, 1])
      call check_geometry("CFGSPROPS", myid, ldiv, iblv=iblv, lbd=lbd)
      ga = gaction(1:4)
      call fail("CFGSPROPS", c_sweep(ga, c_loc(coact), int(nhit, c_int), 1_c_int, c_loc(u), iwlagfib, jrlagfib, &
                                     krlagfib))
   end subroutine sweep

   ! ---- the reference's files (host only: no device, usable before the first operator call) ----

   ! a spinor array the library fills whole: (6, ntotal, 4, 2, 8), both halves of the source alike
   logical function spinor_shape(a)
      real(kind=8), intent(in) :: a(:,:,:,:,:)
      spinor_shape = size(a, 1) == 6 .and. size(a, 3) == 4 .and. size(a, 4) == 2 .and. size(a, 5) == 8
   end function spinor_shape

   ! the single-process array extents the files are laid out in: (nvhalf, ntotal)
   subroutine file_dims(where, nsites, ok)
      character(len=*), intent(in) :: where
      integer, intent(in) :: nsites
      logical, intent(in) :: ok
      integer :: nvh
      nvh = nx*ny*nz*nt/32
      if (.not. ok .or. nsites /= nvh + max(2*nvh/nx, 2*nvh/ny, 2*nvh/nz, 2*nvh/nt)) &
         call fail(where, -1_c_int)
   end subroutine file_dims

   ! GFwrite(u,filename,gaction,beta,alat,tad,icfgsave,myid), gaugetools.f90:1260: fixsu3 on u in
   ! place, rows 1-2 of every link as unformatted records, then the trailer; the file must be new.
   ! The library forgets its device copy of u (re-unitarised in place); the next operator call
   ! that passes u uploads it again.
   subroutine GFwrite(u,filename,gaction,beta,alat,tad,icfgsave,myid)
      real(kind=8), intent(inout), target, contiguous :: u(:,:,:,:,:)
      character(len=*), intent(in) :: filename
      real(kind=8), intent(in) :: beta, alat(:), tad(:)
      integer, intent(in) :: gaction(:), icfgsave, myid
      integer(c_int) :: n(4), ga(4)
      real(c_double) :: al(4), ta(4)
      call file_dims("GAUGETOOLS", size(u, 2), size(u, 1) == 18 .and. size(u, 3) == 4 .and. &
                     size(u, 4) == 2 .and. size(u, 5) == 16)
      n = [nx, ny, nz, nt]
      ga = gaction(1:4)
      al = alat(1:4)
      ta = tad(1:4)
      call fail("GAUGETOOLS", c_gfwrite(n, c_loc(u), trim(filename)//c_null_char, ga, beta, al, ta, &
                                        int(icfgsave, c_int), int(myid, c_int)))
   end subroutine GFwrite

   ! GFread(u,filename,gaction,beta,alat,tad,icfgsave,myidin), gaugetools.f90:1298: rows 1-2 from
   ! the file, row 3 rebuilt by fixsu3.  The library forgets its device copy of u, as after GFwrite.
   subroutine GFread(u,filename,gaction,beta,alat,tad,icfgsave,myidin)
      real(kind=8), intent(inout), target, contiguous :: u(:,:,:,:,:)
      character(len=*), intent(in) :: filename
      real(kind=8), intent(out) :: beta, alat(:), tad(:)
      integer, intent(out) :: gaction(:), icfgsave, myidin
      integer(c_int) :: n(4), ga(4), ic, my
      real(c_double) :: be, al(4), ta(4)
      call file_dims("GAUGETOOLS", size(u, 2), size(u, 1) == 18 .and. size(u, 3) == 4 .and. &
                     size(u, 4) == 2 .and. size(u, 5) == 16)
      n = [nx, ny, nz, nt]
      call fail("GAUGETOOLS", c_gfread(n, c_loc(u), trim(filename)//c_null_char, ga, be, al, ta, ic, my))
      gaction(1:4) = ga
      beta = be
      alat(1:4) = al
      tad(1:4) = ta
      icfgsave = ic
      myidin = my
   end subroutine GFread

   ! FPwrite(xe,xo,vecblinv,filename,kappa,tmqcd,cSW,icsrc,idsrc,icfg,myid), quark.f90:3379: the
   ! propagator by pairs of timeslices, then the trailer; an existing file is replaced.
   subroutine FPwrite(xe,xo,vecblinv,filename,kappa,tmqcd,cSW,icsrc,idsrc,icfg,myid)
      real(kind=8), intent(in), target, contiguous :: xe(:,:,:,:,:), xo(:,:,:,:,:)
      integer, intent(in) :: vecblinv(:,:)
      character(len=*), intent(in) :: filename
      real(kind=8), intent(in) :: kappa, tmqcd, cSW
      integer, intent(in) :: icsrc, idsrc, icfg, myid
      integer(c_int) :: n(4)
      call file_dims("QUARK", size(xe, 2), size(xe, 1) == 6 .and. size(xe, 3) == 4 .and. size(xe, 4) == 2 &
                     .and. size(xe, 5) == 8 .and. all(shape(xo) == shape(xe)) .and. vecblinv_ok(vecblinv))
      n = [nx, ny, nz, nt]
      call fail("QUARK", c_fpwrite(n, c_loc(xe), c_loc(xo), trim(filename)//c_null_char, kappa, tmqcd, cSW, &
                                   int(icsrc, c_int), int(idsrc, c_int), int(icfg, c_int), int(myid, c_int)))
   end subroutine FPwrite

   ! FPread(xe,xo,vecblinv,filename,kappa,tmqcd,cSW,icsrc,idsrc,icfgin,myidin), quark.f90:3438.
   subroutine FPread(xe,xo,vecblinv,filename,kappa,tmqcd,cSW,icsrc,idsrc,icfgin,myidin)
      real(kind=8), intent(inout), target, contiguous :: xe(:,:,:,:,:), xo(:,:,:,:,:)
      integer, intent(in) :: vecblinv(:,:)
      character(len=*), intent(in) :: filename
      real(kind=8), intent(out) :: kappa, tmqcd, cSW
      integer, intent(out) :: icsrc, idsrc, icfgin, myidin
      integer(c_int) :: n(4), ic, id, icf, my
      real(c_double) :: k, t, c
      call file_dims("QUARK", size(xe, 2), size(xe, 1) == 6 .and. size(xe, 3) == 4 .and. size(xe, 4) == 2 &
                     .and. size(xe, 5) == 8 .and. all(shape(xo) == shape(xe)) .and. vecblinv_ok(vecblinv))
      n = [nx, ny, nz, nt]
      call fail("QUARK", c_fpread(n, c_loc(xe), c_loc(xo), trim(filename)//c_null_char, k, t, c, ic, id, icf, my))
      kappa = k
      tmqcd = t
      cSW = c
      icsrc = ic
      idsrc = id
      icfgin = icf
      myidin = my
   end subroutine FPread

   ! vecblinv of one process (the files' block order assumes it)
   logical function vecblinv_ok(vecblinv)
      integer, intent(in) :: vecblinv(:,:)
      integer :: ibl
      vecblinv_ok = size(vecblinv, 1) >= 2 .and. size(vecblinv, 2) >= 16
      if (.not. vecblinv_ok) return
      do ibl = 1, 16
         if (vecblinv(1,ibl) /= 1 + mod(popcnt(ibl - 1), 2) .or. vecblinv(2,ibl) /= (ibl + 1)/2) &
            vecblinv_ok = .false.
      end do
   end function vecblinv_ok

end module qqcd_hip
