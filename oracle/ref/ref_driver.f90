! ref_driver.f90 -- C entry points onto the reference Fortran, for the oracle.
!
! TEST INFRASTRUCTURE ONLY: this project-written module is compiled together
! with the unmodified reference sources (BaylorLQCD "qqcd", compiled from
! /root/reference by oracle/ref/Makefile into oracle/_ref/) and lets the
! golden-vector script and the CPU baseline call the reference's own
! routines through ctypes.  It holds no physics of its own: every entry point
! forwards to one reference procedure with the geometry tables that the
! reference's initlat builds (qqcd-test/cfgsprops/glue/lattice.f90:33-59).
!
! Single process only (npx=npy=npz=npt=1 in the latdims copy the recipe
! writes), myid = 0.

module refdrv
   use iso_c_binding
   use kinds
   use latdims
   use basics
   use lattice
   use diracops
#ifdef WITH_SOLVERS
   use gmresrhs
   use inverters
   use quark
   use gaugetools
   use heatbath
   use lagfib
   use disconloops, only: loopops, momfacs, spacesum, nsaverage, z4source, z2source
#endif
   implicit none
   include 'mpif.h'

   integer(kind=KI), dimension(4)           :: gbc
   integer(kind=KI), dimension(4,2)         :: nn
   logical,          dimension(4)           :: ldiv
   integer(kind=KI), dimension(4)           :: nms
   integer(kind=KI), dimension(nvhalf,4,2)  :: lv, lvbc
   integer(kind=KI), dimension(nbmax,4,2,2) :: ib
   logical,          dimension(16,4)        :: lbd
   integer(kind=KI), dimension(16,4)        :: iblv
   integer(kind=KI), dimension(2,8)         :: vecbl
   integer(kind=KI), dimension(2,16)        :: vecblinv
   character(len=512), dimension(1)         :: rwdir
   integer(kind=KI), parameter              :: myid = 0

contains

   ! Lattice sizes compiled into this build: n(4), nvhalf, ntotal.
   subroutine ref_dims(n, nvh, ntot) bind(C, name="ref_dims")
      integer(c_int), intent(out) :: n(4), nvh, ntot
      n = (/ nx, ny, nz, nt /)
      nvh = nvhalf
      ntot = ntotal
   end subroutine ref_dims

   ! Build the geometry tables for boundary conditions bc(4); dir is the
   ! directory the reference solvers append CFGSPROPS.LOG to.
   subroutine ref_init(bc, dir, ndir) bind(C, name="ref_init")
      integer(c_int), intent(in) :: bc(4)
      integer(c_int), value      :: ndir
      character(kind=c_char), intent(in) :: dir(ndir)
      integer :: k
      gbc = bc
      call initlat(myid, gbc, nn, ldiv, nms, lv, lvbc, ib, lbd, iblv, vecbl, vecblinv)
      rwdir(1) = ' '
      do k = 1, ndir
         rwdir(1)(k:k) = dir(k)
      end do
   end subroutine ref_init

   ! Export the geometry tables (for the layout tests).
   subroutine ref_tables(o_lvbc, o_vecbl, o_vecblinv, o_iblv, o_lbd, o_nms) bind(C, name="ref_tables")
      integer(c_int), intent(out) :: o_lvbc(nvhalf,4,2), o_vecbl(2,8), o_vecblinv(2,16)
      integer(c_int), intent(out) :: o_iblv(16,4), o_lbd(16,4), o_nms(4)
      o_lvbc = lvbc
      o_vecbl = vecbl
      o_vecblinv = vecblinv
      o_iblv = iblv
      o_lbd = merge(1, 0, lbd)
      o_nms = nms
   end subroutine ref_tables

   subroutine ref_hsingle(gout, u, gin, idag, coact, gblclr) bind(C, name="ref_hsingle")
      real(c_double), intent(inout) :: gout(6,ntotal,4,2,8), gin(6,ntotal,4,2,8)
      real(c_double), intent(in)    :: u(18,ntotal,4,2,16), coact(4,4,4)
      integer(c_int), value         :: idag, gblclr
      integer(kind=KI) :: MRT
      MRT = MPI_DOUBLE_PRECISION
      call Hsingle(gout, u, gin, idag, coact, gbc, gblclr, vecbl, vecblinv, myid, nn, &
                   ldiv, nms, lvbc, ib, lbd, iblv, MRT)
   end subroutine ref_hsingle

   subroutine ref_hdouble(he, u, ge, idag, coact, kappa) bind(C, name="ref_hdouble")
      real(c_double), intent(inout) :: he(6,ntotal,4,2,8), ge(6,ntotal,4,2,8)
      real(c_double), intent(in)    :: u(18,ntotal,4,2,16), coact(4,4,4)
      integer(c_int), value         :: idag
      real(c_double), value         :: kappa
      integer(kind=KI) :: MRT
      MRT = MPI_DOUBLE_PRECISION
      call Hdouble(he, u, ge, idag, coact, kappa, gbc, vecbl, vecblinv, myid, nn, ldiv, &
                   nms, lvbc, ib, lbd, iblv, MRT)
   end subroutine ref_hdouble

   subroutine ref_hdbletm(he, u, ge, idag, coact, kappa, iflag) bind(C, name="ref_hdbletm")
      real(c_double), intent(inout) :: he(6,ntotal,4,2,8), ge(6,ntotal,4,2,8)
      real(c_double), intent(in)    :: u(18,ntotal,4,2,16), coact(4,4,4), kappa(2)
      integer(c_int), value         :: idag, iflag
      real(kind=KR) :: gdummy(18,1,1,1,1)
      integer(kind=KI) :: MRT
      MRT = MPI_DOUBLE_PRECISION
      gdummy = 0.0_KR
      call Hdbletm(he, u, gdummy, ge, idag, coact, kappa, iflag, gbc, vecbl, vecblinv, &
                   myid, nn, ldiv, nms, lvbc, ib, lbd, iblv, MRT)
   end subroutine ref_hdbletm

   subroutine ref_gammamult(ein, oin, eout, oout, mu, ieo, ibleo) bind(C, name="ref_gammamult")
      real(c_double), intent(in)    :: ein(6,ntotal,4,2,8), oin(6,ntotal,4,2,8)
      real(c_double), intent(inout) :: eout(6,ntotal,4,2,8), oout(6,ntotal,4,2,8)
      integer(c_int), value         :: mu, ieo, ibleo
      call gammamult(ein, oin, eout, oout, mu, ieo, ibleo)
   end subroutine ref_gammamult

   ! smear (diracops.f90:712-856)
   subroutine ref_smear(he, ho, u, ge, go, asmear, nsmear) bind(C, name="ref_smear")
      real(c_double), intent(inout) :: he(6,ntotal,4,2,8), ho(6,ntotal,4,2,8)
      real(c_double), intent(in)    :: u(18,ntotal,4,2,16), ge(6,ntotal,4,2,8), go(6,ntotal,4,2,8)
      real(c_double), value         :: asmear
      integer(c_int), value         :: nsmear
      integer(kind=KI) :: MRT
      MRT = MPI_DOUBLE_PRECISION
      call smear(he, ho, u, ge, go, asmear, nsmear, gbc, vecbl, vecblinv, myid, nn, ldiv, nms, lvbc, ib, lbd, &
                 iblv, MRT)
   end subroutine ref_smear

   ! twMshift (diracops.f90:1863-1995), iflag = -1; hwhole, gwhole (6,ntotal,4,2,8,2)
   subroutine ref_twmshift(hwhole, u, gwhole, idag, coact, kappa) bind(C, name="ref_twmshift")
      real(c_double), intent(inout) :: hwhole(6,ntotal,4,2,8,2), gwhole(6,ntotal,4,2,8,2)
      real(c_double), intent(in)    :: u(18,ntotal,4,2,16), coact(4,4,4), kappa(2)
      integer(c_int), value         :: idag
      real(kind=KR) :: gdummy(18,1,1,1,1)
      integer(kind=KI) :: MRT
      MRT = MPI_DOUBLE_PRECISION
      gdummy = 0.0_KR
      call twMshift(hwhole, u, gdummy, gwhole, idag, coact, kappa, -1, gbc, vecbl, vecblinv, myid, nn, ldiv, &
                    nms, lvbc, ib, lbd, iblv, MRT)
   end subroutine ref_twmshift

   ! gamma5mult (diracops.f90:1750-1860), iflag = -1
   subroutine ref_gamma5mult(he, u, ge, idag, coact, kappa) bind(C, name="ref_gamma5mult")
      real(c_double), intent(inout) :: he(6,ntotal,4,2,8), ge(6,ntotal,4,2,8)
      real(c_double), intent(in)    :: u(18,ntotal,4,2,16), coact(4,4,4), kappa(2)
      integer(c_int), value         :: idag
      real(kind=KR) :: gdummy(18,1,1,1,1)
      integer(kind=KI) :: MRT
      MRT = MPI_DOUBLE_PRECISION
      gdummy = 0.0_KR
      call gamma5mult(he, u, gdummy, ge, idag, coact, kappa, -1, gbc, vecbl, vecblinv, myid, nn, ldiv, &
                      nms, lvbc, ib, lbd, iblv, MRT)
   end subroutine ref_gamma5mult

   subroutine ref_gam5x(a) bind(C, name="ref_gam5x")
      real(c_double), intent(inout) :: a(6,ntotal,4,2,8)
      call gam5x(a)
   end subroutine ref_gam5x

#ifdef WITH_SOLVERS
   subroutine ref_vecdot(a, b, adb) bind(C, name="ref_vecdot")
      real(c_double), intent(in)  :: a(6,ntotal,4,2,8), b(6,ntotal,4,2,8)
      real(c_double), intent(out) :: adb(2)
      integer(kind=KI) :: MRT2
      MRT2 = MPI_DOUBLE_PRECISION
      call vecdot(a, b, adb, MRT2)
   end subroutine ref_vecdot

   ! gmresdr / psc / bicgstab on the even-odd system phi -> x, iflag=-1.
   subroutine ref_solve(which, phi, x, nk, resmax, itermin, itercount, u, coact, kappa, &
                        isig) bind(C, name="ref_solve")
      integer(c_int), value         :: which, itermin, isig
      real(c_double), intent(inout) :: phi(6,ntotal,4,2,8), x(6,ntotal,4,2,8)
      integer(c_int), intent(in)    :: nk(2)
      real(c_double), value         :: resmax
      integer(c_int), intent(out)   :: itercount
      real(c_double), intent(in)    :: u(18,ntotal,4,2,16), coact(4,4,4), kappa(2)
      real(kind=KR) :: gdummy(18,1,1,1,1)
      integer(kind=KI) :: MRT, MRT2, iflag, gm(2)
      MRT = MPI_DOUBLE_PRECISION
      MRT2 = MPI_DOUBLE_PRECISION
      iflag = -1
      gm = nk
      gdummy = 0.0_KR
      isignal = isig
      select case (which)
      case (0)
         call bicgstab(rwdir, phi, x, resmax, itermin, itercount, u, gdummy, iflag, kappa, &
                       coact, gbc, vecbl, vecblinv, myid, nn, ldiv, nms, lvbc, ib, lbd, iblv, &
                       MRT, MRT2)
      case (5)
         call gmresdr(rwdir, phi, x, gm, resmax, itermin, itercount, u, gdummy, iflag, &
                      kappa, coact, gbc, vecbl, vecblinv, myid, nn, ldiv, nms, lvbc, ib, lbd, &
                      iblv, MRT, MRT2)
      case (6)
         call psc(rwdir, phi, x, gm, resmax, itermin, itercount, u, gdummy, iflag, &
                  kappa, coact, gbc, vecbl, vecblinv, myid, nn, ldiv, nms, lvbc, ib, lbd, &
                  iblv, MRT, MRT2)
      end select
   end subroutine ref_solve

   ! gmresproject(rwdir,b,xshift,GMRES,resmax,itermin,itercount,u,GeeGooinv,iflag,kappa,coact,bc,
   ! ...,MRT,MRT2,isignal,mvp,gdr) called directly (inverters.f90:7183-7271), over the
   ! deflation space a previous gmresdr (through psc with isignal=1) left in module shift:
   ! phi -> xshift(:,:,:,:,:,1..nshift), itercount, mvp and gdr(1:ngdr,1:ncol).
   subroutine ref_gmresproject(phi, xshift, nshift, nk, resmax, itermin, itercount, u, coact, kappa, &
                               isig, mvp, gdr, ngdr, ncol) bind(C, name="ref_gmresproject")
      integer(c_int), value         :: nshift, itermin, isig, ngdr, ncol
      real(c_double), intent(inout) :: phi(6,ntotal,4,2,8), xshift(6,ntotal,4,2,8,nshift)
      integer(c_int), intent(in)    :: nk(2)
      real(c_double), value         :: resmax
      integer(c_int), intent(out)   :: itercount
      integer(c_int), intent(inout) :: mvp
      real(c_double), intent(inout) :: gdr(ngdr,ncol)
      real(c_double), intent(in)    :: u(18,ntotal,4,2,16), coact(4,4,4), kappa(2)
      real(kind=KR) :: gdummy(18,1,1,1,1)
      integer(kind=KI) :: MRT, MRT2, iflag, gm(2), isg
      MRT = MPI_DOUBLE_PRECISION
      MRT2 = MPI_DOUBLE_PRECISION
      iflag = -1
      gm = nk
      isg = isig
      gdummy = 0.0_KR
      call gmresproject(rwdir, phi, xshift, gm, resmax, itermin, itercount, u, gdummy, iflag, kappa, &
                        coact, gbc, vecbl, vecblinv, myid, nn, ldiv, nms, lvbc, ib, lbd, iblv, MRT, MRT2, &
                        isg, mvp, gdr)
   end subroutine ref_gmresproject

   ! fermpropR(...,ir) (quark.f90:846-931): fermprop with coact(3,1,:) negated for ir = -1.
   subroutine ref_fermpropr(be, bo, kappa, inverter, coact, resmax, itermin, nk, u, xe, xo, &
                            isig, ir) bind(C, name="ref_fermpropr")
      real(c_double), intent(inout) :: be(6,ntotal,4,2,8), bo(6,ntotal,4,2,8)
      real(c_double), intent(inout) :: xe(6,ntotal,4,2,8,1), xo(6,ntotal,4,2,8,1)
      real(c_double), intent(in)    :: kappa(2), coact(4,4,4)
      real(c_double), intent(inout) :: u(18,ntotal,4,2,16)
      integer(c_int), value         :: inverter, itermin, isig, ir
      real(c_double), value         :: resmax
      integer(c_int), intent(in)    :: nk(2)
      integer(kind=KI) :: MRT, MRT2, vectype, kg, ntm, hopt, irr
      real(kind=KR2)   :: om
      real(kind=KR)    :: csw
      MRT = MPI_DOUBLE_PRECISION
      MRT2 = MPI_DOUBLE_PRECISION
      vectype = nk(1)
      kg = nk(2)
      ntm = 0
      hopt = 0
      om = 1.0_KR2
      csw = 0.0_KR
      irr = ir
      isignal = isig
      call fermpropR(rwdir, be, bo, -1, kappa, csw, inverter, coact, gbc, resmax, itermin, om, &
                     u, vectype, kg, xe, xo, myid, nn, ldiv, nms, lvbc, lv, ib, lbd, iblv, vecbl, &
                     vecblinv, ntm, hopt, MRT, MRT2, irr)
   end subroutine ref_fermpropr

   ! fermprop for twisted/Wilson (nkappa=-1): be, bo -> xe, xo.
   subroutine ref_fermprop(be, bo, kappa, inverter, coact, resmax, itermin, nk, u, xe, xo, &
                           isig) bind(C, name="ref_fermprop")
      real(c_double), intent(inout) :: be(6,ntotal,4,2,8), bo(6,ntotal,4,2,8)
      real(c_double), intent(inout) :: xe(6,ntotal,4,2,8,1), xo(6,ntotal,4,2,8,1)
      real(c_double), intent(in)    :: kappa(2), coact(4,4,4)
      real(c_double), intent(inout) :: u(18,ntotal,4,2,16)
      integer(c_int), value         :: inverter, itermin, isig
      real(c_double), value         :: resmax
      integer(c_int), intent(in)    :: nk(2)
      integer(kind=KI) :: MRT, MRT2, vectype, kg, ntm, hopt
      real(kind=KR2)   :: om
      real(kind=KR)    :: csw
      MRT = MPI_DOUBLE_PRECISION
      MRT2 = MPI_DOUBLE_PRECISION
      vectype = nk(1)
      kg = nk(2)
      ntm = 0
      hopt = 0
      om = 1.0_KR2
      csw = 0.0_KR
      isignal = isig
      call fermprop(rwdir, be, bo, -1, kappa, csw, inverter, coact, gbc, resmax, itermin, om, &
                    u, vectype, kg, xe, xo, myid, nn, ldiv, nms, lvbc, lv, ib, lbd, iblv, vecbl, &
                    vecblinv, ntm, hopt, MRT, MRT2)
   end subroutine ref_fermprop

   ! pointsource(be,bo,src,icsrc,idsrc,...), quark.f90:46-108
   subroutine ref_pointsource(be, bo, src, icsrc, idsrc) bind(C, name="ref_pointsource")
      real(c_double), intent(inout) :: be(6,ntotal,4,2,8), bo(6,ntotal,4,2,8)
      integer(c_int), intent(in)    :: src(4)
      integer(c_int), value         :: icsrc, idsrc
      integer(kind=KI) :: s(4)
      s = src
      call pointsource(be, bo, s, icsrc, idsrc, myid, iblv, vecblinv)
   end subroutine ref_pointsource

   ! nsweep pseudo-heatbath sweeps at the plaquette action held in coact,
   ! following sweep (qqcd-test/cfgsprops/cfgsprops.f90:1973-2003), with the
   ! lagged-Fibonacci stream seeded by initlagfib(0, iseed).
   subroutine ref_heatbath(u, coact, nsweep, nhit, iseed) bind(C, name="ref_heatbath")
      real(c_double), intent(inout) :: u(18,ntotal,4,2,16)
      real(c_double), intent(in)    :: coact(4,4,4)
      integer(c_int), value         :: nsweep, nhit, iseed
      integer(kind=KI) :: isw, ieo, ibl, mu, MRT
      integer(kind=KI) :: gaction(4)
      real(kind=KR)    :: c(18,nvhalf)
      MRT = MPI_DOUBLE_PRECISION
      gaction = 1
      call initlagfib(myid, iseed)
      do isw = 1, nsweep
         do ieo = 1, 2
            do ibl = 1, 16
               do mu = 1, 4
                  call staple(mu, ieo, ibl, coact, gaction, c, u, myid, nn, ldiv, nms, lv, ib, &
                              lbd, iblv, MRT)
                  call phbup(mu, ieo, ibl, c, u, nhit)
               end do
            end do
         end do
      end do
   end subroutine ref_heatbath

   ! staple(mu,ieo,ibl,coact,gaction,c,u,...) (gaugetools.f90:39-345) of one link set
   subroutine ref_staple(mu, ieo, ibl, coact, gaction, c, u) bind(C, name="ref_staple")
      integer(c_int), value         :: mu, ieo, ibl
      real(c_double), intent(in)    :: coact(4,4,4)
      integer(c_int), intent(in)    :: gaction(4)
      real(c_double), intent(out)   :: c(18,nvhalf)
      real(c_double), intent(inout) :: u(18,ntotal,4,2,16)
      integer(kind=KI) :: MRT, ga(4)
      MRT = MPI_DOUBLE_PRECISION
      ga = gaction
      call staple(mu, ieo, ibl, coact, ga, c, u, myid, nn, ldiv, nms, lv, ib, lbd, iblv, MRT)
   end subroutine ref_staple

   ! The lagged-Fibonacci state (lagfib module variables): initlagfib(myid, iseed), and
   ! get / set of iwlagfib(0:520), jrlagfib, krlagfib.
   subroutine ref_lagfib_init(myidin, iseed, iw, jr, kr) bind(C, name="ref_lagfib_init")
      integer(c_int), value :: myidin, iseed
      integer(c_int), intent(out) :: iw(0:520), jr, kr
      call initlagfib(myidin, iseed)
      iw = iwlagfib
      jr = jrlagfib
      kr = krlagfib
   end subroutine ref_lagfib_init

   subroutine ref_lagfib_get(iw, jr, kr) bind(C, name="ref_lagfib_get")
      integer(c_int), intent(out) :: iw(0:520), jr, kr
      iw = iwlagfib
      jr = jrlagfib
      kr = krlagfib
   end subroutine ref_lagfib_get

   subroutine ref_lagfib_set(iw, jr, kr) bind(C, name="ref_lagfib_set")
      integer(c_int), intent(in) :: iw(0:520)
      integer(c_int), value :: jr, kr
      iwlagfib = iw
      jrlagfib = jr
      krlagfib = kr
   end subroutine ref_lagfib_set

   ! rnds(m, rn) from the current state
   subroutine ref_rnds(m, rn) bind(C, name="ref_rnds")
      integer(c_int), value :: m
      real(c_double), intent(out) :: rn(m)
      call rnds(m, rn)
   end subroutine ref_rnds

   ! z4source / z2source from the current state
   subroutine ref_z4source(be, bo) bind(C, name="ref_z4source")
      real(c_double), intent(out) :: be(6,ntotal,4,2,8), bo(6,ntotal,4,2,8)
      call z4source(be, bo, nvhalf, myid)
   end subroutine ref_z4source

   subroutine ref_z2source(be, bo) bind(C, name="ref_z2source")
      real(c_double), intent(out) :: be(6,ntotal,4,2,8), bo(6,ntotal,4,2,8)
      call z2source(be, bo, nvhalf)
   end subroutine ref_z2source

   ! nsweep sweeps (the loop of sweep, cfgsprops.f90:1973-2003) continuing the current
   ! stream, with the given gaction
   subroutine ref_sweeps(u, coact, gaction, nsweep, nhit) bind(C, name="ref_sweeps")
      real(c_double), intent(inout) :: u(18,ntotal,4,2,16)
      real(c_double), intent(in)    :: coact(4,4,4)
      integer(c_int), intent(in)    :: gaction(4)
      integer(c_int), value         :: nsweep, nhit
      integer(kind=KI) :: isw, ieo, ibl, mu, MRT
      integer(kind=KI) :: ga(4)
      real(kind=KR)    :: c(18,nvhalf)
      MRT = MPI_DOUBLE_PRECISION
      ga = gaction
      do isw = 1, nsweep
         do ieo = 1, 2
            do ibl = 1, 16
               do mu = 1, 4
                  call staple(mu, ieo, ibl, coact, ga, c, u, myid, nn, ldiv, nms, lv, ib, &
                              lbd, iblv, MRT)
                  call phbup(mu, ieo, ibl, c, u, nhit)
               end do
            end do
         end do
      end do
   end subroutine ref_sweeps

   ! average plaquette, aveplaq (qqcd-test/cfgsprops/glue/gaugetools.f90:1436-1474)
   subroutine ref_aveplaq(u, plaq) bind(C, name="ref_aveplaq")
      real(c_double), intent(inout) :: u(18,ntotal,4,2,16)
      real(c_double), intent(out)   :: plaq
      integer(kind=KI) :: MRT
      real(kind=KR)    :: p
      MRT = MPI_DOUBLE_PRECISION
      call aveplaq(p, u, myid, nn, ldiv, nms, lv, ib, lbd, iblv, MRT)
      plaq = p
   end subroutine ref_aveplaq

   ! Gauge-field and propagator files of the reference (GFwrite / GFread,
   ! qqcd-test/cfgsprops/glue/gaugetools.f90:1259-1334; FPwrite / FPread,
   ! qqcd-test/cfgsprops/quark/quark.f90:3379-3481).
   subroutine ref_gfwrite(u, fname, nf, gaction, beta, alat, tad, icfg) bind(C, name="ref_gfwrite")
      real(c_double), intent(inout) :: u(18,ntotal,4,2,16)
      integer(c_int), value         :: nf, icfg
      character(kind=c_char), intent(in) :: fname(nf)
      integer(c_int), intent(in)    :: gaction(4)
      real(c_double), value         :: beta
      real(c_double), intent(in)    :: alat(4), tad(4)
      character(len=512) :: f
      integer(kind=KI) :: ga(4), k
      f = ' '
      do k = 1, nf
         f(k:k) = fname(k)
      end do
      ga = gaction
      call GFwrite(u, f, ga, beta, alat, tad, icfg, myid)
   end subroutine ref_gfwrite

   subroutine ref_gfread(u, fname, nf, gaction, beta, alat, tad, icfg, myidin) bind(C, name="ref_gfread")
      real(c_double), intent(inout) :: u(18,ntotal,4,2,16)
      integer(c_int), value         :: nf
      character(kind=c_char), intent(in) :: fname(nf)
      integer(c_int), intent(out)   :: gaction(4), icfg, myidin
      real(c_double), intent(out)   :: beta, alat(4), tad(4)
      character(len=512) :: f
      integer(kind=KI) :: ga(4), ic, mi, k
      f = ' '
      do k = 1, nf
         f(k:k) = fname(k)
      end do
      call GFread(u, f, ga, beta, alat, tad, ic, mi)
      gaction = ga
      icfg = ic
      myidin = mi
   end subroutine ref_gfread

   subroutine ref_fpwrite(xe, xo, fname, nf, kappa, tmqcd, csw, icsrc, idsrc, icfg) bind(C, name="ref_fpwrite")
      real(c_double), intent(inout) :: xe(6,ntotal,4,2,8), xo(6,ntotal,4,2,8)
      integer(c_int), value         :: nf, icsrc, idsrc, icfg
      character(kind=c_char), intent(in) :: fname(nf)
      real(c_double), value         :: kappa, tmqcd, csw
      character(len=512) :: f
      integer(kind=KI) :: k
      f = ' '
      do k = 1, nf
         f(k:k) = fname(k)
      end do
      call FPwrite(xe, xo, vecblinv, f, kappa, tmqcd, csw, icsrc, idsrc, icfg, myid)
   end subroutine ref_fpwrite

   subroutine ref_fpread(xe, xo, fname, nf, kappa, tmqcd, csw, icsrc, idsrc, icfg, myidin) bind(C, name="ref_fpread")
      real(c_double), intent(inout) :: xe(6,ntotal,4,2,8), xo(6,ntotal,4,2,8)
      integer(c_int), value         :: nf
      character(kind=c_char), intent(in) :: fname(nf)
      real(c_double), intent(out)   :: kappa, tmqcd, csw
      integer(c_int), intent(out)   :: icsrc, idsrc, icfg, myidin
      character(len=512) :: f
      integer(kind=KI) :: a, b, cc, d, k
      f = ' '
      do k = 1, nf
         f(k:k) = fname(k)
      end do
      call FPread(xe, xo, vecblinv, f, kappa, tmqcd, csw, a, b, cc, d)
      icsrc = a
      idsrc = b
      icfg = cc
      myidin = d
   end subroutine ref_fpread

   ! The loop operators of one (sink x, source b) pair on every (ieo, ibleo)
   ! slice: Je/Jo(:,:,:,ieo,ibleo) is loopops' output for that slice
   ! (qqcd-test/cfgsprops/quark/disconloops.f90:7769-7910).
   subroutine ref_loopops(ksub, xe, xo, u, be, bo, Je, Jo) bind(C, name="ref_loopops")
      integer(c_int), value         :: ksub
      real(c_double), intent(inout) :: xe(6,ntotal,4,2,8), xo(6,ntotal,4,2,8)
      real(c_double), intent(inout) :: be(6,ntotal,4,2,8), bo(6,ntotal,4,2,8)
      real(c_double), intent(in)    :: u(18,ntotal,4,2,16)
      real(c_double), intent(out)   :: Je(2,nvhalf,9,2,8), Jo(2,nvhalf,9,2,8)
      real(kind=KR) :: je1(2,nvhalf,9), jo1(2,nvhalf,9)
      integer(kind=KI) :: ieo, ibleo, ks, MRT
      MRT = MPI_DOUBLE_PRECISION
      ks = ksub
      do ibleo = 1, 8
         do ieo = 1, 2
            call loopops(ks, xe, xo, u, be, bo, je1, jo1, ieo, ibleo, gbc, vecbl, vecblinv, &
                         myid, nn, ldiv, nms, lvbc, ib, lbd, iblv, MRT)
            Je(:,:,:,ieo,ibleo) = je1
            Jo(:,:,:,ieo,ibleo) = jo1
         end do
      end do
   end subroutine ref_loopops

   ! Timeslice and momentum sums of loop operators: the momentum factors of
   ! momfacs for origin io(3) (disconloops.f90:8925-9010), then spacesum
   ! (:9012-9124) over every slice into Jtime(2,nt,5,9).
   subroutine ref_spacesum(Je, Jo, io, Jtime) bind(C, name="ref_spacesum")
      real(c_double), intent(in)  :: Je(2,nvhalf,9,2,8), Jo(2,nvhalf,9,2,8)
      integer(c_int), intent(in)  :: io(3)
      real(c_double), intent(out) :: Jtime(2,nt,5,9)
      real(kind=KR), allocatable :: momfac(:,:,:,:,:)
      real(kind=KR) :: je1(2,nvhalf,9), jo1(2,nvhalf,9)
      integer(kind=KI) :: ieo, ibleo, MRT, io3(3)
      MRT = MPI_DOUBLE_PRECISION
      allocate(momfac(nvhalf,2,16,4,9))
      momfac = 0.0_KR
      io3 = io
      call momfacs(io3, momfac, myid, iblv)
      Jtime = 0.0_KR
      do ibleo = 1, 8
         do ieo = 1, 2
            je1 = Je(:,:,:,ieo,ibleo)
            jo1 = Jo(:,:,:,ieo,ibleo)
            call spacesum(Jtime, je1, jo1, ieo, ibleo, momfac, 5, 9, myid, vecbl, vecblinv, MRT)
         end do
      end do
      deallocate(momfac)
   end subroutine ref_spacesum

   ! nsaverage (disconloops.f90:3079-3143) with ESHFES = 1: loopops and
   ! spacesum over every slice with zero momentum factors, currents times kappa.
   subroutine ref_nsaverage(Jtime, u, be, bo, xe, xo, kappa, coact) bind(C, name="ref_nsaverage")
      real(c_double), intent(out)   :: Jtime(2,nt,5,9)
      real(c_double), intent(inout) :: u(18,ntotal,4,2,16)
      real(c_double), intent(inout) :: be(6,ntotal,4,2,8), bo(6,ntotal,4,2,8)
      real(c_double), intent(inout) :: xe(6,ntotal,4,2,8), xo(6,ntotal,4,2,8)
      real(c_double), value         :: kappa
      real(c_double), intent(in)    :: coact(4,4,4)
      integer(kind=KI) :: MRT
      MRT = MPI_DOUBLE_PRECISION
      call nsaverage(Jtime, u, be, bo, xe, xo, kappa, coact, gbc, vecbl, vecblinv, myid, nn, &
                     ldiv, nms, lvbc, ib, lbd, iblv, MRT, 1)
   end subroutine ref_nsaverage
#endif

   ! coact for the isotropic, untadpoled plaquette + Wilson action
   ! (fixaction, qqcd-test/cfgsprops/glue/lattice.f90:358-437)
   subroutine ref_fixaction(beta, coact) bind(C, name="ref_fixaction")
      real(c_double), value       :: beta
      real(c_double), intent(out) :: coact(4,4,4)
      integer(kind=KI) :: gaction(4)
      real(kind=KR)    :: alat(4), tad(4)
      gaction = 1
      alat = 1.0_KR
      tad = 1.0_KR
      call fixaction(gaction, beta, alat, tad, coact)
   end subroutine ref_fixaction

end module refdrv
