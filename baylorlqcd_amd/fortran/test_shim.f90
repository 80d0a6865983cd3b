! test_shim.f90 -- a Fortran host calling the drop-in with the reference's
! argument lists (8^3x16, one process).  Reads u.bin, be.bin, bo.bin (raw
! float64 in the reference order; sink_e.bin, sink_o.bin for the loop
! operators) from the working directory, writes the outputs next to them;
! tests/test_gpu_fortran.py and tests/test_fortran_files.py compare them with
! the oracle and the reference's files.
!
! The first argument selects sections (default: all): "ops" (operators,
! solves, loop operators), "noise" (z4source / z2source from the lagfib
! stream; needs "ops" first), "files" (GFwrite / GFread / FPwrite / FPread of
! the inputs, host only: no device is touched), "sweep" (two pseudo-heatbath
! sweeps of the links from the stream of initlagfib(0, 3), coact read from
! coact_hb.bin), "smear" (smear of be/bo with asmear 0.25, nsmear 3; twMshift
! and gamma5mult at kappa = (0.15, 0.02)).
program test_shim
   use qqcd_hip
   use gmresrhs, only: isignal
   use latdims, only: nx, ny, nz, nt
   use lagfib, only: initlagfib
   implicit none
   integer, parameter :: nvhalf = nx*ny*nz*nt/32
   integer, parameter :: nbmax = max(2*nvhalf/nx, 2*nvhalf/ny, 2*nvhalf/nz, 2*nvhalf/nt)
   integer, parameter :: ntotal = nvhalf + nbmax
   real(kind=8), allocatable :: u(:,:,:,:,:), be(:,:,:,:,:), bo(:,:,:,:,:), g(:,:,:,:,:), h(:,:,:,:,:)
   real(kind=8), allocatable :: xe(:,:,:,:,:,:), xo(:,:,:,:,:,:), ps_e(:,:,:,:,:), ps_o(:,:,:,:,:)
   real(kind=8) :: coact(4,4,4), kappa(2), adb(2), gdummy(18,1,1,1,1)
   real(kind=8), allocatable :: jse(:,:,:), jso(:,:,:), jtime(:,:,:,:)
   integer :: bc(4), vecbl(2,8), vecblinv(2,16), nn(4,2), nms(4), iblv(16,4), ibl, mu, itercount, vectype
   integer, allocatable :: lvbc(:,:,:), ib(:,:,:,:)
   logical :: ldiv(4), lbd(16,4)
   character(len=8) :: rwdir(1)
   character(len=32) :: sel
   logical :: do_ops, do_noise, do_files, do_sweep, do_smear, do_bnd
   integer :: src(4)

   allocate(u(18,ntotal,4,2,16), be(6,ntotal,4,2,8), bo(6,ntotal,4,2,8), g(6,ntotal,4,2,8), h(6,ntotal,4,2,8))
   allocate(xe(6,ntotal,4,2,8,1), xo(6,ntotal,4,2,8,1), ps_e(6,ntotal,4,2,8), ps_o(6,ntotal,4,2,8))
   allocate(lvbc(nvhalf,4,2), ib(nbmax,4,2,2))
   call readbin("u.bin", u, size(u))
   call readbin("be.bin", be, size(be))
   call readbin("bo.bin", bo, size(bo))
   call get_command_argument(1, sel)
   do_ops = len_trim(sel) == 0 .or. index(sel, "ops") > 0
   do_noise = len_trim(sel) == 0 .or. index(sel, "noise") > 0
   do_files = len_trim(sel) == 0 .or. index(sel, "files") > 0
   do_sweep = len_trim(sel) == 0 .or. index(sel, "sweep") > 0
   do_smear = len_trim(sel) == 0 .or. index(sel, "smear") > 0
   do_bnd = len_trim(sel) == 0 .or. index(sel, "boundary") > 0
   ! single-process geometry tables (setblk / gblchker of the reference)
   vecbl(1,:) = [1, 4, 6, 7, 10, 11, 13, 16]
   vecbl(2,:) = [2, 3, 5, 8, 9, 12, 14, 15]
   do ibl = 1, 16
      vecblinv(1,ibl) = 1 + mod(popcnt(ibl - 1), 2)
      vecblinv(2,ibl) = (ibl + 1)/2
      do mu = 1, 4
         iblv(ibl,mu) = 1 + ieor(ibl - 1, ishft(1, mu - 1))
         lbd(ibl,mu) = iand(ishft(ibl - 1, -(mu - 1)), 1) == 0
      end do
   end do
   nn = 0
   nms = 0
   lvbc = 0
   ib = 0
   ldiv = .false.
   bc = [1, 1, 1, -1]
   rwdir(1) = "./"
   coact = 0.0d0
   coact(3,1,:) = 1.0d0
   coact(3,2,:) = 1.0d0
   gdummy = 0.0d0

   if (do_ops) call ops()
   if (do_noise) call noise()
   ! the reference's files: the links (re-unitarised in place by GFwrite) and the input spinor
   ! pair as a propagator, read back into fresh arrays
   if (do_files) call shim_files()
   if (do_sweep) call shim_sweep()
   if (do_smear) call shim_smear()
   if (do_bnd) call shim_boundary()
   if (do_ops .or. do_noise .or. do_sweep .or. do_smear .or. do_bnd) call qqcd_shutdown()
   print *, "test_shim done"

contains

   ! each step is printed and appended to steps.log (opened, written and closed
   ! every time), so a run that stops anywhere names its last step
   subroutine step(name)
      character(len=*), intent(in) :: name
      integer :: unit
      print *, "step ", name
      flush(6)
      open(newunit=unit, file="steps.log", position="append", action="write", form="formatted")
      write(unit, "(a)") name
      close(unit)
   end subroutine step

   subroutine ops()
   call step("operators")
   call Hsingle(g, u, be, 0, coact, bc, 2, vecbl, vecblinv, 0, nn, ldiv, nms, lvbc, ib, lbd, iblv, 0)
   call writebin("hsingle.bin", g, size(g))
   kappa = [0.156d0, 0.005d0]
   call Hdbletm(h, u, gdummy, be, 0, coact, kappa, -1, bc, vecbl, vecblinv, 0, nn, ldiv, nms, lvbc, ib, lbd, iblv, 0)
   call writebin("hdbletm.bin", h, size(h))
   call Hdouble(h, u, be, 1, coact, 0.15d0, bc, vecbl, vecblinv, 0, nn, ldiv, nms, lvbc, ib, lbd, iblv, 0)
   call writebin("hdouble_dag.bin", h, size(h))
   g = be
   call gam5x(g)
   call writebin("gam5x.bin", g, size(g))
   call vecdot(be, bo, adb, 0)
   call writebin("vecdot.bin", adb, 2)

   call step("fermprop")
   ! config-1 solve: point source at the origin, GMRES-DR(40,10), Wilson kappa 0.150
   src = [1, 1, 1, 1]
   call pointsource(ps_e, ps_o, src, 1, 1, 0, iblv, vecblinv)
   kappa = [0.150d0, 0.0d0]
   vectype = 40
   isignal = 1
   call fermprop(rwdir, ps_e, ps_o, -1, kappa, 0.0d0, 5, coact, bc, 1.0d-10, 0, 1.0d0, u, vectype, 10, xe, xo, &
                 0, nn, ldiv, nms, lvbc, lvbc, ib, lbd, iblv, vecbl, vecblinv, 0, 0, 0, 0)
   call writebin("xe.bin", xe, size(xe))
   call writebin("xo.bin", xo, size(xo))
   ! psc: the same source again with isignal=2 runs GMRES-Proj over the deflation space
   isignal = 2
   vectype = 40
   call fermprop(rwdir, ps_e, ps_o, -1, kappa, 0.0d0, 6, coact, bc, 1.0d-10, 0, 1.0d0, u, vectype, 10, xe, xo, &
                 0, nn, ldiv, nms, lvbc, lvbc, ib, lbd, iblv, vecbl, vecblinv, 0, 0, 0, 0)
   call writebin("xe_proj.bin", xe, size(xe))
   call step("mixed")
   ! mixed precision on the even-odd right-hand side of the first solve
   g = 0.0d0
   call qqcd_solve_mixed_f(be, g, [40, 16], 1.0d-10, 2, 12, 0, u, kappa, coact, bc, itercount)
   call writebin("x_mixed.bin", g, size(g))
   call step("loops")
   call readbin("sink_e.bin", g, size(g))
   call readbin("sink_o.bin", h, size(h))
   ! loop operators of sink g/h and source be/bo: slice (ieo, ibleo) = (2, 3)
   ! of loopops with ksub = 1, and nsaverage at kappa 0.15
   allocate(jse(2,nvhalf,9), jso(2,nvhalf,9), jtime(2,nt,5,9))
   call loopops(1, g, h, u, be, bo, jse, jso, 2, 3, bc, vecbl, vecblinv, 0, nn, ldiv, nms, lvbc, ib, lbd, iblv, 0)
   call writebin("loopops_e.bin", jse, size(jse))
   call writebin("loopops_o.bin", jso, size(jso))
   call nsaverage(jtime, u, be, bo, g, h, 0.15d0, coact, bc, vecbl, vecblinv, 0, nn, ldiv, nms, lvbc, ib, lbd, &
                  iblv, 0, 1)
   call writebin("nsaverage.bin", jtime, size(jtime))
   end subroutine ops

   subroutine noise()
   call step("noise")
   ! two Z4 sources of the stream of initlagfib(0, 5): the second one is written out
   call initlagfib(0, 5)
   call z4source(g, h, nvhalf, 0)
   call z4source(g, h, nvhalf, 0)
   call writebin("z4_e.bin", g, size(g))
   call writebin("z4_o.bin", h, size(h))
   ! the next source of the stream as Z2 noise
   call z2source(g, h, nvhalf)
   call writebin("z2_e.bin", g, size(g))
   call writebin("z2_o.bin", h, size(h))
   end subroutine noise

   subroutine shim_sweep()
      real(kind=8) :: cohb(4,4,4)
      call step("sweep")
      call readbin("coact_hb.bin", cohb, size(cohb))
      call initlagfib(0, 3)
      call sweep([1, 1, 1, 1], cohb, 1, u, 0, nn, ldiv, nms, lvbc, ib, lbd, iblv, 0)
      call sweep([1, 1, 1, 1], cohb, 1, u, 0, nn, ldiv, nms, lvbc, ib, lbd, iblv, 0)
      call writebin("u_swept.bin", u, size(u))
      call write_state("lagfib_after.bin")
   end subroutine shim_sweep

   subroutine shim_smear()
      real(kind=8), allocatable :: gw(:,:,:,:,:,:), hw(:,:,:,:,:,:)
      call step("smear")
      call smear(g, h, u, be, bo, 0.25d0, 3, bc, vecbl, vecblinv, 0, nn, ldiv, nms, lvbc, ib, lbd, iblv, 0)
      call writebin("smear_e.bin", g, size(g))
      call writebin("smear_o.bin", h, size(h))
      allocate(gw(6,ntotal,4,2,8,2), hw(6,ntotal,4,2,8,2))
      gw(:,:,:,:,:,1) = be
      gw(:,:,:,:,:,2) = bo
      kappa = [0.15d0, 0.02d0]
      call twMshift(hw, u, gdummy, gw, 0, coact, kappa, -1, bc, vecbl, vecblinv, 0, nn, ldiv, nms, lvbc, ib, lbd, &
                    iblv, 0)
      call writebin("twm.bin", hw, size(hw))
      call gamma5mult(g, u, gdummy, be, 0, coact, kappa, -1, bc, vecbl, vecblinv, 0, nn, ldiv, nms, lvbc, ib, lbd, &
                      iblv, 0)
      call writebin("g5m.bin", g, size(g))
   end subroutine shim_smear

   ! The reference's argument lists of psc, gmresproject (with xshift, isignal, mvp, gdr) and
   ! fermpropR, and GFread of a second configuration into the array the operators used before.
   subroutine shim_boundary()
      real(kind=8), allocatable :: xs(:,:,:,:,:,:), gdr(:,:), p1e(:,:,:,:,:), p1o(:,:,:,:,:), p2e(:,:,:,:,:), &
                                   p2o(:,:,:,:,:), x1(:,:,:,:,:)
      real(kind=8) :: beta, alat(4), tad(4), meta(3)
      integer :: mvp, it1, it2, gaction(4), icfg, myidin
      call step("boundary")
      allocate(xs(6,ntotal,4,2,8,2), gdr(2000,1), p1e(6,ntotal,4,2,8), p1o(6,ntotal,4,2,8), &
               p2e(6,ntotal,4,2,8), p2o(6,ntotal,4,2,8), x1(6,ntotal,4,2,8))
      call pointsource(p1e, p1o, [1, 1, 1, 1], 1, 1, 0, iblv, vecblinv)
      call pointsource(p2e, p2o, [3, 2, 4, 5], 2, 3, 0, iblv, vecblinv)
      call writebin("proj_phi_e.bin", p2e, size(p2e))
      kappa = [0.150d0, 0.0d0]
      ! first right-hand side: psc with isignal = 1 (GMRES-DR(40,10)), the deflation space stays
      isignal = 1
      x1 = 0.0d0
      call psc(rwdir, p1e, x1, [40, 10], 1.0d-11, 0, it1, u, gdummy, -1, kappa, coact, bc, vecbl, vecblinv, 0, &
               nn, ldiv, nms, lvbc, ib, lbd, iblv, 0, 0)
      ! a second one through gmresproject directly, GMRES-Proj(30,10), two shifts
      call step("gmresproject")
      xs = 7.0d0
      gdr = -1.0d0
      mvp = -5
      call gmresproject(rwdir, p2e, xs, [30, 10], 1.0d-10, 0, it2, u, gdummy, -1, kappa, coact, bc, vecbl, &
                        vecblinv, 0, nn, ldiv, nms, lvbc, ib, lbd, iblv, 0, 0, 2, mvp, gdr)
      call writebin("proj_x.bin", xs, size(xs))
      call writebin("proj_gdr.bin", gdr, size(gdr))
      meta = [dble(it1), dble(it2), dble(mvp)]
      call writebin("proj_meta.bin", meta, 3)
      ! isignal = 1: the right-hand side is the resident beg (b of the previous gmresproject)
      xs = 7.0d0
      mvp = -5
      call gmresproject(rwdir, p1e, xs, [30, 10], 1.0d-10, 0, it2, u, gdummy, -1, kappa, coact, bc, vecbl, &
                        vecblinv, 0, nn, ldiv, nms, lvbc, ib, lbd, iblv, 0, 0, 1, mvp, gdr)
      call writebin("proj_beg_x.bin", xs(:,:,:,:,:,1), size(xs(:,:,:,:,:,1)))
      ! fermpropR with Wilson r = -1 on the config-1 point source, GMRES-DR(40,10)
      call step("fermpropR")
      vectype = 40
      isignal = 1
      xe = 0.0d0
      xo = 0.0d0
      call fermpropR(rwdir, p1e, p1o, -1, kappa, 0.0d0, 5, coact, bc, 1.0d-11, 0, 1.0d0, u, vectype, 10, xe, xo, &
                     0, nn, ldiv, nms, lvbc, lvbc, ib, lbd, iblv, vecbl, vecblinv, 0, 0, 0, 0, -1)
      call writebin("fermpropR_xe.bin", xe, size(xe))
      call writebin("fermpropR_xo.bin", xo, size(xo))
      ! a second configuration (cfg2.gf, in the reference's file format) read into u, which the
      ! operators above used and the library holds on the device: Hsingle must see the new links
      call step("gfread")
      g = 0.0d0
      call Hsingle(g, u, be, 0, coact, bc, 2, vecbl, vecblinv, 0, nn, ldiv, nms, lvbc, ib, lbd, iblv, 0)
      call GFread(u, "cfg2.gf", gaction, beta, alat, tad, icfg, myidin)
      call writebin("gfread_u2.bin", u, size(u))
      call Hsingle(g, u, be, 0, coact, bc, 2, vecbl, vecblinv, 0, nn, ldiv, nms, lvbc, ib, lbd, iblv, 0)
      call writebin("hsingle_cfg2.bin", g, size(g))
   end subroutine shim_boundary

   subroutine write_state(name)
      use lagfib, only: iwlagfib, jrlagfib, krlagfib
      character(len=*), intent(in) :: name
      integer :: unit
      open(newunit=unit, file=name, access="stream", form="unformatted", status="replace", action="write")
      write(unit) iwlagfib, jrlagfib, krlagfib
      close(unit)
   end subroutine write_state

   subroutine shim_files()
      real(kind=8), allocatable :: uw(:,:,:,:,:), ur(:,:,:,:,:), fe(:,:,:,:,:), fo(:,:,:,:,:)
      real(kind=8) :: beta, alat(4), tad(4), fk, ft, fc, meta(13)
      integer :: gaction(4), icfg, myidin, icsrc, idsrc
      call step("files")
      allocate(uw(18,ntotal,4,2,16), ur(18,ntotal,4,2,16), fe(6,ntotal,4,2,8), fo(6,ntotal,4,2,8))
      uw = u
      ur = 0.0d0
      call GFwrite(uw, "shim.gf", [1, 1, 1, 2], 6.0d0, [1.0d0, 1.0d0, 1.0d0, 0.5d0], [0.9d0, 0.9d0, 0.9d0, 1.0d0], &
                   7, 0)
      call GFread(ur, "shim.gf", gaction, beta, alat, tad, icfg, myidin)
      call writebin("gfread_u.bin", ur, size(ur))
      call writebin("gfwrite_u.bin", uw, size(uw))
      call FPwrite(be, bo, vecblinv, "shim.fp", 0.156d0, 0.005d0, 0.0d0, 2, 3, 7, 0)
      fe = 0.0d0
      fo = 0.0d0
      call FPread(fe, fo, vecblinv, "shim.fp", fk, ft, fc, icsrc, idsrc, icfg, myidin)
      call writebin("fpread_e.bin", fe, size(fe))
      call writebin("fpread_o.bin", fo, size(fo))
      meta = [dble(gaction), beta, alat(4), tad(1), fk, ft, fc, dble(icsrc), dble(idsrc), dble(icfg)]
      call writebin("files_meta.bin", meta, size(meta))
   end subroutine shim_files

   subroutine readbin(name, a, n)
      character(len=*), intent(in) :: name
      integer, intent(in) :: n
      real(kind=8), intent(out) :: a(n)
      integer :: unit
      open(newunit=unit, file=name, access="stream", form="unformatted", status="old", action="read")
      read(unit) a
      close(unit)
   end subroutine readbin

   subroutine writebin(name, a, n)
      character(len=*), intent(in) :: name
      integer, intent(in) :: n
      real(kind=8), intent(in) :: a(n)
      integer :: unit
      open(newunit=unit, file=name, access="stream", form="unformatted", status="replace", action="write")
      write(unit) a
      close(unit)
   end subroutine writebin

end program test_shim
