/*
 * qqcd_hip.h -- C ABI of the MI355X Wilson / twisted-mass solve path.
 *
 * Drop-in boundary for the Fortran host of BaylorLQCD "qqcd": the entry
 * points below replace the reference's module procedures of the Dirac
 * operator and inverter layer.  Arrays are the reference's own host arrays
 * in its own storage order (Fortran, column-major):
 *   spinor   real(8) g(6, ntotal, 4, 2, 8)     one global checkerboard
 *   links    real(8) u(18, ntotal, 4, 2, 16)
 *   coact    real(8) coact(4, 4, 4)            only coact(3,1:2,mu) is read
 *   bc       integer(4) bc(4)                   1, -1 or 0 per direction
 * with nvhalf = V/32, ntotal = nvhalf + max_mu(2 nvhalf / n_mu) for one
 * process (qqcd-test/cfgsprops/basics.f90:20-30).  Rows nvhalf+1..ntotal are
 * halo scratch in the reference and carry no data here.
 *
 * One context per process, one HIP stream, not thread-safe; every call is
 * synchronous at return.  Every entry point returns 0 on success and a
 * negative code on error; qqcd_last_error() returns the message.
 *
 * The geometry arguments of the reference procedures (vecbl, vecblinv,
 * myid, nn, ldiv, nms, lvbc, ib, lbd, iblv, MRT) are not passed: the
 * Fortran shim checks that they describe one process with the canonical
 * block tables and drops them.
 */
#ifndef QQCD_HIP_H
#define QQCD_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- life cycle ------------------------------------------------------- */

/* Set up the lattice n[4] = (nx, ny, nz, nt) with boundary conditions bc[4]
 * on HIP device `device`.  Replaces the single-process part of initlat
 * (qqcd-test/cfgsprops/glue/lattice.f90:33-59).  Requires every n even and
 * nx, ny, nz, nt divisible by 4 (the reference's rule nx/4/npx integer,
 * qqcd-test/user/latdims.f90:11-13). */
int qqcd_init(const int n[4], const int bc[4], int device);
int qqcd_finalize(void);
const char *qqcd_last_error(void);
/* sizes of the reference arrays: nvhalf, ntotal */
int qqcd_ref_sizes(long *nvhalf, long *ntotal);

/* Upload the link field u (reference order) and coact, and make it the
 * resident field.  Operator and solver calls that pass the same u address
 * reuse the resident copy; call this again after changing u in place
 * (the reference reads u afresh in every Hsingle,
 * qqcd-test/cfgsprops/quark/diracops.f90:554). */
int qqcd_gauge_load(const double *u, const double *coact);
/* Same, from canonical order U[((s*4 + mu)*3 + r)*3 + c] complex (re, im),
 * s = x + nx*(y + ny*(z + nz*t)); for hosts that do not hold the reference
 * layout. */
int qqcd_gauge_load_canonical(const double *U, const double *coact);

/* Seeded SU(3) field generated on the device and made resident (no host
 * array): kind 0 Haar-random, kind 1 reunitarised 1 + eps*Gaussian.  The
 * project's own generator; the reference reads fields from disk or makes
 * them with its heatbath. */
int qqcd_gauge_random(int kind, double eps, unsigned long long seed, const double *coact);
/* Quenched pseudo-heatbath at coupling beta on the device (Cabibbo-Marinari
 * with Kennedy-Pendleton SU(2) updates, the algorithm of the reference's
 * sweep: qqcd-test/cfgsprops/cfgsprops.f90:1973-2003,
 * qqcd-test/cfgsprops/glue/heatbath.f90:33-467), nsweep sweeps from a cold
 * (start 0) or hot (start 1, Haar) field, or continuing the resident one
 * (start -1).  A counter-based generator replaces the reference's
 * lagged-Fibonacci stream (same distribution, different trajectory). */
int qqcd_gauge_heatbath(double beta, int nsweep, unsigned long long seed, int start, const double *coact);
/* Average plaquette (1/3) Re Tr U_P of the resident field, as the
 * reference's aveplaq (qqcd-test/cfgsprops/glue/gaugetools.f90:1436). */
int qqcd_gauge_plaquette(double *plaq);
/* Copy the resident FP64 field back in canonical order (see above). */
int qqcd_gauge_get_canonical(double *U);

/* The resident FP64 links in the reference layout u(18, ntotal, 4, 2, 16)
 * (rows nvhalf+1..ntotal zero): hands a device-generated field (heatbath,
 * random) back to a reference host. */
int qqcd_gauge_get_ref(double *u);

/* ---- operators (qqcd-test/cfgsprops/quark/diracops.f90) ---------------- */

/* Hsingle(gout,u,gin,idag,coact,bc,gblclr,...), diracops.f90:554-599:
 * gout = H_eo gin (gblclr=1) or H_oe gin (gblclr=2); all of gout is written
 * (zero outside rows 1..nvhalf).  idag=1 applies H^dagger = g5 H g5. */
int qqcd_hsingle(double *gout, const double *u, const double *gin, int idag, const double *coact,
                 const int *bc, int gblclr);

/* Hdouble(he,u,ge,idag,coact,kappa,bc,...), diracops.f90:264-346:
 * he = (1/kappa^2 - H_eo H_oe) ge on the gblclr=1 checkerboard. */
int qqcd_hdouble(double *he, const double *u, const double *ge, int idag, const double *coact, double kappa,
                 const int *bc);

/* Hdbletm(he,u,GeeGooinv,ge,idag,coact,kappa,iflag,bc,...),
 * diracops.f90:435-550, Wilson twisted mass (iflag = -1), a = 2 kappa(1)
 * kappa(2): he = (1 + a i g5) ge / kappa(1)^2
 *               - H_eo (1 - a i g5) H_oe ge / (1 + a^2);
 * kappa(2) < 1e-8 (signed test, :467) falls back to Hdouble.  The clover
 * case (iflag = -2) is not supported and returns an error. */
int qqcd_hdbletm(double *he, const double *u, const double *ge, int idag, const double *coact,
                 const double kappa[2], int iflag, const int *bc);

/* smear (qqcd-test/cfgsprops/quark/diracops.f90:712-856): Gaussian smearing,
 * nsmear times h <- (1 - 6 asmear) h + asmear sum_{mu=1..3} (U_mu(x) h(x+mu)
 * + U_mu(x-mu)^dag h(x-mu)) on both checkerboards (ge, go in; he, ho out),
 * bc(mu) on hops across the spatial edge. */
int qqcd_smear(double *he, double *ho, const double *u, const double *ge, const double *go, double asmear,
               int nsmear, const int *bc);

/* twMshift (diracops.f90:1863-1995), iflag = -1: on both checkerboards
 * (hwhole, gwhole = (6, ntotal, 4, 2, 8, 2)) h = g + (-i g5)/(2 kappa mu)
 * (g + kappa H g_other), kappa = (kappa, mu). */
int qqcd_twmshift(double *hwhole, const double *u, const double *gwhole, int idag, const double *coact,
                  const double kappa[2], int iflag, const int *bc);

/* gamma5mult (diracops.f90:1750-1860), iflag = -1: he = (-i g5)/(2 kappa mu)
 * times the Wilson-twisted Schur operator of Hdbletm on ge. */
int qqcd_gamma5mult(double *he, const double *u, const double *ge, int idag, const double *coact,
                    const double kappa[2], int iflag, const int *bc);

/* gammamult(evenIN,oddIN,evenOUT,oddOUT,mu,ieo,ibleo), diracops.f90:105-195:
 * gamma_mu (mu = 1..5) on rows 1..nvhalf of slice (ieo, ibleo) of both
 * arrays; the rest of the outputs is left unchanged. */
int qqcd_gammamult(const double *even_in, const double *odd_in, double *even_out, double *odd_out, int mu,
                   int ieo, int ibleo);

/* gam5x(a), qqcd-test/cfgsprops/basics.f90:408-454: a = gamma_5 a in place on
 * rows 1..nvhalf of every slice. */
int qqcd_gam5x(double *a);

/* vecdot(a,b,adb,MRT2), qqcd-test/cfgsprops/quark/inverters.f90:13485-13530:
 * adb = sum over rows 1..nvhalf of conj(a) b, as (re, im). */
int qqcd_vecdot(const double *a, const double *b, double adb[2]);

/* pointsource(be,bo,src,icsrc,idsrc,...), qqcd-test/cfgsprops/quark/
 * quark.f90:46-108: unit source at 1-based global site src[4], colour icsrc,
 * Dirac idsrc; be and bo are zeroed first. */
int qqcd_pointsource(double *be, double *bo, const int src[4], int icsrc, int idsrc);

/* ---- solvers (qqcd-test/cfgsprops/quark/inverters.f90) ----------------- */

/* Common arguments: phi/b the even-odd right-hand side, x the solution
 * (input: initial guess where the reference uses one), kappa[2] = (kappa,
 * mu), iflag = -1.  Convergence is the reference's: ||r|| / ||r0|| < resmax
 * on the even-odd system with the true residual r = phi - M x recomputed in
 * FP64 at every cycle end.  *itercount receives the number of operator
 * applications in the Krylov iterations, as the reference counts them. */

/* gmresdr(rwdir,phi,x,GMRES,resmax,itermin,itercount,u,GeeGooinv,iflag,
 * kappa,coact,bc,...), inverters.f90:2871-3572: GMRES-DR(m = GMRES[0],
 * k = GMRES[1]), x0 as given.  Keeps the deflation space (V_{k+1}, Hbar_k)
 * of the last cycle resident for qqcd_gmresproject / qqcd_psc. */
int qqcd_gmresdr(const double *phi, double *x, const int GMRES[2], double resmax, int itermin, int *itercount,
                 const double *u, const double *coact, const double kappa[2], int iflag, const int *bc);

/* gmresproject(rwdir,b,xshift,GMRES,resmax,itermin,itercount,u,GeeGooinv,
 * iflag,kappa,coact,bc,...,MRT,MRT2,isignal,mvp,gdr), inverters.f90:
 * 7183-7271 (body to :8304): GMRES-Proj with cycle length GMRES[0] over the
 * resident deflation space of k = GMRES[1] vectors, x0 = 0.
 * xshift: nshift reference-order spinors, all zeroed (:7347), the solution
 * in the first.  isignal == 1: the right-hand side is the resident beg
 * (the last deflation vector after psc's GMRES-DR, :4437; b after a
 * gmresproject with isignal != 1, :7438-7443), b is not read.
 * mvp: operator applications of this call (reset to 0 first, :7385).
 * gdr: column 1 of the reference's gdr(:,:), ldgdr entries: gdr[mvp_j - 1]
 * = the least-squares residual norm after the Arnoldi step that was
 * operator application mvp_j (:7981); other entries are left as they are. */
int qqcd_gmresproject(const double *b, double *xshift, int nshift, const int GMRES[2], double resmax, int itermin,
                      int *itercount, const double *u, const double *coact, const double kappa[2], int iflag,
                      const int *bc, int isignal, int *mvp, double *gdr, int ldgdr);

/* psc(...), inverters.f90:4310-4465: isignal = 1 runs GMRES-DR(m, k) and
 * keeps its deflation space; otherwise GMRES-Proj(m-k, k) with the
 * tolerance the reference hard-codes there (1e-8, :4442) unless
 * resmax_proj > 0 is given. */
int qqcd_psc(const double *b, double *x, const int GMRES[2], double resmax, int itermin, int *itercount,
             const double *u, const double *coact, const double kappa[2], int iflag, const int *bc, int isignal,
             double resmax_proj);

/* bicgstab(...), inverters.f90:302-487, from the given x. */
int qqcd_bicgstab(const double *phi, double *x, double resmax, int itermin, int *itercount, const double *u,
                  const double *coact, const double kappa[2], int iflag, const int *bc);

/* fermprop(...) for nkappa = -1 (Wilson / twisted), qqcd-test/cfgsprops/
 * quark/quark.f90:458-842: even-odd source preparation, solve with
 * inverter 0 (BiCGStab from the constant start of :632-640 when
 * vectype == 1, else from xe), 5 (GMRES-DR) or 6 (psc with isignal), and
 * odd-site reconstruction.  nGMRES = (m, k).  ntmqcd rotations are not
 * applied (ntmqcd = 0). */
int qqcd_fermprop(const double *be, const double *bo, const double kappa[2], int inverter, const double *coact,
                  const int *bc, double resmax, int itermin, const int nGMRES[2], int vectype, const double *u,
                  double *xe, double *xo, int isignal, int *itercount);

/* Mixed-precision solve of the even-odd system phi (gblclr=1 array) to
 * ||phi - M x|| < tol ||phi||: FP64 solution, true residuals and reliable
 * updates; inner GMRES-DR(GMRES[0], GMRES[1]) and then GMRES-Proj over its
 * deflation space on a basis of basis_prec (0 FP64, 1 FP32, 2 FP16: the
 * basis and deflation vectors stored as FP16 fields with per-site norms,
 * every new vector formed and orthogonalised in FP32) with the
 * operator in (op_prec, op_recon) storage (0 FP64, 1 FP32, 2 FP16; 18 or 12
 * link reals).  inner_tol <= 0 picks a default per op_prec.  keep_deflation
 * = 1 reuses the deflation space of the previous mixed solve (same gauge
 * field and kappa): the reference's first-RHS / later-RHS split of psc.  No
 * counterpart in the reference, which is FP64 only. */
int qqcd_solve_mixed(const double *phi, double *x, const int GMRES[2], double tol, int basis_prec, int op_prec,
                     int op_recon, double inner_tol, int keep_deflation, const double *u, const double *coact,
                     const double kappa[2], const int *bc, int *itercount);

/* The log lines of the last gmresdr / psc / bicgstab / fermprop call, as the
 * reference's solvers append them to CFGSPROPS.LOG with format
 * (a12,i9,es17.10) (inverters.f90:3285-3291 "gmresdr-norm", matrix-vector
 * products, ||r||/||r0|| per true residual; :476-480 "bicgstab", iteration,
 * ||r||^2/||r0||): *n lines, the first cap of them as 12-character blank-
 * padded labels (not terminated), counts and values. */
int qqcd_solve_log(int cap, int *n, char *labels, int *counts, double *values);

/* ---- many right-hand sides on one gauge field (disconnected loops) ------ */

/* Z4 noise source {1, -1, i, -i} * scale on every (site, spin, colour) of be
 * (gblclr=1) and bo (gblclr=2), either may be NULL.  Replaces z4source
 * (qqcd-test/cfgsprops/quark/disconloops.f90:4359, active branch
 * :4518-4569), which maps one uniform number per component to the four
 * values with the same thresholds; the uniform numbers here come from a
 * counter-based generator keyed by (seed, rhs, global site, component)
 * instead of the reference's lagged-Fibonacci stream rnds. */
int qqcd_z4source(double *be, double *bo, unsigned long long seed, long rhs, double scale);
/* Z2 noise {1, -1} * scale in the real part of every (site, spin, colour):
 * replaces z2source (qqcd-test/cfgsprops/quark/disconloops.f90:4583-4649,
 * 1 for a uniform number above 0.5, else -1), with the uniform numbers of
 * qqcd_z4source's generator. */
int qqcd_z2source(double *be, double *bo, unsigned long long seed, long rhs, double scale);

/* Solve nrhs systems M x_r = b_r on one gauge field, as the reference's noise
 * loops call fermprop(inverter=6) once per source (testFUNC / discon,
 * qqcd-test/cfgsprops/quark/disconloops.f90:2460-2511, 3566; psc
 * qqcd-test/cfgsprops/quark/inverters.f90:4310-4465): source 1 by the mixed
 * GMRES-DR(GMRES[0], GMRES[1]) keeping its deflation space, sources 2..nrhs
 * together by the batched GMRES-Proj(GMRES[0]-GMRES[1]) over it (FP32
 * right-hand-side-innermost fields, FP64 reliable updates), each to
 * ||b - M x|| < resmax ||b|| on its even-odd system.  be/bo/xe/xo: arrays of
 * nrhs pointers to (6,ntotal,4,2,8) arrays; itercount[r]: Krylov steps
 * spent on source r; rel_res[r] (may be NULL): whole-lattice
 * ||b - M x|| / ||b||. */
int qqcd_fermprop_multi(int nrhs, const double *const *be, const double *const *bo, const double kappa[2],
                        const double *coact, const int *bc, double resmax, const int GMRES[2], const double *u,
                        double *const *xe, double *const *xo, int *itercount, double *rel_res);

/* The noise loop on the device: nrhs Z4 sources (qqcd_z4source with seed,
 * rhs = 0..nrhs-1, scale), solved as in qqcd_fermprop_multi, and the local
 * trace estimators of the disconnected loops (loopops / spacesum,
 * qqcd-test/cfgsprops/quark/disconloops.f90:7769-7910, 9012-9124) for the
 * 16 Dirac matrices Gamma_n (n = 0: 1; 1..4: gamma_mu; 5: gamma_5; 6..9:
 * gamma_mu gamma_5; 10..15: sigma_munu = (i/2)[gamma_mu, gamma_nu] for
 * (1,2),(1,3),(1,4),(2,3),(2,4),(3,4)):
 *   traces[(t*16 + n)*2 + {re, im}] = (1/(nrhs scale^2)) sum_r sum_{x, x_4 = t} eta_r(x)^dag Gamma_n x_r(x),
 * an estimate of Tr(Gamma_n M^-1) on timeslice t.  loops (optional, may be
 * NULL): the reference's own nine loop operators of each source (as
 * qqcd_loopsum with ksub, io) averaged the same way, currents times kappa[0]
 * as nsaverage does (disconloops.f90:3079-3143):
 *   loops[((t*5 + m)*9 + op)*2 + {re, im}], op = J_1..J_4, scalar, V_1..V_4,
 * m = momentum 1..5 (io NULL: momentum 1 only, m > 0 left at zero).
 * u may be NULL to use the resident gauge field.  nb = 0 lets the autotuner
 * pick the RHS block.
 * res_full[r], iters[r]: per source; times = {first solve, batched solves,
 * total} seconds; stats[8] = {outer steps, GMRES cycles, batched operator
 * applications, graph replays, sources finished by the fallback, RHS block,
 * tile order, outer steps of the first solve}.  method 0: the batched path
 * above; 1: one source after the other in plain FP64 (GMRES-DR, then
 * GMRES-Proj per later source: the reference's psc sequence on the GPU);
 * 2: one source after the other with the single-RHS mixed solver.  No
 * reference counterpart as one call (the reference loops over fermprop). */
int qqcd_noise_loop(int nrhs, unsigned long long seed, double scale, const double kappa[2], const double *coact,
                    const int *bc, const double *u, const int GMRES[2], double tol, int nb, int method,
                    double *traces, int ksub, const int *io, double *loops, double *res_full, int *iters,
                    double *times, int *stats);

/* qqcd_noise_loop with the sources of the reference's stream, replacing the
 * per-source loop "call z4source; call fermprop" of nsaverage
 * (qqcd-test/cfgsprops/quark/disconloops.f90:3032-3053): source r is the Z4
 * noise z4source (disconloops.f90:4359-4571, drawing through rnds,
 * lagfib.f90:219-259) makes on its r-th call from the state (iw[521], jr,
 * kr) (384 nvhalf draws each, even checkerboard first), times scale.  The
 * state is advanced past the nrhs sources.  Other arguments and outputs as
 * qqcd_noise_loop. */
int qqcd_noise_loop_lagfib(int nrhs, int *iw, int *jr, int *kr, double scale, const double kappa[2],
                           const double *coact, const int *bc, const double *u, const int GMRES[2], double tol,
                           int nb, int method, double *traces, int ksub, const int *io, double *loops,
                           double *res_full, int *iters, double *times, int *stats);

/* The reference's loop operators of one (sink x, source b) pair, replacing
 * loopops (qqcd-test/cfgsprops/quark/disconloops.f90:7769-7910) called over
 * every (ieo, ibleo) slice: xe/xo the sink, be/bo the source (reference
 * layout, both checkerboards), u the links (NULL: the resident field), bc the
 * boundary conditions (NULL: as set).  Je/Jo(2, nvhalf, 9, 2, 8) receive,
 * slice by slice, what loopops writes into Je/Jo(2, nvhalf, 9) for (ieo,
 * ibleo): the point-split currents J_1..J_4
 *   J_mu(x) = conj( f x(x+mu)^dag (1 - gamma_mu) U_mu(x)^dag b(x)
 *                 - f x(x)^dag (1 + gamma_mu) U_mu(x) b(x+mu) ),
 * the local scalar conj(x^dag b) (+12 when ksub = 1) and the local vector
 * currents conj(x^dag gamma_mu b); f = bc(mu) on hops across the edge. */
int qqcd_loopops(int ksub, const double *xe, const double *xo, const double *u, const double *be, const double *bo,
                 const int *bc, double *Je, double *Jo);

/* The loop operators summed per timeslice with the momentum factors of
 * momfacs for the origin io(3) (1-based; NULL: momentum 1 only), replacing
 * loopops + spacesum over all slices (disconloops.f90:8925-9124) without
 * forming the per-site operators: Jtime(2, nt, 5, 9) in Fortran order, the
 * four currents multiplied by cur_scale (1: spacesum; kappa: nsaverage,
 * :3079-3143).  Arguments otherwise as qqcd_loopops. */
int qqcd_loopsum(int ksub, const double *xe, const double *xo, const double *u, const double *be, const double *bo,
                 const int *bc, const int *io, double cur_scale, double *Jtime);

/* A full propagator on the resident links: nrhs (1..12) point sources at the
 * 1-based site src (spin r/3 + 1, colour r%3 + 1 as pointsource numbers them,
 * qqcd-test/cfgsprops/quark/quark.f90:46-108), each solved to
 * ||b - M x|| < tol ||b|| on the whole lattice (Wilson or twisted mass,
 * kappa = {kappa, mu}).  The first by the mixed solver (inner GMRES-DR(m, k)
 * with op_prec 1 (FP32) / 2 (FP16) storage and op_recon 12 / 18 links, FP32
 * basis, FP64 reliable updates), keeping its deflation space; the rest
 * (method 0) together through the batched GMRES-Proj(m - k) over it, (1) one
 * by one through the mixed solver with GMRES-Proj inner solves over it, or
 * (2) one by one with the learned network as flexible right preconditioner
 * (the unpreconditioned solver takes over when an outer step stalls).
 * Replaces the loop over fermprop (quark.f90:458-842) of a propagator run;
 * the reference solves each right-hand side on its own.  xe/xo: optional
 * arrays of nrhs host pointers (reference layout) for the solutions, may be
 * NULL.  iters[r], res_full[r] per source; times = {first, rest, total}
 * seconds; stats[8] = {outer steps, GMRES cycles, batched operator
 * applications, graph replays, fallbacks, RHS block, preconditioner drops,
 * outer steps of the first solve}. */
int qqcd_propagator(int nrhs, const int src[4], const double kappa[2], const int GMRES[2], double tol, int op_prec,
                    int op_recon, int method, double *const *xe, double *const *xo, int *iters, double *res_full,
                    double *times, int *stats);

/* ---- the reference's lagged-Fibonacci stream ---------------------------
 * State: iw[521], *jr, *kr = iwlagfib(0:520), jrlagfib, krlagfib of module
 * lagfib (qqcd-test/cfgsprops/glue/lagfib.f90:58-61); kr must equal jr - 32
 * modulo 521.  Every call below reads the state and writes the advanced state
 * back, so a host can pass its module variables and keep calling its own
 * rnds afterwards. */

/* initlagfib (lagfib.f90:80-92, with initbirep :96-138 and bigdelay
 * :142-216): the state of process myid for seed iseed.  Host only. */
int qqcd_lagfib_init(int myid, int iseed, int *iw, int *jr, int *kr);

/* rnds (lagfib.f90:219-259): m numbers into rn (host), generated on the
 * device (jump-ahead over chunks), bit-identical to the reference's,
 * including its skip of zero draws. */
int qqcd_rnds(long m, double *rn, int *iw, int *jr, int *kr);

/* No counterpart in the reference: qqcd_rnds for callers that cannot
 * allocate device memory (harnesses that forbid hipMalloc, tight loops).
 * Same arguments, outputs and state update as qqcd_rnds, bit for bit, but
 * every device buffer it needs lies in one static array per GPU (12 MiB,
 * held until the process ends) whose address qqcd_init takes; it calls no
 * hipMalloc or hipFree.  m is limited to QQCD_RNDS_NOALLOC_MAX draws (larger
 * m is an error).  Calls that overlap on one GPU share that array and race. */
#define QQCD_RNDS_NOALLOC_MAX 1048576L
int qqcd_rnds_noalloc(long m, double *rn, int *iw, int *jr, int *kr);

/* z4source / z2source (qqcd-test/cfgsprops/quark/disconloops.f90:4359-4649)
 * from the reference's stream: be, bo (6, ntotal, 4, 2, 8), 384 nvhalf draws;
 * bit-identical to the reference's sources.  Either output may be NULL (its
 * draws are still consumed). */
int qqcd_z4source_lagfib(double *be, double *bo, int *iw, int *jr, int *kr);
int qqcd_z2source_lagfib(double *be, double *bo, int *iw, int *jr, int *kr);

/* sweep (qqcd-test/cfgsprops/cfgsprops.f90:1973-2003): nsweep pseudo-heatbath
 * sweeps of u (18, ntotal, 4, 2, 16), in place, as the reference runs them:
 * link sets (ieo, ibl, mu) in its order, staples (gaugetools.f90:39-345;
 * plaquette terms, gaction = 1 in every direction -- other actions are
 * rejected) weighted by coact(1, mu, nu) of coact (4, 4, 4), phbup
 * (heatbath.f90:33-76: phbsub1-3, su2up, sprd1) drawing from the stream
 * iw/jr/kr, which is advanced past exactly the draws the reference makes.
 * Staples and projections are evaluated in the reference's order of
 * operations; cos/sin/log may differ from the host library's in the last
 * place.  The resident links of the operators are refreshed on the next
 * call that passes u. */
int qqcd_sweep(const int *gaction, const double *coact, int nhit, int nsweep, double *u, int *iw, int *jr, int *kr);

/* ---- the reference's files (host only: n = lattice extents, no device) --- */

/* GFwrite (qqcd-test/cfgsprops/glue/gaugetools.f90:1259-1292): fixsu3 is
 * applied to u in place (rows 1-2 re-orthonormalised, row 3 rebuilt), then
 * rows 1-2 of every link are written as Fortran unformatted records
 * (icri, isite, ibl order, mu and ieo inside), then the trailer gaction(4),
 * beta, alat(4), tad(4), icfgsave, myid.  The file must not exist
 * (status="new").  u: (18, ntotal, 4, 2, 16). */
int qqcd_gfwrite(const int n[4], double *u, const char *filename, const int gaction[4], double beta,
                 const double alat[4], const double tad[4], int icfgsave, int myid);

/* GFread (gaugetools.f90:1296-1334): rows 1-2 from the file, the links
 * completed by fixsu3; the trailer into the outputs. */
int qqcd_gfread(const int n[4], double *u, const char *filename, int gaction[4], double *beta, double alat[4],
                double tad[4], int *icfgsave, int *myidin);

/* FPwrite (qqcd-test/cfgsprops/quark/quark.f90:3379-3429): the propagator
 * xe / xo (6, ntotal, 4, 2, 8) by site (pairs of timeslices) and block, then
 * kappa, tmqcd, cSW, icsrc, idsrc, icfg, myid.  Replaces an existing file. */
int qqcd_fpwrite(const int n[4], const double *xe, const double *xo, const char *filename, double kappa,
                 double tmqcd, double cSW, int icsrc, int idsrc, int icfg, int myid);

/* FPread (quark.f90:3433-3481). */
int qqcd_fpread(const int n[4], double *xe, double *xo, const char *filename, double *kappa, double *tmqcd,
                double *cSW, int *icsrc, int *idsrc, int *icfgin, int *myidin);

/* The 16 local trace estimators of one source: eta and x on both
 * checkerboards (reference layout); traces as in qqcd_noise_loop for
 * nrhs = 1, scale = 1. */
int qqcd_trace16(const double *eta_e, const double *eta_o, const double *x_e, const double *x_o, double *traces);

/* Knobs of the batched solver: reorth 0 (one classical Gram-Schmidt pass),
 * 1 (a second pass for the slots that need it), 2 (always two); use_graph
 * replays the Arnoldi cycle from a captured HIP graph; inner_tol relative
 * tolerance of each FP32 inner solve; inner_cycles GMRES-Proj cycles per
 * inner solve; max_outer FP64 reliable updates before the single-RHS
 * fallback; op_prec 1 (FP32 links) / op_recon 12 or 18 for the inner
 * operator; basis_prec 2 (FP16, the default: unit vectors scaled by
 * sqrt(12 stride), the operator reading an FP32 copy of the newest one) or
 * 1 (FP32) for the inner GMRES's Krylov basis. */
int qqcd_batch_set(int reorth, int use_graph, double inner_tol, int inner_cycles, int max_outer, int op_prec,
                   int op_recon, int basis_prec);

/* Arnoldi steps per Gram-Schmidt pass over an FP16 Krylov basis, in the
 * batched and the single-right-hand-side solvers: 2 (the default) runs the
 * steps in pairs -- u1 = A v_j, u2 = A u1, the dots of both and the updates
 * of both in one pass each over the basis, column j + 1 of the Hessenberg
 * matrix from the Arnoldi relation -- where one Gram-Schmidt pass (and, in
 * the batched solver, the two-slot operator and an even cycle length)
 * allows it; 1 one step at a time.  Tuning aid; no reference counterpart. */
int qqcd_set_arnoldi_steps(int steps);

/* Paired Arnoldi steps of the batched solver: 1 (the default) takes the norm
 * of u1' = u1 - V V^H u1 and <u1', u2'> from those of u1, u2 (formed in the
 * pass of dots) less the projections, so that the update pass also writes
 * the next basis vector and the next operator input; 0 computes them from the
 * updated vectors, with a third pass over the pair.  Tuning aid; no reference
 * counterpart. */
int qqcd_set_pair_norms(int mode);

/* Storage of the batched solver's inner operator's intermediate field (the
 * odd checkerboard between its two hops): 2 (the default) FP16 with one norm
 * per site and right-hand side, where the two-slot FP32 hop runs; 1 FP32.
 * Tuning aid; no reference counterpart. */
int qqcd_set_batch_mid(int prec);
/* Statistics of the last batched solve. */
int qqcd_batch_stats(int *outer, int *cycles, long *slot_matvecs, long *batched_matvecs, int *graph_replays,
                     int *fallback);
/* Per-slot corrections of the last batched solve that held a NaN or an
 * infinity and were not added to x. */
int qqcd_batch_rejected(int *rejected);

/* ---- learned preconditioner (no counterpart in the reference) ---------
 * A gauge-equivariant network of parallel-transport convolution layers:
 *   out_a(x) = sum_{b, p} W[a][b][p] (T_p psi_b)(x),
 * p over {0, +x, -x, +y, -y, +z, -z, +t, -t} (0, 1+2mu, 2+2mu), T_0 = 1,
 * (T_+mu psi)(x) = U_mu(x) psi(x+mu), (T_-mu psi)(x) = U_mu(x-mu)^dag psi(x-mu)
 * (edge hops times bc[mu]); W[a][b][p] a complex 4x4 spin matrix (identity on
 * colour).  Channel counts channels[0..nlayers] with channels[0] =
 * channels[nlayers] = 1 and every count 1 or 8.  Weights W of layer l are
 * complex (re, im) in the order ((a*cin + b)*9 + p)*16 + 4 s + s', layers back
 * to back.  Activations BF16, weights BF16 on the matrix cores, FP32
 * accumulation.  The host weights survive qqcd_finalize / qqcd_init, so a
 * network trained on one lattice can be applied on another. */
/* Start near the point-Jacobi preconditioner scale * 1, plus noise * Gaussian weights. */
int qqcd_nn_init(int nlayers, const int *channels, double scale, double noise, unsigned long long seed);
int qqcd_nn_set_weights(int nlayers, const int *channels, const double *W);
/* nlayers, channels[nlayers + 1]; W may be NULL to query the sizes only */
int qqcd_nn_get_weights(int *nlayers, int *channels, double *W);
/* One layer (adjoint = 1: its adjoint, reversed paths and W^dag) on canonical
 * arrays in[C][V][4][3], out[C'][V][4][3] complex, with the resident links. */
int qqcd_nn_layer_canonical(int layer, int adjoint, const double *in, double *out);
/* dL/dconj(W) of layer l: sum_{x, colour} g_a(x) (T_p psi_b)(x)^dag, complex,
 * weight order; psi: the layer input, g: the gradient at its output. */
int qqcd_nn_wgrad_canonical(int layer, const double *psi, const double *g, double *dW);
/* z = N(r) on the even sites of canonical arrays (the preconditioner as the solver applies it). */
int qqcd_nn_apply_canonical(const double *r, double *z);
/* Train on nconf seeded heatbath configurations of the current lattice:
 * (residual, error) pairs from every Arnoldi step of plain FP64
 * GMRES-DR(m, k) solves of up to nsrc Gaussian sources per configuration
 * (kappa uniform in [kappa_lo, kappa_hi]), max_pairs in all, spread evenly
 * over the configurations and kept in FP16; loss ||N(r) - e||^2 / ||e||^2 with
 * ||r|| = 1; `steps` Adam steps (lr, betas 0.9 / 0.999) on minibatches of
 * `batch` pairs, gradients from the device forward / adjoint / weight-gradient
 * kernels.  keep = 0 starts from the kappa^2-Jacobi initialisation with the
 * given channels.  Every 150 Adam steps the network is checked by the solver
 * itself (flexible GMRES-DR(40,16) to 1e-6 on a point source of the last
 * configuration) and the weights needing the fewest Krylov steps, the
 * starting ones included, are kept.  loss[steps]: minibatch losses; stats[7]
 * = {pairs, harvest seconds, training seconds, mean loss before, mean loss
 * after (over a fixed sample of at most 512 pairs), validation steps of the
 * starting and of the kept weights (plus 1,000,000 for a validation solve
 * that did not converge in 2000 steps)}.  Leaves the
 * last training configuration resident. */
int qqcd_nn_train(int nconf, double beta, int nsweep, double kappa_lo, double kappa_hi, int nsrc, int m, int k,
                  double harvest_tol, int max_pairs, int nlayers, const int *channels, int keep, int steps, int batch,
                  double lr, unsigned long long seed, double *loss, double *stats);
/* Time one solve of M x = point source on the resident field: method 0
 * plain FP64 GMRES-DR(m,k); 1 the mixed solver (qqcd_solve_mixed with an FP32
 * basis and (op_prec, op_recon) operator); 2 the same with the network as
 * the flexible right preconditioner of every inner solve (FGMRES-DR),
 * dropped for the unpreconditioned solver at the first outer step that does
 * not halve the true residual.  nn_steps: Krylov steps taken with the
 * network; dropped: 1 if it was switched off. */
int qqcd_bench_learned(int method, const int src[4], int spin, int colour, const double kappa[2], const int GMRES[2],
                       double tol, int op_prec, int op_recon, double *seconds, int *iters, int *outer,
                       double *true_res, int *nn_steps, int *dropped);

/* Average time (ms, HIP events on the library stream) of one application of
 * the network on the resident field. */
int qqcd_bench_nn(int nwarm, int niter, double *ms);

/* Statistics of the last solver call: final ||r|| / ||r0|| (true residual),
 * Krylov steps, all operator applications, cycles, converged flag. */
int qqcd_last_solve_stats(double *rel_res, int *iters, int *matvecs, int *cycles, int *converged);
/* Drop the resident deflation space (the next psc must use isignal = 1). */
int qqcd_deflation_clear(void);

/* Statistics of the last mixed-precision solve (any entry that runs one):
 * outer FP64 steps; fallback 0 none, 1 operator raised to FP32, 2 finished
 * by plain FP64 GMRES-DR; precond_dropped 1 when the learned preconditioner
 * stalled and was switched off; rejected = inner corrections with a NaN or
 * infinite entry that were not added to x; converged; ||b - A x|| / ||b||.
 * Any pointer may be NULL.  No reference counterpart: the reference stops
 * with INVERTERS.ERROR instead (qqcd-test/cfgsprops/quark/inverters.f90:
 * 13128-13133). */
int qqcd_mixed_stats(int *outer, int *fallback, int *precond_dropped, int *rejected, int *converged,
                     double *rel_res);

/* Failures forced from the host for tests of the fallbacks (process-wide,
 * all zero / -1 to switch off).  precond: the learned preconditioner's
 * output replaced by 1 NaN, 2 zero, 3 random values; corrupt_outer: the
 * inner corrections of the first n outer steps of a mixed solve set to NaN
 * (in the batched solver: the correction of right-hand side batch_rhs);
 * maxiter > 0 caps the Krylov steps of a mixed solve.  Every kernel still
 * runs to completion; only values are replaced.  No reference counterpart. */
int qqcd_set_fault(int precond, int corrupt_outer, int batch_rhs, int maxiter);

/* ---- native (device-resident) interface: benchmarks and tests ---------- */

/* prec: 0 FP64, 1 FP32, 2 FP16 storage; recon: 18 or 12. */

/* Time `niter` back-to-back Dslash launches (after `nwarm`) on seeded random
 * spinors with the resident gauge field in precision (prec, recon).  Returns
 * the average time of one launch in *ms (HIP events on the library stream)
 * and the relative error of the result against the FP64/18 Dslash of the
 * same input in *relerr.  which = 0: one hop H_oe; 1: the full Schur
 * operator (two hops + epilogue). */
int qqcd_bench_dslash(int prec, int recon, int which, int nwarm, int niter, double *ms, double *relerr);
/* The hop H_oe (average ms per launch) cycled over nbuf (input, output)
 * spinor pairs, so that at nbuf >= 2 on 32^3x64 a launch's fields are not
 * left in the last-level cache by the launch before.  Tuning aid; no
 * reference counterpart. */
int qqcd_bench_dslash_rot(int prec, int recon, int nbuf, int nwarm, int niter, double *ms);

/* Apply the hop H (par_out = 0: H_eo, 1: H_oe) to a canonical-order spinor
 * psi[((s*4 + d)*3 + c)] complex, at storage precision (prec, recon);
 * the sites of parity par_out of out are written, the others left alone. */
int qqcd_hop_canonical(const double *psi, double *out, int par_out, int idag, int prec, int recon);

/* Schur operator (Hdbletm semantics) on the even sites of canonical arrays. */
int qqcd_schur_canonical(const double *psi, double *out, const double kappa[2], int idag, int prec, int recon);

/* Time one device-resident solve of M x = point source (1-based site src,
 * Dirac spin, colour) on the resident gauge field: method 0 plain FP64
 * GMRES-DR(m,k), method 1 the mixed solver of qqcd_solve_mixed (FP32
 * basis), method 2 the same with the FP16-stored basis.  *seconds
 * covers source preparation, solve and reconstruction; *true_res is the
 * FP64 ||b - M x|| / ||b|| of the full system. */
int qqcd_bench_solve(int method, const int src[4], int spin, int colour, const double kappa[2], const int GMRES[2],
                     double tol, int op_prec, int op_recon, double inner_tol, int keep_deflation, double *seconds,
                     int *iters, int *outer, double *true_res);

/* Multi-right-hand-side hop / Schur operator on nrhs canonical arrays (each
 * as in qqcd_hop_canonical, back to back) through the RHS-innermost fields
 * (prec 0 FP64 or 1 FP32; RHS block nb, 0 = the power of two >= nrhs, at
 * most 64; order 0/1 tile order).  Only the first nact right-hand sides are
 * active: the others are left as they are in out.  The FP32 Schur operator
 * keeps its intermediate field as qqcd_set_batch_mid says (as the batched
 * solver's inner operator does). */
int qqcd_hop_multi_canonical(int nrhs, int nact, int nb, int prec, int recon, int idag, int par_out, int order,
                             const double *in, double *out);
int qqcd_schur_multi_canonical(int nrhs, int nact, int nb, int prec, int recon, const double kappa[2], int idag,
                               int order, const double *in, double *out);
/* Time the multi-RHS hop (average ms per launch) with nact of nrhs slots active. */
int qqcd_bench_dslash_multi(int prec, int recon, int nrhs, int nact, int nb, int order, int nwarm, int niter,
                            double *ms);
/* The autotuner's choice of RHS block and tile order for nrhs right-hand
 * sides on the resident gauge field, and its time per launch. */
int qqcd_autotune_multi(int prec, int recon, int nrhs, int *nb, int *order, double *ms);

/* Kernel time (ms) of one smearing iteration (both checkerboards, FP64) on
 * the resident links; of generating m draws of the lagged-Fibonacci stream
 * in one launch of chunks of `chunk` draws (0: the generator's own choice;
 * m is cut to 2048 chunks). */
int qqcd_bench_smear(int nwarm, int niter, double *ms);
int qqcd_bench_lagfib(long m, int niter, long chunk, double *ms);

/* Time a plain device-to-device copy kernel (read + write `bytes` each
 * launch): the calibration point for hardware byte counters. */
int qqcd_bench_copy(double bytes, int iters, double *ms);

/* Kernel time (ms, mean of niter) of one fused Krylov pass over k (1..16)
 * checkerboard vectors of the current lattice in prec 0 (FP64) / 1 (FP32) /
 * 2 (FP16 vectors with per-site norms, an FP32 w):
 * op 0 the multi-dot <v_i, w>, op 1 the multi-axpy w -= sum h_i v_i with
 * ||w||^2; grid workgroups of 256 threads.  Tuning aid; no reference
 * counterpart. */
int qqcd_bench_blas(int prec, int k, int op, int grid, int niter, double *ms);

/* Override the launch shape of every multi-right-hand-side hop (-1 keeps
 * the caller's or the autotuner's choice): kernel 0 (hops one after the
 * other), 1 (pipelined hops, one contiguous tile range per XCD, sites in
 * blocks of bz z-planes swept in t, persistent grid), 2 (as 1 with one
 * workgroup per tile), 3 or 4 (FP32 spinors: two right-hand sides per lane,
 * one workgroup per tile, two or one hops' loads in flight; FP64: as 1; 3 is
 * the default), bz, nt 0/1 (links and output streamed past the L2).  Tuning
 * aid; no reference counterpart. */
int qqcd_set_multi_tune(int kernel, int bz, int nt);

/* Dynamic LDS per workgroup of the one-workgroup-per-tile multi-RHS hop
 * (kernel 2; -1 restores the default): unused by the kernel, it caps the
 * workgroups resident on a CU and so the sites in flight on an XCD.  Tuning
 * aid; no reference counterpart. */
int qqcd_set_multi_lds(int bytes);

/* Replay of repeated groups of Krylov steps from HIP graphs in the
 * single-right-hand-side solvers: 0 never, 1 in every solve (a graph on a
 * fresh basis replays about three times, too few to repay its capture), 2
 * (the default) only in the mixed solver's later outer steps, whose
 * GMRES-Proj keeps its basis and graphs from one outer step to the next.
 * Tuning aid; no reference counterpart. */
int qqcd_set_step_graphs(int mode);

/* Toggle the XCD-aware workgroup order of the Dslash (default on). */
int qqcd_set_xcd_remap(int on);

/* Launch shape of the single-right-hand-side Dslash (projector form): depth
 * 0 computes the eight hops one after the other, 1/2/4 issues that many
 * hops' loads ahead; order 1 sweeps blocks of bz z-planes in t, order 2
 * blocks of bz t-slices in z (bz 0: the largest of 16, 8, 4, 2 dividing the
 * blocked extent; order 0: the checkerboard index); nt 1
 * streams links and output past the L2.  Default (1, 0, 2, 1).  Tuning aid;
 * no reference counterpart. */
int qqcd_set_dslash_tune(int order, int bz, int depth, int nt);

/* ---- host dense linear algebra of the solvers (verification hooks) ------
 * Column-major complex matrices as interleaved (re, im) doubles.  They
 * replace the reference's pseudolapack routines used by gmresdr
 * (qqcd-test/cfgsprops/quark/pseudolapack.f90: hessenberg :700, qgenerator
 * :774, evalues/QRdouble :868/:2196, orgschur/schurswap :252/:389). */
/* A = Q T Q^H, T upper triangular (complex Schur form) */
int qqcd_dense_schur(int n, const double *a, int lda, double *t, double *q);
/* move the k eigenvalues of smallest modulus to the leading positions */
int qqcd_dense_reorder(int n, double *t, double *q, int k);
/* min ||c - H y|| for rows x cols H; *resid = the minimum */
int qqcd_dense_lsq(int rows, int cols, const double *h, int ldh, const double *c, double *y, double *resid);
/* GMRES-DR harmonic Ritz step: k Schur vectors (m x k) of
 * H_m + |h_{m+1,m}|^2 H_m^{-H} e_m e_m^T for the k values of smallest modulus */
int qqcd_dense_harmonic_ritz(int m, const double *hbar, int ldh, int k, double *p, double *theta);

#ifdef __cplusplus
}
#endif
#endif
