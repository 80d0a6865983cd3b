// qqcd_capi_solvers.cpp -- extern "C" solver entry points on reference-order
// host arrays: gmresdr, gmresproject, psc, bicgstab, fermprop.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/qqcd_hip.h"
#include "qqcd_solvers.h"

using namespace qqcd;
using qqcd::dense::zd;

#define API_BEGIN try {
#define API_END                     \
  }                                 \
  catch (const std::exception &e) { \
    set_error(e.what());            \
    return -1;                      \
  }                                 \
  return 0;

namespace {

// The deflation space kept between calls, as the reference keeps vtemp and
// hcnew in module shift (qqcd-test/cfgsprops/quark/shift.f90:20-45).
Deflation *g_defl = nullptr;
Deflation *g_defl_mixed = nullptr;  // deflation space of the mixed solver
SolveStats g_last;

// the reference's cycle limit kcyclim (qqcd-test/user/latdims.f90:70)
constexpr int KCYCLIM = 5000;

Deflation &defl() {
  if (!g_defl) g_defl = new Deflation();
  return *g_defl;
}

void to_device(Context &c, const double *h, Spinor &s, int par) {
  const size_t bytes = sizeof(double) * 6 * ref_ntotal(c.g) * 64;
  QQCD_HIP(hipMemcpyAsync(c.dref[2], h, bytes, hipMemcpyHostToDevice, c.stream));
  ref_to_spinor(c.g, c.dref[2], s, par, c.stream);
}

void to_host(Context &c, const Spinor &s, double *h, int par) {
  const size_t bytes = sizeof(double) * 6 * ref_ntotal(c.g) * 64;
  spinor_to_ref(c.g, s, c.dref[2], par, c.stream);
  QQCD_HIP(hipMemcpyAsync(h, c.dref[2], bytes, hipMemcpyDeviceToHost, c.stream));
  QQCD_HIP(hipStreamSynchronize(c.stream));
}

void prepare(Context &c, const double *u, const double *coact, const int *bc, int iflag) {
  if (iflag != -1) throw std::runtime_error("only iflag = -1 (Wilson / twisted mass) is supported");
  set_bc(c, bc);
  ensure_gauge(c, u, coact);
}

void check_mk(int m, int k) {
  if (m < 2 || k < 0 || k >= m || m > 120) throw std::runtime_error("GMRES sizes need 0 <= k < m <= 120");
}

// The lines the reference's solvers append to CFGSPROPS.LOG, format
// (a12,i9,es17.10): GMRES-DR at every true residual "gmresdr-norm", matrix-
// vector products, ||r|| / ||r0|| (inverters.f90:3285-3291); BiCGStab at
// every iteration "bicgstab", iterations, ||r||^2 / ||r0|| as written there
// (:476-480).  Cleared by every solve; read back with qqcd_solve_log.
struct LogLine {
  char label[13];
  int count;
  double value;
};
std::vector<LogLine> g_log;

void log_line(const char *label, int count, double value) {
  LogLine l{};
  std::memset(l.label, ' ', 12);
  std::memcpy(l.label, label, std::min<size_t>(12, std::strlen(label)));
  l.count = count;
  l.value = value;
  g_log.push_back(l);
}

const IterHook &bicgstab_log() {
  static const IterHook h = [](int it, double rn, double r0) { log_line("bicgstab", it, rn * rn / r0); };
  return h;
}

SolveStats run_gmresdr(Context &c, const Spinor &b, Spinor &x, const int GMRES[2], double resmax, int itermin,
                       const double kappa[2]) {
  check_mk(GMRES[0], GMRES[1]);
  SchurOp A(c, PREC_DOUBLE, 18, kappa);
  LinOp op = A.op();
  const CycleHook hook = [](const Spinor &, const Spinor &, const SolveStats &st) {
    log_line("gmresdr-norm", st.matvecs, st.rel_res);
  };
  return gmresdr(c, op, b, x, GMRES[0], GMRES[1], resmax, itermin, KCYCLIM * GMRES[0], &defl(), 1, nullptr, &hook);
}

SolveStats run_gmresproj(Context &c, const Spinor &b, Spinor &x, int m, int k, double resmax, int itermin,
                         const double kappa[2], ResHistory *hist = nullptr) {
  Deflation &d = defl();
  if (d.k == 0 || (int)d.V.size() != d.k + 1)
    throw std::runtime_error("GMRES-Proj needs the deflation space of an earlier GMRES-DR solve");
  if (k != d.k) throw std::runtime_error("GMRES-Proj k differs from the stored deflation space");
  if (m < 1 || m > 120) throw std::runtime_error("GMRES-Proj cycle length must be 1..120");
  SchurOp A(c, PREC_DOUBLE, 18, kappa);
  LinOp op = A.op();
  return gmresproj(c, op, b, x, m, d, resmax, itermin, KCYCLIM * m, 1, hist);
}

// The right-hand side gmresproject solves when called with isignal = 1: the
// reference's module variable beg (shift.f90), which psc sets after its
// GMRES-DR to the last vector of the deflation space, vtemp(:,:,:,:,:,k+1)
// (inverters.f90:4437), and gmresproject with isignal /= 1 sets to its
// initial residual b - M 0 = b (:7438-7443).
Spinor g_beg;

void set_beg(Context &c, const Spinor &v) {
  if (!g_beg.data) g_beg = alloc_spinor(c.g, PREC_DOUBLE);
  spinor_copy_convert(c.g, v, g_beg, c.stream);
}

void beg_after_gmresdr(Context &c) {
  Deflation &d = defl();
  if (d.k > 0 && (int)d.V.size() == d.k + 1) set_beg(c, d.V[d.k]);
}

}  // namespace

extern "C" int qqcd_gmresdr(const double *phi, double *x, const int GMRES[2], double resmax, int itermin,
                            int *itercount, const double *u, const double *coact, const double kappa[2], int iflag,
                            const int *bc) {
  API_BEGIN
  g_log.clear();
  Context &c = ctx();
  prepare(c, u, coact, bc, iflag);
  SpinorBuf b(c.g, PREC_DOUBLE), X(c.g, PREC_DOUBLE);
  to_device(c, phi, b.s, 0);
  to_device(c, x, X.s, 0);
  g_last = run_gmresdr(c, b.s, X.s, GMRES, resmax, itermin, kappa);
  to_host(c, X.s, x, 0);
  *itercount = g_last.iters;
  API_END
}

extern "C" int qqcd_gmresproject(const double *bh, double *xshift, int nshift, const int GMRES[2], double resmax,
                                 int itermin, int *itercount, const double *u, const double *coact,
                                 const double kappa[2], int iflag, const int *bc, int isignal, int *mvp, double *gdr,
                                 int ldgdr) {
  API_BEGIN
  g_log.clear();
  Context &c = ctx();
  if (nshift < 1) throw std::runtime_error("gmresproject: xshift needs at least one shift");
  if (!mvp || (ldgdr > 0 && !gdr)) throw std::runtime_error("gmresproject: mvp and gdr are outputs");
  prepare(c, u, coact, bc, iflag);
  SpinorBuf b(c.g, PREC_DOUBLE), X(c.g, PREC_DOUBLE);
  if (isignal == 1) {
    // inverters.f90:7438-7440: the right-hand side is the resident beg
    if (!g_beg.data) throw std::runtime_error("gmresproject with isignal = 1 needs a beg from an earlier psc");
    spinor_copy_convert(c.g, g_beg, b.s, c.stream);
  } else {
    to_device(c, bh, b.s, 0);
    set_beg(c, b.s);
  }
  ResHistory hist;
  g_last = run_gmresproj(c, b.s, X.s, GMRES[0], GMRES[1], resmax, itermin, kappa, &hist);
  // every shift of xshift is zeroed (the reference's xshift = 0), the solution is shift 1
  const size_t slen = (size_t)6 * ref_ntotal(c.g) * 64;
  std::memset(xshift, 0, slen * nshift * sizeof(double));
  to_host(c, X.s, xshift, 0);
  *itercount = g_last.iters;
  *mvp = g_last.matvecs;
  for (const auto &h : hist)
    if (h.first >= 1 && h.first <= ldgdr) gdr[h.first - 1] = h.second;
  API_END
}

extern "C" int qqcd_psc(const double *bh, double *x, const int GMRES[2], double resmax, int itermin,
                        int *itercount, const double *u, const double *coact, const double kappa[2], int iflag,
                        const int *bc, int isignal, double resmax_proj) {
  API_BEGIN
  g_log.clear();
  Context &c = ctx();
  prepare(c, u, coact, bc, iflag);
  check_mk(GMRES[0], GMRES[1]);
  SpinorBuf b(c.g, PREC_DOUBLE), X(c.g, PREC_DOUBLE);
  to_device(c, bh, b.s, 0);
  if (isignal == 1) {
    to_device(c, x, X.s, 0);
    g_last = run_gmresdr(c, b.s, X.s, GMRES, resmax, itermin, kappa);
    beg_after_gmresdr(c);
  } else {
    // inverters.f90:4439-4458: GMRES-Proj(m-k, k), tolerance 1e-8 hard-coded
    const double tol = resmax_proj > 0 ? resmax_proj : 1e-8;
    g_last = run_gmresproj(c, b.s, X.s, GMRES[0] - GMRES[1], GMRES[1], tol, itermin, kappa);
  }
  to_host(c, X.s, x, 0);
  *itercount = g_last.iters;
  API_END
}

extern "C" int qqcd_bicgstab(const double *phi, double *x, double resmax, int itermin, int *itercount,
                             const double *u, const double *coact, const double kappa[2], int iflag, const int *bc) {
  API_BEGIN
  g_log.clear();
  Context &c = ctx();
  prepare(c, u, coact, bc, iflag);
  SpinorBuf b(c.g, PREC_DOUBLE), X(c.g, PREC_DOUBLE);
  to_device(c, phi, b.s, 0);
  to_device(c, x, X.s, 0);
  SchurOp A(c, PREC_DOUBLE, 18, kappa);
  LinOp op = A.op();
  g_last = bicgstab(c, op, b.s, X.s, resmax, itermin, 1000000, &bicgstab_log());
  to_host(c, X.s, x, 0);
  *itercount = g_last.iters;
  API_END
}

// fermprop for nkappa = -1, qqcd-test/cfgsprops/quark/quark.f90:570-617
// (source), :630-692 (solve), :786-819 (reconstruction).  a = 2 kappa mu is
// taken with the sign of mu as given, while the operator falls back to
// Wilson for mu < 1e-8 (diracops.f90:467), exactly as in the reference.
extern "C" int qqcd_fermprop(const double *be, const double *bo, const double kappa[2], int inverter,
                             const double *coact, const int *bc, double resmax, int itermin, const int nGMRES[2],
                             int vectype, const double *u, double *xe, double *xo, int isignal, int *itercount) {
  API_BEGIN
  g_log.clear();
  Context &c = ctx();
  prepare(c, u, coact, bc, -1);
  const double k = kappa[0], a = 2.0 * kappa[0] * kappa[1];
  SpinorBuf b_e(c.g, PREC_DOUBLE), b_o(c.g, PREC_DOUBLE), t_o(c.g, PREC_DOUBLE), src(c.g, PREC_DOUBLE);
  SpinorBuf x_e(c.g, PREC_DOUBLE), x_o(c.g, PREC_DOUBLE);
  to_device(c, be, b_e.s, 0);
  to_device(c, bo, b_o.s, 1);
  // vecsrc = b_e/kappa^2 + H_eo (1 - a i g5) b_o / (kappa (1 + a^2))
  local_spin(c.g, b_o.s, t_o.s, 1.0, -a, c.stream);
  Epilogue es;
  es.a0 = 1.0 / (k * k);
  es.b0 = 1.0 / (k * (1.0 + a * a));
  dslash(c.g, c.U64, t_o.s, src.s, 0, 0, es, &b_e.s, c.stream);
  // initial guess (quark.f90:630-640)
  if (inverter > 3) {
    Vec xv = as_vec(x_e.s);
    c.blas->zero(xv);
  } else if (inverter == 0) {
    if (vectype == 1) fill_spinor(c.g, x_e.s, 1.0, 0.0, c.stream);
    else to_device(c, xe, x_e.s, 0);
  } else {
    throw std::runtime_error("fermprop: inverter must be 0 (BiCGStab), 5 (GMRES-DR) or 6 (psc)");
  }
  if (inverter == 0) {
    SchurOp A(c, PREC_DOUBLE, 18, kappa);
    LinOp op = A.op();
    g_last = bicgstab(c, op, src.s, x_e.s, resmax, itermin, 1000000, &bicgstab_log());
  } else if (inverter == 5) {
    g_last = run_gmresdr(c, src.s, x_e.s, nGMRES, resmax, itermin, kappa);
  } else if (inverter == 6) {
    if (isignal == 1) {
      g_last = run_gmresdr(c, src.s, x_e.s, nGMRES, resmax, itermin, kappa);
      beg_after_gmresdr(c);
    } else {
      check_mk(nGMRES[0], nGMRES[1]);
      g_last = run_gmresproj(c, src.s, x_e.s, nGMRES[0] - nGMRES[1], nGMRES[1], 1e-8, itermin, kappa);
    }
  } else {
    throw std::runtime_error("fermprop: inverter must be 0 (BiCGStab), 5 (GMRES-DR) or 6 (psc)");
  }
  // x_o = (1 - a i g5)(b_o + kappa H_oe x_e) / (1 + a^2)
  Epilogue er;
  const double f = 1.0 / (1.0 + a * a);
  er.a0 = f;
  er.a1 = -a * f;
  er.b0 = k * f;
  er.b1 = -a * k * f;
  dslash(c.g, c.U64, x_e.s, x_o.s, 1, 0, er, &b_o.s, c.stream);
  to_host(c, x_e.s, xe, 0);
  to_host(c, x_o.s, xo, 1);
  *itercount = g_last.iters;
  API_END
}

extern "C" int qqcd_last_solve_stats(double *rel_res, int *iters, int *matvecs, int *cycles, int *converged) {
  API_BEGIN
  *rel_res = g_last.rel_res;
  *iters = g_last.iters;
  *matvecs = g_last.matvecs;
  *cycles = g_last.cycles;
  *converged = g_last.converged;
  API_END
}

extern "C" int qqcd_set_step_graphs(int mode) {
  API_BEGIN
  set_step_graphs(mode);
  API_END
}

extern "C" int qqcd_deflation_clear(void) {
  API_BEGIN
  if (g_defl) {
    delete g_defl;
    g_defl = nullptr;
  }
  if (g_defl_mixed) {
    delete g_defl_mixed;
    g_defl_mixed = nullptr;
  }
  free_spinor(g_beg);
  API_END
}

// ---- host dense linear algebra of the solvers, exported for verification ----
extern "C" int qqcd_dense_schur(int n, const double *a, int lda, double *t, double *q) {
  API_BEGIN
  if (!dense::schur(n, (const zd *)a, lda, (zd *)t, (zd *)q)) throw std::runtime_error("schur: QR did not converge");
  API_END
}

extern "C" int qqcd_dense_reorder(int n, double *t, double *q, int k) {
  API_BEGIN
  dense::reorder_smallest(n, (zd *)t, (zd *)q, k);
  API_END
}

extern "C" int qqcd_dense_lsq(int rows, int cols, const double *h, int ldh, const double *c, double *y,
                              double *resid) {
  API_BEGIN
  *resid = dense::lsq(rows, cols, (const zd *)h, ldh, (const zd *)c, (zd *)y);
  API_END
}

extern "C" int qqcd_dense_harmonic_ritz(int m, const double *hbar, int ldh, int k, double *p, double *theta) {
  API_BEGIN
  if (!dense::harmonic_ritz(m, (const zd *)hbar, ldh, k, (zd *)p, (zd *)theta))
    throw std::runtime_error("harmonic_ritz failed");
  API_END
}

// ---- mixed precision ----------------------------------------------------

MixedParams qqcd::mixed_params(const int GMRES[2], double tol, int basis, int op_prec, int op_recon,
                               double inner_tol) {
  check_mk(GMRES[0], GMRES[1]);
  if (basis < 0 || basis > 2) throw std::runtime_error("basis precision must be 0 (FP64), 1 (FP32) or 2 (FP16)");
  if (op_prec < 0 || op_prec > 2 || (op_recon != 18 && op_recon != 12))
    throw std::runtime_error("bad operator precision / link format");
  if (op_prec == 0 && basis != 0) throw std::runtime_error("an FP64 operator needs an FP64 basis");
  MixedParams p;
  p.m = GMRES[0];
  p.k = GMRES[1];
  p.tol = tol;
  p.basis = (Prec)basis;
  p.op_prec = (Prec)op_prec;
  p.op_recon = op_recon;
  if (inner_tol > 0) p.inner_tol = inner_tol;
  else p.inner_tol = op_prec == 2 ? 1e-2 : (op_prec == 1 ? 1e-6 : 1e-10);
  return p;
}

// unit source at 1-based global site src, Dirac d, colour col (pointsource)
void qqcd::point_source_dev(Context &c, Spinor &b_e, Spinor &b_o, const int src[4], int d, int col) {
  const Geom &g = c.g;
  int x[4];
  for (int m = 0; m < 4; m++) {
    if (src[m] < 1 || src[m] > g.n[m]) throw std::runtime_error("source site outside the lattice");
    x[m] = src[m] - 1;
  }
  if (d < 1 || d > 4 || col < 1 || col > 3) throw std::runtime_error("bad Dirac / colour index");
  const int par = (x[0] + x[1] + x[2] + x[3]) & 1;
  const long cb = (x[0] + (long)g.n[0] * (x[1] + (long)g.n[1] * (x[2] + (long)g.n[2] * x[3]))) / 2;
  Spinor &b = par == 0 ? b_e : b_o;
  const double one[2] = {1.0, 0.0};
  const long k = 3 * (d - 1) + (col - 1);
  QQCD_HIP(hipMemcpyAsync((double2 *)b.data + k * b.stride + cb, one, sizeof(one), hipMemcpyHostToDevice, c.stream));
  QQCD_HIP(hipStreamSynchronize(c.stream));
}


extern "C" int qqcd_solve_mixed(const double *phi, double *x, const int GMRES[2], double tol, int basis_prec,
                                int op_prec, int op_recon, double inner_tol, int keep_deflation, const double *u,
                                const double *coact, const double kappa[2], const int *bc, int *itercount) {
  API_BEGIN
  Context &c = ctx();
  prepare(c, u, coact, bc, -1);
  MixedParams p = mixed_params(GMRES, tol, basis_prec, op_prec, op_recon, inner_tol);
  SpinorBuf b(c.g, PREC_DOUBLE), X(c.g, PREC_DOUBLE);
  to_device(c, phi, b.s, 0);
  to_device(c, x, X.s, 0);
  if (!keep_deflation && g_defl_mixed) {
    delete g_defl_mixed;
    g_defl_mixed = nullptr;
  }
  if (!g_defl_mixed) g_defl_mixed = new Deflation();
  MixedStats ms = solve_mixed(c, kappa, b.s, X.s, p, g_defl_mixed);
  to_host(c, X.s, x, 0);
  g_last = ms.inner;
  g_last.rel_res = ms.rel_res;
  g_last.converged = ms.converged;
  g_last.cycles = ms.outer;
  *itercount = ms.inner.iters;
  API_END
}

// Device-resident solve of M x = (point source) for timing: method 0 is the
// plain FP64 GMRES-DR(m,k), method 1 the mixed-precision solver.  The timed
// region covers source preparation, the solve and odd-site reconstruction
// (inputs resident, no host transfers); the true full-system residual is
// computed afterwards in FP64.
extern "C" int qqcd_bench_solve(int method, const int src[4], int spin, int colour, const double kappa[2],
                                const int GMRES[2], double tol, int op_prec, int op_recon, double inner_tol,
                                int keep_deflation, double *seconds, int *iters, int *outer, double *true_res) {
  API_BEGIN
  Context &c = ctx();
  if (!c.have_gauge) throw std::runtime_error("no gauge field loaded");
  SpinorBuf b_e(c.g, PREC_DOUBLE), b_o(c.g, PREC_DOUBLE), src_e(c.g, PREC_DOUBLE);
  SpinorBuf x_e(c.g, PREC_DOUBLE), x_o(c.g, PREC_DOUBLE);
  point_source_dev(c, b_e.s, b_o.s, src, spin, colour);
  if (!keep_deflation && g_defl_mixed) {
    delete g_defl_mixed;
    g_defl_mixed = nullptr;
  }
  if (!g_defl_mixed) g_defl_mixed = new Deflation();
  QQCD_HIP(hipDeviceSynchronize());
  auto t0 = std::chrono::steady_clock::now();
  prep_source(c, kappa, b_e.s, b_o.s, src_e.s);
  if (method == 0) {
    check_mk(GMRES[0], GMRES[1]);
    SchurOp A(c, PREC_DOUBLE, 18, kappa);
    LinOp op = A.op();
    // x0 = 0, so ||r0|| = ||src||; with b_o = 0 the even-odd relative
    // residual equals the full-system one
    SolveStats st = gmresdr(c, op, src_e.s, x_e.s, GMRES[0], GMRES[1], tol, 0, KCYCLIM * GMRES[0], nullptr);
    *iters = st.iters;
    *outer = st.cycles;
  } else if (method == 1 || method == 2) {
    MixedParams p = mixed_params(GMRES, tol, method == 1 ? 1 : 2, op_prec, op_recon, inner_tol);
    MixedStats ms = solve_mixed(c, kappa, src_e.s, x_e.s, p, g_defl_mixed);
    *iters = ms.inner.iters;
    *outer = ms.outer;
  } else {
    throw std::runtime_error("bench_solve: method must be 0 (FP64 GMRES-DR), 1 (mixed) or 2 (mixed, FP16 basis)");
  }
  reconstruct_odd(c, kappa, b_o.s, x_e.s, x_o.s);
  QQCD_HIP(hipDeviceSynchronize());
  auto t1 = std::chrono::steady_clock::now();
  *seconds = std::chrono::duration<double>(t1 - t0).count();
  *true_res = full_residual(c, kappa, b_e.s, b_o.s, x_e.s, x_o.s);
  API_END
}

extern "C" int qqcd_set_fault(int precond, int corrupt_outer, int batch_rhs, int maxiter) {
  API_BEGIN
  if (precond < 0 || precond > 3 || corrupt_outer < 0 || maxiter < 0)
    throw std::runtime_error("qqcd_set_fault: precond 0..3, corrupt_outer >= 0, maxiter >= 0");
  FaultPlan f;
  f.precond = precond;
  f.corrupt_outer = corrupt_outer;
  f.batch_rhs = batch_rhs;
  f.maxiter = maxiter;
  fault_plan() = f;
  API_END
}

extern "C" int qqcd_mixed_stats(int *outer, int *fallback, int *precond_dropped, int *rejected, int *converged,
                                double *rel_res) {
  API_BEGIN
  const MixedStats &m = last_mixed_stats();
  if (outer) *outer = m.outer;
  if (fallback) *fallback = m.fallback;
  if (precond_dropped) *precond_dropped = m.precond_dropped;
  if (rejected) *rejected = m.rejected;
  if (converged) *converged = m.converged;
  if (rel_res) *rel_res = m.rel_res;
  API_END
}

extern "C" int qqcd_solve_log(int cap, int *n, char *labels, int *counts, double *values) {
  API_BEGIN
  if (!n) throw std::runtime_error("qqcd_solve_log: null count");
  *n = (int)g_log.size();
  for (int i = 0; i < std::min(cap, *n); i++) {
    if (labels) std::memcpy(labels + 12 * i, g_log[i].label, 12);
    if (counts) counts[i] = g_log[i].count;
    if (values) values[i] = g_log[i].value;
  }
  API_END
}
