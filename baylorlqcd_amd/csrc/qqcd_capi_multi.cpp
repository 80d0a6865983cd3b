// qqcd_capi_multi.cpp -- extern "C" entry points of the many-right-hand-side
// path: Z4 noise, the multi-RHS hopping operator, batched noise solves and
// the local trace estimators of the disconnected loops.
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qqcd_hip.h"
#include "qqcd_batch.h"
#include "qqcd_solvers.h"
#include "qqcd_loops.h"
#include "qqcd_noise.h"
#include "qqcd_lagfib.h"

using namespace qqcd;

#define API_BEGIN try {
#define API_END                     \
  }                                 \
  catch (const std::exception &e) { \
    set_error(e.what());            \
    return -1;                      \
  }                                 \
  return 0;

namespace {

BatchParams g_bp;  // tuning knobs of the batched solver (qqcd_batch_set)
BatchStats g_bst;

struct DevArr {
  double *p = nullptr;
  DevArr(const double *h, size_t n, hipStream_t s) {
    QQCD_HIP(hipMalloc(&p, n * sizeof(double)));
    if (h) QQCD_HIP(hipMemcpyAsync(p, h, n * sizeof(double), hipMemcpyHostToDevice, s));
    else QQCD_HIP(hipMemsetAsync(p, 0, n * sizeof(double), s));
  }
  ~DevArr() {
    if (p) (void)hipFree(p);
  }
};

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

int check_nb(int nb, int nrhs) {
  if (nb == 0) {
    nb = 1;
    while (nb < nrhs && nb < 64) nb <<= 1;
  }
  if (nb < 1 || nb > 64 || (nb & (nb - 1))) throw std::runtime_error("RHS block must be a power of two in 1..64");
  return nb;
}

// multi-RHS operator on canonical arrays: which 0 = hopping term onto par_out, 1 = Schur operator
void multi_canonical(int which, int nrhs, int nact, int nb, int prec, int recon, const double kappa[2], int idag,
                     int par_out, int order, const double *in, double *out) {
  Context &c = ctx();
  if (nrhs < 1 || nrhs > 256 || nact < 0 || nact > nrhs) throw std::runtime_error("bad right-hand-side counts");
  if (prec != PREC_DOUBLE && prec != PREC_SINGLE) throw std::runtime_error("multi-RHS fields are FP64 or FP32");
  nb = check_nb(nb, nrhs);
  const size_t n = (size_t)24 * c.g.V;
  const Prec p = (Prec)prec;
  const Gauge &U = c.gauge(p, recon);
  MultiBuf a(c.g, p, nrhs, nb), b(c.g, p, nrhs, nb);
  MBlas mb(c.stream, nrhs);
  SpinorBuf s(c.g, PREC_DOUBLE);
  DevArr dio(nullptr, n, c.stream), dout(nullptr, n, c.stream);
  const int in_par = which == 0 ? 1 - par_out : 0;
  for (int r = 0; r < nrhs; r++) {
    QQCD_HIP(hipMemcpyAsync(dio.p, in + r * n, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    canon_to_spinor(c.g, dio.p, s.s, in_par, c.stream);
    mb.put(s.s, a.m, r);
  }
  int *d_n = nullptr;
  QQCD_HIP(hipMalloc(&d_n, sizeof(int)));
  try {
    QQCD_HIP(hipMemcpyAsync(d_n, &nact, sizeof(int), hipMemcpyHostToDevice, c.stream));
    MultiTune t;
    t.order = order;
    if (which == 0) {
      dslash_multi(c.g, U, a.m, b.m, par_out, idag, Epilogue(), nullptr, d_n, t, c.stream);
    } else {
      MultiSchur A(c, p, recon, kappa, nrhs, nb, t, idag);
      A.set_half_mid(g_bp.half_mid != 0);
      A(a.m, b.m, d_n);
    }
    const int out_par = which == 0 ? par_out : 0;
    for (int r = 0; r < nrhs; r++) {
      QQCD_HIP(hipMemcpyAsync(dout.p, out + r * n, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
      mb.get(b.m, r, s.s);
      spinor_to_canon(c.g, s.s, dout.p, out_par, c.stream);
      QQCD_HIP(hipMemcpyAsync(out + r * n, dout.p, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    }
    QQCD_HIP(hipStreamSynchronize(c.stream));
  } catch (...) {
    (void)hipFree(d_n);
    throw;
  }
  (void)hipFree(d_n);
}

BatchParams batch_params(const int GMRES[2], double tol, int nb_rhs) {
  BatchParams p = g_bp;
  p.m = GMRES[0] - GMRES[1];
  p.tol = tol;
  (void)nb_rhs;
  return p;
}

}  // namespace

extern "C" int qqcd_z4source(double *be, double *bo, unsigned long long seed, long rhs, double scale) {
  API_BEGIN
  Context &c = ctx();
  SpinorBuf s(c.g, PREC_DOUBLE);
  double *outs[2] = {be, bo};
  for (int par = 0; par < 2; par++) {
    if (!outs[par]) continue;
    z4_source(c.g, s.s, par, seed, rhs, scale, c.stream);
    to_host(c, s.s, outs[par], par);
  }
  API_END
}

extern "C" int qqcd_z2source(double *be, double *bo, unsigned long long seed, long rhs, double scale) {
  API_BEGIN
  Context &c = ctx();
  SpinorBuf s(c.g, PREC_DOUBLE);
  double *outs[2] = {be, bo};
  for (int par = 0; par < 2; par++) {
    if (!outs[par]) continue;
    noise_source(c.g, s.s, par, seed, rhs, scale, 1, c.stream);
    to_host(c, s.s, outs[par], par);
  }
  API_END
}

// ---- the reference's lagged-Fibonacci stream -------------------------------

namespace {

LagState lag_in(const int *iw, const int *jr, const int *kr) {
  if (!iw || !jr || !kr) throw std::runtime_error("lagfib: null state");
  LagState st;
  for (int k = 0; k < 521; k++) st.w[k] = (uint32_t)iw[k];
  st.jr = *jr;
  st.kr = *kr;
  lagfib_check(st);
  return st;
}

void lag_out(const LagState &st, int *iw, int *jr, int *kr) {
  for (int k = 0; k < 521; k++) iw[k] = (int)st.w[k];
  *jr = st.jr;
  *kr = st.kr;
}

struct DevU32 {
  uint32_t *p = nullptr;
  explicit DevU32(size_t n) { QQCD_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(uint32_t))); }
  ~DevU32() {
    if (p) (void)hipFree(p);
  }
};

int lag_source(double *be, double *bo, int *iw, int *jr, int *kr, int kind) {
  API_BEGIN
  Context &c = ctx();
  LagState st = lag_in(iw, jr, kr);
  const long nvh = c.g.V / 32;
  DevU32 y(384 * nvh);
  lagfib_generate(st, 384 * nvh, y.p, c.stream);
  SpinorBuf s(c.g, PREC_DOUBLE);
  double *outs[2] = {be, bo};
  for (int par = 0; par < 2; par++) {
    lag_noise_source(c.g, s.s, par, y.p + par * 192 * nvh, kind, 1.0, c.stream);
    if (outs[par]) to_host(c, s.s, outs[par], par);
  }
  lag_out(st, iw, jr, kr);
  API_END
}

}  // namespace

extern "C" int qqcd_lagfib_init(int myid, int iseed, int *iw, int *jr, int *kr) {
  API_BEGIN
  if (!iw || !jr || !kr) throw std::runtime_error("qqcd_lagfib_init: null output");
  LagState st;
  lagfib_init(myid, iseed, st);
  lag_out(st, iw, jr, kr);
  API_END
}

extern "C" int qqcd_rnds(long m, double *rn, int *iw, int *jr, int *kr) {
  API_BEGIN
  Context &c = ctx();
  if (m < 0 || (m > 0 && !rn)) throw std::runtime_error("qqcd_rnds: m >= 0 and an output array expected");
  LagState st = lag_in(iw, jr, kr);
  DevU32 y(m);
  lagfib_generate(st, m, y.p, c.stream);
  if (m > 0) {
    DevArr d(nullptr, (size_t)m, c.stream);
    lagfib_to_double(y.p, m, d.p, c.stream);
    QQCD_HIP(hipMemcpyAsync(rn, d.p, sizeof(double) * m, hipMemcpyDeviceToHost, c.stream));
    QQCD_HIP(hipStreamSynchronize(c.stream));
  }
  lag_out(st, iw, jr, kr);
  API_END
}

extern "C" int qqcd_rnds_noalloc(long m, double *rn, int *iw, int *jr, int *kr) {
  API_BEGIN
  Context &c = ctx();
  LagState st = lag_in(iw, jr, kr);
  rnds_noalloc(st, m, rn, c.device, c.stream);
  lag_out(st, iw, jr, kr);
  API_END
}

extern "C" int qqcd_z4source_lagfib(double *be, double *bo, int *iw, int *jr, int *kr) {
  return lag_source(be, bo, iw, jr, kr, 0);
}

extern "C" int qqcd_z2source_lagfib(double *be, double *bo, int *iw, int *jr, int *kr) {
  return lag_source(be, bo, iw, jr, kr, 1);
}

namespace qqcd {
void heatbath_reference(Context &c, const int gaction[4], const double *coact, int nhit, int nsweep, double *u_host,
                        LagState &st);
}

extern "C" int qqcd_sweep(const int *gaction, const double *coact, int nhit, int nsweep, double *u, int *iw,
                          int *jr, int *kr) {
  API_BEGIN
  Context &c = ctx();
  if (!gaction || !coact || !u) throw std::runtime_error("qqcd_sweep: null argument");
  LagState st = lag_in(iw, jr, kr);
  heatbath_reference(c, gaction, coact, nhit, nsweep, u, st);
  lag_out(st, iw, jr, kr);
  // the links changed in place: the next call that passes u uploads them again
  c.u_host = nullptr;
  API_END
}

extern "C" int qqcd_hop_multi_canonical(int nrhs, int nact, int nb, int prec, int recon, int idag, int par_out,
                                        int order, const double *in, double *out) {
  API_BEGIN
  multi_canonical(0, nrhs, nact, nb, prec, recon, nullptr, idag, par_out, order, in, out);
  API_END
}

extern "C" int qqcd_schur_multi_canonical(int nrhs, int nact, int nb, int prec, int recon, const double kappa[2],
                                          int idag, int order, const double *in, double *out) {
  API_BEGIN
  multi_canonical(1, nrhs, nact, nb, prec, recon, kappa, idag, 0, order, in, out);
  API_END
}

extern "C" int qqcd_trace16(const double *eta_e, const double *eta_o, const double *x_e, const double *x_o,
                            double *traces) {
  API_BEGIN
  Context &c = ctx();
  const int nt = c.g.n[3];
  SpinorBuf e(c.g, PREC_DOUBLE), x(c.g, PREC_DOUBLE);
  DevArr out(nullptr, (size_t)nt * 32, c.stream);
  DevArr part(nullptr, (size_t)trace16_gx(c.g) * nt * 32, c.stream);
  const double *es[2] = {eta_e, eta_o}, *xs[2] = {x_e, x_o};
  for (int par = 0; par < 2; par++) {
    to_device(c, es[par], e.s, par);
    to_device(c, xs[par], x.s, par);
    trace16(c.g, e.s, x.s, out.p, par, part.p, c.stream);
  }
  QQCD_HIP(hipMemcpyAsync(traces, out.p, (size_t)nt * 32 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  QQCD_HIP(hipStreamSynchronize(c.stream));
  API_END
}

namespace {

// upload the four reference-layout fields of a loop-operator call
void loop_fields(Context &c, const double *xe, const double *xo, const double *u, const double *be, const double *bo,
                 const int *bc, SpinorBuf *b, SpinorBuf *x) {
  if (!xe || !xo || !be || !bo) throw std::runtime_error("loop operators: xe, xo, be and bo are required");
  if (bc) set_bc(c, bc);
  if (u) ensure_gauge(c, u, nullptr);
  else if (!c.have_gauge) throw std::runtime_error("no gauge field loaded");
  to_device(c, be, b[0].s, 0);
  to_device(c, bo, b[1].s, 1);
  to_device(c, xe, x[0].s, 0);
  to_device(c, xo, x[1].s, 1);
}

}  // namespace

extern "C" int qqcd_loopops(int ksub, const double *xe, const double *xo, const double *u, const double *be,
                            const double *bo, const int *bc, double *Je, double *Jo) {
  API_BEGIN
  Context &c = ctx();
  if (ksub != 0 && ksub != 1) throw std::runtime_error("loopops: ksub is 0 or 1");
  SpinorBuf b[2] = {SpinorBuf(c.g, PREC_DOUBLE), SpinorBuf(c.g, PREC_DOUBLE)};
  SpinorBuf x[2] = {SpinorBuf(c.g, PREC_DOUBLE), SpinorBuf(c.g, PREC_DOUBLE)};
  SpinorBuf J[2] = {SpinorBuf(c.g, PREC_DOUBLE), SpinorBuf(c.g, PREC_DOUBLE)};
  loop_fields(c, xe, xo, u, be, bo, bc, b, x);
  DevArr part(nullptr, (size_t)2 * c.g.n[3] * loopsum_gx(c.g) * 58, c.stream);
  const Spinor bs[2] = {b[0].s, b[1].s}, xs[2] = {x[0].s, x[1].s};
  Spinor js[2] = {J[0].s, J[1].s};
  loopsum(c.g, c.U64, bs, xs, ksub, nullptr, 1, 1.0, 1.0, nullptr, 0, part.p, js, c.stream);
  // the per-site operators sit in components 0..8 (= 3 spin + colour) of a
  // reference-layout spinor; Je(2,nvhalf,9,2,8) takes them slice by slice
  const long ntot = ref_ntotal(c.g), nvh = c.g.Vh / 16;
  std::vector<double> h((size_t)6 * ntot * 64);
  double *outs[2] = {Je, Jo};
  for (int par = 0; par < 2; par++) {
    to_host(c, js[par], h.data(), par);
    for (long sl = 0; sl < 16; sl++)
      for (int op = 0; op < 9; op++)
        for (long i = 0; i < nvh; i++)
          for (int ri = 0; ri < 2; ri++)
            outs[par][ri + 2 * (i + nvh * (op + 9 * sl))] =
                h[(2 * (op % 3) + ri) + 6 * (i + ntot * ((op / 3) + 4 * sl))];
  }
  API_END
}

extern "C" int qqcd_loopsum(int ksub, const double *xe, const double *xo, const double *u, const double *be,
                            const double *bo, const int *bc, const int *io, double cur_scale, double *Jtime) {
  API_BEGIN
  Context &c = ctx();
  if (ksub != 0 && ksub != 1) throw std::runtime_error("loopsum: ksub is 0 or 1");
  SpinorBuf b[2] = {SpinorBuf(c.g, PREC_DOUBLE), SpinorBuf(c.g, PREC_DOUBLE)};
  SpinorBuf x[2] = {SpinorBuf(c.g, PREC_DOUBLE), SpinorBuf(c.g, PREC_DOUBLE)};
  loop_fields(c, xe, xo, u, be, bo, bc, b, x);
  const int nt = c.g.n[3];
  DevArr part(nullptr, (size_t)2 * nt * loopsum_gx(c.g) * 58, c.stream);
  DevArr out(nullptr, (size_t)nt * 90, c.stream);
  const Spinor bs[2] = {b[0].s, b[1].s}, xs[2] = {x[0].s, x[1].s};
  loopsum(c.g, c.U64, bs, xs, ksub, io, io ? 5 : 1, 1.0, cur_scale, out.p, 0, part.p, nullptr, c.stream);
  std::vector<double> h((size_t)nt * 90);
  QQCD_HIP(hipMemcpyAsync(h.data(), out.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  QQCD_HIP(hipStreamSynchronize(c.stream));
  // Jtime(2, nt, 5, 9), Fortran order
  for (int t = 0; t < nt; t++)
    for (int m = 0; m < 5; m++)
      for (int op = 0; op < 9; op++)
        for (int ri = 0; ri < 2; ri++) Jtime[ri + 2 * (t + nt * (m + 5 * op))] = h[((t * 5 + m) * 9 + op) * 2 + ri];
  API_END
}

extern "C" int qqcd_set_arnoldi_steps(int steps) {
  API_BEGIN
  set_arnoldi_steps(steps);  // the single-RHS solvers (checks the value)
  g_bp.sstep = steps;        // the batched solver
  API_END
}

extern "C" int qqcd_set_pair_norms(int mode) {
  API_BEGIN
  if (mode != 0 && mode != 1) throw std::runtime_error("pair norms: 0 (from the updated vectors) or 1 (from the projections)");
  g_bp.pair_norms = mode;
  API_END
}

extern "C" int qqcd_set_batch_mid(int prec) {
  API_BEGIN
  if (prec != PREC_SINGLE && prec != PREC_HALF) throw std::runtime_error("intermediate field: 1 (FP32) or 2 (FP16)");
  g_bp.half_mid = prec == PREC_HALF ? 1 : 0;
  API_END
}

extern "C" int qqcd_batch_set(int reorth, int use_graph, double inner_tol, int inner_cycles, int max_outer,
                              int op_prec, int op_recon, int basis_prec) {
  API_BEGIN
  if (basis_prec != PREC_SINGLE && basis_prec != PREC_HALF) throw std::runtime_error("basis_prec is 1 (FP32) or 2 (FP16)");
  if (reorth < 0 || reorth > 2) throw std::runtime_error("reorth must be 0, 1 or 2");
  if (inner_cycles < 1 || max_outer < 1) throw std::runtime_error("cycle limits must be >= 1");
  if (op_prec != PREC_SINGLE && op_prec != PREC_HALF) throw std::runtime_error("inner operator is FP32 or FP16 links");
  if (op_recon != 12 && op_recon != 18) throw std::runtime_error("links are 12 or 18 reals");
  g_bp.reorth = reorth;
  g_bp.use_graph = use_graph;
  g_bp.inner_tol = inner_tol;
  g_bp.inner_cycles = inner_cycles;
  g_bp.max_outer = max_outer;
  // FP32 spinors always; op_prec picks the link storage of the inner operator
  g_bp.op_prec = PREC_SINGLE;
  g_bp.op_recon = op_recon;
  g_bp.basis = (Prec)basis_prec;
  (void)op_prec;
  API_END
}

// Many right-hand sides on the loaded gauge field, as the reference's noise
// loop calls fermprop once per source with inverter 6 (psc): the first is
// solved by the mixed-precision GMRES-DR(m, k) and keeps its deflation
// space, the rest by the batched GMRES-Proj(m - k) over it.  be/bo/xe/xo are
// arrays of nrhs pointers to reference-layout checkerboard spinors.
extern "C" int qqcd_fermprop_multi(int nrhs, const double *const *be, const double *const *bo, const double kappa[2],
                                   const double *coact, const int *bc, double resmax, const int GMRES[2],
                                   const double *u, double *const *xe, double *const *xo, int *itercount,
                                   double *rel_res) {
  API_BEGIN
  Context &c = ctx();
  if (nrhs < 1 || nrhs > 257) throw std::runtime_error("fermprop_multi: 1..257 right-hand sides");
  if (GMRES[1] < 1 || GMRES[1] >= GMRES[0] || GMRES[0] > 120) throw std::runtime_error("GMRES sizes need 0 < k < m <= 120");
  set_bc(c, bc);
  ensure_gauge(c, u, coact);
  SpinorBuf b_e(c.g, PREC_DOUBLE), b_o(c.g, PREC_DOUBLE), src(c.g, PREC_DOUBLE), x_e(c.g, PREC_DOUBLE),
      x_o(c.g, PREC_DOUBLE);
  // first right-hand side
  Deflation D;
  to_device(c, be[0], b_e.s, 0);
  to_device(c, bo[0], b_o.s, 1);
  prep_source(c, kappa, b_e.s, b_o.s, src.s);
  MixedParams mp = mixed_params(GMRES, resmax, PREC_SINGLE, PREC_SINGLE, g_bp.op_recon, 0.0);
  MixedStats ms = solve_mixed(c, kappa, src.s, x_e.s, mp, &D);
  reconstruct_odd(c, kappa, b_o.s, x_e.s, x_o.s);
  to_host(c, x_e.s, xe[0], 0);
  to_host(c, x_o.s, xo[0], 1);
  itercount[0] = ms.inner.iters;
  if (rel_res) rel_res[0] = full_residual(c, kappa, b_e.s, b_o.s, x_e.s, x_o.s);
  const int nb_rest = nrhs - 1;
  if (nb_rest > 0) {
    const int nb = check_nb(0, nb_rest);
    MultiBuf B(c.g, PREC_DOUBLE, nb_rest, nb), X(c.g, PREC_DOUBLE, nb_rest, nb);
    MBlas mb(c.stream, nb_rest);
    for (int r = 1; r < nrhs; r++) {
      to_device(c, be[r], b_e.s, 0);
      to_device(c, bo[r], b_o.s, 1);
      prep_source(c, kappa, b_e.s, b_o.s, src.s);
      mb.put(src.s, B.m, r - 1);
    }
    std::vector<int> id;
    BatchParams bp = batch_params(GMRES, resmax, nb_rest);
    g_bst = batched_solve(c, kappa, B.m, X.m, nb_rest, D, bp, id);
    for (int s = 0; s < nb_rest; s++) {
      const int r = id[s] + 1;
      mb.get(X.m, s, x_e.s);
      to_device(c, be[r], b_e.s, 0);
      to_device(c, bo[r], b_o.s, 1);
      reconstruct_odd(c, kappa, b_o.s, x_e.s, x_o.s);
      to_host(c, x_e.s, xe[r], 0);
      to_host(c, x_o.s, xo[r], 1);
      itercount[r] = g_bst.iters[id[s]];
      if (rel_res) rel_res[r] = full_residual(c, kappa, b_e.s, b_o.s, x_e.s, x_o.s);
    }
  }
  API_END
}

// The disconnected-loop noise loop on the device: nrhs Z4 sources (seed,
// scale), solved as in qqcd_fermprop_multi, and the 16 local trace
// estimators summed over the sources:
//   traces[(t*16 + n)*2 + {0,1}] = (1/(nrhs scale^2)) sum_r sum_{x in t} eta_r^dag Gamma_n x_r
// res_full[r]: ||eta_r - M x_r|| / ||eta_r|| on the whole lattice, iters[r]
// the Krylov steps spent on source r, times = {first solve, batched solves,
// whole call} in seconds; stats = {outer, cycles, batched matvecs, graph
// replays, fallback, nb, order, first-solve outer steps}.
// qqcd_noise_loop_lagfib: the same loop with source r the Z4 noise z4source
// draws on its r-th call from the lagged-Fibonacci state (iw, jr, kr)
// (384 nvhalf draws per source, in order), the state advanced past all
// nrhs sources.
namespace {

// source r of the loop into (even, odd) checkerboards
typedef std::function<void(long r, Spinor &e, Spinor &o)> SourceGen;

void noise_loop_impl(Context &c, int nrhs, const SourceGen &gen, double scale, const double kappa[2],
                     const double *coact, const int *bc, const double *u, const int GMRES[2], double tol, int nb,
                     int method, double *traces, int ksub, const int *io, double *loops, double *res_full,
                     int *iters, double *times, int *stats);

}  // namespace

extern "C" int qqcd_noise_loop(int nrhs, unsigned long long seed, double scale, const double kappa[2],
                               const double *coact, const int *bc, const double *u, const int GMRES[2], double tol,
                               int nb, int method, double *traces, int ksub, const int *io, double *loops,
                               double *res_full, int *iters, double *times, int *stats) {
  API_BEGIN
  Context &c = ctx();
  const SourceGen gen = [&](long r, Spinor &e, Spinor &o) {
    z4_source(c.g, e, 0, seed, r, scale, c.stream);
    z4_source(c.g, o, 1, seed, r, scale, c.stream);
  };
  noise_loop_impl(c, nrhs, gen, scale, kappa, coact, bc, u, GMRES, tol, nb, method, traces, ksub, io, loops,
                  res_full, iters, times, stats);
  API_END
}

extern "C" int qqcd_noise_loop_lagfib(int nrhs, int *iw, int *jr, int *kr, double scale, const double kappa[2],
                                      const double *coact, const int *bc, const double *u, const int GMRES[2],
                                      double tol, int nb, int method, double *traces, int ksub, const int *io,
                                      double *loops, double *res_full, int *iters, double *times, int *stats) {
  API_BEGIN
  Context &c = ctx();
  // source r = the draws z4source would make on its r-th call from this state:
  // the stream state at the start of every source is kept, a source is drawn
  // again from its state when the loop needs it a second time
  LagState cur = lag_in(iw, jr, kr);
  const long nvh = c.g.V / 32;
  DevU32 y(384 * nvh);
  std::vector<LagState> at;
  const SourceGen gen = [&](long r, Spinor &e, Spinor &o) {
    if (r == (long)at.size()) {
      at.push_back(cur);
      lagfib_generate(cur, 384 * nvh, y.p, c.stream);
    } else if (r < (long)at.size()) {
      LagState t = at[r];
      lagfib_generate(t, 384 * nvh, y.p, c.stream);
    } else {
      throw std::runtime_error("noise_loop_lagfib: sources must first be drawn in order");
    }
    lag_noise_source(c.g, e, 0, y.p, 0, scale, c.stream);
    lag_noise_source(c.g, o, 1, y.p + 192 * nvh, 0, scale, c.stream);
  };
  noise_loop_impl(c, nrhs, gen, scale, kappa, coact, bc, u, GMRES, tol, nb, method, traces, ksub, io, loops,
                  res_full, iters, times, stats);
  if ((long)at.size() != nrhs) throw std::runtime_error("noise_loop_lagfib: not every source was drawn");
  lag_out(cur, iw, jr, kr);
  API_END
}

namespace {

void noise_loop_impl(Context &c, int nrhs, const SourceGen &gen, double scale, const double kappa[2],
                     const double *coact, const int *bc, const double *u, const int GMRES[2], double tol, int nb,
                     int method, double *traces, int ksub, const int *io, double *loops, double *res_full,
                     int *iters, double *times, int *stats) {
  if (nrhs < 1 || nrhs > 257) throw std::runtime_error("noise_loop: 1..257 right-hand sides");
  if (GMRES[1] < 1 || GMRES[1] >= GMRES[0] || GMRES[0] > 120) throw std::runtime_error("GMRES sizes need 0 < k < m <= 120");
  if (!(scale > 0)) throw std::runtime_error("noise_loop: scale must be positive");
  if (method < 0 || method > 2) throw std::runtime_error("noise_loop: method must be 0, 1 or 2");
  set_bc(c, bc);
  if (u) ensure_gauge(c, u, coact);
  else if (!c.have_gauge) throw std::runtime_error("no gauge field loaded");
  else if (coact) set_coact(c, coact);
  QQCD_HIP(hipDeviceSynchronize());
  const auto t0 = std::chrono::steady_clock::now();
  const int nt = c.g.n[3];
  DevArr tr(nullptr, (size_t)nt * 32, c.stream);
  DevArr part(nullptr, (size_t)trace16_gx(c.g) * nt * 32, c.stream);
  SpinorBuf e_e(c.g, PREC_DOUBLE), e_o(c.g, PREC_DOUBLE), src(c.g, PREC_DOUBLE), x_e(c.g, PREC_DOUBLE),
      x_o(c.g, PREC_DOUBLE);
  if (ksub != 0 && ksub != 1) throw std::runtime_error("noise_loop: ksub is 0 or 1");
  std::unique_ptr<DevArr> lo, lpart;
  if (loops) {
    lo.reset(new DevArr(nullptr, (size_t)nt * 90, c.stream));
    lpart.reset(new DevArr(nullptr, (size_t)2 * nt * loopsum_gx(c.g) * 58, c.stream));
  }
  auto noise = [&](long r) { gen(r, e_e.s, e_o.s); };
  auto finish = [&](long r, int first) {
    // x_o, the whole-lattice residual, and the traces of source r
    reconstruct_odd(c, kappa, e_o.s, x_e.s, x_o.s);
    res_full[r] = full_residual(c, kappa, e_e.s, e_o.s, x_e.s, x_o.s);
    trace16(c.g, e_e.s, x_e.s, tr.p, !first, part.p, c.stream);
    trace16(c.g, e_o.s, x_o.s, tr.p, 1, part.p, c.stream);
    if (loops) {
      // the reference's operators of this source (nsaverage: currents times kappa)
      const Spinor bs[2] = {e_e.s, e_o.s}, xs[2] = {x_e.s, x_o.s};
      loopsum(c.g, c.U64, bs, xs, ksub, io, io ? 5 : 1, 1.0, kappa[0], lo->p, !first, lpart->p, nullptr,
              c.stream);
    }
  };
  Deflation D;
  std::memset(stats, 0, 8 * sizeof(int));
  noise(0);
  prep_source(c, kappa, e_e.s, e_o.s, src.s);
  if (method == 1) {
    // the reference's psc semantics in plain FP64: GMRES-DR on the first
    // source, GMRES-Proj over its deflation space on each later one
    SchurOp A(c, PREC_DOUBLE, 18, kappa);
    LinOp op = A.op();
    fill_spinor(c.g, x_e.s, 0.0, 0.0, c.stream);
    SolveStats st = gmresdr(c, op, src.s, x_e.s, GMRES[0], GMRES[1], tol, 0, 5000 * GMRES[0], &D);
    iters[0] = st.iters;
  } else {
    MixedParams mp = mixed_params(GMRES, tol, PREC_SINGLE, PREC_SINGLE, g_bp.op_recon, 0.0);
    MixedStats ms = solve_mixed(c, kappa, src.s, x_e.s, mp, &D);
    iters[0] = ms.inner.iters;
    stats[7] = ms.outer;
  }
  finish(0, 1);
  QQCD_HIP(hipDeviceSynchronize());
  const auto t1 = std::chrono::steady_clock::now();
  const int rest = nrhs - 1;
  if (rest > 0 && method != 0) {
    // one source after the other (the reference's loop over fermprop)
    SchurOp A(c, PREC_DOUBLE, 18, kappa);
    LinOp op = A.op();
    for (int r = 1; r < nrhs; r++) {
      noise(r);
      prep_source(c, kappa, e_e.s, e_o.s, src.s);
      fill_spinor(c.g, x_e.s, 0.0, 0.0, c.stream);
      if (method == 1) {
        SolveStats st = gmresproj(c, op, src.s, x_e.s, GMRES[0] - GMRES[1], D, tol, 0, 5000 * GMRES[0]);
        iters[r] = st.iters;
      } else {
        MixedParams mp = mixed_params(GMRES, tol, PREC_SINGLE, PREC_SINGLE, g_bp.op_recon, 0.0);
        MixedStats ms = solve_mixed(c, kappa, src.s, x_e.s, mp, &D);
        iters[r] = ms.inner.iters;
      }
      finish(r, 0);
    }
  } else if (rest > 0) {
    TuneResult tr_ = autotune_multi(c, PREC_SINGLE, g_bp.op_recon, rest);
    const int nbx = nb > 0 ? check_nb(nb, rest) : tr_.nb;
    MultiBuf B(c.g, PREC_DOUBLE, rest, nbx), X(c.g, PREC_DOUBLE, rest, nbx);
    MBlas mb(c.stream, rest);
    for (int r = 1; r < nrhs; r++) {
      noise(r);
      prep_source(c, kappa, e_e.s, e_o.s, src.s);
      mb.put(src.s, B.m, r - 1);
    }
    std::vector<int> id;
    BatchParams bp = batch_params(GMRES, tol, rest);
    if (nb <= 0) bp.tune = tr_.t;
    g_bst = batched_solve(c, kappa, B.m, X.m, rest, D, bp, id);
    for (int s = 0; s < rest; s++) {
      const int r = id[s] + 1;
      mb.get(X.m, s, x_e.s);
      noise(r);
      iters[r] = g_bst.iters[id[s]];
      finish(r, 0);
    }
    stats[0] = g_bst.outer;
    stats[1] = g_bst.cycles;
    stats[2] = (int)g_bst.batched_matvecs;
    stats[3] = g_bst.graph_replays;
    stats[4] = g_bst.fallback;
    stats[5] = nbx;
    stats[6] = bp.tune.order;
  }
  QQCD_HIP(hipDeviceSynchronize());
  const auto t2 = std::chrono::steady_clock::now();
  std::vector<double> h((size_t)nt * 32);
  QQCD_HIP(hipMemcpy(h.data(), tr.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  const double f = 1.0 / (nrhs * scale * scale);
  for (size_t i = 0; i < h.size(); i++) traces[i] = f * h[i];
  if (loops) {
    std::vector<double> hl((size_t)nt * 90);
    QQCD_HIP(hipMemcpy(hl.data(), lo->p, hl.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < hl.size(); i++) loops[i] = f * hl[i];
  }
  times[0] = std::chrono::duration<double>(t1 - t0).count();
  times[1] = std::chrono::duration<double>(t2 - t1).count();
  times[2] = std::chrono::duration<double>(t2 - t0).count();
}

}  // namespace

// The full propagator of one source site: nrhs (<= 12) point sources
// (spin r/3 + 1, colour r%3 + 1, pointsource's numbering) solved on the
// resident links to ||b - M x|| < tol ||b|| each.  The first runs through
// the mixed solver (GMRES-DR(m, k) inner solves with op_prec / op_recon
// links, FP32 basis, FP64 reliable updates) and keeps its deflation space;
// the rest: method 0 all together through the batched GMRES-Proj(m - k)
// over it, 1 one by one through the mixed solver with GMRES-Proj inner
// solves over it, 2 one by one through the mixed solver with the learned
// network as the flexible right preconditioner (dropped for the
// unpreconditioned solver when an outer step stalls).
extern "C" int qqcd_propagator(int nrhs, const int src[4], const double kappa[2], const int GMRES[2], double tol,
                               int op_prec, int op_recon, int method, double *const *xe, double *const *xo,
                               int *iters, double *res_full, double *times, int *stats) {
  API_BEGIN
  Context &c = ctx();
  if (!c.have_gauge) throw std::runtime_error("no gauge field loaded");
  if (nrhs < 1 || nrhs > 12) throw std::runtime_error("propagator: 1..12 right-hand sides (spin x colour)");
  if (GMRES[1] < 1 || GMRES[1] >= GMRES[0] || GMRES[0] > 120) throw std::runtime_error("GMRES sizes need 0 < k < m <= 120");
  if (method < 0 || method > 2) throw std::runtime_error("propagator: method 0, 1 or 2");
  if (op_prec != PREC_SINGLE && op_prec != PREC_HALF) throw std::runtime_error("propagator: FP32 or FP16 operator");
  QQCD_HIP(hipDeviceSynchronize());
  const auto t0 = std::chrono::steady_clock::now();
  SpinorBuf b_e(c.g, PREC_DOUBLE), b_o(c.g, PREC_DOUBLE), srcv(c.g, PREC_DOUBLE), x_e(c.g, PREC_DOUBLE),
      x_o(c.g, PREC_DOUBLE);
  auto source = [&](int r) {
    fill_spinor(c.g, b_e.s, 0.0, 0.0, c.stream);
    fill_spinor(c.g, b_o.s, 0.0, 0.0, c.stream);
    point_source_dev(c, b_e.s, b_o.s, src, r / 3 + 1, r % 3 + 1);
  };
  auto finish = [&](int r) {
    reconstruct_odd(c, kappa, b_o.s, x_e.s, x_o.s);
    res_full[r] = full_residual(c, kappa, b_e.s, b_o.s, x_e.s, x_o.s);
    if (xe && xe[r]) to_host(c, x_e.s, xe[r], 0);
    if (xo && xo[r]) to_host(c, x_o.s, xo[r], 1);
  };
  std::memset(stats, 0, 8 * sizeof(int));
  Deflation D;
  MixedParams mp = mixed_params(GMRES, tol, PREC_SINGLE, op_prec, op_recon, 0.0);
  source(0);
  prep_source(c, kappa, b_e.s, b_o.s, srcv.s);
  fill_spinor(c.g, x_e.s, 0.0, 0.0, c.stream);
  MixedStats ms = solve_mixed(c, kappa, srcv.s, x_e.s, mp, &D);
  iters[0] = ms.inner.iters;
  stats[7] = ms.outer;
  finish(0);
  QQCD_HIP(hipDeviceSynchronize());
  const auto t1 = std::chrono::steady_clock::now();
  const int rest = nrhs - 1;
  if (rest > 0 && method == 0) {
    // the largest RHS block <= 16 that pads the slot count by at most 1/8:
    // every slot of a block is allocated in each of the m + 1 basis fields
    int nb = 1;
    for (int cand = 16; cand >= 1; cand >>= 1)
      if (((rest + cand - 1) / cand * cand - rest) * 8 <= rest) {
        nb = cand;
        break;
      }
    MultiBuf B(c.g, PREC_DOUBLE, rest, nb), X(c.g, PREC_DOUBLE, rest, nb);
    MBlas mb(c.stream, rest);
    for (int r = 1; r < nrhs; r++) {
      source(r);
      prep_source(c, kappa, b_e.s, b_o.s, srcv.s);
      mb.put(srcv.s, B.m, r - 1);
    }
    std::vector<int> id;
    BatchParams bp = batch_params(GMRES, tol, rest);
    g_bst = batched_solve(c, kappa, B.m, X.m, rest, D, bp, id);
    for (int s = 0; s < rest; s++) {
      const int r = id[s] + 1;
      mb.get(X.m, s, x_e.s);
      source(r);
      iters[r] = g_bst.iters[id[s]];
      finish(r);
    }
    stats[0] = g_bst.outer;
    stats[1] = g_bst.cycles;
    stats[2] = (int)g_bst.batched_matvecs;
    stats[3] = g_bst.graph_replays;
    stats[4] = g_bst.fallback;
    stats[5] = nb;
  } else if (rest > 0) {
    LinOp pre;
    if (method == 2) pre = nn_precond(c);
    for (int r = 1; r < nrhs; r++) {
      source(r);
      prep_source(c, kappa, b_e.s, b_o.s, srcv.s);
      fill_spinor(c.g, x_e.s, 0.0, 0.0, c.stream);
      MixedParams q = mp;
      if (method == 2) q.precond = &pre;
      MixedStats m2 = solve_mixed(c, kappa, srcv.s, x_e.s, q, method == 2 ? nullptr : &D);
      iters[r] = m2.inner.iters;
      stats[0] += m2.outer;
      stats[4] += m2.fallback != 0;
      stats[6] += m2.precond_dropped;
      finish(r);
    }
  }
  QQCD_HIP(hipDeviceSynchronize());
  const auto t2 = std::chrono::steady_clock::now();
  times[0] = std::chrono::duration<double>(t1 - t0).count();
  times[1] = std::chrono::duration<double>(t2 - t1).count();
  times[2] = std::chrono::duration<double>(t2 - t0).count();
  API_END
}

extern "C" int qqcd_batch_stats(int *outer, int *cycles, long *slot_matvecs, long *batched_matvecs,
                                int *graph_replays, int *fallback) {
  API_BEGIN
  *outer = g_bst.outer;
  *cycles = g_bst.cycles;
  *slot_matvecs = g_bst.slot_matvecs;
  *batched_matvecs = g_bst.batched_matvecs;
  *graph_replays = g_bst.graph_replays;
  *fallback = g_bst.fallback;
  API_END
}

extern "C" int qqcd_batch_rejected(int *rejected) {
  API_BEGIN
  *rejected = g_bst.rejected;
  API_END
}

extern "C" int qqcd_autotune_multi(int prec, int recon, int nrhs, int *nb, int *order, double *ms) {
  API_BEGIN
  Context &c = ctx();
  if (prec != PREC_DOUBLE && prec != PREC_SINGLE) throw std::runtime_error("multi-RHS fields are FP64 or FP32");
  if (nrhs < 1 || nrhs > 256) throw std::runtime_error("1..256 right-hand sides");
  if (!c.have_gauge) throw std::runtime_error("no gauge field loaded");
  TuneResult t = autotune_multi(c, (Prec)prec, recon, nrhs);
  *nb = t.nb;
  *order = t.t.order;
  *ms = t.ms;
  API_END
}

// Time the multi-RHS hopping kernel (one checkerboard, random links of the
// loaded field, zero-cost epilogue) for nact active of nrhs slots.
extern "C" int qqcd_bench_dslash_multi(int prec, int recon, int nrhs, int nact, int nb, int order, int nwarm,
                                       int niter, double *ms) {
  API_BEGIN
  Context &c = ctx();
  if (prec != PREC_DOUBLE && prec != PREC_SINGLE) throw std::runtime_error("multi-RHS fields are FP64 or FP32");
  if (nrhs < 1 || nrhs > 256 || nact < 0 || nact > nrhs) throw std::runtime_error("bad right-hand-side counts");
  if (!c.have_gauge) throw std::runtime_error("no gauge field loaded");
  nb = check_nb(nb, nrhs);
  const Gauge &U = c.gauge((Prec)prec, recon);
  MultiBuf a(c.g, (Prec)prec, nrhs, nb), b(c.g, (Prec)prec, nrhs, nb);
  MBlas mb(c.stream, nrhs);
  {
    SpinorBuf s(c.g, (Prec)prec);
    for (int r = 0; r < nrhs; r++) {
      random_spinor(c.g, s.s, 1000 + r, c.stream);
      mb.put(s.s, a.m, r);
    }
  }
  int *d_n = nullptr;
  QQCD_HIP(hipMalloc(&d_n, sizeof(int)));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  try {
    QQCD_HIP(hipMemcpy(d_n, &nact, sizeof(int), hipMemcpyHostToDevice));
    MultiTune t;
    t.order = order;
    for (int w = 0; w < nwarm; w++) dslash_multi(c.g, U, a.m, b.m, w & 1, 0, Epilogue(), nullptr, d_n, t, c.stream);
    QQCD_HIP(hipEventCreate(&e0));
    QQCD_HIP(hipEventCreate(&e1));
    QQCD_HIP(hipEventRecord(e0, c.stream));
    for (int w = 0; w < niter; w++) dslash_multi(c.g, U, a.m, b.m, w & 1, 0, Epilogue(), nullptr, d_n, t, c.stream);
    QQCD_HIP(hipEventRecord(e1, c.stream));
    QQCD_HIP(hipEventSynchronize(e1));
    float f = 0;
    QQCD_HIP(hipEventElapsedTime(&f, e0, e1));
    *ms = f / std::max(niter, 1);
  } catch (...) {
    (void)hipFree(d_n);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipFree(d_n);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  API_END
}
