// qqcd_capi.cpp -- extern "C" entry points: life cycle, reference-order
// operators and the device-resident benchmark/test interface.
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qqcd_hip.h"
#include "qqcd_context.h"
#include "qqcd_lagfib.h"

namespace qqcd {
static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }
}  // namespace qqcd

using namespace qqcd;

#define API_BEGIN try {
#define API_END                       \
  }                                   \
  catch (const std::exception &e) {   \
    set_error(e.what());              \
    return -1;                        \
  }                                   \
  return 0;

extern "C" const char *qqcd_last_error(void) { return g_err.c_str(); }

extern "C" int qqcd_init(const int n[4], const int bc[4], int device) {
  API_BEGIN
  qqcd_deflation_clear();
  if (g_ctx) {
    delete g_ctx;
    g_ctx = nullptr;
  }
  for (int m = 0; m < 4; m++)
    if (n[m] <= 0 || n[m] % 4 != 0)
      throw std::runtime_error("lattice extents must be positive multiples of 4 (2 blocks x 2 checkerboards)");
  int ndev = 0;
  QQCD_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) throw std::runtime_error("no such HIP device");
  QQCD_HIP(hipSetDevice(device));
  lagfib_noalloc_setup(device);
  Context *c = new Context();
  c->device = device;
  Geom &g = c->g;
  for (int m = 0; m < 4; m++) {
    g.n[m] = n[m];
    g.c1[m] = 1.0;
    g.c2[m] = 1.0;
  }
  g.V = (long)n[0] * n[1] * n[2] * n[3];
  if (g.V / 2 >= (1L << 30)) throw std::runtime_error("lattice too large for 32-bit site indices");
  g.Vh = g.V / 2;
  g.stride = (g.Vh + 63) / 64 * 64;
  set_bc(*c, bc);
  QQCD_HIP(hipStreamCreate(&c->stream));
  c->blas.reset(new Blas(c->stream));
  for (double *&p : c->dref) QQCD_HIP(hipMalloc(&p, sizeof(double) * 6 * ref_ntotal(g) * 64));
  g_ctx = c;
  API_END
}

namespace qqcd {
void nn_release_device();

void forget_host_gauge(const void *u) {
  if (g_ctx && g_ctx->u_host == u) g_ctx->u_host = nullptr;
}

}  // namespace qqcd

extern "C" int qqcd_finalize(void) {
  API_BEGIN
  qqcd_deflation_clear();
  nn_release_device();
  lagfib_release_device();
  delete g_ctx;
  g_ctx = nullptr;
  API_END
}

extern "C" int qqcd_ref_sizes(long *nvhalf, long *ntotal) {
  API_BEGIN
  Context &c = ctx();
  *nvhalf = ref_nvhalf(c.g);
  *ntotal = ref_ntotal(c.g);
  API_END
}

extern "C" int qqcd_gauge_load(const double *u, const double *coact) {
  API_BEGIN
  Context &c = ctx();
  c.have_gauge = 0;
  c.u_host = nullptr;
  ensure_gauge(c, u, coact);
  API_END
}

extern "C" int qqcd_gauge_load_canonical(const double *U, const double *coact) {
  API_BEGIN
  Context &c = ctx();
  set_coact(c, coact);
  const size_t n = (size_t)2 * 36 * c.g.V;
  double *d = nullptr;
  QQCD_HIP(hipMalloc(&d, n * sizeof(double)));
  QQCD_HIP(hipMemcpyAsync(d, U, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  if (!c.U64.data) c.U64 = alloc_gauge(c.g, PREC_DOUBLE, 18);
  canon_to_gauge(c.g, d, c.U64, c.stream);
  QQCD_HIP(hipStreamSynchronize(c.stream));
  (void)hipFree(d);
  c.u_host = nullptr;
  c.have_gauge = 1;
  c.gauge_gen++;
  API_END
}

static size_t ref_spinor_bytes(const Context &c) { return sizeof(double) * 6 * ref_ntotal(c.g) * 64; }

static void upload_ref(Context &c, const double *h, double *d) {
  QQCD_HIP(hipMemcpyAsync(d, h, ref_spinor_bytes(c), hipMemcpyHostToDevice, c.stream));
}
static void download_ref(Context &c, const double *d, double *h) {
  QQCD_HIP(hipMemcpyAsync(h, d, ref_spinor_bytes(c), hipMemcpyDeviceToHost, c.stream));
  QQCD_HIP(hipStreamSynchronize(c.stream));
}

extern "C" int qqcd_hsingle(double *gout, const double *u, const double *gin, int idag, const double *coact,
                            const int *bc, int gblclr) {
  API_BEGIN
  Context &c = ctx();
  if (gblclr != 1 && gblclr != 2) throw std::runtime_error("gblclr must be 1 or 2");
  if (idag != 0 && idag != 1) throw std::runtime_error("idag must be 0 or 1");
  set_bc(c, bc);
  ensure_gauge(c, u, coact);
  const int pout = gblclr - 1, pin = 1 - pout;
  SpinorBuf in(c.g, PREC_DOUBLE), out(c.g, PREC_DOUBLE);
  upload_ref(c, gin, c.dref[0]);
  ref_to_spinor(c.g, c.dref[0], in.s, pin, c.stream);
  dslash(c.g, c.U64, in.s, out.s, pout, idag, Epilogue(), nullptr, c.stream);
  spinor_to_ref(c.g, out.s, c.dref[1], pout, c.stream);
  download_ref(c, c.dref[1], gout);
  API_END
}

static void schur_ref(Context &c, double *he, const double *ge, int idag, const double kappa[2]) {
  SpinorBuf in(c.g, PREC_DOUBLE), out(c.g, PREC_DOUBLE), tmp(c.g, PREC_DOUBLE);
  upload_ref(c, ge, c.dref[0]);
  ref_to_spinor(c.g, c.dref[0], in.s, 0, c.stream);
  schur_apply(c, c.U64, kappa, idag, in.s, out.s, tmp.s);
  spinor_to_ref(c.g, out.s, c.dref[1], 0, c.stream);
  download_ref(c, c.dref[1], he);
}

extern "C" int qqcd_hdouble(double *he, const double *u, const double *ge, int idag, const double *coact,
                            double kappa, const int *bc) {
  API_BEGIN
  Context &c = ctx();
  set_bc(c, bc);
  ensure_gauge(c, u, coact);
  const double k[2] = {kappa, 0.0};
  schur_ref(c, he, ge, idag, k);
  API_END
}

extern "C" int qqcd_hdbletm(double *he, const double *u, const double *ge, int idag, const double *coact,
                            const double kappa[2], int iflag, const int *bc) {
  API_BEGIN
  Context &c = ctx();
  if (iflag != -1) throw std::runtime_error("Hdbletm: only iflag = -1 (Wilson twisted mass) is supported");
  set_bc(c, bc);
  ensure_gauge(c, u, coact);
  schur_ref(c, he, ge, idag, kappa);
  API_END
}

namespace qqcd {
void smear(Context &c, Spinor h[2], Spinor t[2], double asmear, int nsmear);
}

extern "C" int qqcd_smear(double *he, double *ho, const double *u, const double *ge, const double *go, double asmear,
                          int nsmear, const int *bc) {
  API_BEGIN
  Context &c = ctx();
  if (!he || !ho || !ge || !go) throw std::runtime_error("smear: null field");
  set_bc(c, bc);
  ensure_gauge(c, u, nullptr);
  SpinorBuf h0(c.g, PREC_DOUBLE), h1(c.g, PREC_DOUBLE), t0(c.g, PREC_DOUBLE), t1(c.g, PREC_DOUBLE);
  upload_ref(c, ge, c.dref[0]);
  ref_to_spinor(c.g, c.dref[0], h0.s, 0, c.stream);
  upload_ref(c, go, c.dref[1]);
  ref_to_spinor(c.g, c.dref[1], h1.s, 1, c.stream);
  Spinor h[2] = {h0.s, h1.s}, t[2] = {t0.s, t1.s};
  smear(c, h, t, asmear, nsmear);
  spinor_to_ref(c.g, h[0], c.dref[0], 0, c.stream);
  download_ref(c, c.dref[0], he);
  spinor_to_ref(c.g, h[1], c.dref[1], 1, c.stream);
  download_ref(c, c.dref[1], ho);
  API_END
}

// twMshift / gamma5mult share the Wilson-twisted checks of the reference (iflag = -1 only)
static void shift_checks(int iflag, int idag, const double kappa[2]) {
  if (iflag != -1) throw std::runtime_error("multiple shifts need the Wilson operator (iflag = -1), no clover term");
  if (idag != 0 && idag != 1) throw std::runtime_error("idag must be 0 or 1");
  if (!kappa || kappa[0] == 0.0 || kappa[1] == 0.0)
    throw std::runtime_error("twMshift / gamma5mult divide by 2 kappa mu: kappa(1) and kappa(2) must be nonzero");
}

extern "C" int qqcd_twmshift(double *hwhole, const double *u, const double *gwhole, int idag, const double *coact,
                             const double kappa[2], int iflag, const int *bc) {
  API_BEGIN
  Context &c = ctx();
  shift_checks(iflag, idag, kappa);
  set_bc(c, bc);
  ensure_gauge(c, u, coact);
  const size_t half = 6 * (size_t)ref_ntotal(c.g) * 64;
  SpinorBuf g0(c.g, PREC_DOUBLE), g1(c.g, PREC_DOUBLE), out(c.g, PREC_DOUBLE);
  upload_ref(c, gwhole, c.dref[0]);
  ref_to_spinor(c.g, c.dref[0], g0.s, 0, c.stream);
  upload_ref(c, gwhole + half, c.dref[1]);
  ref_to_spinor(c.g, c.dref[1], g1.s, 1, c.stream);
  const Spinor *g[2] = {&g0.s, &g1.s};
  // h_p = g_p + fac2 (-i g5) g_p + fac1 (-i g5) H g_other (diracops.f90:1925-1987)
  Epilogue e;
  e.a0 = 1.0;
  e.a1 = -1.0 / (2.0 * kappa[0] * kappa[1]);
  e.b0 = 0.0;
  e.b1 = -1.0 / (2.0 * kappa[1]);
  for (int p = 0; p < 2; p++) {
    dslash(c.g, c.U64, *g[1 - p], out.s, p, idag, e, g[p], c.stream);
    spinor_to_ref(c.g, out.s, c.dref[2], p, c.stream);
    download_ref(c, c.dref[2], hwhole + p * half);
  }
  API_END
}

extern "C" int qqcd_gamma5mult(double *he, const double *u, const double *ge, int idag, const double *coact,
                               const double kappa[2], int iflag, const int *bc) {
  API_BEGIN
  Context &c = ctx();
  shift_checks(iflag, idag, kappa);
  set_bc(c, bc);
  ensure_gauge(c, u, coact);
  SpinorBuf g0(c.g, PREC_DOUBLE), t(c.g, PREC_DOUBLE), out(c.g, PREC_DOUBLE);
  upload_ref(c, ge, c.dref[0]);
  ref_to_spinor(c.g, c.dref[0], g0.s, 0, c.stream);
  const double k = kappa[0], mu = kappa[1], a = 2.0 * k * mu;
  // htmp = (1 + 2 kappa mu (-i g5)) H_oe g (diracops.f90:1800-1812)
  Epilogue e1;
  e1.b0 = 1.0;
  e1.b1 = -a;
  dslash(c.g, c.U64, g0.s, t.s, 1, idag, e1, nullptr, c.stream);
  // he = fac1 g + fac4 fac1 (-i g5) g - fac4 fac3 (-i g5) H_eo htmp (:1815-1836)
  const double fac1 = 1.0 / (k * k), fac3 = 1.0 / (1.0 + a * a), fac4 = 1.0 / a;
  Epilogue e2;
  e2.a0 = fac1;
  e2.a1 = -fac4 * fac1;
  e2.b0 = 0.0;
  e2.b1 = fac4 * fac3;
  dslash(c.g, c.U64, t.s, out.s, 0, idag, e2, &g0.s, c.stream);
  spinor_to_ref(c.g, out.s, c.dref[1], 0, c.stream);
  download_ref(c, c.dref[1], he);
  API_END
}

extern "C" int qqcd_gammamult(const double *even_in, const double *odd_in, double *even_out, double *odd_out,
                              int mu, int ieo, int ibleo) {
  API_BEGIN
  Context &c = ctx();
  if (ieo < 1 || ieo > 2 || ibleo < 1 || ibleo > 8) throw std::runtime_error("gammamult: bad (ieo, ibleo)");
  const int sl = (ieo - 1) + 2 * (ibleo - 1);
  const double *ins[2] = {even_in, odd_in};
  double *outs[2] = {even_out, odd_out};
  for (int q = 0; q < 2; q++) {
    upload_ref(c, ins[q], c.dref[0]);
    upload_ref(c, outs[q], c.dref[1]);
    gamma_ref(c.g, c.dref[0], c.dref[1], mu, 1, &sl, c.stream);
    download_ref(c, c.dref[1], outs[q]);
  }
  API_END
}

extern "C" int qqcd_gam5x(double *a) {
  API_BEGIN
  Context &c = ctx();
  int sl[16];
  for (int k = 0; k < 16; k++) sl[k] = k;
  upload_ref(c, a, c.dref[0]);
  gamma_ref(c.g, c.dref[0], c.dref[0], 5, 16, sl, c.stream);
  download_ref(c, c.dref[0], a);
  API_END
}

extern "C" int qqcd_vecdot(const double *a, const double *b, double adb[2]) {
  API_BEGIN
  Context &c = ctx();
  SpinorBuf sa(c.g, PREC_DOUBLE), sb(c.g, PREC_DOUBLE);
  upload_ref(c, a, c.dref[0]);
  upload_ref(c, b, c.dref[1]);
  ref_to_spinor(c.g, c.dref[0], sa.s, 0, c.stream);
  ref_to_spinor(c.g, c.dref[1], sb.s, 0, c.stream);
  std::complex<double> r = c.blas->dot(as_vec(sa.s), as_vec(sb.s));
  adb[0] = r.real();
  adb[1] = r.imag();
  API_END
}

extern "C" int qqcd_pointsource(double *be, double *bo, const int src[4], int icsrc, int idsrc) {
  API_BEGIN
  Context &c = ctx();
  const Geom &g = c.g;
  int x[4];
  for (int m = 0; m < 4; m++) {
    if (src[m] < 1 || src[m] > g.n[m]) throw std::runtime_error("pointsource: site outside the lattice");
    x[m] = src[m] - 1;
  }
  if (icsrc < 1 || icsrc > 3 || idsrc < 1 || idsrc > 4) throw std::runtime_error("pointsource: bad colour/Dirac");
  const long ntot = ref_ntotal(g);
  const size_t len = (size_t)6 * ntot * 64;
  std::memset(be, 0, len * sizeof(double));
  std::memset(bo, 0, len * sizeof(double));
  int b[4], s[4];
  for (int m = 0; m < 4; m++) {
    b[m] = x[m] & 1;
    s[m] = x[m] >> 1;
  }
  const long L = s[0] + (long)(g.n[0] / 2) * (s[1] + (long)(g.n[1] / 2) * (s[2] + (long)(g.n[2] / 2) * s[3]));
  const long i = L / 2;
  const int ieo = (s[0] + s[1] + s[2] + s[3]) & 1;
  const int ibl = b[0] + 2 * b[1] + 4 * b[2] + 8 * b[3];
  const int gbl = (b[0] + b[1] + b[2] + b[3]) & 1;
  double *dst = gbl == 0 ? be : bo;
  dst[(2 * icsrc - 2) + 6 * (i + ntot * ((idsrc - 1) + 4 * (ieo + 2 * (long)(ibl >> 1))))] = 1.0;
  API_END
}

// ---- native interface ----------------------------------------------------

extern "C" int qqcd_set_xcd_remap(int on) {
  API_BEGIN
  set_xcd_remap(on ? 1 : 0);
  API_END
}

extern "C" int qqcd_set_dslash_tune(int order, int bz, int depth, int nt) {
  API_BEGIN
  DslashTune t;
  t.order = order;
  t.bz = bz;
  t.depth = depth;
  t.nt = nt;
  set_dslash_tune(t);
  API_END
}

extern "C" int qqcd_set_multi_tune(int kernel, int bz, int nt) {
  API_BEGIN
  set_multi_override(kernel, bz, nt);
  API_END
}

extern "C" int qqcd_set_multi_lds(int bytes) {
  API_BEGIN
  set_multi_lds(bytes);
  API_END
}

extern "C" int qqcd_bench_dslash(int prec, int recon, int which, int nwarm, int niter, double *ms, double *relerr) {
  API_BEGIN
  Context &c = ctx();
  if (prec < 0 || prec > 2 || (recon != 18 && recon != 12)) throw std::runtime_error("bad precision / recon");
  if (niter < 1) throw std::runtime_error("niter must be >= 1");
  const Prec p = (Prec)prec;
  const Gauge &U = c.gauge(p, recon);
  SpinorBuf in64(c.g, PREC_DOUBLE), ref(c.g, PREC_DOUBLE), tmp64(c.g, PREC_DOUBLE);
  SpinorBuf in(c.g, p), out(c.g, p), tmp(c.g, p), out64(c.g, PREC_DOUBLE);
  random_spinor(c.g, in64.s, 12345, c.stream);
  spinor_copy_convert(c.g, in64.s, in.s, c.stream);
  const double kap[2] = {0.156, 0.005};
  auto apply = [&](const Gauge &G, const Spinor &a, Spinor &o, Spinor &t) {
    if (which == 0) dslash(c.g, G, a, o, 1, 0, Epilogue(), nullptr, c.stream);
    else schur_apply(c, G, kap, 0, a, o, t);
  };
  for (int k = 0; k < nwarm; k++) apply(U, in.s, out.s, tmp.s);
  hipEvent_t e0, e1;
  QQCD_HIP(hipEventCreate(&e0));
  QQCD_HIP(hipEventCreate(&e1));
  QQCD_HIP(hipEventRecord(e0, c.stream));
  for (int k = 0; k < niter; k++) apply(U, in.s, out.s, tmp.s);
  QQCD_HIP(hipEventRecord(e1, c.stream));
  QQCD_HIP(hipEventSynchronize(e1));
  float t = 0;
  QQCD_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = (double)t / niter / (which == 0 ? 1 : 2);
  // accuracy against FP64 / 18 on the same (FP64) input
  apply(c.U64, in64.s, ref.s, tmp64.s);
  spinor_copy_convert(c.g, out.s, out64.s, c.stream);
  Vec r = as_vec(ref.s), o = as_vec(out64.s);
  const double nr = c.blas->norm2(r);
  const double nd = c.blas->axpby(-1.0, r, 1.0, o, true);
  *relerr = std::sqrt(nd / nr);
  API_END
}

// The hop H_oe timed over nbuf (input, output) pairs in turn, so that with
// nbuf >= 2 at 32^3x64 the fields of one launch were last touched several
// hundred MB of traffic earlier and do not sit in the 256 MB last-level cache:
// the state a solver's Dslash meets.
extern "C" int qqcd_bench_dslash_rot(int prec, int recon, int nbuf, int nwarm, int niter, double *ms) {
  API_BEGIN
  Context &c = ctx();
  if (prec < 0 || prec > 2 || (recon != 18 && recon != 12)) throw std::runtime_error("bad precision / recon");
  if (niter < 1 || nbuf < 1 || nbuf > 16) throw std::runtime_error("niter >= 1 and 1 <= nbuf <= 16");
  const Prec p = (Prec)prec;
  const Gauge &U = c.gauge(p, recon);
  SpinorBuf in64(c.g, PREC_DOUBLE);
  random_spinor(c.g, in64.s, 12345, c.stream);
  std::vector<std::unique_ptr<SpinorBuf>> in, out;
  for (int b = 0; b < nbuf; b++) {
    in.emplace_back(new SpinorBuf(c.g, p));
    out.emplace_back(new SpinorBuf(c.g, p));
    spinor_copy_convert(c.g, in64.s, in.back()->s, c.stream);
  }
  for (int k = 0; k < nwarm; k++) dslash(c.g, U, in[k % nbuf]->s, out[k % nbuf]->s, 1, 0, Epilogue(), nullptr, c.stream);
  hipEvent_t e0, e1;
  QQCD_HIP(hipEventCreate(&e0));
  QQCD_HIP(hipEventCreate(&e1));
  QQCD_HIP(hipEventRecord(e0, c.stream));
  for (int k = 0; k < niter; k++) dslash(c.g, U, in[k % nbuf]->s, out[k % nbuf]->s, 1, 0, Epilogue(), nullptr, c.stream);
  QQCD_HIP(hipEventRecord(e1, c.stream));
  QQCD_HIP(hipEventSynchronize(e1));
  float t = 0;
  QQCD_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = (double)t / niter;
  API_END
}

static void canon_io(Context &c, const double *h, size_t n, double **d) {
  QQCD_HIP(hipMalloc(d, n * sizeof(double)));
  QQCD_HIP(hipMemcpyAsync(*d, h, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
}

extern "C" int qqcd_hop_canonical(const double *psi, double *out, int par_out, int idag, int prec, int recon) {
  API_BEGIN
  Context &c = ctx();
  const size_t n = (size_t)24 * c.g.V;
  double *dp = nullptr, *dout = nullptr;
  canon_io(c, psi, n, &dp);
  canon_io(c, out, n, &dout);
  try {
    const Prec p = (Prec)prec;
    const Gauge &U = c.gauge(p, recon);
    SpinorBuf a(c.g, p), b(c.g, p);
    canon_to_spinor(c.g, dp, a.s, 1 - par_out, c.stream);
    dslash(c.g, U, a.s, b.s, par_out, idag, Epilogue(), nullptr, c.stream);
    spinor_to_canon(c.g, b.s, dout, par_out, c.stream);
    QQCD_HIP(hipMemcpyAsync(out, dout, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    QQCD_HIP(hipStreamSynchronize(c.stream));
  } catch (...) {
    (void)hipFree(dp);
    (void)hipFree(dout);
    throw;
  }
  (void)hipFree(dp);
  (void)hipFree(dout);
  API_END
}

extern "C" int qqcd_schur_canonical(const double *psi, double *out, const double kappa[2], int idag, int prec,
                                    int recon) {
  API_BEGIN
  Context &c = ctx();
  const size_t n = (size_t)24 * c.g.V;
  double *dp = nullptr, *dout = nullptr;
  canon_io(c, psi, n, &dp);
  canon_io(c, out, n, &dout);
  try {
    const Prec p = (Prec)prec;
    const Gauge &U = c.gauge(p, recon);
    SpinorBuf a(c.g, p), b(c.g, p), t(c.g, p);
    canon_to_spinor(c.g, dp, a.s, 0, c.stream);
    schur_apply(c, U, kappa, idag, a.s, b.s, t.s);
    spinor_to_canon(c.g, b.s, dout, 0, c.stream);
    QQCD_HIP(hipMemcpyAsync(out, dout, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    QQCD_HIP(hipStreamSynchronize(c.stream));
  } catch (...) {
    (void)hipFree(dp);
    (void)hipFree(dout);
    throw;
  }
  (void)hipFree(dp);
  (void)hipFree(dout);
  API_END
}

namespace qqcd {
void gauge_random(const Geom &g, Gauge &U, int kind, double eps, uint64_t seed, hipStream_t s);
void gauge_to_canon(const Geom &g, const Gauge &U, double *dcn, hipStream_t s);
}  // namespace qqcd

extern "C" int qqcd_gauge_random(int kind, double eps, unsigned long long seed, const double *coact) {
  API_BEGIN
  Context &c = ctx();
  if (kind != 0 && kind != 1) throw std::runtime_error("gauge_random: kind must be 0 (Haar) or 1 (near identity)");
  set_coact(c, coact);
  if (!c.U64.data) c.U64 = alloc_gauge(c.g, PREC_DOUBLE, 18);
  gauge_random(c.g, c.U64, kind, eps, seed, c.stream);
  QQCD_HIP(hipStreamSynchronize(c.stream));
  c.u_host = nullptr;
  c.have_gauge = 1;
  c.gauge_gen++;
  API_END
}

extern "C" int qqcd_gauge_get_canonical(double *U) {
  API_BEGIN
  Context &c = ctx();
  if (!c.have_gauge) throw std::runtime_error("no gauge field loaded");
  const size_t n = (size_t)2 * 36 * c.g.V;
  double *d = nullptr;
  QQCD_HIP(hipMalloc(&d, n * sizeof(double)));
  gauge_to_canon(c.g, c.U64, d, c.stream);
  QQCD_HIP(hipMemcpyAsync(U, d, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  QQCD_HIP(hipStreamSynchronize(c.stream));
  (void)hipFree(d);
  API_END
}

namespace qqcd {
void gauge_to_ref(const Geom &g, const Gauge &U, double *dref, hipStream_t s);
}

extern "C" int qqcd_gauge_get_ref(double *u) {
  API_BEGIN
  Context &c = ctx();
  if (!c.have_gauge) throw std::runtime_error("no gauge field loaded");
  const size_t n = (size_t)18 * ref_ntotal(c.g) * 128;
  double *d = nullptr;
  QQCD_HIP(hipMalloc(&d, n * sizeof(double)));
  gauge_to_ref(c.g, c.U64, d, c.stream);
  QQCD_HIP(hipMemcpyAsync(u, d, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  QQCD_HIP(hipStreamSynchronize(c.stream));
  (void)hipFree(d);
  API_END
}

namespace qqcd {
void heatbath(Context &c, double beta, int nsweep, uint64_t seed, int hot_start);
double plaquette(Context &c);
}  // namespace qqcd

extern "C" int qqcd_gauge_heatbath(double beta, int nsweep, unsigned long long seed, int start,
                                   const double *coact) {
  API_BEGIN
  Context &c = ctx();
  if (beta <= 0 || nsweep < 0) throw std::runtime_error("heatbath: need beta > 0 and nsweep >= 0");
  if (start < -1 || start > 1) throw std::runtime_error("heatbath: start is -1 (continue), 0 (cold) or 1 (hot)");
  set_coact(c, coact);
  heatbath(c, beta, nsweep, seed, start);
  API_END
}

extern "C" int qqcd_gauge_plaquette(double *plaq) {
  API_BEGIN
  Context &c = ctx();
  if (!c.have_gauge) throw std::runtime_error("no gauge field loaded");
  *plaq = plaquette(c);
  API_END
}

namespace qqcd {
double bench_copy(size_t bytes, int iters, hipStream_t s);
}

extern "C" int qqcd_bench_blas(int prec, int k, int op, int grid, int niter, double *ms) {
  API_BEGIN
  Context &c = ctx();
  *ms = bench_blas(12 * c.g.stride, (Prec)prec, k, op, grid, niter, c.stream);
  API_END
}

namespace qqcd {
double bench_smear(Context &c, int nwarm, int niter);
}

extern "C" int qqcd_bench_smear(int nwarm, int niter, double *ms) {
  API_BEGIN
  Context &c = ctx();
  if (!c.have_gauge) throw std::runtime_error("no gauge field loaded");
  *ms = bench_smear(c, nwarm, niter);
  API_END
}

extern "C" int qqcd_bench_lagfib(long m, int niter, long chunk, double *ms) {
  API_BEGIN
  Context &c = ctx();
  *ms = bench_lagfib(m, niter, chunk, c.stream);
  API_END
}

extern "C" int qqcd_bench_copy(double bytes, int iters, double *ms) {
  API_BEGIN
  Context &c = ctx();
  if (bytes < 16 || iters < 1) throw std::runtime_error("bench_copy: bad size");
  *ms = bench_copy((size_t)bytes, iters, c.stream);
  API_END
}
