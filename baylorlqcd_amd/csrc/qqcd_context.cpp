// qqcd_context.cpp -- context lifetime, field allocation, gauge residency
// and the Schur operators built from two Dslash launches.
#include "qqcd_context.h"

#include <cmath>
#include <cstring>
#include <stdexcept>

namespace qqcd {

Context *g_ctx = nullptr;

Context &ctx() {
  if (!g_ctx) throw std::runtime_error("library not initialised: call qqcd_init first");
  return *g_ctx;
}

static size_t spinor_bytes(Prec p, long stride) {
  return (size_t)12 * stride * (p == PREC_DOUBLE ? 16 : p == PREC_SINGLE ? 8 : 4);
}

size_t Spinor::bytes() const { return spinor_bytes(prec, stride); }

size_t Gauge::bytes() const {
  return (size_t)8 * (recon == 18 ? 9 : 6) * stride * (prec == PREC_DOUBLE ? 16 : prec == PREC_SINGLE ? 8 : 4);
}

Spinor alloc_spinor(const Geom &g, Prec p) {
  Spinor s;
  s.prec = p;
  s.stride = g.stride;
  QQCD_HIP(hipMalloc(&s.data, s.bytes()));
  QQCD_HIP(hipMemset(s.data, 0, s.bytes()));
  if (p == PREC_HALF) {
    QQCD_HIP(hipMalloc(&s.norm, sizeof(float) * g.stride));
    QQCD_HIP(hipMemset(s.norm, 0, sizeof(float) * g.stride));
  }
  return s;
}

void free_spinor(Spinor &s) {
  if (s.data) (void)hipFree(s.data);
  if (s.norm) (void)hipFree(s.norm);
  s.data = nullptr;
  s.norm = nullptr;
}

Gauge alloc_gauge(const Geom &g, Prec p, int recon) {
  Gauge u;
  u.prec = p;
  u.recon = recon;
  u.stride = g.stride;
  QQCD_HIP(hipMalloc(&u.data, u.bytes()));
  QQCD_HIP(hipMemset(u.data, 0, u.bytes()));
  return u;
}

void free_gauge(Gauge &u) {
  if (u.data) (void)hipFree(u.data);
  u.data = nullptr;
}

Context::~Context() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto &kv : copies) free_gauge(kv.second.U);
  free_gauge(U64);
  for (double *&p : dref)
    if (p) (void)hipFree(p);
  blas.reset();
  if (stream) (void)hipStreamDestroy(stream);
}

// Lower-precision or 12-real copies of the master field, rebuilt lazily.
// The 12-real form drops row 2 and the kernels rebuild it as
// conj(row0 x row1): exact only for SU(3) links.
const Gauge &Context::gauge(Prec p, int recon) {
  if (!have_gauge) throw std::runtime_error("no gauge field loaded");
  if (p == PREC_DOUBLE && recon == 18) return U64;
  Copy &cp = copies[(int)p * 100 + recon];
  if (cp.gen != gauge_gen) {
    if (!cp.U.data) cp.U = alloc_gauge(g, p, recon);
    // convert through the canonical FP64 staging on the device
    double *tmp = nullptr;
    QQCD_HIP(hipMalloc(&tmp, sizeof(double) * 2 * 36 * g.V));
    // unpack the master into canonical order, then repack
    extern void gauge_to_canon(const Geom &, const Gauge &, double *, hipStream_t);
    gauge_to_canon(g, U64, tmp, stream);
    canon_to_gauge(g, tmp, cp.U, stream);
    QQCD_HIP(hipStreamSynchronize(stream));
    (void)hipFree(tmp);
    cp.gen = gauge_gen;
  }
  return cp.U;
}

void set_coact(Context &c, const double *coact) {
  if (!coact) return;
  for (int mu = 0; mu < 4; mu++) {
    c.g.c1[mu] = coact[2 + 4 * (0 + 4 * mu)];
    c.g.c2[mu] = coact[2 + 4 * (1 + 4 * mu)];
  }
  std::memcpy(c.coact, coact, sizeof(c.coact));
}

void set_bc(Context &c, const int *bc) {
  if (!bc) return;
  for (int mu = 0; mu < 4; mu++) {
    if (bc[mu] != 1 && bc[mu] != -1 && bc[mu] != 0) throw std::runtime_error("bc must be 1, -1 or 0");
    c.g.bc[mu] = bc[mu];
  }
}

// Upload the reference-order link array when it is not the resident one.
// The resident field is identified by the host address it came from and the
// coact it was loaded with; qqcd_gauge_load forces a reload after the host
// changed the links in place.
void ensure_gauge(Context &c, const double *u_ref, const double *coact) {
  set_coact(c, coact);
  if (!u_ref) {
    if (!c.have_gauge) throw std::runtime_error("no gauge field loaded");
    return;
  }
  if (c.have_gauge && c.u_host == (const void *)u_ref) return;
  const size_t n = (size_t)18 * ref_ntotal(c.g) * 128;
  double *d = nullptr;
  QQCD_HIP(hipMalloc(&d, n * sizeof(double)));
  QQCD_HIP(hipMemcpyAsync(d, u_ref, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  if (!c.U64.data) c.U64 = alloc_gauge(c.g, PREC_DOUBLE, 18);
  ref_to_gauge(c.g, d, c.U64, c.stream);
  QQCD_HIP(hipStreamSynchronize(c.stream));
  (void)hipFree(d);
  c.u_host = u_ref;
  c.have_gauge = 1;
  c.gauge_gen++;
}

void schur_apply(Context &c, const Gauge &U, const double kappa[2], int idag, const Spinor &in, Spinor &out,
                 Spinor &tmp_odd, Spinor *out32, bool only32) {
  const double k = kappa[0], mu = kappa[1];
  Epilogue e1, e2;
  if (mu < 1e-8) {
    // the reference's signed test: mu <= 0 is plain Wilson (diracops.f90:467)
    e2.a0 = 1.0 / (k * k);
    e2.b0 = -1.0;
  } else {
    const double a = 2.0 * k * mu;
    e1.b1 = -a;  // (1 - a i g5) H_oe in
    e2.a0 = 1.0 / (k * k);
    e2.a1 = 2.0 * mu / k;
    e2.b0 = -1.0 / (1.0 + a * a);
  }
  dslash(c.g, U, in, tmp_odd, 1, idag, e1, nullptr, c.stream);
  dslash(c.g, U, tmp_odd, out, 0, idag, e2, &in, c.stream, out32, only32);
}

}  // namespace qqcd
