// qqcd_context.h -- the library's single per-process state.
//
// One context per process, one HIP stream, synchronous at every C entry
// point: the reference is single-threaded per rank and keeps its solver
// state in module globals (qqcd-test/cfgsprops/quark/shift.f90:20-45,
// gmresrhs.f90:10), and the library mirrors that contract.
#pragma once
#include <map>
#include <memory>
#include <vector>
#include "qqcd_blas.h"

namespace qqcd {

struct Context {
  Geom g{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::unique_ptr<Blas> blas;

  // master gauge field: FP64, 18 reals, resident
  Gauge U64;
  uint64_t gauge_gen = 0;  // bumped on every load
  const void *u_host = nullptr;
  double coact[64] = {0};
  int have_gauge = 0;
  int hb_sweep = 0;  // sweeps of the current heatbath chain (keys its random stream)
  // lower-precision / compressed copies, keyed by prec*100+recon
  struct Copy {
    Gauge U;
    uint64_t gen = ~0ull;
  };
  std::map<int, Copy> copies;

  // staging for reference-layout spinors (6*ntotal*64 doubles each)
  double *dref[3] = {nullptr, nullptr, nullptr};

  ~Context();
  const Gauge &gauge(Prec p, int recon);
};

extern Context *g_ctx;
Context &ctx();

Spinor alloc_spinor(const Geom &g, Prec p);
void free_spinor(Spinor &s);
Gauge alloc_gauge(const Geom &g, Prec p, int recon);
void free_gauge(Gauge &u);

// scratch spinors that free themselves
struct SpinorBuf {
  Spinor s;
  SpinorBuf(const Geom &g, Prec p) : s(alloc_spinor(g, p)) {}
  ~SpinorBuf() { free_spinor(s); }
  SpinorBuf(const SpinorBuf &) = delete;
  SpinorBuf &operator=(const SpinorBuf &) = delete;
  SpinorBuf(SpinorBuf &&o) noexcept : s(o.s) { o.s.data = nullptr; o.s.norm = nullptr; }
};

// geometry updates from the reference arguments
void set_coact(Context &c, const double *coact);
void set_bc(Context &c, const int *bc);
void ensure_gauge(Context &c, const double *u_ref, const double *coact);

// Schur operators on device spinors of checkerboard 0:
//   Wilson   (mu < 1e-8): out = in/k^2 - H_eo H_oe in
//   twisted  (a=2 k mu):  out = (1 + a i g5) in/k^2 - H_eo (1 - a i g5) H_oe in/(1+a^2)
// (Hdouble / Hdbletm, qqcd-test/cfgsprops/quark/diracops.f90:264-346, 435-550)
// out32 (FP16 spinors): the result also as FP32, from the same launch where
// the pipelined kernel runs (dslash's out32 / only32).
void schur_apply(Context &c, const Gauge &U, const double kappa[2], int idag, const Spinor &in, Spinor &out,
                 Spinor &tmp_odd, Spinor *out32 = nullptr, bool only32 = false);

void canon_to_spinor(const Geom &g, const double *dcn, Spinor &out, int par, hipStream_t s);
void spinor_to_canon(const Geom &g, const Spinor &in, double *dcn, int par, hipStream_t s);
void gamma_ref(const Geom &g, const double *in, double *out, int mu, int nslices, const int *slices,
               hipStream_t s);
void random_spinor(const Geom &g, Spinor &s, uint64_t seed, hipStream_t st);
// every component of the lattice sites set to re + i im, any precision (padding left as it is)
void fill_spinor(const Geom &g, Spinor &s, double re, double im, hipStream_t st);
// gamma_mu (mu = 1..5) of the reference basis as a dense 4x4 complex matrix
void gamma_dense(int mu, double re[4][4], double im[4][4]);
void set_xcd_remap(int v);


}  // namespace qqcd
