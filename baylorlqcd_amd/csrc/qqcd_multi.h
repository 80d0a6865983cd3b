// qqcd_multi.h -- multi-right-hand-side fields and their level-1 kernels.
//
// A MultiSpinor (qqcd_internal.h) holds `cap` right-hand-side slots; the
// active slots are [0, nact) with nact kept in device memory, so kernels
// launched (or captured in a graph) once keep working while converged
// right-hand sides are retired by swapping them behind the active ones.
// Per-slot coefficients and results live in device arrays indexed
// [q * cap + slot] (complex as double2), so a whole Arnoldi cycle runs
// without a host round trip.
#pragma once
#include <complex>
#include <vector>

#include "qqcd_context.h"

namespace qqcd {

// norms = false for PREC_HALF: a Krylov basis field, unit vectors stored
// times sqrt(12 stride) without per-site norms (read by dot and axpy only)
MultiSpinor alloc_multi(const Geom &g, Prec p, int cap, int nb, bool norms = true);
void free_multi(MultiSpinor &m);

struct MultiBuf {
  MultiSpinor m;
  MultiBuf(const Geom &g, Prec p, int cap, int nb, bool norms = true) : m(alloc_multi(g, p, cap, nb, norms)) {}
  ~MultiBuf() { free_multi(m); }
  MultiBuf(const MultiBuf &) = delete;
  MultiBuf &operator=(const MultiBuf &) = delete;
};

// device scratch of the multi kernels
class MBlas {
 public:
  MBlas(hipStream_t s, int cap);
  ~MBlas();
  MBlas(const MBlas &) = delete;
  MBlas &operator=(const MBlas &) = delete;

  // res[q*cap + r] = <V_q(r), W(r)> for q < nq (V_q multi, FP64, FP32 or an
  // FP16 basis) -- device result; with `any` given, the pass is skipped when
  // *any == 0; self: also res[nq*cap + r] = <W(r), W(r)> from the same pass
  void dot(int nq, const MultiSpinor *V, const MultiSpinor &W, const int *nact, double *res,
           const int *any = nullptr, bool self = false);
  // res[q*cap + r] = <D_q, W(r)> for single-RHS vectors D_q shared by all slots
  void dot_shared(int nq, const Spinor *D, const MultiSpinor &W, const int *nact, double *res);
  // W(r) += sign * sum_q coef[(q0+q)*cap + r] V_q(r); if n2 != null, n2[r] = ||W(r)||^2.
  // With `flag` given, slots with flag[r] == 0 are left alone (and *any == 0 skips the pass).
  void axpy(int nq, const MultiSpinor *V, const double *coef, int q0, double sign, MultiSpinor &W,
            const int *nact, double *n2, const int *flag = nullptr, const int *any = nullptr);
  // W(r) += sign * sum_q coef[(q0+q)*cap + r] D_q with shared single-RHS D_q
  void axpy_shared(int nq, const Spinor *D, const double *coef, int q0, double sign, MultiSpinor &W,
                   const int *nact, double *n2);
  // n2[r] = ||W(r)||^2
  void norm2(const MultiSpinor &W, const int *nact, double *n2);
  // Y(r) = a X(r) + b Y(r) for every active slot (any precisions, same layout), n2 optional
  void axpby(double a, const MultiSpinor &X, double b, MultiSpinor &Y, const int *nact, double *n2);
  // W(r) = X(r) * s[r] where s[r] = 1/sqrt(n2[r]) (0 when n2[r] == 0); X may be W
  void scale_inv_norm(const MultiSpinor &X, MultiSpinor &W, const double *n2, const int *nact);
  // the same into Q and into the FP16 basis field H (X may be Q)
  void scale_inv_norm2(const MultiSpinor &X, MultiSpinor &Q, MultiSpinor &H, const double *n2, const int *nact);
  // Two-vector passes of the two-step Arnoldi block (FP32 W1, W2; FP16
  // basis fields V): g1[q*cap + r] = <V_q(r), W1(r)>, g2[...] = <V_q(r),
  // W2(r)> in one pass over V (per-lane sums in FP32, across lanes in FP64);
  // self (optional): self[c*cap + r] = ||W1||^2, Re <W1, W2>, Im <W1, W2>,
  // ||W2||^2 for c = 0..3 from the same pass.
  void dot2(int nq, const MultiSpinor *V, const MultiSpinor &W1, const MultiSpinor &W2, const int *nact, double *g1,
            double *g2, double *self = nullptr);
  // W1(r) -= sum_q g1[q*cap + r] V_q(r), W2(r) -= sum_q g2[...] V_q(r) in one
  // pass over V; then sc[c*cap + r] = ||W1||^2, Re <W1, W2>, Im <W1, W2>,
  // ||W2||^2 for c = 0..3
  void axpy2(int nq, const MultiSpinor *V, const double *g1, const double *g2, MultiSpinor &W1, MultiSpinor &W2,
             const int *nact, double *sc);
  // R(r) = b[4r] W2(r) + (b[4r+1] + i b[4r+2]) W1(r) (FP32) and H(r) =
  // b[4r+3] W1(r) as an FP16 basis field; nr[r] = ||R(r)||^2
  void next2(const MultiSpinor &W1, const MultiSpinor &W2, MultiSpinor &R, MultiSpinor &H, const double *b,
             const int *nact, double *nr);
  // axpy2 and next2 in one pass: the updated W1, W2 stay in registers (W1, W2
  // keep the partial updates of earlier chunks when nq exceeds one pass), R
  // and H are formed from them with the coefficients b, nr[r] = ||R(r)||^2.
  void axpy2_next(int nq, const MultiSpinor *V, const double *g1, const double *g2, MultiSpinor &W1, MultiSpinor &W2,
                  const double *b, MultiSpinor &R, MultiSpinor &H, const int *nact, double *nr);
  // swap slots a and b of W
  void swap(MultiSpinor &W, int a, int b);
  // single <-> slot, with precision conversion
  void put(const Spinor &s, MultiSpinor &W, int slot);
  void get(const MultiSpinor &W, int slot, Spinor &s);
  void zero(MultiSpinor &W);

  hipStream_t stream;
  int cap;

 private:
  double *part = nullptr;
  size_t part_cap = 0;

 public:
  // size the partial-sum scratch before a graph capture (no allocation inside one)
  void reserve(size_t n);
  size_t need() const;
};

// Element-wise per-slot Arnoldi bookkeeping (qqcd_mblas.hip):
// flag[r] = n2[r] < 0.5 * w2[r] (a second Gram-Schmidt pass is needed; every
// active slot when force), any = OR(flag); w2 is complex ([2r] read)
void arnoldi_flags(const double *n2, const double *w2, int *flag, int *any, const int *nact, int cap, int force,
                   hipStream_t s);
// flag[r] = 0 for every slot, any = 0
void clear_flags(int *flag, int *any, int cap, hipStream_t s);
// Hessenberg column j (complex, [(j*ldh + q)*cap + r]): h1 + (flag ? h2 : 0) for
// q <= j, and sqrt(nf) below the diagonal with nf[r] = flag ? n2b[r] : n2a[r]
// (the squared norm after the last Gram-Schmidt pass)
// Columns j and j+1 of the batched Hessenberg matrix for a two-step block
// from g1 = V_j^H A v_j, g2 = V_j^H A^2 v_j and the four scalars of axpy2;
// b receives the per-slot coefficients next2 needs.
void sstep_columns(double *H, int ldh, int j, const double *g1, const double *g2, const double *sc, double *b,
                   const int *nact, int cap, double hs, hipStream_t s);
// The four scalars of axpy2 before the pass, from those of W1, W2 (dot2's
// self) and the projections g1, g2 over nq orthonormal basis fields.
void pythag_scalars(int nq, const double *g1, const double *g2, const double *self, double *sc, const int *nact,
                    int cap, hipStream_t s);
// H[q, j] = sqrt(n2[r]) for one row q of column j (the subdiagonal entry after next2)
void set_subdiag(double *H, int ldh, int j, const double *n2, const int *nact, int cap, hipStream_t s);
void arnoldi_column(double *H, int ldh, int j, const double *h1, const double *h2, const int *flag,
                    const double *n2a, const double *n2b, double *nf, const int *nact, int cap, hipStream_t s);

}  // namespace qqcd
