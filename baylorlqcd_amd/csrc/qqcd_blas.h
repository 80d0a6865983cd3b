// qqcd_blas.h -- level-1 operations on device vectors (see qqcd_blas.hip).
#pragma once
#include <complex>
#include "qqcd_internal.h"

namespace qqcd {

// A flat run of n = 12 * stride complex numbers (the 12 component runs of a
// checkerboard field) in FP64, FP32 or FP16.  FP16 vectors are the FP16
// spinors of the operator: element (k, site) is its half-precision pair times
// norm[site]; they are read by every operation and written only through the
// site-wise conversion (spinor_copy_convert), which forms the norms.
struct Vec {
  void *data = nullptr;
  Prec prec = PREC_DOUBLE;
  long n = 0;
  const float *norm = nullptr;  // FP16 only
  size_t bytes() const { return (size_t)n * (prec == PREC_DOUBLE ? 16 : (prec == PREC_SINGLE ? 8 : 4)); }
};

inline Vec as_vec(const Spinor &s) {
  Vec v;
  v.data = s.data;
  v.prec = s.prec;
  v.n = 12 * s.stride;
  v.norm = s.norm;
  return v;
}

// Column j + 1 of a paired step from the device Hessenberg (see k_pair_columns)
void pair_columns(double *d, int ld, int m, int j, const double *g2, const double *sc, double *b, hipStream_t s);

class Blas {
 public:
  explicit Blas(hipStream_t s);
  ~Blas();
  Blas(const Blas &) = delete;
  Blas &operator=(const Blas &) = delete;

  // out[i] = <v_i, w> = sum conj(v_i) w
  void multi_dot(int k, const Vec *v, const Vec &w, std::complex<double> *out);
  std::complex<double> dot(const Vec &a, const Vec &b);
  double norm2(const Vec &a);
  // w -= sum_i h_i v_i ; returns ||w||^2 when want_norm
  double multi_axpy(int k, const std::complex<double> *h, const Vec *v, Vec &w, bool want_norm);
  // Device-side variants (no host synchronisation): dout[2i], dout[2i+1] =
  // <v_i, w>; w -= sum_i h_i v_i with h_i = dcoef[2i] + i dcoef[2i+1] read
  // from device memory, ||w||^2 into *dnorm2 (if given); w *= 1/sqrt(*dnorm2).
  void multi_dot_dev(int k, const Vec *v, const Vec &w, double *dout);
  void multi_axpy_dev(int k, const double *dcoef, const Vec *v, Vec &w, double *dnorm2);
  void rscale_dev(Vec &w, const double *dnorm2);
  // Paired Arnoldi steps over an FP16 basis with per-site norms (FP32 w1,
  // w2, device-side): g1[2i..] = <v_i, w1>, g2[2i..] = <v_i, w2> in one pass;
  // w1 -= sum g1_i v_i, w2 -= sum g2_i v_i in one pass with sc[0..3] =
  // ||w1||^2, Re <w1, w2>, Im <w1, w2>, ||w2||^2; w2 = b[0] w2 + (b[1] +
  // i b[2]) w1 with ||w2||^2 into *dnorm2.
  void pair_dot_dev(int k, const Vec *v, const Vec &w1, const Vec &w2, double *g1, double *g2);
  void pair_axpy_dev(int k, const double *g1, const double *g2, const Vec *v, Vec &w1, Vec &w2, double *sc);
  void pair_comb_dev(const double *b, const Vec &w1, Vec &w2, double *dnorm2);
  // pair_comb_dev and the two conversions into an FP16 basis with per-site
  // norms in one pass over the Vh sites: v1 = w1 / sqrt(*n1), v2 = r / ||r||
  // with r = b0 w2 + (b1 + i b2) w1 (w2 is left as it is), *dnorm2 = ||r||^2
  void pair_comb_basis(long Vh, const double *b, const Spinor &w1, const Spinor &w2, const double *n1, Spinor &v1,
                       Spinor &v2, double *dnorm2);
  // y = a x + b y ; returns ||y||^2 when want_norm
  double axpby(std::complex<double> a, const Vec &x, std::complex<double> b, Vec &y, bool want_norm = false);
  void copy(const Vec &x, Vec &y);
  void zero(Vec &x);
  // Y_j (+)= sum_{i<m} V_i Z(i, j), j < k; Z column-major with leading dim ldz
  void combine(int m, const Vec *V, int k, const std::complex<double> *Z, int ldz, Vec *Y, bool accumulate);

  hipStream_t stream;

 private:
  void fetch(int nq);
  double *part = nullptr;
  double *res = nullptr;
  double *hres = nullptr;
  double *zbuf = nullptr;
  void **vptr = nullptr;
  static constexpr long zcap = 128 * 128;

 public:
  static constexpr int kmax = 136;  // vectors per multi_dot
};

// Kernel time (ms) of one multi-dot (op 0) or multi-axpy (op 1) pass over k
// vectors of n complex numbers with `grid` workgroups (tuning).
double bench_blas(long n, Prec p, int k, int op, int grid, int niter, hipStream_t s);

}  // namespace qqcd
