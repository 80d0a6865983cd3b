// qqcd_layout.hip -- moving fields between the reference's storage order,
// the canonical lexicographic order and the device even-odd SoA order, and
// between storage precisions.
//
// Reference order (single process): a spinor on global checkerboard gblclr
// is g(a, i, d, ieo, ibleo) of shape (6, ntotal, 4, 2, 8), a link field
// u(a, i, mu, ieo, ibl) of shape (18, ntotal, 4, 2, 16), Fortran order.  For a
// site with coordinates x_mu: block bit b = x mod 2, s = x/2,
// ibl = 1 + bx + 2by + 4bz + 8bt, ieo = 1 + (sum s) mod 2,
// i = 1 + floor((sx + nsx(sy + nsy(sz + nsz st)))/2), ibleo = ceil(ibl/2)
// (qqcd-test/cfgsprops/glue/lattice.f90:140-354; closed form in
// qqcd-test/cfgsprops/quark/quark.f90:82-105).  Colour c of Dirac d is
// (Re, Im) = (g(2c-1), g(2c)); U_rc = (u(6r+2c-1), u(6r+2c)) with 1-based r, c
// (qqcd-test/cfgsprops/cfgsprops.f90:113-117).
#include <stdexcept>
#include "qqcd_internal.h"

namespace qqcd {

typedef f16 h2 __attribute__((ext_vector_type(2)));

long ref_nvhalf(const Geom &g) { return g.V / 32; }
long ref_ntotal(const Geom &g) {
  long nvh = ref_nvhalf(g), nb = 0;
  for (int m = 0; m < 4; m++) {
    long b = 2 * nvh / g.n[m];
    if (b > nb) nb = b;
  }
  return nvh + nb;
}

struct Coord {
  int x[4];
};

__device__ __forceinline__ Coord coords_of(const int n[4], int par, int i) {
  Coord c;
  const int r = 2 * i;
  const int x0 = r % n[0];
  int q = r / n[0];
  c.x[1] = q % n[1];
  q /= n[1];
  c.x[2] = q % n[2];
  c.x[3] = q / n[2];
  c.x[0] = x0 + ((par + c.x[1] + c.x[2] + c.x[3]) & 1);
  return c;
}

// 0-based (i, ieo, ibl, ibleo) of the reference slot of a site
__device__ __forceinline__ void ref_slot(const int n[4], const Coord &c, long &i, int &ieo, int &ibl, int &ibleo) {
  int b[4], s[4];
  for (int m = 0; m < 4; m++) {
    b[m] = c.x[m] & 1;
    s[m] = c.x[m] >> 1;
  }
  const long L = s[0] + (long)(n[0] / 2) * (s[1] + (long)(n[1] / 2) * (s[2] + (long)(n[2] / 2) * s[3]));
  i = L / 2;
  ieo = (s[0] + s[1] + s[2] + s[3]) & 1;
  ibl = b[0] + 2 * b[1] + 4 * b[2] + 8 * b[3];
  ibleo = ibl >> 1;
}

template <typename S> struct Put;
template <> struct Put<double> {
  static __device__ void put(void *b, float *, long stride, int i, const double *v /*24*/) {
    double2 *p = (double2 *)b;
    for (int k = 0; k < 12; k++) p[(long)k * stride + i] = make_double2(v[2 * k], v[2 * k + 1]);
  }
  static __device__ void get(const void *b, const float *, long stride, int i, double *v) {
    const double2 *p = (const double2 *)b;
    for (int k = 0; k < 12; k++) {
      double2 w = p[(long)k * stride + i];
      v[2 * k] = w.x;
      v[2 * k + 1] = w.y;
    }
  }
};
template <> struct Put<float> {
  static __device__ void put(void *b, float *, long stride, int i, const double *v) {
    float2 *p = (float2 *)b;
    for (int k = 0; k < 12; k++) p[(long)k * stride + i] = make_float2((float)v[2 * k], (float)v[2 * k + 1]);
  }
  static __device__ void get(const void *b, const float *, long stride, int i, double *v) {
    const float2 *p = (const float2 *)b;
    for (int k = 0; k < 12; k++) {
      float2 w = p[(long)k * stride + i];
      v[2 * k] = w.x;
      v[2 * k + 1] = w.y;
    }
  }
};
template <> struct Put<f16> {
  static __device__ void put(void *b, float *nrm, long stride, int i, const double *v) {
    double m = 0;
    for (int k = 0; k < 24; k++) m = fmax(m, fabs(v[k]));
    nrm[i] = (float)m;
    const double inv = m > 0 ? 1.0 / m : 0.0;
    h2 *p = (h2 *)b;
    for (int k = 0; k < 12; k++) {
      h2 w;
      w.x = (f16)(float)(v[2 * k] * inv);
      w.y = (f16)(float)(v[2 * k + 1] * inv);
      p[(long)k * stride + i] = w;
    }
  }
  static __device__ void get(const void *b, const float *nrm, long stride, int i, double *v) {
    const h2 *p = (const h2 *)b;
    const double m = nrm[i];
    for (int k = 0; k < 12; k++) {
      h2 w = p[(long)k * stride + i];
      v[2 * k] = m * (double)(float)w.x;
      v[2 * k + 1] = m * (double)(float)w.y;
    }
  }
};

struct LayoutArgs {
  int n[4];
  int Vh;
  int par;
  long stride;
  long ntotal;
};

template <typename S>
__global__ void k_ref_to_spinor(const LayoutArgs a, const double *ref, void *out, float *nrm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.Vh) return;
  const Coord c = coords_of(a.n, a.par, i);
  long ri;
  int ieo, ibl, ibleo;
  ref_slot(a.n, c, ri, ieo, ibl, ibleo);
  double v[24];
  for (int d = 0; d < 4; d++) {
    const double *p = ref + 6 * (ri + a.ntotal * (d + 4 * (ieo + 2 * (long)ibleo)));
    for (int q = 0; q < 6; q++) v[6 * d + q] = p[q];
  }
  Put<S>::put(out, nrm, a.stride, i, v);
}

template <typename S>
__global__ void k_spinor_to_ref(const LayoutArgs a, const void *in, const float *nrm, double *ref) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.Vh) return;
  const Coord c = coords_of(a.n, a.par, i);
  long ri;
  int ieo, ibl, ibleo;
  ref_slot(a.n, c, ri, ieo, ibl, ibleo);
  double v[24];
  Put<S>::get(in, nrm, a.stride, i, v);
  for (int d = 0; d < 4; d++) {
    double *p = ref + 6 * (ri + a.ntotal * (d + 4 * (ieo + 2 * (long)ibleo)));
    for (int q = 0; q < 6; q++) p[q] = v[6 * d + q];
  }
}

// canonical spinor psi[(s*4 + d)*3 + c] complex, s = x + nx(y + ny(z + nz t))
template <typename S>
__global__ void k_canon_to_spinor(const LayoutArgs a, const double *cn, void *out, float *nrm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.Vh) return;
  const Coord c = coords_of(a.n, a.par, i);
  const long s = c.x[0] + (long)a.n[0] * (c.x[1] + (long)a.n[1] * (c.x[2] + (long)a.n[2] * c.x[3]));
  double v[24];
  for (int q = 0; q < 24; q++) v[q] = cn[s * 24 + q];
  Put<S>::put(out, nrm, a.stride, i, v);
}

template <typename S>
__global__ void k_spinor_to_canon(const LayoutArgs a, const void *in, const float *nrm, double *cn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.Vh) return;
  const Coord c = coords_of(a.n, a.par, i);
  const long s = c.x[0] + (long)a.n[0] * (c.x[1] + (long)a.n[1] * (c.x[2] + (long)a.n[2] * c.x[3]));
  double v[24];
  Put<S>::get(in, nrm, a.stride, i, v);
  for (int q = 0; q < 24; q++) cn[s * 24 + q] = v[q];
}

template <typename L> __device__ void put_c(void *b, long off, double re, double im);
template <> __device__ void put_c<double>(void *b, long off, double re, double im) {
  ((double2 *)b)[off] = make_double2(re, im);
}
template <> __device__ void put_c<float>(void *b, long off, double re, double im) {
  ((float2 *)b)[off] = make_float2((float)re, (float)im);
}
template <> __device__ void put_c<f16>(void *b, long off, double re, double im) {
  h2 w;
  w.x = (f16)(float)re;
  w.y = (f16)(float)im;
  ((h2 *)b)[off] = w;
}

// links of (parity par, mu) from the reference (src_ref=1) or canonical
// U[(s*4 + mu)*9 + r*3 + c] (src_ref=0) order.
template <typename L>
__global__ void k_gauge(const LayoutArgs a, int src_ref, int nc, const double *src, void *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (i >= a.Vh) return;
  const Coord c = coords_of(a.n, a.par, i);
  const long base = (long)((a.par * 4 + mu) * nc) * a.stride + i;
  if (src_ref) {
    long ri;
    int ieo, ibl, ibleo;
    ref_slot(a.n, c, ri, ieo, ibl, ibleo);
    const double *p = src + 18 * (ri + a.ntotal * (mu + 4 * (ieo + 2 * (long)ibl)));
    for (int k = 0; k < nc; k++) put_c<L>(out, base + (long)k * a.stride, p[2 * k], p[2 * k + 1]);
  } else {
    const long s = c.x[0] + (long)a.n[0] * (c.x[1] + (long)a.n[1] * (c.x[2] + (long)a.n[2] * c.x[3]));
    const double *p = src + ((s * 4 + mu) * 9) * 2;
    for (int k = 0; k < nc; k++) put_c<L>(out, base + (long)k * a.stride, p[2 * k], p[2 * k + 1]);
  }
}

template <typename A, typename B>
__global__ void k_convert(int Vh, long stride, const void *in, const float *ninn, void *out, float *nout,
                          double scale, const double *dn2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Vh) return;
  double v[24];
  Put<A>::get(in, ninn, stride, i, v);
  if (dn2) {
    // scale by 1 / sqrt(*dn2), a norm left in device memory by a reduction
    const double q = *dn2;
    scale *= q > 0 ? 1.0 / sqrt(q) : 0.0;
  }
  if (scale != 1.0)
    for (int k = 0; k < 24; k++) v[k] *= scale;
  Put<B>::put(out, nout, stride, i, v);
}

static LayoutArgs largs(const Geom &g, int par) {
  LayoutArgs a{};
  for (int m = 0; m < 4; m++) a.n[m] = g.n[m];
  a.Vh = (int)g.Vh;
  a.par = par;
  a.stride = g.stride;
  a.ntotal = ref_ntotal(g);
  return a;
}

#define DISPATCH_S(prec, KER, ...)                                                        \
  do {                                                                                    \
    if ((prec) == PREC_DOUBLE)                                                            \
      hipLaunchKernelGGL(KER<double>, dim3(nb), dim3(256), 0, s, __VA_ARGS__);            \
    else if ((prec) == PREC_SINGLE)                                                       \
      hipLaunchKernelGGL(KER<float>, dim3(nb), dim3(256), 0, s, __VA_ARGS__);             \
    else                                                                                  \
      hipLaunchKernelGGL(KER<f16>, dim3(nb), dim3(256), 0, s, __VA_ARGS__);               \
    QQCD_HIP(hipGetLastError());                                                          \
  } while (0)

void ref_to_spinor(const Geom &g, const double *dref, Spinor &out, int par, hipStream_t s) {
  const int nb = (int)((g.Vh + 255) / 256);
  DISPATCH_S(out.prec, k_ref_to_spinor, largs(g, par), dref, out.data, out.norm);
}

void spinor_to_ref(const Geom &g, const Spinor &in, double *dref, int par, hipStream_t s) {
  const int nb = (int)((g.Vh + 255) / 256);
  QQCD_HIP(hipMemsetAsync(dref, 0, sizeof(double) * 6 * ref_ntotal(g) * 64, s));
  DISPATCH_S(in.prec, k_spinor_to_ref, largs(g, par), in.data, in.norm, dref);
}

void canon_to_spinor(const Geom &g, const double *dcn, Spinor &out, int par, hipStream_t s) {
  const int nb = (int)((g.Vh + 255) / 256);
  DISPATCH_S(out.prec, k_canon_to_spinor, largs(g, par), dcn, out.data, out.norm);
}

void spinor_to_canon(const Geom &g, const Spinor &in, double *dcn, int par, hipStream_t s) {
  const int nb = (int)((g.Vh + 255) / 256);
  DISPATCH_S(in.prec, k_spinor_to_canon, largs(g, par), in.data, in.norm, dcn);
}

// the FP64/18 device field back to the reference layout u(18, ntotal, 4, 2, 16)
__global__ void k_gauge_to_ref(const LayoutArgs a, const double2 *U, double *dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (i >= a.Vh) return;
  const Coord c = coords_of(a.n, a.par, i);
  long ri;
  int ieo, ibl, ibleo;
  ref_slot(a.n, c, ri, ieo, ibl, ibleo);
  const long base = (long)((a.par * 4 + mu) * 9) * a.stride + i;
  double *p = dst + 18 * (ri + a.ntotal * (mu + 4 * (ieo + 2 * (long)ibl)));
  for (int k = 0; k < 9; k++) {
    const double2 v = U[base + (long)k * a.stride];
    p[2 * k] = v.x;
    p[2 * k + 1] = v.y;
  }
}

void gauge_to_ref(const Geom &g, const Gauge &U, double *dref, hipStream_t s) {
  if (U.prec != PREC_DOUBLE || U.recon != 18) throw std::runtime_error("gauge_to_ref: master must be FP64/18");
  QQCD_HIP(hipMemsetAsync(dref, 0, sizeof(double) * 18 * ref_ntotal(g) * 128, s));
  const int nb = (int)((g.Vh + 255) / 256);
  for (int par = 0; par < 2; par++) {
    LayoutArgs a = largs(g, par);
    a.stride = U.stride;
    hipLaunchKernelGGL(k_gauge_to_ref, dim3(nb, 4), dim3(256), 0, s, a, (const double2 *)U.data, dref);
    QQCD_HIP(hipGetLastError());
  }
}

static void gauge_any(const Geom &g, int src_ref, const double *src, Gauge &U, hipStream_t s) {
  const int nc = U.recon == 18 ? 9 : 6;
  const int nb = (int)((g.Vh + 255) / 256);
  for (int par = 0; par < 2; par++) {
    LayoutArgs a = largs(g, par);
    a.stride = U.stride;
    if (U.prec == PREC_DOUBLE)
      hipLaunchKernelGGL(k_gauge<double>, dim3(nb, 4), dim3(256), 0, s, a, src_ref, nc, src, U.data);
    else if (U.prec == PREC_SINGLE)
      hipLaunchKernelGGL(k_gauge<float>, dim3(nb, 4), dim3(256), 0, s, a, src_ref, nc, src, U.data);
    else
      hipLaunchKernelGGL(k_gauge<f16>, dim3(nb, 4), dim3(256), 0, s, a, src_ref, nc, src, U.data);
    QQCD_HIP(hipGetLastError());
  }
}

void ref_to_gauge(const Geom &g, const double *dref, Gauge &U, hipStream_t s) { gauge_any(g, 1, dref, U, s); }
void canon_to_gauge(const Geom &g, const double *dcn, Gauge &U, hipStream_t s) { gauge_any(g, 0, dcn, U, s); }

void spinor_copy_convert(const Geom &g, const Spinor &in, Spinor &out, hipStream_t s, double scale) {
  spinor_convert_rnorm(g, in, out, nullptr, s, scale);
}

void spinor_convert_rnorm(const Geom &g, const Spinor &in, Spinor &out, const double *dn2, hipStream_t s,
                          double scale) {
  const int nb = (int)((g.Vh + 255) / 256);
  const int Vh = (int)g.Vh;
#define CV(A, B)                                                                                                  \
  hipLaunchKernelGGL((k_convert<A, B>), dim3(nb), dim3(256), 0, s, Vh, g.stride, in.data, in.norm, out.data, out.norm, \
                     scale, dn2)
  if (in.prec == PREC_DOUBLE) {
    if (out.prec == PREC_DOUBLE) CV(double, double);
    else if (out.prec == PREC_SINGLE) CV(double, float);
    else CV(double, f16);
  } else if (in.prec == PREC_SINGLE) {
    if (out.prec == PREC_DOUBLE) CV(float, double);
    else if (out.prec == PREC_SINGLE) CV(float, float);
    else CV(float, f16);
  } else {
    if (out.prec == PREC_DOUBLE) CV(f16, double);
    else if (out.prec == PREC_SINGLE) CV(f16, float);
    else CV(f16, f16);
  }
#undef CV
  QQCD_HIP(hipGetLastError());
}

}  // namespace qqcd
