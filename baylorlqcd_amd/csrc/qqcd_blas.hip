// qqcd_blas.hip -- level-1 kernels of the Krylov solvers.
//
// Vectors are flat runs of n complex numbers (a Spinor's 12 component runs
// back to back; padding sites are zero and stay zero).  Element type double
// or float; every reduction accumulates in FP64 and is deterministic: each
// workgroup reduces its lanes with cross-lane shuffles and LDS in a fixed
// order, writes one partial, and a single-workgroup pass sums the partials
// in a fixed order.
//
// Replaces the reference's per-site loops in vecdot (qqcd-test/cfgsprops/
// quark/inverters.f90:13485-13530), the MGS dot/axpy pairs of gmresdr
// (:3120-3152), the per-step solution update (:3242-3255) and the restart
// V <- V Z (:3461-3485) with fused, bandwidth-bound passes: one multi-dot
// pass reads w and up to 16 basis vectors, one multi-axpy pass updates w and
// returns its norm; every lane moves 16 bytes per vector and element pair.
#include <stdexcept>
#include <type_traits>
#include <vector>
#include "qqcd_blas.h"

namespace qqcd {

static constexpr int RB = 256;   // threads per reduction workgroup
static constexpr int RG = 1024;  // reduction workgroups (grid-stride)

// Element access in units of U consecutive elements (U = 4 for FP16
// vectors, else 2), so that every lane moves at least 16 bytes per vector
// and pass: two double2, one float4, or four half-precision complex numbers
// (16 bytes) with the float4 of their four sites' norms.  Element e belongs to
// site e % stride; the stride is a multiple of 64, so a unit never leaves a
// component run.
template <typename T> struct C2;
template <> struct C2<double> {
  template <int U>
  static __device__ __forceinline__ void ld(const void *p, const float *, unsigned, long i, double2 (&e)[U]) {
#pragma unroll
    for (int u = 0; u < U; u++) e[u] = ((const double2 *)p)[U * i + u];
  }
  template <int U> static __device__ __forceinline__ void st(void *p, long i, const double2 (&e)[U]) {
#pragma unroll
    for (int u = 0; u < U; u++) ((double2 *)p)[U * i + u] = e[u];
  }
};
template <> struct C2<float> {
  template <int U>
  static __device__ __forceinline__ void ld(const void *p, const float *, unsigned, long i, double2 (&e)[U]) {
#pragma unroll
    for (int u = 0; u < U; u += 2) {
      const float4 v = ((const float4 *)p)[(U * i + u) / 2];
      e[u] = make_double2(v.x, v.y);
      e[u + 1] = make_double2(v.z, v.w);
    }
  }
  template <int U> static __device__ __forceinline__ void st(void *p, long i, const double2 (&e)[U]) {
#pragma unroll
    for (int u = 0; u < U; u += 2)
      ((float4 *)p)[(U * i + u) / 2] =
          make_float4((float)e[u].x, (float)e[u].y, (float)e[u + 1].x, (float)e[u + 1].y);
  }
};
typedef f16 h8 __attribute__((ext_vector_type(8)));
template <> struct C2<f16> {
  template <int U>
  static __device__ __forceinline__ void ld(const void *p, const float *nrm, unsigned stride, long i,
                                            double2 (&e)[U]) {
    static_assert(U == 4, "FP16 vectors are read four elements at a time");
    const h8 v = ((const h8 *)p)[i];
    const unsigned site = (4u * (unsigned)i) % stride;
    const float4 m = *(const float4 *)(nrm + site);
    e[0] = make_double2((double)(m.x * (float)v[0]), (double)(m.x * (float)v[1]));
    e[1] = make_double2((double)(m.y * (float)v[2]), (double)(m.y * (float)v[3]));
    e[2] = make_double2((double)(m.z * (float)v[4]), (double)(m.z * (float)v[5]));
    e[3] = make_double2((double)(m.w * (float)v[6]), (double)(m.w * (float)v[7]));
  }
};
template <typename T> struct UnitOf {
  static constexpr int U = std::is_same<T, f16>::value ? 4 : 2;
};

// FP16 vectors are walked site-major: a lane takes four consecutive sites
// (q) and all 12 component runs of them, so each vector's per-site norms are
// read once per pass (component-major units would read them once per run).
__device__ __forceinline__ void ld_h4(const void *p, float4 m, long i, double2 (&e)[4]) {
  const h8 v = ((const h8 *)p)[i];
  e[0] = make_double2((double)(m.x * (float)v[0]), (double)(m.x * (float)v[1]));
  e[1] = make_double2((double)(m.y * (float)v[2]), (double)(m.y * (float)v[3]));
  e[2] = make_double2((double)(m.z * (float)v[4]), (double)(m.z * (float)v[5]));
  e[3] = make_double2((double)(m.w * (float)v[6]), (double)(m.w * (float)v[7]));
}

// the same four elements in single precision, and four elements of an FP32
// or FP64 vector as single-precision pairs (the FP16-basis passes accumulate
// within a lane in FP32: fewer registers, and the basis carries 11 bits)
__device__ __forceinline__ void ld_h4f(const void *p, float4 m, long i, float2 (&e)[4]) {
  const h8 v = ((const h8 *)p)[i];
  e[0] = make_float2(m.x * (float)v[0], m.x * (float)v[1]);
  e[1] = make_float2(m.y * (float)v[2], m.y * (float)v[3]);
  e[2] = make_float2(m.z * (float)v[4], m.z * (float)v[5]);
  e[3] = make_float2(m.w * (float)v[6], m.w * (float)v[7]);
}
template <typename T> __device__ __forceinline__ void ld4f(const void *p, long i, float2 (&e)[4]);
template <> __device__ __forceinline__ void ld4f<float>(const void *p, long i, float2 (&e)[4]) {
  const float4 a = ((const float4 *)p)[2 * i], b = ((const float4 *)p)[2 * i + 1];
  e[0] = make_float2(a.x, a.y);
  e[1] = make_float2(a.z, a.w);
  e[2] = make_float2(b.x, b.y);
  e[3] = make_float2(b.z, b.w);
}
template <> __device__ __forceinline__ void ld4f<double>(const void *p, long i, float2 (&e)[4]) {
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const double2 d = ((const double2 *)p)[4 * i + u];
    e[u] = make_float2((float)d.x, (float)d.y);
  }
}
template <typename T> __device__ __forceinline__ void st4f(void *p, long i, const float2 (&e)[4]);
template <> __device__ __forceinline__ void st4f<float>(void *p, long i, const float2 (&e)[4]) {
  ((float4 *)p)[2 * i] = make_float4(e[0].x, e[0].y, e[1].x, e[1].y);
  ((float4 *)p)[2 * i + 1] = make_float4(e[2].x, e[2].y, e[3].x, e[3].y);
}
template <> __device__ __forceinline__ void st4f<double>(void *p, long i, const float2 (&e)[4]) {
#pragma unroll
  for (int u = 0; u < 4; u++) ((double2 *)p)[4 * i + u] = make_double2(e[u].x, e[u].y);
}

// conj(a) b accumulated into (re, im)
__device__ __forceinline__ void cdot_acc(double2 a, double2 b, double &re, double &im) {
  re += a.x * b.x + a.y * b.y;
  im += a.x * b.y - a.y * b.x;
}
// b - h a
__device__ __forceinline__ double2 csub_mul(double2 b, double hr, double hi, double2 a) {
  return make_double2(b.x - (hr * a.x - hi * a.y), b.y - (hr * a.y + hi * a.x));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum NV doubles per lane over the workgroup; lane 0 of wave 0 gets totals.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *out) {
  __shared__ double sm[RB / 64][NV];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const double s = wave_sum(v[k]);
    if (lane == 0) sm[w][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0;
#pragma unroll
    for (int q = 0; q < RB / 64; q++) s += sm[q][threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// ---- multi-dot: out[k] = sum conj(v_k) w, k < K (<= KMAXP), one pass ----
// nu = n / U units (vector lengths are 12 runs of a stride that is a multiple of 64)
static constexpr int KMAXP = 16;
struct PtrPack {
  const void *p[KMAXP];
  const float *n[KMAXP];  // norms of FP16 vectors
};

template <typename TV, typename TW, int K>
__global__ void __launch_bounds__(RB) k_mdot(long nu, PtrPack v, const void *w, const float *wn, unsigned stride,
                                            double *part) {
  constexpr int U = UnitOf<TV>::U;
  double acc[2 * K];
#pragma unroll
  for (int k = 0; k < 2 * K; k++) acc[k] = 0;
  if constexpr (std::is_same<TV, f16>::value) {
    const long nq = nu / 12;  // site quads
    float af[2 * K];
#pragma unroll
    for (int k = 0; k < 2 * K; k++) af[k] = 0.f;
    for (long q = blockIdx.x * (long)RB + threadIdx.x; q < nq; q += (long)gridDim.x * RB) {
      float4 m[K];
#pragma unroll
      for (int k = 0; k < K; k++) m[k] = *(const float4 *)(v.n[k] + 4 * q);
      for (int c = 0; c < 12; c++) {
        const long i = c * nq + q;
        float2 b[4];
        ld4f<TW>(w, i, b);
#pragma unroll
        for (int k = 0; k < K; k++) {
          float2 a[4];
          ld_h4f(v.p[k], m[k], i, a);
#pragma unroll
          for (int u = 0; u < 4; u++) {
            af[2 * k] += a[u].x * b[u].x + a[u].y * b[u].y;
            af[2 * k + 1] += a[u].x * b[u].y - a[u].y * b[u].x;
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < 2 * K; k++) acc[k] = (double)af[k];
  } else {
    for (long i = blockIdx.x * (long)RB + threadIdx.x; i < nu; i += (long)gridDim.x * RB) {
      double2 b[U];
      C2<TW>::template ld<U>(w, wn, stride, i, b);
#pragma unroll
      for (int k = 0; k < K; k++) {
        double2 a[U];
        C2<TV>::template ld<U>(v.p[k], v.n[k], stride, i, a);
#pragma unroll
        for (int u = 0; u < U; u++) cdot_acc(a[u], b[u], acc[2 * k], acc[2 * k + 1]);
      }
    }
  }
  block_sum<2 * K>(acc, part + (long)blockIdx.x * 2 * K);
}


// final pass: res[q] = sum_b part[b*nq + q], one workgroup per q, fixed order
__global__ void __launch_bounds__(RB) k_final(int nb, int nq, const double *part, double *res) {
  const int q = blockIdx.x;
  double s = 0;
  for (int b = threadIdx.x; b < nb; b += RB) s += part[(long)b * nq + q];
  double v[1] = {s};
  __shared__ double o[1];
  block_sum<1>(v, o);
  __syncthreads();
  if (threadIdx.x == 0) res[q] = o[0];
}

// res[q] = sum_b part[b*ncol + c0 + q]: columns c0.. of rows ncol wide
__global__ void __launch_bounds__(RB) k_final_cols(int nb, int ncol, int c0, const double *part, double *res) {
  const int q = blockIdx.x;
  double s = 0;
  for (int b = threadIdx.x; b < nb; b += RB) s += part[(long)b * ncol + c0 + q];
  double v[1] = {s};
  __shared__ double o[1];
  block_sum<1>(v, o);
  __syncthreads();
  if (threadIdx.x == 0) res[q] = o[0];
}

// ---- paired Arnoldi steps over an FP16 basis with per-site norms ----
// at most KMAXP2 basis vectors per pass: the two work vectors' accumulators
// or coefficients and the basis norms double the registers of a lane, and a
// 16-vector pass drops to two waves per SIMD
static constexpr int KMAXP2 = 8;
// Two FP32 vectors w1 = A v_j, w2 = A w1 against K basis vectors in one pass:
// part rows of 4K: the K complex dots with w1, then the K with w2.
template <int K>
__global__ void __launch_bounds__(RB) k_mdot2(long nu, PtrPack v, const void *w1, const void *w2, double *part) {
  const long nq = nu / 12;  // site quads
  float af[4 * K];
#pragma unroll
  for (int k = 0; k < 4 * K; k++) af[k] = 0.f;
  for (long q = blockIdx.x * (long)RB + threadIdx.x; q < nq; q += (long)gridDim.x * RB) {
    float4 m[K];
#pragma unroll
    for (int k = 0; k < K; k++) m[k] = *(const float4 *)(v.n[k] + 4 * q);
    for (int c = 0; c < 12; c++) {
      const long i = c * nq + q;
      float2 b1[4], b2[4];
      ld4f<float>(w1, i, b1);
      ld4f<float>(w2, i, b2);
#pragma unroll
      for (int k = 0; k < K; k++) {
        float2 a[4];
        ld_h4f(v.p[k], m[k], i, a);
#pragma unroll
        for (int u = 0; u < 4; u++) {
          af[2 * k] += a[u].x * b1[u].x + a[u].y * b1[u].y;
          af[2 * k + 1] += a[u].x * b1[u].y - a[u].y * b1[u].x;
          af[2 * K + 2 * k] += a[u].x * b2[u].x + a[u].y * b2[u].y;
          af[2 * K + 2 * k + 1] += a[u].x * b2[u].y - a[u].y * b2[u].x;
        }
      }
    }
  }
  double acc[4 * K];
#pragma unroll
  for (int k = 0; k < 4 * K; k++) acc[k] = (double)af[k];
  block_sum<4 * K>(acc, part + (long)blockIdx.x * 4 * K);
}

// w1 -= sum_k g1_k v_k, w2 -= sum_k g2_k v_k (coefficients in device memory);
// LAST: partials of ||w1||^2, Re <w1, w2>, Im <w1, w2>, ||w2||^2
template <int K, bool LAST>
__global__ void __launch_bounds__(RB) k_maxpy2(long nu, PtrPack v, const double *g1, const double *g2, void *w1,
                                              void *w2, double *part) {
  const long nq = nu / 12;
  float c1r[K], c1i[K], c2r[K], c2i[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    c1r[k] = (float)g1[2 * k];
    c1i[k] = (float)g1[2 * k + 1];
    c2r[k] = (float)g2[2 * k];
    c2i[k] = (float)g2[2 * k + 1];
  }
  float s4[4] = {0.f, 0.f, 0.f, 0.f};
  for (long q = blockIdx.x * (long)RB + threadIdx.x; q < nq; q += (long)gridDim.x * RB) {
    float4 m[K];
#pragma unroll
    for (int k = 0; k < K; k++) m[k] = *(const float4 *)(v.n[k] + 4 * q);
    for (int c = 0; c < 12; c++) {
      const long i = c * nq + q;
      float2 b1[4], b2[4];
      ld4f<float>(w1, i, b1);
      ld4f<float>(w2, i, b2);
#pragma unroll
      for (int k = 0; k < K; k++) {
        float2 a[4];
        ld_h4f(v.p[k], m[k], i, a);
#pragma unroll
        for (int u = 0; u < 4; u++) {
          b1[u] = make_float2(b1[u].x - (c1r[k] * a[u].x - c1i[k] * a[u].y), b1[u].y - (c1r[k] * a[u].y + c1i[k] * a[u].x));
          b2[u] = make_float2(b2[u].x - (c2r[k] * a[u].x - c2i[k] * a[u].y), b2[u].y - (c2r[k] * a[u].y + c2i[k] * a[u].x));
        }
      }
      st4f<float>(w1, i, b1);
      st4f<float>(w2, i, b2);
      if (LAST)
#pragma unroll
        for (int u = 0; u < 4; u++) {
          s4[0] += b1[u].x * b1[u].x + b1[u].y * b1[u].y;
          s4[1] += b1[u].x * b2[u].x + b1[u].y * b2[u].y;
          s4[2] += b1[u].x * b2[u].y - b1[u].y * b2[u].x;
          s4[3] += b2[u].x * b2[u].x + b2[u].y * b2[u].y;
        }
    }
  }
  if (LAST) {
    double acc[4] = {s4[0], s4[1], s4[2], s4[3]};
    block_sum<4>(acc, part + (long)blockIdx.x * 4);
  }
}

// w2 = b[0] w2 + (b[1] + i b[2]) w1 (FP32, in place), partials of ||w2||^2
__global__ void __launch_bounds__(RB) k_comb2(long nu, const double *b, const void *w1, void *w2, double *part) {
  const float s2 = (float)b[0], s1r = (float)b[1], s1i = (float)b[2];
  float nr = 0.f;
  for (long i = blockIdx.x * (long)RB + threadIdx.x; i < nu; i += (long)gridDim.x * RB) {
    float2 u1[4], u2[4];
    ld4f<float>(w1, i, u1);
    ld4f<float>(w2, i, u2);
#pragma unroll
    for (int u = 0; u < 4; u++) {
      u2[u] = make_float2(s2 * u2[u].x + (s1r * u1[u].x - s1i * u1[u].y), s2 * u2[u].y + (s1r * u1[u].y + s1i * u1[u].x));
      nr += u2[u].x * u2[u].x + u2[u].y * u2[u].y;
    }
    st4f<float>(w2, i, u2);
  }
  double acc[1] = {(double)nr};
  block_sum<1>(acc, part + blockIdx.x);
}

// The end of a pair of Arnoldi steps over an FP16 basis with per-site norms,
// one thread per site (grid-stride): r = b0 u2 + (b1 + i b2) u1 from the FP32
// u1, u2 (12 component runs of stride elements), v1 = u1 / sqrt(*n1) and v2 = r
// in the basis's FP16 format (the site's 24 reals over their largest
// magnitude, that magnitude as the site's norm), partials of ||r||^2.  v2 is
// r unnormalised: its norms are scaled by 1 / ||r|| afterwards (k_scale_norms),
// the FP16 values of a site being the same for r and r / ||r||.
__global__ void __launch_bounds__(RB) k_comb2_basis(int Vh, long stride, const double *b, const double *n1,
                                                    const void *u1, const void *u2, void *v1, float *v1n, void *v2,
                                                    float *v2n, double *part) {
  const float s2 = (float)b[0], s1r = (float)b[1], s1i = (float)b[2];
  const double q = *n1;
  const float sv = q > 0 ? (float)(1.0 / sqrt(q)) : 0.f;
  const float2 *a1 = (const float2 *)u1, *a2 = (const float2 *)u2;
  typedef f16 hv2 __attribute__((ext_vector_type(2)));
  hv2 *o1 = (hv2 *)v1, *o2 = (hv2 *)v2;
  float nr = 0.f;
  for (long i = blockIdx.x * (long)RB + threadIdx.x; i < Vh; i += (long)gridDim.x * RB) {
    float2 x[12], r[12];
    float m1 = 0.f, m2 = 0.f;
#pragma unroll
    for (int k = 0; k < 12; k++) {
      x[k] = a1[k * stride + i];
      const float2 y = a2[k * stride + i];
      r[k] = make_float2(s2 * y.x + (s1r * x[k].x - s1i * x[k].y), s2 * y.y + (s1r * x[k].y + s1i * x[k].x));
      nr += r[k].x * r[k].x + r[k].y * r[k].y;
      m1 = fmaxf(m1, fmaxf(fabsf(x[k].x), fabsf(x[k].y)));
      m2 = fmaxf(m2, fmaxf(fabsf(r[k].x), fabsf(r[k].y)));
    }
    const float i1 = m1 > 0.f ? 1.f / m1 : 0.f, i2 = m2 > 0.f ? 1.f / m2 : 0.f;
#pragma unroll
    for (int k = 0; k < 12; k++) {
      hv2 h1, h2;
      h1.x = (f16)(x[k].x * i1);
      h1.y = (f16)(x[k].y * i1);
      h2.x = (f16)(r[k].x * i2);
      h2.y = (f16)(r[k].y * i2);
      o1[k * stride + i] = h1;
      o2[k * stride + i] = h2;
    }
    v1n[i] = m1 * sv;
    v2n[i] = m2;
  }
  double acc[1] = {(double)nr};
  block_sum<1>(acc, part + blockIdx.x);
}

// norm[i] *= 1 / sqrt(*n2) (0 where *n2 is not positive)
__global__ void __launch_bounds__(RB) k_scale_norms(int Vh, float *norm, const double *n2) {
  const double q = *n2;
  const float s = q > 0 ? (float)(1.0 / sqrt(q)) : 0.f;
  for (long i = blockIdx.x * (long)RB + threadIdx.x; i < Vh; i += (long)gridDim.x * RB) norm[i] *= s;
}

// Paired steps j, j + 1 on the device Hessenberg of the single-RHS solvers
// (columns of ld doubles: the complex <v_i, w>, then m squared norms): column
// j holds g1 = V_j^H A v_j (from pair_dot_dev); with nu^2 = ||u1'||^2 = sc[0],
// column j + 1 follows from the Arnoldi relation A v_{j+1} = (A^2 v_j -
// V_{j+1} H_j g1) / nu: H[k, j+1] = (g2[k] - c[k]) / nu for k <= j, (<u1',
// u2'> / nu - c[j+1]) / nu on the diagonal, c = H_j g1; b = (1/nu, -alpha/nu)
// with alpha = <u1', u2'> / nu^2, the coefficients of the part of A v_{j+1}
// orthogonal to v_0..v_{j+1}.  One thread.
__global__ void k_pair_columns(double *d, int ld, int m, int j, const double *g2, const double *sc, double *b) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double *n2 = d + (long)ld * m;
  const double *g1 = d + (long)ld * j;
  double *c1 = d + (long)ld * (j + 1);
  const double n1 = sc[0];
  n2[j] = n1;
  const double nu = n1 > 0.0 ? sqrt(n1) : 0.0;
  if (!(nu > 0.0)) {
    for (int k = 0; k <= j + 1; k++) c1[2 * k] = c1[2 * k + 1] = 0.0;
    b[0] = b[1] = b[2] = 0.0;
    return;
  }
  const double inv = 1.0 / nu;
  for (int k = 0; k <= j + 1; k++) {
    double cr = 0.0, ci = 0.0;
    for (int i = k > 0 ? k - 1 : 0; i <= j; i++) {
      double hr, hi;
      if (k <= i) {
        hr = d[(long)ld * i + 2 * k];
        hi = d[(long)ld * i + 2 * k + 1];
      } else {  // k == i + 1: the subdiagonal entry
        hr = n2[i] > 0.0 ? sqrt(n2[i]) : 0.0;
        hi = 0.0;
      }
      cr += hr * g1[2 * i] - hi * g1[2 * i + 1];
      ci += hr * g1[2 * i + 1] + hi * g1[2 * i];
    }
    if (k <= j) {
      c1[2 * k] = (g2[2 * k] - cr) * inv;
      c1[2 * k + 1] = (g2[2 * k + 1] - ci) * inv;
    } else {
      c1[2 * k] = (sc[1] * inv - cr) * inv;
      c1[2 * k + 1] = (sc[2] * inv - ci) * inv;
    }
  }
  b[0] = inv;
  b[1] = -sc[1] * inv * inv * inv;
  b[2] = -sc[2] * inv * inv * inv;
}

// ---- w = w - sum_k h_k v_k, returns ||w||^2 (optional) ----
struct CoefPack {
  double re[KMAXP], im[KMAXP];
};

template <typename TV, typename TW, int K, bool NORM>
__global__ void __launch_bounds__(RB) k_maxpy(long nu, PtrPack v, CoefPack h, void *w, unsigned stride,
                                             double *part) {
  constexpr int U = UnitOf<TV>::U;
  double acc[1] = {0};
  if constexpr (std::is_same<TV, f16>::value) {
    const long nq = nu / 12;  // site quads
    float fr[K], fi[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      fr[k] = (float)(h.re[k]);
      fi[k] = (float)(h.im[k]);
    }
    float nf = 0.f;
    for (long q = blockIdx.x * (long)RB + threadIdx.x; q < nq; q += (long)gridDim.x * RB) {
      float4 m[K];
#pragma unroll
      for (int k = 0; k < K; k++) m[k] = *(const float4 *)(v.n[k] + 4 * q);
      for (int c = 0; c < 12; c++) {
        const long i = c * nq + q;
        float2 b[4];
        ld4f<TW>(w, i, b);
#pragma unroll
        for (int k = 0; k < K; k++) {
          float2 a[4];
          ld_h4f(v.p[k], m[k], i, a);
#pragma unroll
          for (int u = 0; u < 4; u++)
            b[u] = make_float2(b[u].x - (fr[k] * a[u].x - fi[k] * a[u].y), b[u].y - (fr[k] * a[u].y + fi[k] * a[u].x));
        }
        st4f<TW>(w, i, b);
        if (NORM)
#pragma unroll
          for (int u = 0; u < 4; u++) nf += b[u].x * b[u].x + b[u].y * b[u].y;
      }
    }
    if (NORM) acc[0] = (double)nf;
  } else {
    for (long i = blockIdx.x * (long)RB + threadIdx.x; i < nu; i += (long)gridDim.x * RB) {
      double2 b[U];
      C2<TW>::template ld<U>(w, nullptr, stride, i, b);
#pragma unroll
      for (int k = 0; k < K; k++) {
        double2 a[U];
        C2<TV>::template ld<U>(v.p[k], v.n[k], stride, i, a);
#pragma unroll
        for (int u = 0; u < U; u++) b[u] = csub_mul(b[u], h.re[k], h.im[k], a[u]);
      }
      C2<TW>::template st<U>(w, i, b);
      if (NORM)
#pragma unroll
        for (int u = 0; u < U; u++) acc[0] += b[u].x * b[u].x + b[u].y * b[u].y;
    }
  }
  if (NORM) block_sum<1>(acc, part + blockIdx.x);
}

// the same with the coefficients read from device memory, h_k = coef[2k] + i coef[2k+1]
// (the multi-dot's results), so that no host round trip separates the two passes
template <typename TV, typename TW, int K, bool NORM>
__global__ void __launch_bounds__(RB) k_maxpy_d(long nu, PtrPack v, const double *coef, void *w, unsigned stride,
                                               double *part) {
  constexpr int U = UnitOf<TV>::U;
  double hr[K], hi[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    hr[k] = coef[2 * k];
    hi[k] = coef[2 * k + 1];
  }
  double acc[1] = {0};
  if constexpr (std::is_same<TV, f16>::value) {
    const long nq = nu / 12;  // site quads
    float fr[K], fi[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      fr[k] = (float)(hr[k]);
      fi[k] = (float)(hi[k]);
    }
    float nf = 0.f;
    for (long q = blockIdx.x * (long)RB + threadIdx.x; q < nq; q += (long)gridDim.x * RB) {
      float4 m[K];
#pragma unroll
      for (int k = 0; k < K; k++) m[k] = *(const float4 *)(v.n[k] + 4 * q);
      for (int c = 0; c < 12; c++) {
        const long i = c * nq + q;
        float2 b[4];
        ld4f<TW>(w, i, b);
#pragma unroll
        for (int k = 0; k < K; k++) {
          float2 a[4];
          ld_h4f(v.p[k], m[k], i, a);
#pragma unroll
          for (int u = 0; u < 4; u++)
            b[u] = make_float2(b[u].x - (fr[k] * a[u].x - fi[k] * a[u].y), b[u].y - (fr[k] * a[u].y + fi[k] * a[u].x));
        }
        st4f<TW>(w, i, b);
        if (NORM)
#pragma unroll
          for (int u = 0; u < 4; u++) nf += b[u].x * b[u].x + b[u].y * b[u].y;
      }
    }
    if (NORM) acc[0] = (double)nf;
  } else {
    for (long i = blockIdx.x * (long)RB + threadIdx.x; i < nu; i += (long)gridDim.x * RB) {
      double2 b[U];
      C2<TW>::template ld<U>(w, nullptr, stride, i, b);
#pragma unroll
      for (int k = 0; k < K; k++) {
        double2 a[U];
        C2<TV>::template ld<U>(v.p[k], v.n[k], stride, i, a);
#pragma unroll
        for (int u = 0; u < U; u++) b[u] = csub_mul(b[u], hr[k], hi[k], a[u]);
      }
      C2<TW>::template st<U>(w, i, b);
      if (NORM)
#pragma unroll
        for (int u = 0; u < U; u++) acc[0] += b[u].x * b[u].x + b[u].y * b[u].y;
    }
  }
  if (NORM) block_sum<1>(acc, part + blockIdx.x);
}

// w *= 1 / sqrt(*n2) (0 when *n2 == 0), the norm read from device memory
template <typename T>
__global__ void __launch_bounds__(RB) k_rscale(long nu, void *w, const double *n2, unsigned stride) {
  constexpr int U = UnitOf<T>::U;
  const double q = *n2;
  const double s = q > 0 ? 1.0 / sqrt(q) : 0.0;
  for (long i = blockIdx.x * (long)RB + threadIdx.x; i < nu; i += (long)gridDim.x * RB) {
    double2 b[U];
    C2<T>::template ld<U>(w, nullptr, stride, i, b);
#pragma unroll
    for (int u = 0; u < U; u++) b[u] = make_double2(s * b[u].x, s * b[u].y);
    C2<T>::template st<U>(w, i, b);
  }
}

// ---- y = a x + b y (+ norm of y) ----
__device__ __forceinline__ double2 caxpby(double ar, double ai, double2 a, double br, double bi, double2 b) {
  return make_double2(ar * a.x - ai * a.y + br * b.x - bi * b.y, ar * a.y + ai * a.x + br * b.y + bi * b.x);
}

template <typename TX, typename TY, bool NORM>
__global__ void __launch_bounds__(RB) k_axpby(long nu, double ar, double ai, const void *x, const float *xn,
                                             double br, double bi, void *y, unsigned stride, double *part) {
  constexpr int U = UnitOf<TX>::U;
  double acc[1] = {0};
  const bool ready = br != 0.0 || bi != 0.0;
  for (long i = blockIdx.x * (long)RB + threadIdx.x; i < nu; i += (long)gridDim.x * RB) {
    double2 a[U], b[U];
    C2<TX>::template ld<U>(x, xn, stride, i, a);
    if (ready) {
      C2<TY>::template ld<U>(y, nullptr, stride, i, b);
    } else {
#pragma unroll
      for (int u = 0; u < U; u++) b[u] = make_double2(0, 0);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      b[u] = caxpby(ar, ai, a[u], br, bi, b[u]);
      if (NORM) acc[0] += b[u].x * b[u].x + b[u].y * b[u].y;
    }
    C2<TY>::template st<U>(y, i, b);
  }
  if (NORM) block_sum<1>(acc, part + blockIdx.x);
}

// ---- Y_j = sum_i V_i Z(i,j) (+ add to Y if accumulate) -----------------
// Z is column-major m x k complex on the device.  Each lane walks its unit
// over the m inputs once per chunk of up to 8 outputs.
struct OutPack {
  void *p[8];
};
template <typename TV, typename TY, int KO>
__global__ void __launch_bounds__(RB) k_combine(long nu, int m, const void *const *V, const float *const *VN,
                                               const double *Z, int ldz, int j0, OutPack Y, int accumulate,
                                               unsigned stride) {
  constexpr int U = UnitOf<TV>::U;
  for (long i = blockIdx.x * (long)RB + threadIdx.x; i < nu; i += (long)gridDim.x * RB) {
    double2 s[KO][U];
#pragma unroll
    for (int q = 0; q < KO; q++)
#pragma unroll
      for (int u = 0; u < U; u++) s[q][u] = make_double2(0, 0);
    for (int r = 0; r < m; r++) {
      double2 a[U];
      C2<TV>::template ld<U>(V[r], VN[r], stride, i, a);
#pragma unroll
      for (int q = 0; q < KO; q++) {
        const double zr = Z[2 * (r + (long)ldz * (j0 + q))], zi = Z[2 * (r + (long)ldz * (j0 + q)) + 1];
#pragma unroll
        for (int u = 0; u < U; u++) s[q][u] = caxpby(zr, zi, a[u], 1.0, 0.0, s[q][u]);
      }
    }
#pragma unroll
    for (int q = 0; q < KO; q++) {
      if (accumulate) {
        double2 b[U];
        C2<TY>::template ld<U>(Y.p[q], nullptr, stride, i, b);
#pragma unroll
        for (int u = 0; u < U; u++) s[q][u] = make_double2(b[u].x + s[q][u].x, b[u].y + s[q][u].y);
      }
      C2<TY>::template st<U>(Y.p[q], i, s[q]);
    }
  }
}


// ===================================================================== //
// host side

Blas::Blas(hipStream_t s) : stream(s) {
  // partials: RG rows of 2 per vector (multi_dot), 4 per vector (pair_dot_dev)
  QQCD_HIP(hipMalloc(&part, sizeof(double) * RG * 4 * kmax));
  // one 2*KMAXP slab of results per pass of up to KMAXP vectors
  const long nres = 2L * KMAXP * ((kmax + KMAXP - 1) / KMAXP);
  QQCD_HIP(hipMalloc(&res, sizeof(double) * nres));
  QQCD_HIP(hipHostMalloc(&hres, sizeof(double) * nres, hipHostMallocDefault));
  QQCD_HIP(hipMalloc(&zbuf, sizeof(double) * 2 * zcap));
  QQCD_HIP(hipMalloc(&vptr, sizeof(void *) * 256));  // 128 vector and 128 norm pointers
}

Blas::~Blas() {
  (void)hipFree(part);
  (void)hipFree(res);
  (void)hipHostFree(hres);
  (void)hipFree(zbuf);
  (void)hipFree(vptr);
}

void Blas::fetch(int nq) {
  hipLaunchKernelGGL(k_final, dim3(nq), dim3(RB), 0, stream, RG, nq, part, res);
  QQCD_HIP(hipGetLastError());
  QQCD_HIP(hipMemcpyAsync(hres, res, sizeof(double) * nq, hipMemcpyDeviceToHost, stream));
  QQCD_HIP(hipStreamSynchronize(stream));
}

// call f(TA*, TB*) with the element types of a precision pair
// (FP16 only as the read side: FP16 vectors are written by the site-wise conversion)
template <typename F> static void tsw(Prec pa, Prec pb, F &&f) {
  if (pa == PREC_DOUBLE && pb == PREC_DOUBLE) f((double *)nullptr, (double *)nullptr);
  else if (pa == PREC_SINGLE && pb == PREC_DOUBLE) f((float *)nullptr, (double *)nullptr);
  else if (pa == PREC_DOUBLE && pb == PREC_SINGLE) f((double *)nullptr, (float *)nullptr);
  else if (pa == PREC_SINGLE && pb == PREC_SINGLE) f((float *)nullptr, (float *)nullptr);
  else if (pa == PREC_HALF && pb == PREC_SINGLE) f((f16 *)nullptr, (float *)nullptr);
  else if (pa == PREC_HALF && pb == PREC_DOUBLE) f((f16 *)nullptr, (double *)nullptr);
  else throw std::runtime_error("blas: unsupported precision pair");
}

static unsigned stride_of(const Vec &v) {
  if (v.n % 12 || v.n / 12 > 0x7FFFFFFFL) throw std::runtime_error("blas: vector is not a checkerboard field");
  return (unsigned)(v.n / 12);
}
static void check_half(const Vec &v) {
  if (v.prec == PREC_HALF && !v.norm) throw std::runtime_error("blas: FP16 vector without norms");
}
#define TYPES_OF(ta, tb)                                   \
  typedef typename std::remove_pointer<decltype(ta)>::type TA; \
  typedef typename std::remove_pointer<decltype(tb)>::type TB;

// kernel units of a vector of n elements read as TA
template <typename TA> static long units(long n) {
  if (n % UnitOf<TA>::U) throw std::runtime_error("blas: vector length is not a whole number of units");
  return n / UnitOf<TA>::U;
}

// f(std::integral_constant<int, K>) for K = k (1..16)
template <typename F> static void with_k(int k, F &&f) {
  switch (k) {
#define KC(K) \
  case K: f(std::integral_constant<int, K>()); break;
    KC(1) KC(2) KC(3) KC(4) KC(5) KC(6) KC(7) KC(8) KC(9) KC(10) KC(11) KC(12) KC(13) KC(14) KC(15) KC(16)
#undef KC
    default: throw std::runtime_error("blas: 1..16 vectors per pass");
  }
}

template <int K, typename TA, typename TB>
static void launch_mdot(hipStream_t s, const Vec &w, const PtrPack &pp, double *part) {
  hipLaunchKernelGGL((k_mdot<TA, TB, K>), dim3(RG), dim3(RB), 0, s, units<TA>(w.n), pp, w.data, w.norm, stride_of(w),
                     part);
}

void Blas::multi_dot(int k, const Vec *v, const Vec &w, std::complex<double> *out) {
  if (k > kmax) throw std::runtime_error("multi_dot: too many vectors");
  // every chunk of <= 8 vectors writes its partials into its own slab of
  // part (RG rows of 2*kk); one final pass and one host synchronisation
  long off = 0;
  std::vector<long> offs;
  std::vector<int> kks;
  for (int k0 = 0; k0 < k; k0 += KMAXP) {
    const int kk = std::min(KMAXP, k - k0);
    PtrPack pp{};
    for (int q = 0; q < kk; q++) {
      if (v[k0 + q].n != w.n || v[k0 + q].prec != v[k0].prec) throw std::runtime_error("multi_dot: mixed vectors");
      check_half(v[k0 + q]);
      pp.p[q] = v[k0 + q].data;
      pp.n[q] = v[k0 + q].norm;
    }
    double *pt = part + off;
    tsw(v[k0].prec, w.prec, [&](auto ta, auto tb) {
      TYPES_OF(ta, tb)
      with_k(kk, [&](auto K) { launch_mdot<decltype(K)::value, TA, TB>(stream, w, pp, pt); });
    });
    QQCD_HIP(hipGetLastError());
    offs.push_back(off);
    kks.push_back(kk);
    off += (long)RG * 2 * kk;
  }
  for (size_t c = 0; c < offs.size(); c++)
    hipLaunchKernelGGL(k_final, dim3(2 * kks[c]), dim3(RB), 0, stream, RG, 2 * kks[c], part + offs[c],
                       res + 2 * KMAXP * c);
  QQCD_HIP(hipGetLastError());
  QQCD_HIP(hipMemcpyAsync(hres, res, sizeof(double) * 2 * KMAXP * offs.size(), hipMemcpyDeviceToHost, stream));
  QQCD_HIP(hipStreamSynchronize(stream));
  for (size_t c = 0; c < offs.size(); c++)
    for (int q = 0; q < kks[c]; q++)
      out[KMAXP * c + q] = std::complex<double>(hres[2 * KMAXP * c + 2 * q], hres[2 * KMAXP * c + 2 * q + 1]);
}

std::complex<double> Blas::dot(const Vec &a, const Vec &b) {
  std::complex<double> r;
  multi_dot(1, &a, b, &r);
  return r;
}

double Blas::norm2(const Vec &a) { return dot(a, a).real(); }

template <int K, bool NORM, typename TA, typename TB>
static void launch_maxpy(hipStream_t s, Vec &w, const PtrPack &pp, const CoefPack &cp, double *part) {
  hipLaunchKernelGGL((k_maxpy<TA, TB, K, NORM>), dim3(RG), dim3(RB), 0, s, units<TA>(w.n), pp, cp, w.data,
                     stride_of(w), part);
}

double Blas::multi_axpy(int k, const std::complex<double> *h, const Vec *v, Vec &w, bool want_norm) {
  double nrm = 0;
  for (int k0 = 0; k0 < k; k0 += KMAXP) {
    const int kk = std::min(KMAXP, k - k0);
    const bool last = k0 + KMAXP >= k;
    PtrPack pp{};
    CoefPack cp{};
    for (int q = 0; q < kk; q++) {
      if (v[k0 + q].n != w.n || v[k0 + q].prec != v[k0].prec) throw std::runtime_error("multi_axpy: mixed vectors");
      check_half(v[k0 + q]);
      pp.p[q] = v[k0 + q].data;
      pp.n[q] = v[k0 + q].norm;
      cp.re[q] = h[k0 + q].real();
      cp.im[q] = h[k0 + q].imag();
    }
    const bool nm = want_norm && last;
    tsw(v[k0].prec, w.prec, [&](auto ta, auto tb) {
      TYPES_OF(ta, tb)
      with_k(kk, [&](auto K) {
        if (nm) launch_maxpy<decltype(K)::value, true, TA, TB>(stream, w, pp, cp, part);
        else launch_maxpy<decltype(K)::value, false, TA, TB>(stream, w, pp, cp, part);
      });
    });
    QQCD_HIP(hipGetLastError());
    if (nm) {
      fetch(1);
      nrm = hres[0];
    }
  }
  if (want_norm && k == 0) nrm = norm2(w);
  return nrm;
}

void Blas::multi_dot_dev(int k, const Vec *v, const Vec &w, double *dout) {
  if (k > kmax) throw std::runtime_error("multi_dot: too many vectors");
  long off = 0;
  for (int k0 = 0; k0 < k; k0 += KMAXP) {
    const int kk = std::min(KMAXP, k - k0);
    PtrPack pp{};
    for (int q = 0; q < kk; q++) {
      if (v[k0 + q].n != w.n || v[k0 + q].prec != v[k0].prec) throw std::runtime_error("multi_dot: mixed vectors");
      check_half(v[k0 + q]);
      pp.p[q] = v[k0 + q].data;
      pp.n[q] = v[k0 + q].norm;
    }
    double *pt = part + off;
    tsw(v[k0].prec, w.prec, [&](auto ta, auto tb) {
      TYPES_OF(ta, tb)
      with_k(kk, [&](auto K) { launch_mdot<decltype(K)::value, TA, TB>(stream, w, pp, pt); });
    });
    QQCD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_final, dim3(2 * kk), dim3(RB), 0, stream, RG, 2 * kk, pt, dout + 2 * k0);
    QQCD_HIP(hipGetLastError());
    off += (long)RG * 2 * kk;
  }
}

void Blas::multi_axpy_dev(int k, const double *dcoef, const Vec *v, Vec &w, double *dnorm2) {
  for (int k0 = 0; k0 < k; k0 += KMAXP) {
    const int kk = std::min(KMAXP, k - k0);
    const bool last = k0 + KMAXP >= k;
    PtrPack pp{};
    for (int q = 0; q < kk; q++) {
      if (v[k0 + q].n != w.n || v[k0 + q].prec != v[k0].prec) throw std::runtime_error("multi_axpy: mixed vectors");
      check_half(v[k0 + q]);
      pp.p[q] = v[k0 + q].data;
      pp.n[q] = v[k0 + q].norm;
    }
    const bool nm = dnorm2 && last;
    const double *cf = dcoef + 2 * k0;
    tsw(v[k0].prec, w.prec, [&](auto ta, auto tb) {
      TYPES_OF(ta, tb)
      with_k(kk, [&](auto K) {
        constexpr int KK = decltype(K)::value;
        if (nm)
          hipLaunchKernelGGL((k_maxpy_d<TA, TB, KK, true>), dim3(RG), dim3(RB), 0, stream, units<TA>(w.n), pp, cf,
                             w.data, stride_of(w), part);
        else
          hipLaunchKernelGGL((k_maxpy_d<TA, TB, KK, false>), dim3(RG), dim3(RB), 0, stream, units<TA>(w.n), pp, cf,
                             w.data, stride_of(w), part);
      });
    });
    QQCD_HIP(hipGetLastError());
    if (nm) {
      hipLaunchKernelGGL(k_final, dim3(1), dim3(RB), 0, stream, RG, 1, part, dnorm2);
      QQCD_HIP(hipGetLastError());
    }
  }
}

void Blas::pair_dot_dev(int k, const Vec *v, const Vec &w1, const Vec &w2, double *g1, double *g2) {
  if (k > kmax) throw std::runtime_error("pair_dot: too many vectors");
  if (w1.prec != PREC_SINGLE || w2.prec != PREC_SINGLE || w1.n != w2.n) throw std::runtime_error("pair_dot: FP32 w1, w2");
  long off = 0;
  for (int k0 = 0; k0 < k; k0 += KMAXP2) {
    const int kk = std::min(KMAXP2, k - k0);
    PtrPack pp{};
    for (int q = 0; q < kk; q++) {
      if (v[k0 + q].n != w1.n || v[k0 + q].prec != PREC_HALF) throw std::runtime_error("pair_dot: FP16 basis vectors");
      check_half(v[k0 + q]);
      pp.p[q] = v[k0 + q].data;
      pp.n[q] = v[k0 + q].norm;
    }
    double *pt = part + off;
    with_k(kk, [&](auto K) {
      hipLaunchKernelGGL((k_mdot2<decltype(K)::value>), dim3(RG), dim3(RB), 0, stream, units<f16>(w1.n), pp, w1.data,
                         w2.data, pt);
    });
    QQCD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_final_cols, dim3(2 * kk), dim3(RB), 0, stream, RG, 4 * kk, 0, pt, g1 + 2 * k0);
    hipLaunchKernelGGL(k_final_cols, dim3(2 * kk), dim3(RB), 0, stream, RG, 4 * kk, 2 * kk, pt, g2 + 2 * k0);
    QQCD_HIP(hipGetLastError());
    off += (long)RG * 4 * kk;
  }
}

void Blas::pair_axpy_dev(int k, const double *g1, const double *g2, const Vec *v, Vec &w1, Vec &w2, double *sc) {
  if (w1.prec != PREC_SINGLE || w2.prec != PREC_SINGLE || w1.n != w2.n || w1.data == w2.data)
    throw std::runtime_error("pair_axpy: distinct FP32 w1, w2");
  if (k < 1) throw std::runtime_error("pair_axpy: at least one vector");
  for (int k0 = 0; k0 < k; k0 += KMAXP2) {
    const int kk = std::min(KMAXP2, k - k0);
    const bool last = k0 + KMAXP2 >= k;
    PtrPack pp{};
    for (int q = 0; q < kk; q++) {
      if (v[k0 + q].n != w1.n || v[k0 + q].prec != PREC_HALF) throw std::runtime_error("pair_axpy: FP16 basis vectors");
      check_half(v[k0 + q]);
      pp.p[q] = v[k0 + q].data;
      pp.n[q] = v[k0 + q].norm;
    }
    with_k(kk, [&](auto K) {
      constexpr int KK = decltype(K)::value;
      if (last)
        hipLaunchKernelGGL((k_maxpy2<KK, true>), dim3(RG), dim3(RB), 0, stream, units<f16>(w1.n), pp, g1 + 2 * k0,
                           g2 + 2 * k0, w1.data, w2.data, part);
      else
        hipLaunchKernelGGL((k_maxpy2<KK, false>), dim3(RG), dim3(RB), 0, stream, units<f16>(w1.n), pp, g1 + 2 * k0,
                           g2 + 2 * k0, w1.data, w2.data, part);
    });
    QQCD_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_final, dim3(4), dim3(RB), 0, stream, RG, 4, part, sc);
  QQCD_HIP(hipGetLastError());
}

void Blas::pair_comb_dev(const double *b, const Vec &w1, Vec &w2, double *dnorm2) {
  if (w1.prec != PREC_SINGLE || w2.prec != PREC_SINGLE || w1.n != w2.n) throw std::runtime_error("pair_comb: FP32 w1, w2");
  hipLaunchKernelGGL(k_comb2, dim3(RG), dim3(RB), 0, stream, units<f16>(w1.n), b, w1.data, w2.data, part);
  QQCD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RB), 0, stream, RG, 1, part, dnorm2);
  QQCD_HIP(hipGetLastError());
}

void Blas::pair_comb_basis(long Vh, const double *b, const Spinor &w1, const Spinor &w2, const double *n1, Spinor &v1,
                           Spinor &v2, double *dnorm2) {
  if (w1.prec != PREC_SINGLE || w2.prec != PREC_SINGLE || v1.prec != PREC_HALF || v2.prec != PREC_HALF ||
      !v1.norm || !v2.norm || w1.stride != w2.stride || v1.stride != w1.stride || v2.stride != w1.stride)
    throw std::runtime_error("pair_comb_basis: FP32 w1, w2 and FP16 v1, v2 with norms, one layout");
  if (Vh < 1 || Vh > w1.stride) throw std::runtime_error("pair_comb_basis: site count");
  hipLaunchKernelGGL(k_comb2_basis, dim3(RG), dim3(RB), 0, stream, (int)Vh, w1.stride, b, n1, w1.data, w2.data,
                     v1.data, v1.norm, v2.data, v2.norm, part);
  QQCD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RB), 0, stream, RG, 1, part, dnorm2);
  QQCD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_scale_norms, dim3(RG), dim3(RB), 0, stream, (int)Vh, v2.norm, dnorm2);
  QQCD_HIP(hipGetLastError());
}

void pair_columns(double *d, int ld, int m, int j, const double *g2, const double *sc, double *b, hipStream_t s) {
  hipLaunchKernelGGL(k_pair_columns, dim3(1), dim3(64), 0, s, d, ld, m, j, g2, sc, b);
  QQCD_HIP(hipGetLastError());
}

void Blas::rscale_dev(Vec &w, const double *dnorm2) {
  if (w.prec == PREC_DOUBLE)
    hipLaunchKernelGGL(k_rscale<double>, dim3(RG), dim3(RB), 0, stream, units<double>(w.n), w.data, dnorm2,
                       stride_of(w));
  else if (w.prec == PREC_SINGLE)
    hipLaunchKernelGGL(k_rscale<float>, dim3(RG), dim3(RB), 0, stream, units<float>(w.n), w.data, dnorm2,
                       stride_of(w));
  else
    throw std::runtime_error("rscale: FP64 or FP32 vector expected");
  QQCD_HIP(hipGetLastError());
}

double Blas::axpby(std::complex<double> a, const Vec &x, std::complex<double> b, Vec &y, bool want_norm) {
  if (x.n != y.n) throw std::runtime_error("axpby: length mismatch");
  check_half(x);
  const unsigned st = stride_of(x);
  tsw(x.prec, y.prec, [&](auto ta, auto tb) {
    TYPES_OF(ta, tb)
    if (want_norm)
      hipLaunchKernelGGL((k_axpby<TA, TB, true>), dim3(RG), dim3(RB), 0, stream, units<TA>(x.n), a.real(), a.imag(),
                         x.data, x.norm, b.real(), b.imag(), y.data, st, part);
    else
      hipLaunchKernelGGL((k_axpby<TA, TB, false>), dim3(RG), dim3(RB), 0, stream, units<TA>(x.n), a.real(), a.imag(),
                         x.data, x.norm, b.real(), b.imag(), y.data, st, part);
  });
  QQCD_HIP(hipGetLastError());
  if (!want_norm) return 0;
  fetch(1);
  return hres[0];
}

void Blas::copy(const Vec &x, Vec &y) {
  if (x.prec == y.prec) {
    if (x.n != y.n) throw std::runtime_error("copy: length mismatch");
    QQCD_HIP(hipMemcpyAsync(y.data, x.data, x.bytes(), hipMemcpyDeviceToDevice, stream));
    if (x.prec == PREC_HALF) {
      check_half(x);
      check_half(y);
      QQCD_HIP(hipMemcpyAsync((void *)y.norm, x.norm, sizeof(float) * stride_of(x), hipMemcpyDeviceToDevice, stream));
    }
  } else {
    axpby(1.0, x, 0.0, y, false);
  }
}

// (an FP16 vector with zero halves is zero whatever its norms)
void Blas::zero(Vec &x) { QQCD_HIP(hipMemsetAsync(x.data, 0, x.bytes(), stream)); }

template <int KO, typename TA, typename TB>
static void launch_combine(hipStream_t s, const Vec &y, int m, void **vp, void **np, const double *Z, int ldz,
                           int j0, const OutPack &op, int acc) {
  hipLaunchKernelGGL((k_combine<TA, TB, KO>), dim3(RG), dim3(RB), 0, s, units<TA>(y.n), m, (const void *const *)vp,
                     (const float *const *)np, Z, ldz, j0, op, acc, stride_of(y));
}

void Blas::combine(int m, const Vec *V, int k, const std::complex<double> *Z, int ldz, Vec *Y, bool accumulate) {
  if (m == 0 || k == 0) return;
  if ((long)ldz * k > zcap || m > 128) throw std::runtime_error("combine: coefficient block too large");
  std::vector<void *> hp(2 * m);
  for (int r = 0; r < m; r++) {
    if (V[r].prec != V[0].prec || V[r].n != Y[0].n) throw std::runtime_error("combine: mixed vectors");
    check_half(V[r]);
    hp[r] = V[r].data;
    hp[m + r] = (void *)V[r].norm;
  }
  QQCD_HIP(hipMemcpyAsync(vptr, hp.data(), sizeof(void *) * 2 * m, hipMemcpyHostToDevice, stream));
  QQCD_HIP(hipMemcpyAsync(zbuf, Z, sizeof(double) * 2 * ldz * k, hipMemcpyHostToDevice, stream));
  for (int j0 = 0; j0 < k; j0 += 8) {
    const int kk = std::min(8, k - j0);
    OutPack op{};
    for (int q = 0; q < kk; q++) op.p[q] = Y[j0 + q].data;
    tsw(V[0].prec, Y[0].prec, [&](auto ta, auto tb) {
      TYPES_OF(ta, tb)
      switch (kk) {
        case 1: launch_combine<1, TA, TB>(stream, Y[0], m, vptr, vptr + m, zbuf, ldz, j0, op, accumulate); break;
        case 2: launch_combine<2, TA, TB>(stream, Y[0], m, vptr, vptr + m, zbuf, ldz, j0, op, accumulate); break;
        case 3: launch_combine<3, TA, TB>(stream, Y[0], m, vptr, vptr + m, zbuf, ldz, j0, op, accumulate); break;
        case 4: launch_combine<4, TA, TB>(stream, Y[0], m, vptr, vptr + m, zbuf, ldz, j0, op, accumulate); break;
        case 5: launch_combine<5, TA, TB>(stream, Y[0], m, vptr, vptr + m, zbuf, ldz, j0, op, accumulate); break;
        case 6: launch_combine<6, TA, TB>(stream, Y[0], m, vptr, vptr + m, zbuf, ldz, j0, op, accumulate); break;
        case 7: launch_combine<7, TA, TB>(stream, Y[0], m, vptr, vptr + m, zbuf, ldz, j0, op, accumulate); break;
        default: launch_combine<8, TA, TB>(stream, Y[0], m, vptr, vptr + m, zbuf, ldz, j0, op, accumulate); break;
      }
    });
    QQCD_HIP(hipGetLastError());
  }
  // the host copies of the pointer and coefficient tables go out of scope
  QQCD_HIP(hipStreamSynchronize(stream));
}

// Time one fused pass over k lattice vectors of precision p (FP64, FP32, or
// FP16 with an FP32 w) and grid workgroups: op 0 the multi-dot, 1 the
// multi-axpy with norm.
// Average kernel time in ms over niter launches (tuning only).
double bench_blas(long n, Prec p, int k, int op, int grid, int niter, hipStream_t s) {
  if (k < 1 || k > KMAXP || grid < 1 || grid > 8192 || niter < 1) throw std::runtime_error("bench_blas: bad sizes");
  if (p != PREC_DOUBLE && p != PREC_SINGLE && p != PREC_HALF)
    throw std::runtime_error("bench_blas: FP64, FP32 or FP16 (with an FP32 w)");
  // FP16: the k basis vectors in FP16 with per-site norms, w in FP32
  const Prec pw = p == PREC_HALF ? PREC_SINGLE : p;
  const size_t bytes = (size_t)n * (p == PREC_DOUBLE ? 16 : p == PREC_SINGLE ? 8 : 4);
  const size_t wbytes = (size_t)n * (pw == PREC_DOUBLE ? 16 : 8);
  std::vector<void *> v(k + 1, nullptr), nv(k, nullptr);
  double *part = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0;
  try {
    for (int q = 0; q <= k; q++) {
      const size_t b = q < k ? bytes : wbytes;
      QQCD_HIP(hipMalloc(&v[q], b));
      QQCD_HIP(hipMemsetAsync(v[q], 0, b, s));
      if (q < k && p == PREC_HALF) {
        QQCD_HIP(hipMalloc(&nv[q], sizeof(float) * (n / 12)));
        QQCD_HIP(hipMemsetAsync(nv[q], 0, sizeof(float) * (n / 12), s));
      }
    }
    QQCD_HIP(hipMalloc(&part, sizeof(double) * grid * 2 * KMAXP));
    PtrPack pp{};
    CoefPack cp{};
    for (int q = 0; q < k; q++) {
      pp.p[q] = v[q];
      pp.n[q] = (const float *)nv[q];
      cp.re[q] = 1e-3;
    }
    QQCD_HIP(hipEventCreate(&e0));
    QQCD_HIP(hipEventCreate(&e1));
    auto run = [&]() {
      tsw(p, pw, [&](auto ta, auto tb) {
        TYPES_OF(ta, tb)
        with_k(k, [&](auto K) {
          constexpr int KK = decltype(K)::value;
          if (op == 0)
            hipLaunchKernelGGL((k_mdot<TA, TB, KK>), dim3(grid), dim3(RB), 0, s, units<TA>(n), pp, v[k], nullptr,
                               (unsigned)(n / 12), part);
          else
            hipLaunchKernelGGL((k_maxpy<TA, TB, KK, true>), dim3(grid), dim3(RB), 0, s, units<TA>(n), pp, cp, v[k],
                               (unsigned)(n / 12), part);
        });
      });
      QQCD_HIP(hipGetLastError());
    };
    run();
    QQCD_HIP(hipEventRecord(e0, s));
    for (int it = 0; it < niter; it++) run();
    QQCD_HIP(hipEventRecord(e1, s));
    QQCD_HIP(hipEventSynchronize(e1));
    QQCD_HIP(hipEventElapsedTime(&ms, e0, e1));
  } catch (...) {
    for (auto &q : v)
      if (q) (void)hipFree(q);
    for (auto &q : nv)
      if (q) (void)hipFree(q);
    if (part) (void)hipFree(part);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    throw;
  }
  for (auto &q : v) (void)hipFree(q);
  for (auto &q : nv)
    if (q) (void)hipFree(q);
  (void)hipFree(part);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / niter;
}

}  // namespace qqcd
