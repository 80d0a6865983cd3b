// qqcd_mblas.hip -- level-1 kernels on multi-right-hand-side fields.
//
// The batched counterparts of qqcd_blas.hip for the noise solves (many
// right-hand sides on one gauge field; the reference runs them one after
// the other, qqcd-test/cfgsprops/quark/disconloops.f90:2460-2511).  Every
// per-slot dot product and norm accumulates in FP64 and is deterministic: a
// workgroup of 256 lanes covers 256/nb sites x nb slots, the lanes of one
// slot are combined by cross-lane shuffles over the site bits and a fixed
// LDS pass, and one final pass sums the workgroup partials in a fixed order.
// Coefficients are read from device arrays [q * cap + slot] and the active
// count from device memory, so the kernels can be replayed from a graph.
#include <algorithm>
#include <cmath>
#include <functional>
#include <stdexcept>
#include <type_traits>

#include "qqcd_multi.h"

namespace qqcd {

static constexpr int MB = 256;  // threads per workgroup

template <typename T> struct E2;
// ld: (re, im) in FP64; raw / ldr / cvt: the stored pair as loaded, and its
// conversion, so that a pass can issue all its loads before the arithmetic
template <> struct E2<double> {
  typedef double raw __attribute__((ext_vector_type(2)));
  static __device__ __forceinline__ raw ldr(const void *p, long i) { return ((const raw *)p)[i]; }
  static __device__ __forceinline__ double2 cvt(raw v) { return make_double2(v[0], v[1]); }
  static __device__ __forceinline__ double2 ld(const void *p, long i) { return ((const double2 *)p)[i]; }
  // stores (re, im) and returns the stored value
  static __device__ __forceinline__ double2 st(void *p, long i, double re, double im) {
    ((double2 *)p)[i] = make_double2(re, im);
    return make_double2(re, im);
  }
};
template <> struct E2<float> {
  typedef float raw __attribute__((ext_vector_type(2)));
  static __device__ __forceinline__ raw ldr(const void *p, long i) { return ((const raw *)p)[i]; }
  static __device__ __forceinline__ double2 cvt(raw v) { return make_double2(v[0], v[1]); }
  static __device__ __forceinline__ double2 ld(const void *p, long i) {
    const float2 v = ((const float2 *)p)[i];
    return make_double2(v.x, v.y);
  }
  static __device__ __forceinline__ double2 st(void *p, long i, double re, double im) {
    const float2 v = make_float2((float)re, (float)im);
    ((float2 *)p)[i] = v;
    return make_double2(v.x, v.y);
  }
};

// FP16 Krylov basis (no per-site norms): unit vectors stored times
// hs = sqrt(12 stride), so their elements are of order one; E2<f16> moves the
// stored values, and the passes that read such vectors scale by 1/hs
template <> struct E2<f16> {
  typedef f16 h2 __attribute__((ext_vector_type(2)));
  typedef h2 raw;
  static __device__ __forceinline__ raw ldr(const void *p, long i) { return ((const h2 *)p)[i]; }
  static __device__ __forceinline__ double2 cvt(raw v) { return make_double2((double)(float)v[0], (double)(float)v[1]); }
  static __device__ __forceinline__ double2 ld(const void *p, long i) {
    const h2 v = ((const h2 *)p)[i];
    return make_double2((double)(float)v[0], (double)(float)v[1]);
  }
  static __device__ __forceinline__ double2 st(void *p, long i, double re, double im) {
    h2 v;
    v[0] = (f16)(float)re;
    v[1] = (f16)(float)im;
    ((h2 *)p)[i] = v;
    return make_double2((double)(float)v[0], (double)(float)v[1]);
  }
};

struct MShape {
  long rows;       // 12 * stride
  long blk_elems;  // rows * nb
  const int *nact;
  int lognb;
  int cap;
  int gx;         // workgroups per RHS block
  double hs, hinv;  // FP16 basis scale and its inverse
};

// at most MAXK vectors per pass: a pass over q vectors reads (and, for the
// multi-axpy, writes) the working field once, so longer chunks cut those
// re-reads.  The slot sums go through LDS in groups of 18 values.
static constexpr int MAXK = 24;
// the two-vector passes of the two-step block: at most MAXK2 basis fields
static constexpr int MAXK2 = 24;
struct MPtr {
  const void *p[MAXK];
};

// Sum NV per-lane values over the lanes of each slot in the workgroup;
// thread (k, r) with k < NV, r < nb then writes its total to
// part[((k * cap + slot0 + r) * gx + bx].
template <int NV, typename T>
__device__ __forceinline__ void slot_sum(T (&v)[NV], int lognb, double *part, int cap, int slot0, int gx) {
  constexpr int G = NV < 18 ? NV : 18;  // values per round through LDS
  __shared__ double sm[MB / 64][G][64];
  const int nb = 1 << lognb;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k0 = 0; k0 < NV; k0 += G) {
    if (k0 > 0) __syncthreads();  // the previous round's reads are done
#pragma unroll
    for (int k = k0; k < (k0 + G < NV ? k0 + G : NV); k++) {
      double s = (double)v[k];
      for (int o = 32; o >= nb; o >>= 1) s += __shfl_xor(s, o, 64);
      if (lane < nb) sm[w][k - k0][lane] = s;
    }
    __syncthreads();
    const int ng = k0 + G < NV ? G : NV - k0;
    for (int t = threadIdx.x; t < ng * nb; t += MB) {
      const int k = t >> lognb, r = t & (nb - 1);
      // the last block may hold fewer than nb slots
      if (slot0 + r >= cap) continue;
      double s = 0;
#pragma unroll
      for (int q = 0; q < MB / 64; q++) s += sm[q][k][r];
      part[((long)(k0 + k) * cap + slot0 + r) * gx + blockIdx.x] = s;
    }
  }
}

// The K basis values of one element pass through an empty asm statement (no
// instruction is emitted) before they are used, so all K loads are issued
// back to back and are in flight together: left to itself the compiler
// interleaved loads and arithmetic for K >= 15 and those passes ran at half
// speed (24^3x48, 63 slots: 7.2-8.2 ms for 15 or 16 FP16 vectors against
// 3.6 ms for 14).
template <typename R, int K> __device__ __forceinline__ void pin_loads(R (&x)[K]) {
#pragma unroll
  for (int k = 0; k < K; k++) asm volatile("" : "+v"(x[k]));
}

// partials: part[((k * cap + slot) * gx + bx] for workgroup bx of the slot's block;
// SELF: two more components, <W, W> (real, imaginary 0) after the K dots
template <typename TV, typename TW, int K, bool SHARED, bool SELF>
__global__ void __launch_bounds__(MB) k_mdot(MShape a, MPtr v, const void *w, double *part, const int *any) {
  if (any && !*any) return;
  const int nact = *a.nact, nb = 1 << a.lognb, blk = blockIdx.y;
  if ((blk << a.lognb) >= nact) return;
  const int rr = threadIdx.x & (nb - 1), slot = (blk << a.lognb) + rr;
  constexpr int NV = 2 * K + (SELF ? 2 : 0);
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) acc[k] = 0;
  if (slot < nact) {
    const int per = MB >> a.lognb;
    const long base = (long)blk * a.blk_elems;
    for (long p = (long)blockIdx.x * per + (threadIdx.x >> a.lognb); p < a.rows; p += (long)a.gx * per) {
      const long e = base + (p << a.lognb) + rr;
      const double2 b = E2<TW>::ld(w, e);
      typename E2<TV>::raw xr[K];
#pragma unroll
      for (int k = 0; k < K; k++) xr[k] = E2<TV>::ldr(v.p[k], SHARED ? p : e);
      pin_loads(xr);
#pragma unroll
      for (int k = 0; k < K; k++) {
        const double2 x = E2<TV>::cvt(xr[k]);
        acc[2 * k] += x.x * b.x + x.y * b.y;
        acc[2 * k + 1] += x.x * b.y - x.y * b.x;
      }
      if (SELF) acc[2 * K] += b.x * b.x + b.y * b.y;
    }
  }
  if (std::is_same<TV, f16>::value) {
#pragma unroll
    for (int k = 0; k < 2 * K; k++) acc[k] *= a.hinv;
  }
  slot_sum<NV>(acc, a.lognb, part, a.cap, blk << a.lognb, a.gx);
}

// One wavefront per (component c, slot): the sum over bx of the partials
// part[((c * cap + slot) * gx + bx], lane-strided then folded by shuffles
// (a fixed order).  Complex results (nv = 2 nq components) go to
// res[2 ((q0 + c/2) * cap + slot) + c%2], real ones to res[(q0 + c) * cap + slot].
__global__ void __launch_bounds__(64) k_mfinal(int nv, int cap, int gx, const int *nact, const double *part,
                                               double *res, int q0, int is_complex) {
  const int slot = blockIdx.x, c = blockIdx.y;
  if (slot >= *nact) return;
  const double *p = part + ((long)c * cap + slot) * gx;
  double s = 0;
  for (int b = threadIdx.x; b < gx; b += 64) s += p[b];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (threadIdx.x == 0) {
    if (is_complex) res[2 * ((long)(q0 + c / 2) * cap + slot) + (c & 1)] = s;
    else res[(long)(q0 + c) * cap + slot] = s;
  }
}

// W += sign * sum_k c_k V_k  per slot (optionally masked by flag), optional norm partials
template <typename TV, typename TW, int K, bool SHARED, bool NORM>
__global__ void __launch_bounds__(MB) k_maxpy(MShape a, MPtr v, const double *coef, int q0, double sign, void *w,
                                              double *part, const int *flag, const int *any) {
  if (any && !*any) return;
  const int nact = *a.nact, nb = 1 << a.lognb, blk = blockIdx.y;
  if ((blk << a.lognb) >= nact) return;
  const int rr = threadIdx.x & (nb - 1), slot = (blk << a.lognb) + rr;
  double acc[1] = {0};
  bool act = slot < nact;
  if (act && flag) act = flag[slot] != 0;
  if (act) {
    double cr[K], ci[K];
    const double f = std::is_same<TV, f16>::value ? sign * a.hinv : sign;
#pragma unroll
    for (int k = 0; k < K; k++) {
      cr[k] = f * coef[2 * ((long)(q0 + k) * a.cap + slot)];
      ci[k] = f * coef[2 * ((long)(q0 + k) * a.cap + slot) + 1];
    }
    const int per = MB >> a.lognb;
    const long base = (long)blk * a.blk_elems;
    for (long p = (long)blockIdx.x * per + (threadIdx.x >> a.lognb); p < a.rows; p += (long)a.gx * per) {
      const long e = base + (p << a.lognb) + rr;
      double2 b = E2<TW>::ld(w, e);
      typename E2<TV>::raw xr[K];
#pragma unroll
      for (int k = 0; k < K; k++) xr[k] = E2<TV>::ldr(v.p[k], SHARED ? p : e);
      pin_loads(xr);
#pragma unroll
      for (int k = 0; k < K; k++) {
        const double2 x = E2<TV>::cvt(xr[k]);
        b.x += cr[k] * x.x - ci[k] * x.y;
        b.y += cr[k] * x.y + ci[k] * x.x;
      }
      // the norm of what was stored (FP32 fields: of the rounded value)
      const double2 s = E2<TW>::st(w, e, b.x, b.y);
      if (NORM) acc[0] += s.x * s.x + s.y * s.y;
    }
  }
  if (NORM) slot_sum<1>(acc, a.lognb, part, a.cap, blk << a.lognb, a.gx);
}

template <typename TX, typename TY, bool NORM>
__global__ void __launch_bounds__(MB) k_maxpby(MShape a, double ca, const void *x, double cb, void *y, double *part) {
  const int nact = *a.nact, nb = 1 << a.lognb, blk = blockIdx.y;
  if ((blk << a.lognb) >= nact) return;
  const int rr = threadIdx.x & (nb - 1), slot = (blk << a.lognb) + rr;
  double acc[1] = {0};
  if (slot < nact) {
    const int per = MB >> a.lognb;
    const long base = (long)blk * a.blk_elems;
    for (long p = (long)blockIdx.x * per + (threadIdx.x >> a.lognb); p < a.rows; p += (long)a.gx * per) {
      const long e = base + (p << a.lognb) + rr;
      const double2 u = E2<TX>::ld(x, e);
      double2 b = make_double2(0, 0);
      if (cb != 0.0) b = E2<TY>::ld(y, e);
      const double2 s = E2<TY>::st(y, e, ca * u.x + cb * b.x, ca * u.y + cb * b.y);
      if (NORM) acc[0] += s.x * s.x + s.y * s.y;
    }
  }
  if (NORM) slot_sum<1>(acc, a.lognb, part, a.cap, blk << a.lognb, a.gx);
}

template <typename T>
__global__ void __launch_bounds__(MB) k_mnorm(MShape a, const void *w, double *part) {
  const int nact = *a.nact, nb = 1 << a.lognb, blk = blockIdx.y;
  if ((blk << a.lognb) >= nact) return;
  const int rr = threadIdx.x & (nb - 1), slot = (blk << a.lognb) + rr;
  double acc[1] = {0};
  if (slot < nact) {
    const int per = MB >> a.lognb;
    const long base = (long)blk * a.blk_elems;
    for (long p = (long)blockIdx.x * per + (threadIdx.x >> a.lognb); p < a.rows; p += (long)a.gx * per) {
      const double2 s = E2<T>::ld(w, base + (p << a.lognb) + rr);
      acc[0] += s.x * s.x + s.y * s.y;
    }
  }
  slot_sum<1>(acc, a.lognb, part, a.cap, blk << a.lognb, a.gx);
}

template <typename TX, typename TW>
__global__ void __launch_bounds__(MB) k_mscale(MShape a, const void *x, void *w, const double *n2) {
  const int nact = *a.nact, nb = 1 << a.lognb, blk = blockIdx.y;
  if ((blk << a.lognb) >= nact) return;
  const int rr = threadIdx.x & (nb - 1), slot = (blk << a.lognb) + rr;
  if (slot >= nact) return;
  const double n = n2[slot];
  const double s = n > 0.0 ? 1.0 / sqrt(n) : 0.0;
  const int per = MB >> a.lognb;
  const long base = (long)blk * a.blk_elems;
  for (long p = (long)blockIdx.x * per + (threadIdx.x >> a.lognb); p < a.rows; p += (long)a.gx * per) {
    const long e = base + (p << a.lognb) + rr;
    const double2 u = E2<TX>::ld(x, e);
    E2<TW>::st(w, e, s * u.x, s * u.y);
  }
}

// Q(r) = X(r) / sqrt(n2[r]) and H(r) = hs Q(r) in FP16 (a new basis vector:
// Q feeds the operator, H the Gram-Schmidt passes)
template <typename TX, typename TQ>
__global__ void __launch_bounds__(MB) k_mscale2(MShape a, const void *x, void *q, void *h, const double *n2) {
  const int nact = *a.nact, nb = 1 << a.lognb, blk = blockIdx.y;
  if ((blk << a.lognb) >= nact) return;
  const int rr = threadIdx.x & (nb - 1), slot = (blk << a.lognb) + rr;
  if (slot >= nact) return;
  const double n = n2[slot];
  const double s = n > 0.0 ? 1.0 / sqrt(n) : 0.0;
  const int per = MB >> a.lognb;
  const long base = (long)blk * a.blk_elems;
  for (long p = (long)blockIdx.x * per + (threadIdx.x >> a.lognb); p < a.rows; p += (long)a.gx * per) {
    const long e = base + (p << a.lognb) + rr;
    const double2 u = E2<TX>::ld(x, e);
    const double2 v = E2<TQ>::st(q, e, s * u.x, s * u.y);
    E2<f16>::st(h, e, a.hs * v.x, a.hs * v.y);
  }
}

template <typename T>
__global__ void __launch_bounds__(MB) k_mswap(long rows, int lognb, long blk_elems, void *w, int sa, int sb) {
  const long p = blockIdx.x * (long)MB + threadIdx.x;
  if (p >= rows) return;
  const int nb = 1 << lognb;
  const long ea = (long)(sa >> lognb) * blk_elems + (p << lognb) + (sa & (nb - 1));
  const long eb = (long)(sb >> lognb) * blk_elems + (p << lognb) + (sb & (nb - 1));
  const double2 u = E2<T>::ld(w, ea), v = E2<T>::ld(w, eb);
  E2<T>::st(w, ea, v.x, v.y);
  E2<T>::st(w, eb, u.x, u.y);
}

template <typename TS, typename TW>
__global__ void __launch_bounds__(MB) k_mput(long rows, int lognb, long blk_elems, const void *s, void *w, int slot) {
  const long p = blockIdx.x * (long)MB + threadIdx.x;
  if (p >= rows) return;
  const int nb = 1 << lognb;
  const long e = (long)(slot >> lognb) * blk_elems + (p << lognb) + (slot & (nb - 1));
  const double2 u = E2<TS>::ld(s, p);
  E2<TW>::st(w, e, u.x, u.y);
}

template <typename TW, typename TS>
__global__ void __launch_bounds__(MB) k_mget(long rows, int lognb, long blk_elems, const void *w, void *s, int slot) {
  const long p = blockIdx.x * (long)MB + threadIdx.x;
  if (p >= rows) return;
  const int nb = 1 << lognb;
  const long e = (long)(slot >> lognb) * blk_elems + (p << lognb) + (slot & (nb - 1));
  const double2 u = E2<TW>::ld(w, e);
  E2<TS>::st(s, p, u.x, u.y);
}

__global__ void k_arnoldi_flags(const double *n2, const double *w2, int *flag, int *any, const int *nact, int cap,
                                int force) {
  const int r = threadIdx.x;
  int f = 0;
  if (r < cap) {
    if (nact && r < *nact) f = force ? 1 : n2[r] < 0.5 * w2[2 * r];
    flag[r] = f;
  }
  const int a = __syncthreads_or(f);
  if (r == 0) *any = a;
}

// column j of the batched Hessenberg matrix, and the final squared norm in nf
__global__ void k_arnoldi_column(double *H, int ldh, int j, const double *h1, const double *h2, const int *flag,
                                 const double *n2a, const double *n2b, double *nf, const int *nact, int cap) {
  const int t = blockIdx.x * MB + threadIdx.x;
  const int q = t / cap, r = t - q * cap;
  if (q > j + 1 || r >= *nact) return;
  const int f = flag[r];
  double *h = H + 2 * (((long)j * ldh + q) * cap + r);
  if (q <= j) {
    double re = h1[2 * ((long)q * cap + r)], im = h1[2 * ((long)q * cap + r) + 1];
    if (f) {
      re += h2[2 * ((long)q * cap + r)];
      im += h2[2 * ((long)q * cap + r) + 1];
    }
    h[0] = re;
    h[1] = im;
  } else {
    const double n = f ? n2b[r] : n2a[r];
    nf[r] = n;
    h[0] = n > 0.0 ? sqrt(n) : 0.0;
    h[1] = 0.0;
  }
}

// ---- two-step Arnoldi block (FP16 basis V, FP32 W1 = A v_j, W2 = A W1) ----
// Per-lane sums in FP32 (the basis carries 11 bits), across lanes in FP64.

// part[((c * cap + slot) * gx + bx]: c < 2K <V_k, W1>, then 2K + c <V_k, W2>;
// SELF: then ||W1||^2, Re <W1, W2>, Im <W1, W2>, ||W2||^2
template <int K, bool SELF>
__global__ void __launch_bounds__(MB) k_mdot2(MShape a, MPtr v, const void *w1, const void *w2, double *part) {
  const int nact = *a.nact, nb = 1 << a.lognb, blk = blockIdx.y;
  if ((blk << a.lognb) >= nact) return;
  const int rr = threadIdx.x & (nb - 1), slot = (blk << a.lognb) + rr;
  constexpr int NV = 4 * K + (SELF ? 4 : 0);
  float af[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) af[k] = 0.f;
  if (slot < nact) {
    const int per = MB >> a.lognb;
    const long base = (long)blk * a.blk_elems;
    for (long p = (long)blockIdx.x * per + (threadIdx.x >> a.lognb); p < a.rows; p += (long)a.gx * per) {
      const long e = base + (p << a.lognb) + rr;
      const typename E2<float>::raw b1 = E2<float>::ldr(w1, e), b2 = E2<float>::ldr(w2, e);
      typename E2<f16>::raw xr[K];
#pragma unroll
      for (int k = 0; k < K; k++) xr[k] = E2<f16>::ldr(v.p[k], e);
      pin_loads(xr);
#pragma unroll
      for (int k = 0; k < K; k++) {
        const float xx = (float)xr[k][0], xy = (float)xr[k][1];
        af[2 * k] += xx * b1[0] + xy * b1[1];
        af[2 * k + 1] += xx * b1[1] - xy * b1[0];
        af[2 * K + 2 * k] += xx * b2[0] + xy * b2[1];
        af[2 * K + 2 * k + 1] += xx * b2[1] - xy * b2[0];
      }
      if (SELF) {
        af[4 * K] += b1[0] * b1[0] + b1[1] * b1[1];
        af[4 * K + 1] += b1[0] * b2[0] + b1[1] * b2[1];
        af[4 * K + 2] += b1[0] * b2[1] - b1[1] * b2[0];
        af[4 * K + 3] += b2[0] * b2[0] + b2[1] * b2[1];
      }
    }
  }
  const float hinv = (float)a.hinv;
#pragma unroll
  for (int k = 0; k < 4 * K; k++) af[k] *= hinv;
  slot_sum<NV>(af, a.lognb, part, a.cap, blk << a.lognb, a.gx);
}

// W1 -= sum_k g1_k V_k, W2 -= sum_k g2_k V_k (coefficients g[2 ((q0 + k) cap + slot)]);
// LAST: partials of ||W1||^2, Re <W1, W2>, Im <W1, W2>, ||W2||^2 of the stored values
template <int K, bool LAST>
__global__ void __launch_bounds__(MB) k_maxpy2(MShape a, MPtr v, const double *g1, const double *g2, int q0, void *w1,
                                               void *w2, double *part) {
  const int nact = *a.nact, nb = 1 << a.lognb, blk = blockIdx.y;
  if ((blk << a.lognb) >= nact) return;
  const int rr = threadIdx.x & (nb - 1), slot = (blk << a.lognb) + rr;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  if (slot < nact) {
    float c1r[K], c1i[K], c2r[K], c2i[K];
    const double f = a.hinv;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const long o = 2 * ((long)(q0 + k) * a.cap + slot);
      c1r[k] = (float)(f * g1[o]);
      c1i[k] = (float)(f * g1[o + 1]);
      c2r[k] = (float)(f * g2[o]);
      c2i[k] = (float)(f * g2[o + 1]);
    }
    const int per = MB >> a.lognb;
    const long base = (long)blk * a.blk_elems;
    typedef typename E2<float>::raw f2;
    for (long p = (long)blockIdx.x * per + (threadIdx.x >> a.lognb); p < a.rows; p += (long)a.gx * per) {
      const long e = base + (p << a.lognb) + rr;
      f2 b1 = E2<float>::ldr(w1, e), b2 = E2<float>::ldr(w2, e);
      typename E2<f16>::raw xr[K];
#pragma unroll
      for (int k = 0; k < K; k++) xr[k] = E2<f16>::ldr(v.p[k], e);
      pin_loads(xr);
#pragma unroll
      for (int k = 0; k < K; k++) {
        const float xx = (float)xr[k][0], xy = (float)xr[k][1];
        b1[0] -= c1r[k] * xx - c1i[k] * xy;
        b1[1] -= c1r[k] * xy + c1i[k] * xx;
        b2[0] -= c2r[k] * xx - c2i[k] * xy;
        b2[1] -= c2r[k] * xy + c2i[k] * xx;
      }
      ((f2 *)w1)[e] = b1;
      ((f2 *)w2)[e] = b2;
      if (LAST) {
        acc[0] += b1[0] * b1[0] + b1[1] * b1[1];
        acc[1] += b1[0] * b2[0] + b1[1] * b2[1];
        acc[2] += b1[0] * b2[1] - b1[1] * b2[0];
        acc[3] += b2[0] * b2[0] + b2[1] * b2[1];
      }
    }
  }
  if (LAST) slot_sum<4>(acc, a.lognb, part, a.cap, blk << a.lognb, a.gx);
}

// The last chunk of axpy2 with next2 folded in: W1' = W1 - sum_k g1_k V_k and
// W2' likewise stay in registers, R = b0 W2' + (b1 + i b2) W1' (FP32) and
// H = b3 W1' (FP16) are stored, with partials of ||R||^2 of the stored values.
// The coefficients b come before the pass, from norms and inner products of
// W1', W2' formed from those of W1, W2 and the projections (pythag_scalars).
template <int K>
__global__ void __launch_bounds__(MB) k_maxpy2n(MShape a, MPtr v, const double *g1, const double *g2, int q0,
                                                const void *w1, const void *w2, void *rout, void *hout,
                                                const double *b, double *part) {
  const int nact = *a.nact, nb = 1 << a.lognb, blk = blockIdx.y;
  if ((blk << a.lognb) >= nact) return;
  const int rr = threadIdx.x & (nb - 1), slot = (blk << a.lognb) + rr;
  double acc[1] = {0.0};
  if (slot < nact) {
    float c1r[K], c1i[K], c2r[K], c2i[K];
    const double f = a.hinv;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const long o = 2 * ((long)(q0 + k) * a.cap + slot);
      c1r[k] = (float)(f * g1[o]);
      c1i[k] = (float)(f * g1[o + 1]);
      c2r[k] = (float)(f * g2[o]);
      c2i[k] = (float)(f * g2[o + 1]);
    }
    const float s2 = (float)b[4 * slot], s1r = (float)b[4 * slot + 1], s1i = (float)b[4 * slot + 2],
                sv = (float)b[4 * slot + 3];
    const int per = MB >> a.lognb;
    const long base = (long)blk * a.blk_elems;
    typedef typename E2<float>::raw f2;
    typedef typename E2<f16>::raw h2;
    float nr = 0.f;
    for (long p = (long)blockIdx.x * per + (threadIdx.x >> a.lognb); p < a.rows; p += (long)a.gx * per) {
      const long e = base + (p << a.lognb) + rr;
      f2 b1 = E2<float>::ldr(w1, e), b2 = E2<float>::ldr(w2, e);
      typename E2<f16>::raw xr[K];
#pragma unroll
      for (int k = 0; k < K; k++) xr[k] = E2<f16>::ldr(v.p[k], e);
      pin_loads(xr);
#pragma unroll
      for (int k = 0; k < K; k++) {
        const float xx = (float)xr[k][0], xy = (float)xr[k][1];
        b1[0] -= c1r[k] * xx - c1i[k] * xy;
        b1[1] -= c1r[k] * xy + c1i[k] * xx;
        b2[0] -= c2r[k] * xx - c2i[k] * xy;
        b2[1] -= c2r[k] * xy + c2i[k] * xx;
      }
      f2 r;
      r[0] = s2 * b2[0] + (s1r * b1[0] - s1i * b1[1]);
      r[1] = s2 * b2[1] + (s1r * b1[1] + s1i * b1[0]);
      ((f2 *)rout)[e] = r;
      h2 hv;
      hv[0] = (f16)(sv * b1[0]);
      hv[1] = (f16)(sv * b1[1]);
      ((h2 *)hout)[e] = hv;
      nr += r[0] * r[0] + r[1] * r[1];
    }
    acc[0] = (double)nr;
  }
  slot_sum<1>(acc, a.lognb, part, a.cap, blk << a.lognb, a.gx);
}

// One thread per slot: sc = (||W1'||^2, Re <W1', W2'>, Im <W1', W2'>,
// ||W2'||^2) of W1' = W1 - V g1, W2' = W2 - V g2 from self = the same four of
// W1, W2 and the projections g = <V_k, W> (V orthonormal: ||W1'||^2 = ||W1||^2
// - sum |g1_k|^2, <W1', W2'> = <W1, W2> - sum conj(g1_k) g2_k).
__global__ void k_pythag_scalars(int nq, const double *g1, const double *g2, const double *self, double *sc,
                                 const int *nact, int cap) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= cap || r >= *nact) return;
  double n1 = 0.0, cr = 0.0, ci = 0.0, n2 = 0.0;
  for (int q = 0; q < nq; q++) {
    const long o = 2 * ((long)q * cap + r);
    const double ar = g1[o], ai = g1[o + 1], br = g2[o], bi = g2[o + 1];
    n1 += ar * ar + ai * ai;
    n2 += br * br + bi * bi;
    cr += ar * br + ai * bi;
    ci += ar * bi - ai * br;
  }
  const double p1 = self[r] - n1;
  sc[r] = p1 > 0.0 ? p1 : 0.0;
  sc[cap + r] = self[cap + r] - cr;
  sc[2 * cap + r] = self[2 * cap + r] - ci;
  sc[3 * cap + r] = self[3 * cap + r] - n2;
}

// R = b0 W2 + (b1 + i b2) W1 in FP32, H = b3 W1 in FP16 (b per slot, 4 doubles),
// partials of ||R||^2 of the stored values
__global__ void __launch_bounds__(MB) k_next2(MShape a, const void *w1, const void *w2, void *rout, void *hout,
                                              const double *b, double *part) {
  const int nact = *a.nact, nb = 1 << a.lognb, blk = blockIdx.y;
  if ((blk << a.lognb) >= nact) return;
  const int rr = threadIdx.x & (nb - 1), slot = (blk << a.lognb) + rr;
  double acc[1] = {0.0};
  if (slot < nact) {
    const float s2 = (float)b[4 * slot], s1r = (float)b[4 * slot + 1], s1i = (float)b[4 * slot + 2],
                sv = (float)b[4 * slot + 3];
    const int per = MB >> a.lognb;
    const long base = (long)blk * a.blk_elems;
    typedef typename E2<float>::raw f2;
    typedef typename E2<f16>::raw h2;
    float nr = 0.f;
    for (long p = (long)blockIdx.x * per + (threadIdx.x >> a.lognb); p < a.rows; p += (long)a.gx * per) {
      const long e = base + (p << a.lognb) + rr;
      const f2 u1 = E2<float>::ldr(w1, e), u2 = E2<float>::ldr(w2, e);
      f2 r;
      r[0] = s2 * u2[0] + (s1r * u1[0] - s1i * u1[1]);
      r[1] = s2 * u2[1] + (s1r * u1[1] + s1i * u1[0]);
      ((f2 *)rout)[e] = r;
      h2 hv;
      hv[0] = (f16)(sv * u1[0]);
      hv[1] = (f16)(sv * u1[1]);
      ((h2 *)hout)[e] = hv;
      nr += r[0] * r[0] + r[1] * r[1];
    }
    acc[0] = (double)nr;
  }
  slot_sum<1>(acc, a.lognb, part, a.cap, blk << a.lognb, a.gx);
}

// One thread per (row k, slot): column j of H from g1 and nu = ||W1'||,
// column j+1 from the Arnoldi relation A v_{j+1} = (A^2 v_j - V_{j+1} H_j g1)
// / nu, and (row 0) the coefficients of R = (W2' - alpha W1') / nu, alpha =
// <W1', W2'> / nu^2, with W1' = W1 - V g1, W2' = W2 - V g2 as axpy2 left them
// (so R is the part of A v_{j+1} orthogonal to v_0 .. v_{j+1});
// H[j+2, j+1] = ||R|| follows from next2 (set_subdiag).
__global__ void k_sstep_columns(double *H, int ldh, int j, const double *g1, const double *g2, const double *sc,
                                double *b, const int *nact, int cap, double hs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = t / cap, r = t - k * cap;
  if (k > j + 1 || r >= *nact) return;
#define HH(q, c) H[2 * (((long)(c) * ldh + (q)) * cap + r)]
#define HI(q, c) H[2 * (((long)(c) * ldh + (q)) * cap + r) + 1]
  const double n1 = sc[r];
  const double nu = n1 > 0.0 ? sqrt(n1) : 0.0;
  // row k of column j: (g1, nu)
  const double hjr = k <= j ? g1[2 * ((long)k * cap + r)] : nu, hji = k <= j ? g1[2 * ((long)k * cap + r) + 1] : 0.0;
  HH(k, j) = hjr;
  HI(k, j) = hji;
  if (!(nu > 0.0)) {
    HH(k, j + 1) = 0.0;
    HI(k, j + 1) = 0.0;
    if (k == 0) b[4 * r] = b[4 * r + 1] = b[4 * r + 2] = b[4 * r + 3] = 0.0;
    return;
  }
  const double inv = 1.0 / nu;
  // c_k = sum_i H[k, i] g1_i over the columns i >= k - 1 where H[k, i] can be
  // nonzero; column j from the values above (another thread writes it)
  const double gjr = g1[2 * ((long)j * cap + r)], gji = g1[2 * ((long)j * cap + r) + 1];
  double cr = hjr * gjr - hji * gji, ci = hjr * gji + hji * gjr;
  for (int i = k > 0 ? k - 1 : 0; i < j; i++) {
    const double hr = HH(k, i), hi = HI(k, i);
    const double gr = g1[2 * ((long)i * cap + r)], gi = g1[2 * ((long)i * cap + r) + 1];
    cr += hr * gr - hi * gi;
    ci += hr * gi + hi * gr;
  }
  if (k <= j) {
    HH(k, j + 1) = (g2[2 * ((long)k * cap + r)] - cr) * inv;
    HI(k, j + 1) = (g2[2 * ((long)k * cap + r) + 1] - ci) * inv;
  } else {
    HH(k, j + 1) = (sc[cap + r] * inv - cr) * inv;
    HI(k, j + 1) = (sc[2 * cap + r] * inv - ci) * inv;
  }
  if (k == 0) {
    const double ar = sc[cap + r] * inv * inv, ai = sc[2 * cap + r] * inv * inv;
    b[4 * r] = inv;
    b[4 * r + 1] = -ar * inv;
    b[4 * r + 2] = -ai * inv;
    b[4 * r + 3] = hs * inv;
  }
#undef HH
#undef HI
}

__global__ void k_set_subdiag(double *H, int ldh, int j, const double *n2, const int *nact, int cap) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= cap || r >= *nact) return;
  const double n = n2[r];
  H[2 * (((long)j * ldh + j + 1) * cap + r)] = n > 0.0 ? sqrt(n) : 0.0;
  H[2 * (((long)j * ldh + j + 1) * cap + r) + 1] = 0.0;
}

// ===================================================================== //

static size_t esize(Prec p) { return p == PREC_DOUBLE ? 16 : p == PREC_SINGLE ? 8 : 4; }

MultiSpinor alloc_multi(const Geom &g, Prec p, int cap, int nb, bool norms) {
  if (nb < 1 || nb > 64 || (nb & (nb - 1))) throw std::runtime_error("alloc_multi: RHS block must be 1..64, a power of two");
  if (cap < 1 || cap > 256) throw std::runtime_error("alloc_multi: 1..256 right-hand sides");
  MultiSpinor m;
  m.prec = p;
  m.cap = cap;
  m.nb = nb;
  m.lognb = 0;
  while ((1 << m.lognb) < nb) m.lognb++;
  m.stride = g.stride;
  QQCD_HIP(hipMalloc(&m.data, (size_t)m.elems() * esize(p)));
  QQCD_HIP(hipMemset(m.data, 0, (size_t)m.elems() * esize(p)));
  if (p == PREC_HALF && norms) {
    const size_t nn = (size_t)m.stride * nb * m.nblk();
    QQCD_HIP(hipMalloc(&m.norm, nn * sizeof(float)));
    QQCD_HIP(hipMemset(m.norm, 0, nn * sizeof(float)));
  }
  return m;
}

void free_multi(MultiSpinor &m) {
  if (m.data) (void)hipFree(m.data);
  if (m.norm) (void)hipFree(m.norm);
  m.data = nullptr;
  m.norm = nullptr;
}

MBlas::MBlas(hipStream_t s, int cap_) : stream(s), cap(cap_) {}
MBlas::~MBlas() {
  if (part) (void)hipFree(part);
}

void MBlas::reserve(size_t n) {
  if (n <= part_cap) return;
  // only outside graph capture: the first call of each shape sizes the buffer
  if (part) {
    QQCD_HIP(hipStreamSynchronize(stream));
    (void)hipFree(part);
  }
  QQCD_HIP(hipMalloc(&part, n * sizeof(double)));
  part_cap = n;
}

static int gx_of(long rows, int nb, int nblk) {
  // about 2048 workgroups in all, at most one row group per lane
  const long groups = (rows * nb + MB - 1) / MB;
  return (int)std::max(1L, std::min(groups, (long)std::max(16, 2048 / nblk)));
}

// gx <= 2048 and at most max(2 MAXK + 2, 4 MAXK2) partial values per slot
// (MAXK complex dots and a norm, or MAXK2 complex dots with each of two
// fields): sized once, before any capture, so that no later pass moves the
// buffer a captured graph names
size_t MBlas::need() const { return (size_t)2048 * std::max(2 * MAXK + 2, 4 * MAXK2 + 4) * cap + 64; }

static MShape shape(const MultiSpinor &W, const int *nact) {
  MShape a;
  a.rows = 12 * W.stride;
  a.blk_elems = W.blk_elems();
  a.nact = nact;
  a.lognb = W.lognb;
  a.cap = W.cap;
  a.gx = gx_of(a.rows, W.nb, W.nblk());
  a.hs = std::sqrt(12.0 * (double)W.stride);
  a.hinv = 1.0 / a.hs;
  return a;
}

// a scaled FP16 basis field: PREC_HALF without per-site norms
static bool basis16(const MultiSpinor &V) { return V.prec == PREC_HALF && !V.norm; }

static void need_fp(const MultiSpinor &W, const char *who) {
  if (W.prec == PREC_HALF) throw std::runtime_error(std::string(who) + ": FP16 multi-RHS fields are hopping-only");
}

template <typename F> static void two_types(Prec a, Prec b, F &&f) {
  if (a == PREC_DOUBLE && b == PREC_DOUBLE) f((double *)0, (double *)0);
  else if (a == PREC_DOUBLE && b == PREC_SINGLE) f((double *)0, (float *)0);
  else if (a == PREC_SINGLE && b == PREC_DOUBLE) f((float *)0, (double *)0);
  else if (a == PREC_SINGLE && b == PREC_SINGLE) f((float *)0, (float *)0);
  else throw std::runtime_error("multi BLAS: FP64 / FP32 fields only");
}

// the Gram-Schmidt passes also read scaled FP16 basis vectors (a = PREC_HALF)
template <typename F> static void basis_types(Prec a, Prec b, F &&f) {
  if (a == PREC_HALF && b == PREC_SINGLE) f((f16 *)0, (float *)0);
  else if (a == PREC_HALF && b == PREC_DOUBLE) f((f16 *)0, (double *)0);
  else two_types(a, b, f);
}

#define T_OF(p) typename std::remove_pointer<decltype(p)>::type

// f(std::integral_constant<int, K>()) for a run-time K in 1..KK
template <int KK, typename F> static void with_k(int K, F &&f) {
  if constexpr (KK >= 1) {
    if (K == KK) f(std::integral_constant<int, KK>());
    else with_k<KK - 1>(K, f);
  }
}

// length of the chunk starting at q0 when nq vectors are split into the
// fewest chunks of at most MAXK, as even as possible
static int chunk_len(int nq, int q0, int maxk = MAXK) {
  const int n = (nq + maxk - 1) / maxk, base = nq / n, extra = nq % n;
  // chunks 0 .. extra-1 hold base + 1 vectors, the rest base
  const int big = extra * (base + 1);
  return q0 < big ? base + 1 : base;
}

template <bool SHARED>
static void dot_impl(MBlas &mb, int nq, const void *const *vp, Prec vprec, const MultiSpinor &W, const int *nact,
                     double *res, double *&part, const int *any, std::function<void(size_t)> reserve,
                     bool self = false) {
  const MShape a = shape(W, nact);
  const int nblk = W.nblk();
  reserve((size_t)a.gx * (2 * MAXK + 2) * W.cap);
  for (int q0 = 0, K; q0 < nq; q0 += K) {
    K = chunk_len(nq, q0);
    // the last chunk also forms <W, W> when asked (res entry nq)
    const bool sf = self && q0 + K == nq;
    MPtr pk{};
    for (int k = 0; k < K; k++) pk.p[k] = vp[q0 + k];
    basis_types(vprec, W.prec, [&](auto *tv, auto *tw) {
      typedef T_OF(tv) TV;
      typedef T_OF(tw) TW;
      dim3 grid(a.gx, nblk);
#define L_(KK)                                                                                                    \
  do {                                                                                                            \
    if (sf)                                                                                                       \
      hipLaunchKernelGGL((k_mdot<TV, TW, KK, SHARED, true>), grid, dim3(MB), 0, mb.stream, a, pk, W.data, part, any); \
    else                                                                                                          \
      hipLaunchKernelGGL((k_mdot<TV, TW, KK, SHARED, false>), grid, dim3(MB), 0, mb.stream, a, pk, W.data, part, any); \
  } while (0)
      with_k<MAXK>(K, [&](auto kk) { L_(decltype(kk)::value); });
#undef L_
    });
    QQCD_HIP(hipGetLastError());
    const int nv = 2 * K + (sf ? 2 : 0);
    hipLaunchKernelGGL(k_mfinal, dim3(W.cap, nv), dim3(64), 0, mb.stream, nv, W.cap, a.gx, nact,
                       (const double *)part, res, q0, 1);
    QQCD_HIP(hipGetLastError());
  }
}

void MBlas::dot(int nq, const MultiSpinor *V, const MultiSpinor &W, const int *nact, double *res, const int *any,
                bool self) {
  need_fp(W, "MBlas::dot");
  std::vector<const void *> vp(nq);
  for (int q = 0; q < nq; q++) {
    if (V[q].prec != V[0].prec || V[q].nb != W.nb || V[q].cap != W.cap || V[q].stride != W.stride)
      throw std::runtime_error("MBlas::dot: layouts differ");
    if (V[q].prec == PREC_HALF && !basis16(V[q])) throw std::runtime_error("MBlas::dot: FP16 fields with norms");
    vp[q] = V[q].data;
  }
  if (nq == 0) return;
  dot_impl<false>(*this, nq, vp.data(), V[0].prec, W, nact, res, part, any, [this](size_t n) { reserve(n); }, self);
}

void MBlas::dot_shared(int nq, const Spinor *D, const MultiSpinor &W, const int *nact, double *res) {
  need_fp(W, "MBlas::dot_shared");
  std::vector<const void *> vp(nq);
  for (int q = 0; q < nq; q++) {
    if (D[q].prec != D[0].prec || D[q].stride != W.stride) throw std::runtime_error("MBlas::dot_shared: layouts differ");
    vp[q] = D[q].data;
  }
  if (nq == 0) return;
  dot_impl<true>(*this, nq, vp.data(), D[0].prec, W, nact, res, part, nullptr, [this](size_t n) { reserve(n); });
}

template <bool SHARED>
static void axpy_impl(MBlas &mb, int nq, const void *const *vp, Prec vprec, const double *coef, int q0g,
                      double sign, MultiSpinor &W, const int *nact, double *n2, const int *flag, const int *any,
                      double *part) {
  const MShape a = shape(W, nact);
  const int nblk = W.nblk();
  for (int q0 = 0, K; q0 < nq; q0 += K) {
    K = chunk_len(nq, q0);
    const bool last = q0 + K >= nq;
    MPtr pk{};
    for (int k = 0; k < K; k++) pk.p[k] = vp[q0 + k];
    basis_types(vprec, W.prec, [&](auto *tv, auto *tw) {
      typedef T_OF(tv) TV;
      typedef T_OF(tw) TW;
      dim3 grid(a.gx, nblk);
#define L_(KK, NR) \
  hipLaunchKernelGGL((k_maxpy<TV, TW, KK, SHARED, NR>), grid, dim3(MB), 0, mb.stream, a, pk, coef, q0g + q0, sign, W.data, part, flag, any)
#define LK_(NR) with_k<MAXK>(K, [&](auto kk) { L_(decltype(kk)::value, NR); });
      if (n2 && last) {
        LK_(true)
      } else {
        LK_(false)
      }
#undef LK_
#undef L_
    });
    QQCD_HIP(hipGetLastError());
  }
  if (n2) {
    hipLaunchKernelGGL(k_mfinal, dim3(W.cap, 1), dim3(64), 0, mb.stream, 1, W.cap, a.gx, nact,
                       (const double *)part, n2, 0, 0);
    QQCD_HIP(hipGetLastError());
  }
}

void MBlas::axpy(int nq, const MultiSpinor *V, const double *coef, int q0, double sign, MultiSpinor &W,
                 const int *nact, double *n2, const int *flag, const int *any) {
  need_fp(W, "MBlas::axpy");
  std::vector<const void *> vp(nq);
  for (int q = 0; q < nq; q++) {
    if (V[q].prec != V[0].prec || V[q].nb != W.nb || V[q].cap != W.cap || V[q].stride != W.stride)
      throw std::runtime_error("MBlas::axpy: layouts differ");
    if (V[q].prec == PREC_HALF && !basis16(V[q])) throw std::runtime_error("MBlas::axpy: FP16 fields with norms");
    vp[q] = V[q].data;
  }
  if (nq == 0) {
    if (n2) norm2(W, nact, n2);
    return;
  }
  reserve((size_t)shape(W, nact).gx * 16 * W.cap);
  axpy_impl<false>(*this, nq, vp.data(), V[0].prec, coef, q0, sign, W, nact, n2, flag, any, part);
}

void MBlas::axpy_shared(int nq, const Spinor *D, const double *coef, int q0, double sign, MultiSpinor &W,
                        const int *nact, double *n2) {
  need_fp(W, "MBlas::axpy_shared");
  std::vector<const void *> vp(nq);
  for (int q = 0; q < nq; q++) {
    if (D[q].prec != D[0].prec || D[q].stride != W.stride)
      throw std::runtime_error("MBlas::axpy_shared: layouts differ");
    vp[q] = D[q].data;
  }
  if (nq == 0) {
    if (n2) norm2(W, nact, n2);
    return;
  }
  reserve((size_t)shape(W, nact).gx * 16 * W.cap);
  axpy_impl<true>(*this, nq, vp.data(), D[0].prec, coef, q0, sign, W, nact, n2, nullptr, nullptr, part);
}

void MBlas::norm2(const MultiSpinor &W, const int *nact, double *n2) {
  need_fp(W, "MBlas::norm2");
  const MShape a = shape(W, nact);
  reserve((size_t)a.gx * 16 * W.cap);
  if (W.prec == PREC_DOUBLE)
    hipLaunchKernelGGL(k_mnorm<double>, dim3(a.gx, W.nblk()), dim3(MB), 0, stream, a, W.data, part);
  else
    hipLaunchKernelGGL(k_mnorm<float>, dim3(a.gx, W.nblk()), dim3(MB), 0, stream, a, W.data, part);
  QQCD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mfinal, dim3(W.cap, 1), dim3(64), 0, stream, 1, W.cap, a.gx, nact, (const double *)part,
                     n2, 0, 0);
  QQCD_HIP(hipGetLastError());
}

void MBlas::axpby(double ca, const MultiSpinor &X, double cb, MultiSpinor &Y, const int *nact, double *n2) {
  need_fp(Y, "MBlas::axpby");
  if (X.nb != Y.nb || X.cap != Y.cap || X.stride != Y.stride) throw std::runtime_error("MBlas::axpby: layouts differ");
  const MShape a = shape(Y, nact);
  reserve((size_t)a.gx * 16 * Y.cap);
  two_types(X.prec, Y.prec, [&](auto *tx, auto *ty) {
    typedef T_OF(tx) TX;
    typedef T_OF(ty) TY;
    dim3 grid(a.gx, Y.nblk());
    if (n2)
      hipLaunchKernelGGL((k_maxpby<TX, TY, true>), grid, dim3(MB), 0, stream, a, ca, X.data, cb, Y.data, part);
    else
      hipLaunchKernelGGL((k_maxpby<TX, TY, false>), grid, dim3(MB), 0, stream, a, ca, X.data, cb, Y.data, part);
  });
  QQCD_HIP(hipGetLastError());
  if (n2) {
    hipLaunchKernelGGL(k_mfinal, dim3(Y.cap, 1), dim3(64), 0, stream, 1, Y.cap, a.gx, nact, (const double *)part,
                       n2, 0, 0);
    QQCD_HIP(hipGetLastError());
  }
}

void MBlas::scale_inv_norm(const MultiSpinor &X, MultiSpinor &W, const double *n2, const int *nact) {
  need_fp(W, "MBlas::scale_inv_norm");
  const MShape a = shape(W, nact);
  two_types(X.prec, W.prec, [&](auto *tx, auto *tw) {
    typedef T_OF(tx) TX;
    typedef T_OF(tw) TW;
    hipLaunchKernelGGL((k_mscale<TX, TW>), dim3(a.gx, W.nblk()), dim3(MB), 0, stream, a, X.data, W.data, n2);
  });
  QQCD_HIP(hipGetLastError());
}

void MBlas::scale_inv_norm2(const MultiSpinor &X, MultiSpinor &Q, MultiSpinor &H, const double *n2,
                            const int *nact) {
  need_fp(Q, "MBlas::scale_inv_norm2");
  if (!basis16(H)) throw std::runtime_error("MBlas::scale_inv_norm2: H must be an FP16 basis field");
  if (X.nb != Q.nb || X.cap != Q.cap || X.stride != Q.stride || H.nb != Q.nb || H.cap != Q.cap ||
      H.stride != Q.stride)
    throw std::runtime_error("MBlas::scale_inv_norm2: layouts differ");
  const MShape a = shape(Q, nact);
  two_types(X.prec, Q.prec, [&](auto *tx, auto *tq) {
    typedef T_OF(tx) TX;
    typedef T_OF(tq) TQ;
    hipLaunchKernelGGL((k_mscale2<TX, TQ>), dim3(a.gx, Q.nblk()), dim3(MB), 0, stream, a, X.data, Q.data, H.data,
                       n2);
  });
  QQCD_HIP(hipGetLastError());
}

// ---- two-step Arnoldi block ----
static void same_layout2(const MultiSpinor &A, const MultiSpinor &B, const char *who) {
  if (A.nb != B.nb || A.cap != B.cap || A.stride != B.stride) throw std::runtime_error(std::string(who) + ": layouts differ");
}
static void basis_fields(int nq, const MultiSpinor *V, const MultiSpinor &W, std::vector<const void *> &vp,
                         const char *who) {
  vp.resize(nq);
  for (int q = 0; q < nq; q++) {
    if (!basis16(V[q])) throw std::runtime_error(std::string(who) + ": V must be FP16 basis fields");
    same_layout2(V[q], W, who);
    vp[q] = V[q].data;
  }
}

void MBlas::dot2(int nq, const MultiSpinor *V, const MultiSpinor &W1, const MultiSpinor &W2, const int *nact,
                 double *g1, double *g2, double *self) {
  if (W1.prec != PREC_SINGLE || W2.prec != PREC_SINGLE) throw std::runtime_error("MBlas::dot2: FP32 W1, W2");
  same_layout2(W1, W2, "MBlas::dot2");
  std::vector<const void *> vp;
  basis_fields(nq, V, W1, vp, "MBlas::dot2");
  const MShape a = shape(W1, nact);
  reserve((size_t)a.gx * (4 * MAXK2 + 4) * W1.cap);
  for (int q0 = 0, K; q0 < nq; q0 += K) {
    K = chunk_len(nq, q0, MAXK2);
    const bool with_self = self && q0 == 0;
    MPtr pk{};
    for (int k = 0; k < K; k++) pk.p[k] = vp[q0 + k];
    dim3 grid(a.gx, W1.nblk());
    with_k<MAXK2>(K, [&](auto kk) {
      constexpr int KK = decltype(kk)::value;
      if (with_self)
        hipLaunchKernelGGL((k_mdot2<KK, true>), grid, dim3(MB), 0, stream, a, pk, W1.data, W2.data, part);
      else
        hipLaunchKernelGGL((k_mdot2<KK, false>), grid, dim3(MB), 0, stream, a, pk, W1.data, W2.data, part);
    });
    QQCD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_mfinal, dim3(W1.cap, 2 * K), dim3(64), 0, stream, 2 * K, W1.cap, a.gx, nact,
                       (const double *)part, g1, q0, 1);
    hipLaunchKernelGGL(k_mfinal, dim3(W1.cap, 2 * K), dim3(64), 0, stream, 2 * K, W1.cap, a.gx, nact,
                       (const double *)part + (size_t)2 * K * W1.cap * a.gx, g2, q0, 1);
    if (with_self)
      hipLaunchKernelGGL(k_mfinal, dim3(W1.cap, 4), dim3(64), 0, stream, 4, W1.cap, a.gx, nact,
                         (const double *)part + (size_t)4 * K * W1.cap * a.gx, self, 0, 0);
    QQCD_HIP(hipGetLastError());
  }
}

void MBlas::axpy2_next(int nq, const MultiSpinor *V, const double *g1, const double *g2, MultiSpinor &W1,
                       MultiSpinor &W2, const double *b, MultiSpinor &R, MultiSpinor &H, const int *nact, double *nr) {
  if (W1.prec != PREC_SINGLE || W2.prec != PREC_SINGLE || R.prec != PREC_SINGLE)
    throw std::runtime_error("MBlas::axpy2_next: FP32 W1, W2, R");
  if (W1.data == W2.data) throw std::runtime_error("MBlas::axpy2_next: W1 and W2 alias");
  if (!basis16(H)) throw std::runtime_error("MBlas::axpy2_next: H must be an FP16 basis field");
  same_layout2(W1, W2, "MBlas::axpy2_next");
  same_layout2(W1, R, "MBlas::axpy2_next");
  same_layout2(W1, H, "MBlas::axpy2_next");
  if (R.data == W1.data || R.data == W2.data) throw std::runtime_error("MBlas::axpy2_next: R aliases W1 or W2");
  if (nq < 1) throw std::runtime_error("MBlas::axpy2_next: at least one basis field");
  std::vector<const void *> vp;
  basis_fields(nq, V, W1, vp, "MBlas::axpy2_next");
  for (int q = 0; q < nq; q++)
    if (vp[q] == H.data) throw std::runtime_error("MBlas::axpy2_next: H is one of the basis fields read");
  const MShape a = shape(W1, nact);
  reserve((size_t)a.gx * W1.cap);
  for (int q0 = 0, K; q0 < nq; q0 += K) {
    K = chunk_len(nq, q0, MAXK2);
    const bool last = q0 + K >= nq;
    MPtr pk{};
    for (int k = 0; k < K; k++) pk.p[k] = vp[q0 + k];
    dim3 grid(a.gx, W1.nblk());
    with_k<MAXK2>(K, [&](auto kk) {
      constexpr int KK = decltype(kk)::value;
      if (last)
        hipLaunchKernelGGL((k_maxpy2n<KK>), grid, dim3(MB), 0, stream, a, pk, g1, g2, q0, W1.data, W2.data, R.data,
                           H.data, b, part);
      else
        hipLaunchKernelGGL((k_maxpy2<KK, false>), grid, dim3(MB), 0, stream, a, pk, g1, g2, q0, W1.data, W2.data,
                           part);
    });
    QQCD_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_mfinal, dim3(W1.cap, 1), dim3(64), 0, stream, 1, W1.cap, a.gx, nact, (const double *)part, nr,
                     0, 0);
  QQCD_HIP(hipGetLastError());
}

void pythag_scalars(int nq, const double *g1, const double *g2, const double *self, double *sc, const int *nact,
                    int cap, hipStream_t s) {
  hipLaunchKernelGGL(k_pythag_scalars, dim3((cap + 63) / 64), dim3(64), 0, s, nq, g1, g2, self, sc, nact, cap);
  QQCD_HIP(hipGetLastError());
}

void MBlas::axpy2(int nq, const MultiSpinor *V, const double *g1, const double *g2, MultiSpinor &W1,
                  MultiSpinor &W2, const int *nact, double *sc) {
  if (W1.prec != PREC_SINGLE || W2.prec != PREC_SINGLE) throw std::runtime_error("MBlas::axpy2: FP32 W1, W2");
  if (W1.data == W2.data) throw std::runtime_error("MBlas::axpy2: W1 and W2 alias");
  same_layout2(W1, W2, "MBlas::axpy2");
  if (nq < 1) throw std::runtime_error("MBlas::axpy2: at least one basis field");
  std::vector<const void *> vp;
  basis_fields(nq, V, W1, vp, "MBlas::axpy2");
  const MShape a = shape(W1, nact);
  reserve((size_t)a.gx * 4 * W1.cap);
  for (int q0 = 0, K; q0 < nq; q0 += K) {
    K = chunk_len(nq, q0, MAXK2);
    const bool last = q0 + K >= nq;
    MPtr pk{};
    for (int k = 0; k < K; k++) pk.p[k] = vp[q0 + k];
    dim3 grid(a.gx, W1.nblk());
    with_k<MAXK2>(K, [&](auto kk) {
      constexpr int KK = decltype(kk)::value;
      if (last)
        hipLaunchKernelGGL((k_maxpy2<KK, true>), grid, dim3(MB), 0, stream, a, pk, g1, g2, q0, W1.data, W2.data,
                           part);
      else
        hipLaunchKernelGGL((k_maxpy2<KK, false>), grid, dim3(MB), 0, stream, a, pk, g1, g2, q0, W1.data, W2.data,
                           part);
    });
    QQCD_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_mfinal, dim3(W1.cap, 4), dim3(64), 0, stream, 4, W1.cap, a.gx, nact, (const double *)part, sc,
                     0, 0);
  QQCD_HIP(hipGetLastError());
}

void MBlas::next2(const MultiSpinor &W1, const MultiSpinor &W2, MultiSpinor &R, MultiSpinor &H, const double *b,
                  const int *nact, double *nr) {
  if (W1.prec != PREC_SINGLE || W2.prec != PREC_SINGLE || R.prec != PREC_SINGLE)
    throw std::runtime_error("MBlas::next2: FP32 W1, W2, R");
  if (!basis16(H)) throw std::runtime_error("MBlas::next2: H must be an FP16 basis field");
  same_layout2(W1, W2, "MBlas::next2");
  same_layout2(W1, R, "MBlas::next2");
  same_layout2(W1, H, "MBlas::next2");
  if (R.data == W1.data || R.data == W2.data) throw std::runtime_error("MBlas::next2: R aliases W1 or W2");
  const MShape a = shape(W1, nact);
  reserve((size_t)a.gx * W1.cap);
  hipLaunchKernelGGL(k_next2, dim3(a.gx, W1.nblk()), dim3(MB), 0, stream, a, W1.data, W2.data, R.data, H.data, b,
                     part);
  QQCD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mfinal, dim3(W1.cap, 1), dim3(64), 0, stream, 1, W1.cap, a.gx, nact, (const double *)part, nr,
                     0, 0);
  QQCD_HIP(hipGetLastError());
}

void MBlas::swap(MultiSpinor &W, int sa, int sb) {
  need_fp(W, "MBlas::swap");
  if (sa == sb) return;
  if (sa < 0 || sb < 0 || sa >= W.cap || sb >= W.cap) throw std::runtime_error("MBlas::swap: slot out of range");
  const long rows = 12 * W.stride;
  const dim3 grid((unsigned)((rows + MB - 1) / MB));
  if (W.prec == PREC_DOUBLE)
    hipLaunchKernelGGL(k_mswap<double>, grid, dim3(MB), 0, stream, rows, W.lognb, W.blk_elems(), W.data, sa, sb);
  else
    hipLaunchKernelGGL(k_mswap<float>, grid, dim3(MB), 0, stream, rows, W.lognb, W.blk_elems(), W.data, sa, sb);
  QQCD_HIP(hipGetLastError());
}

void MBlas::put(const Spinor &s, MultiSpinor &W, int slot) {
  need_fp(W, "MBlas::put");
  if (slot < 0 || slot >= W.cap || s.stride != W.stride) throw std::runtime_error("MBlas::put: bad slot or layout");
  const long rows = 12 * W.stride;
  two_types(s.prec, W.prec, [&](auto *ts, auto *tw) {
    typedef T_OF(ts) TS;
    typedef T_OF(tw) TW;
    hipLaunchKernelGGL((k_mput<TS, TW>), dim3((unsigned)((rows + MB - 1) / MB)), dim3(MB), 0, stream, rows, W.lognb,
                       W.blk_elems(), s.data, W.data, slot);
  });
  QQCD_HIP(hipGetLastError());
}

void MBlas::get(const MultiSpinor &W, int slot, Spinor &s) {
  need_fp(W, "MBlas::get");
  if (slot < 0 || slot >= W.cap || s.stride != W.stride) throw std::runtime_error("MBlas::get: bad slot or layout");
  const long rows = 12 * W.stride;
  two_types(W.prec, s.prec, [&](auto *tw, auto *ts) {
    typedef T_OF(tw) TW;
    typedef T_OF(ts) TS;
    hipLaunchKernelGGL((k_mget<TW, TS>), dim3((unsigned)((rows + MB - 1) / MB)), dim3(MB), 0, stream, rows, W.lognb,
                       W.blk_elems(), W.data, s.data, slot);
  });
  QQCD_HIP(hipGetLastError());
}

void MBlas::zero(MultiSpinor &W) {
  QQCD_HIP(hipMemsetAsync(W.data, 0, (size_t)W.elems() * esize(W.prec), stream));
}

void arnoldi_flags(const double *n2, const double *w2, int *flag, int *any, const int *nact, int cap, int force,
                   hipStream_t s) {
  if (cap > MB) throw std::runtime_error("arnoldi_flags: at most 256 slots");
  hipLaunchKernelGGL(k_arnoldi_flags, dim3(1), dim3(MB), 0, s, n2, w2, flag, any, nact, cap, force);
  QQCD_HIP(hipGetLastError());
}

void clear_flags(int *flag, int *any, int cap, hipStream_t s) {
  if (cap > MB) throw std::runtime_error("clear_flags: at most 256 slots");
  hipLaunchKernelGGL(k_arnoldi_flags, dim3(1), dim3(MB), 0, s, nullptr, nullptr, flag, any, nullptr, cap, 0);
  QQCD_HIP(hipGetLastError());
}

void sstep_columns(double *H, int ldh, int j, const double *g1, const double *g2, const double *sc, double *b,
                   const int *nact, int cap, double hs, hipStream_t s) {
  const int nt = (j + 2) * cap;
  hipLaunchKernelGGL(k_sstep_columns, dim3((nt + 63) / 64), dim3(64), 0, s, H, ldh, j, g1, g2, sc, b, nact, cap, hs);
  QQCD_HIP(hipGetLastError());
}

void set_subdiag(double *H, int ldh, int j, const double *n2, const int *nact, int cap, hipStream_t s) {
  hipLaunchKernelGGL(k_set_subdiag, dim3((cap + 63) / 64), dim3(64), 0, s, H, ldh, j, n2, nact, cap);
  QQCD_HIP(hipGetLastError());
}

void arnoldi_column(double *H, int ldh, int j, const double *h1, const double *h2, const int *flag,
                    const double *n2a, const double *n2b, double *nf, const int *nact, int cap, hipStream_t s) {
  const int nt = (j + 2) * cap;
  hipLaunchKernelGGL(k_arnoldi_column, dim3((nt + MB - 1) / MB), dim3(MB), 0, s, H, ldh, j, h1, h2, flag, n2a, n2b,
                     nf, nact, cap);
  QQCD_HIP(hipGetLastError());
}

}  // namespace qqcd
