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
// pass reads w and up to 8 basis vectors, one multi-axpy pass updates w and
// returns its norm.
#include <stdexcept>
#include <type_traits>
#include <vector>
#include "qqcd_blas.h"

namespace qqcd {

static constexpr int RB = 256;   // threads per reduction workgroup
static constexpr int RG = 1024;  // reduction workgroups (grid-stride)

template <typename T> struct C2;
template <> struct C2<double> {
  typedef double2 V;
  static __device__ __forceinline__ double2 ld(const void *p, long i) { return ((const double2 *)p)[i]; }
  static __device__ __forceinline__ void st(void *p, long i, double re, double im) {
    ((double2 *)p)[i] = make_double2(re, im);
  }
};
template <> struct C2<float> {
  typedef float2 V;
  static __device__ __forceinline__ double2 ld(const void *p, long i) {
    const float2 v = ((const float2 *)p)[i];
    return make_double2(v.x, v.y);
  }
  static __device__ __forceinline__ void st(void *p, long i, double re, double im) {
    ((float2 *)p)[i] = make_float2((float)re, (float)im);
  }
};

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

// ---- multi-dot: out[k] = sum conj(v_k) w, k < K (<= 8), one pass ----
struct PtrPack {
  const void *p[8];
};

template <typename TV, typename TW, int K>
__global__ void __launch_bounds__(RB) k_mdot(long n, PtrPack v, const void *w, double *part) {
  double acc[2 * K];
#pragma unroll
  for (int k = 0; k < 2 * K; k++) acc[k] = 0;
  for (long i = blockIdx.x * (long)RB + threadIdx.x; i < n; i += (long)gridDim.x * RB) {
    const double2 b = C2<TW>::ld(w, i);
#pragma unroll
    for (int k = 0; k < K; k++) {
      const double2 a = C2<TV>::ld(v.p[k], i);
      acc[2 * k] += a.x * b.x + a.y * b.y;
      acc[2 * k + 1] += a.x * b.y - a.y * b.x;
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

// ---- w = w - sum_k h_k v_k, returns ||w||^2 (optional) ----
struct CoefPack {
  double re[8], im[8];
};

template <typename TV, typename TW, int K, bool NORM>
__global__ void __launch_bounds__(RB) k_maxpy(long n, PtrPack v, CoefPack h, void *w, double *part) {
  double acc[1] = {0};
  for (long i = blockIdx.x * (long)RB + threadIdx.x; i < n; i += (long)gridDim.x * RB) {
    double2 b = C2<TW>::ld(w, i);
#pragma unroll
    for (int k = 0; k < K; k++) {
      const double2 a = C2<TV>::ld(v.p[k], i);
      b.x -= h.re[k] * a.x - h.im[k] * a.y;
      b.y -= h.re[k] * a.y + h.im[k] * a.x;
    }
    C2<TW>::st(w, i, b.x, b.y);
    if (NORM) acc[0] += b.x * b.x + b.y * b.y;
  }
  if (NORM) block_sum<1>(acc, part + blockIdx.x);
}

// ---- y = a x + b y (+ norm of y) ----
template <typename TX, typename TY, bool NORM>
__global__ void __launch_bounds__(RB) k_axpby(long n, double ar, double ai, const void *x, double br, double bi,
                                             void *y, double *part) {
  double acc[1] = {0};
  for (long i = blockIdx.x * (long)RB + threadIdx.x; i < n; i += (long)gridDim.x * RB) {
    const double2 a = C2<TX>::ld(x, i);
    double2 b = make_double2(0, 0);
    if (br != 0.0 || bi != 0.0) b = C2<TY>::ld(y, i);
    const double re = ar * a.x - ai * a.y + br * b.x - bi * b.y;
    const double im = ar * a.y + ai * a.x + br * b.y + bi * b.x;
    C2<TY>::st(y, i, re, im);
    if (NORM) acc[0] += re * re + im * im;
  }
  if (NORM) block_sum<1>(acc, part + blockIdx.x);
}

// ---- Y_j = sum_i V_i Z(i,j) (+ add to Y if accumulate) -----------------
// Z is column-major m x k complex on the device.  Each lane walks its site
// over the m inputs once per chunk of up to 8 outputs.
struct OutPack {
  void *p[8];
};
template <typename TV, typename TY, int KO>
__global__ void __launch_bounds__(RB) k_combine(long n, int m, const void *const *V, const double *Z, int ldz,
                                               int j0, OutPack Y, int accumulate) {
  for (long i = blockIdx.x * (long)RB + threadIdx.x; i < n; i += (long)gridDim.x * RB) {
    double sr[KO], si[KO];
#pragma unroll
    for (int q = 0; q < KO; q++) {
      sr[q] = 0;
      si[q] = 0;
    }
    for (int r = 0; r < m; r++) {
      const double2 a = C2<TV>::ld(V[r], i);
#pragma unroll
      for (int q = 0; q < KO; q++) {
        const double zr = Z[2 * (r + (long)ldz * (j0 + q))], zi = Z[2 * (r + (long)ldz * (j0 + q)) + 1];
        sr[q] += a.x * zr - a.y * zi;
        si[q] += a.x * zi + a.y * zr;
      }
    }
#pragma unroll
    for (int q = 0; q < KO; q++) {
      if (accumulate) {
        const double2 b = C2<TY>::ld(Y.p[q], i);
        C2<TY>::st(Y.p[q], i, b.x + sr[q], b.y + si[q]);
      } else {
        C2<TY>::st(Y.p[q], i, sr[q], si[q]);
      }
    }
  }
}

// ===================================================================== //
// host side

Blas::Blas(hipStream_t s) : stream(s) {
  QQCD_HIP(hipMalloc(&part, sizeof(double) * RG * 2 * kmax));
  QQCD_HIP(hipMalloc(&res, sizeof(double) * 2 * kmax));
  QQCD_HIP(hipHostMalloc(&hres, sizeof(double) * 2 * kmax, hipHostMallocDefault));
  QQCD_HIP(hipMalloc(&zbuf, sizeof(double) * 2 * zcap));
  QQCD_HIP(hipMalloc(&vptr, sizeof(void *) * 128));
}

Blas::~Blas() {
  hipFree(part);
  hipFree(res);
  hipHostFree(hres);
  hipFree(zbuf);
  hipFree(vptr);
}

void Blas::fetch(int nq) {
  hipLaunchKernelGGL(k_final, dim3(nq), dim3(RB), 0, stream, RG, nq, part, res);
  QQCD_HIP(hipGetLastError());
  QQCD_HIP(hipMemcpyAsync(hres, res, sizeof(double) * nq, hipMemcpyDeviceToHost, stream));
  QQCD_HIP(hipStreamSynchronize(stream));
}

// call f(TA*, TB*) with the element types of a precision pair
template <typename F> static void tsw(Prec pa, Prec pb, F &&f) {
  if (pa == PREC_DOUBLE && pb == PREC_DOUBLE) f((double *)nullptr, (double *)nullptr);
  else if (pa == PREC_SINGLE && pb == PREC_DOUBLE) f((float *)nullptr, (double *)nullptr);
  else if (pa == PREC_DOUBLE && pb == PREC_SINGLE) f((double *)nullptr, (float *)nullptr);
  else if (pa == PREC_SINGLE && pb == PREC_SINGLE) f((float *)nullptr, (float *)nullptr);
  else throw std::runtime_error("blas: unsupported precision pair");
}
#define TYPES_OF(ta, tb)                                   \
  typedef typename std::remove_pointer<decltype(ta)>::type TA; \
  typedef typename std::remove_pointer<decltype(tb)>::type TB;

template <int K, typename TA, typename TB>
static void launch_mdot(hipStream_t s, long n, const PtrPack &pp, const void *w, double *part) {
  hipLaunchKernelGGL((k_mdot<TA, TB, K>), dim3(RG), dim3(RB), 0, s, n, pp, w, part);
}

void Blas::multi_dot(int k, const Vec *v, const Vec &w, std::complex<double> *out) {
  if (k > kmax) throw std::runtime_error("multi_dot: too many vectors");
  // every chunk of <= 8 vectors writes its partials into its own slab of
  // part (RG rows of 2*kk); one final pass and one host synchronisation
  long off = 0;
  std::vector<long> offs;
  std::vector<int> kks;
  for (int k0 = 0; k0 < k; k0 += 8) {
    const int kk = std::min(8, k - k0);
    PtrPack pp{};
    for (int q = 0; q < kk; q++) {
      if (v[k0 + q].n != w.n) throw std::runtime_error("multi_dot: length mismatch");
      pp.p[q] = v[k0 + q].data;
    }
    double *pt = part + off;
    tsw(v[k0].prec, w.prec, [&](auto ta, auto tb) {
      TYPES_OF(ta, tb)
      switch (kk) {
        case 1: launch_mdot<1, TA, TB>(stream, w.n, pp, w.data, pt); break;
        case 2: launch_mdot<2, TA, TB>(stream, w.n, pp, w.data, pt); break;
        case 3: launch_mdot<3, TA, TB>(stream, w.n, pp, w.data, pt); break;
        case 4: launch_mdot<4, TA, TB>(stream, w.n, pp, w.data, pt); break;
        case 5: launch_mdot<5, TA, TB>(stream, w.n, pp, w.data, pt); break;
        case 6: launch_mdot<6, TA, TB>(stream, w.n, pp, w.data, pt); break;
        case 7: launch_mdot<7, TA, TB>(stream, w.n, pp, w.data, pt); break;
        default: launch_mdot<8, TA, TB>(stream, w.n, pp, w.data, pt); break;
      }
    });
    QQCD_HIP(hipGetLastError());
    offs.push_back(off);
    kks.push_back(kk);
    off += (long)RG * 2 * kk;
  }
  for (size_t c = 0; c < offs.size(); c++)
    hipLaunchKernelGGL(k_final, dim3(2 * kks[c]), dim3(RB), 0, stream, RG, 2 * kks[c], part + offs[c],
                       res + 16 * c);
  QQCD_HIP(hipGetLastError());
  QQCD_HIP(hipMemcpyAsync(hres, res, sizeof(double) * 16 * offs.size(), hipMemcpyDeviceToHost, stream));
  QQCD_HIP(hipStreamSynchronize(stream));
  for (size_t c = 0; c < offs.size(); c++)
    for (int q = 0; q < kks[c]; q++)
      out[8 * c + q] = std::complex<double>(hres[16 * c + 2 * q], hres[16 * c + 2 * q + 1]);
}

std::complex<double> Blas::dot(const Vec &a, const Vec &b) {
  std::complex<double> r;
  multi_dot(1, &a, b, &r);
  return r;
}

double Blas::norm2(const Vec &a) { return dot(a, a).real(); }

template <int K, bool NORM, typename TA, typename TB>
static void launch_maxpy(hipStream_t s, long n, const PtrPack &pp, const CoefPack &cp, void *w, double *part) {
  hipLaunchKernelGGL((k_maxpy<TA, TB, K, NORM>), dim3(RG), dim3(RB), 0, s, n, pp, cp, w, part);
}

double Blas::multi_axpy(int k, const std::complex<double> *h, const Vec *v, Vec &w, bool want_norm) {
  double nrm = 0;
  for (int k0 = 0; k0 < k; k0 += 8) {
    const int kk = std::min(8, k - k0);
    const bool last = k0 + 8 >= k;
    PtrPack pp{};
    CoefPack cp{};
    for (int q = 0; q < kk; q++) {
      pp.p[q] = v[k0 + q].data;
      cp.re[q] = h[k0 + q].real();
      cp.im[q] = h[k0 + q].imag();
    }
    const bool nm = want_norm && last;
    tsw(v[k0].prec, w.prec, [&](auto ta, auto tb) {
      TYPES_OF(ta, tb)
#define MX(KK)                                                                   \
  if (nm) launch_maxpy<KK, true, TA, TB>(stream, w.n, pp, cp, w.data, part);    \
  else launch_maxpy<KK, false, TA, TB>(stream, w.n, pp, cp, w.data, part);
      switch (kk) {
        case 1: MX(1); break;
        case 2: MX(2); break;
        case 3: MX(3); break;
        case 4: MX(4); break;
        case 5: MX(5); break;
        case 6: MX(6); break;
        case 7: MX(7); break;
        default: MX(8); break;
      }
#undef MX
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

double Blas::axpby(std::complex<double> a, const Vec &x, std::complex<double> b, Vec &y, bool want_norm) {
  if (x.n != y.n) throw std::runtime_error("axpby: length mismatch");
  tsw(x.prec, y.prec, [&](auto ta, auto tb) {
    TYPES_OF(ta, tb)
    if (want_norm)
      hipLaunchKernelGGL((k_axpby<TA, TB, true>), dim3(RG), dim3(RB), 0, stream, x.n, a.real(), a.imag(), x.data,
                         b.real(), b.imag(), y.data, part);
    else
      hipLaunchKernelGGL((k_axpby<TA, TB, false>), dim3(RG), dim3(RB), 0, stream, x.n, a.real(), a.imag(), x.data,
                         b.real(), b.imag(), y.data, part);
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
  } else {
    axpby(1.0, x, 0.0, y, false);
  }
}

void Blas::zero(Vec &x) { QQCD_HIP(hipMemsetAsync(x.data, 0, x.bytes(), stream)); }

template <int KO, typename TA, typename TB>
static void launch_combine(hipStream_t s, long n, int m, void **vp, const double *Z, int ldz, int j0,
                           const OutPack &op, int acc) {
  hipLaunchKernelGGL((k_combine<TA, TB, KO>), dim3(RG), dim3(RB), 0, s, n, m, (const void *const *)vp, Z, ldz, j0,
                     op, acc);
}

void Blas::combine(int m, const Vec *V, int k, const std::complex<double> *Z, int ldz, Vec *Y, bool accumulate) {
  if (m == 0 || k == 0) return;
  if ((long)ldz * k > zcap || m > 128) throw std::runtime_error("combine: coefficient block too large");
  std::vector<void *> hp(m);
  for (int r = 0; r < m; r++) hp[r] = V[r].data;
  QQCD_HIP(hipMemcpyAsync(vptr, hp.data(), sizeof(void *) * m, hipMemcpyHostToDevice, stream));
  QQCD_HIP(hipMemcpyAsync(zbuf, Z, sizeof(double) * 2 * ldz * k, hipMemcpyHostToDevice, stream));
  for (int j0 = 0; j0 < k; j0 += 8) {
    const int kk = std::min(8, k - j0);
    OutPack op{};
    for (int q = 0; q < kk; q++) op.p[q] = Y[j0 + q].data;
    tsw(V[0].prec, Y[0].prec, [&](auto ta, auto tb) {
      TYPES_OF(ta, tb)
      switch (kk) {
        case 1: launch_combine<1, TA, TB>(stream, Y[0].n, m, vptr, zbuf, ldz, j0, op, accumulate); break;
        case 2: launch_combine<2, TA, TB>(stream, Y[0].n, m, vptr, zbuf, ldz, j0, op, accumulate); break;
        case 3: launch_combine<3, TA, TB>(stream, Y[0].n, m, vptr, zbuf, ldz, j0, op, accumulate); break;
        case 4: launch_combine<4, TA, TB>(stream, Y[0].n, m, vptr, zbuf, ldz, j0, op, accumulate); break;
        case 5: launch_combine<5, TA, TB>(stream, Y[0].n, m, vptr, zbuf, ldz, j0, op, accumulate); break;
        case 6: launch_combine<6, TA, TB>(stream, Y[0].n, m, vptr, zbuf, ldz, j0, op, accumulate); break;
        case 7: launch_combine<7, TA, TB>(stream, Y[0].n, m, vptr, zbuf, ldz, j0, op, accumulate); break;
        default: launch_combine<8, TA, TB>(stream, Y[0].n, m, vptr, zbuf, ldz, j0, op, accumulate); break;
      }
    });
    QQCD_HIP(hipGetLastError());
  }
  // the host copies of the pointer and coefficient tables go out of scope
  QQCD_HIP(hipStreamSynchronize(stream));
}

}  // namespace qqcd
