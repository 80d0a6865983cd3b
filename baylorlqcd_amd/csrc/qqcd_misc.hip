// qqcd_misc.hip -- small device kernels: gamma matrices on reference-order
// slices, gauge unpacking, seeded random spinors.
#include <stdexcept>
#include "qqcd_context.h"

namespace qqcd {

typedef f16 h2 __attribute__((ext_vector_type(2)));

// gamma_mu (mu = 1..5) of qqcd-test/cfgsprops/quark/diracops.f90:61-71: every
// row has one non-zero entry, col[mu][row] with phase (re, im).
struct GammaRow {
  int col[5][4];
  float re[5][4], im[5][4];
};

static GammaRow gamma_rows() {
  GammaRow g{};
  // gamma_1 = antidiag(1,1,1,1)
  const int c1[4] = {3, 2, 1, 0};
  const float r1[4] = {1, 1, 1, 1}, i1[4] = {0, 0, 0, 0};
  // gamma_2: rows (0,0,0,-i) (0,0,i,0) (0,-i,0,0) (i,0,0,0)
  const int c2[4] = {3, 2, 1, 0};
  const float r2[4] = {0, 0, 0, 0}, i2[4] = {-1, 1, -1, 1};
  // gamma_3: rows (0,0,1,0) (0,0,0,-1) (1,0,0,0) (0,-1,0,0)
  const int c3[4] = {2, 3, 0, 1};
  const float r3[4] = {1, -1, 1, -1}, i3[4] = {0, 0, 0, 0};
  // gamma_4 = diag(1,1,-1,-1)
  const int c4[4] = {0, 1, 2, 3};
  const float r4[4] = {1, 1, -1, -1}, i4[4] = {0, 0, 0, 0};
  // gamma_5 = gamma_1 gamma_2 gamma_3 gamma_4: rows (0,0,-i,0) (0,0,0,-i) (i,0,0,0) (0,i,0,0)
  const int c5[4] = {2, 3, 0, 1};
  const float r5[4] = {0, 0, 0, 0}, i5[4] = {-1, -1, 1, 1};
  const int *cs[5] = {c1, c2, c3, c4, c5};
  const float *rs[5] = {r1, r2, r3, r4, r5}, *is[5] = {i1, i2, i3, i4, i5};
  for (int m = 0; m < 5; m++)
    for (int r = 0; r < 4; r++) {
      g.col[m][r] = cs[m][r];
      g.re[m][r] = rs[m][r];
      g.im[m][r] = is[m][r];
    }
  return g;
}

void gamma_dense(int mu, double re[4][4], double im[4][4]) {
  if (mu < 1 || mu > 5) throw std::runtime_error("gamma_dense: mu must be 1..5");
  static const GammaRow gr = gamma_rows();
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) re[r][c] = im[r][c] = 0.0;
  for (int r = 0; r < 4; r++) {
    re[r][gr.col[mu - 1][r]] = gr.re[mu - 1][r];
    im[r][gr.col[mu - 1][r]] = gr.im[mu - 1][r];
  }
}

struct SliceList {
  int n;
  int sl[16];  // slice index ieo + 2*ibleo (0-based)
};

// out(:, i, :, slice) = gamma_mu in(:, i, :, slice) for rows i < nvhalf of
// the listed (ieo, ibleo) slices of a (6, ntotal, 4, 2, 8) array.  in may
// equal out (every lane reads its 4 spins before writing).
__global__ void k_gamma_ref(long nvhalf, long ntotal, GammaRow gr, int mu, SliceList sl, const double *in,
                            double *out) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int s = sl.sl[blockIdx.y];
  if (i >= nvhalf) return;
  double v[4][6];
  for (int d = 0; d < 4; d++)
    for (int q = 0; q < 6; q++) v[d][q] = in[q + 6 * (i + ntotal * (d + 4 * (long)s))];
  for (int r = 0; r < 4; r++) {
    const int c = gr.col[mu][r];
    const double pr = gr.re[mu][r], pi = gr.im[mu][r];
    for (int q = 0; q < 3; q++) {
      const double a = v[c][2 * q], b = v[c][2 * q + 1];
      out[2 * q + 6 * (i + ntotal * (r + 4 * (long)s))] = pr * a - pi * b;
      out[2 * q + 1 + 6 * (i + ntotal * (r + 4 * (long)s))] = pr * b + pi * a;
    }
  }
}

void gamma_ref(const Geom &g, const double *in, double *out, int mu, int nslices, const int *slices,
               hipStream_t s) {
  if (mu < 1 || mu > 5) throw std::runtime_error("gammamult: mu must be 1..5");
  const long nvh = ref_nvhalf(g), ntot = ref_ntotal(g);
  SliceList sl{};
  sl.n = nslices;
  for (int k = 0; k < nslices; k++) sl.sl[k] = slices[k];
  static const GammaRow gr = gamma_rows();
  hipLaunchKernelGGL(k_gamma_ref, dim3((unsigned)((nvh + 255) / 256), nslices), dim3(256), 0, s, nvh, ntot, gr,
                     mu - 1, sl, in, out);
  QQCD_HIP(hipGetLastError());
}

// master FP64 18-real gauge -> canonical U[(s*4 + mu)*9 + r*3 + c]
__global__ void k_gauge_to_canon(int n0, int n1, int n2, int Vh, int par, long stride, const double2 *U,
                                 double2 *cn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y;
  if (i >= Vh) return;
  const int r2 = 2 * i;
  const int x0 = r2 % n0;
  int q = r2 / n0;
  const int y = q % n1;
  q /= n1;
  const int z = q % n2, t = q / n2;
  const int x = x0 + ((par + y + z + t) & 1);
  const long s = x + (long)n0 * (y + (long)n1 * (z + (long)n2 * t));
  for (int k = 0; k < 9; k++) cn[(s * 4 + mu) * 9 + k] = U[((long)(par * 4 + mu) * 9 + k) * stride + i];
}

void gauge_to_canon(const Geom &g, const Gauge &U, double *dcn, hipStream_t s) {
  if (U.prec != PREC_DOUBLE || U.recon != 18) throw std::runtime_error("gauge_to_canon: master must be FP64/18");
  const int nb = (int)((g.Vh + 255) / 256);
  for (int par = 0; par < 2; par++)
    hipLaunchKernelGGL(k_gauge_to_canon, dim3(nb, 4), dim3(256), 0, s, g.n[0], g.n[1], g.n[2], (int)g.Vh, par,
                       U.stride, (const double2 *)U.data, (double2 *)dcn);
  QQCD_HIP(hipGetLastError());
}

// seeded, counter-based normal deviates (splitmix64 + Box-Muller)
__device__ __forceinline__ uint64_t smix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename S>
__global__ void k_random(int Vh, long stride, uint64_t seed, void *out, float *nrm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Vh) return;
  double v[24];
  for (int k = 0; k < 12; k++) {
    const uint64_t a = smix(seed ^ smix((uint64_t)i * 12 + k));
    const double u1 = ((a >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = ((smix(a) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double r = sqrt(-2.0 * log(u1));
    v[2 * k] = r * cos(6.283185307179586 * u2);
    v[2 * k + 1] = r * sin(6.283185307179586 * u2);
  }
  if (sizeof(S) == 8) {
    double2 *p = (double2 *)out;
    for (int k = 0; k < 12; k++) p[(long)k * stride + i] = make_double2(v[2 * k], v[2 * k + 1]);
  } else if (sizeof(S) == 4) {
    float2 *p = (float2 *)out;
    for (int k = 0; k < 12; k++) p[(long)k * stride + i] = make_float2((float)v[2 * k], (float)v[2 * k + 1]);
  } else {
    double m = 0;
    for (int k = 0; k < 24; k++) m = fmax(m, fabs(v[k]));
    nrm[i] = (float)m;
    h2 *p = (h2 *)out;
    for (int k = 0; k < 12; k++) {
      h2 w;
      w.x = (f16)(float)(v[2 * k] / m);
      w.y = (f16)(float)(v[2 * k + 1] / m);
      p[(long)k * stride + i] = w;
    }
  }
}

void random_spinor(const Geom &g, Spinor &s, uint64_t seed, hipStream_t st) {
  const int nb = (int)((g.Vh + 255) / 256);
  if (s.prec == PREC_DOUBLE)
    hipLaunchKernelGGL(k_random<double>, dim3(nb), dim3(256), 0, st, (int)g.Vh, s.stride, seed, s.data, s.norm);
  else if (s.prec == PREC_SINGLE)
    hipLaunchKernelGGL(k_random<float>, dim3(nb), dim3(256), 0, st, (int)g.Vh, s.stride, seed, s.data, s.norm);
  else
    hipLaunchKernelGGL(k_random<f16>, dim3(nb), dim3(256), 0, st, (int)g.Vh, s.stride, seed, s.data, s.norm);
  QQCD_HIP(hipGetLastError());
}

// every component of the Vh sites set to re + i im (FP16: norm 1); padding untouched
template <typename S> __global__ void k_fill(int Vh, long stride, double re, double im, void *out, float *nrm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Vh) return;
  if (sizeof(S) == 8) {
    double2 *p = (double2 *)out;
    for (int k = 0; k < 12; k++) p[(long)k * stride + i] = make_double2(re, im);
  } else if (sizeof(S) == 4) {
    float2 *p = (float2 *)out;
    for (int k = 0; k < 12; k++) p[(long)k * stride + i] = make_float2((float)re, (float)im);
  } else {
    nrm[i] = 1.f;
    h2 *p = (h2 *)out;
    h2 w;
    w.x = (f16)(float)re;
    w.y = (f16)(float)im;
    for (int k = 0; k < 12; k++) p[(long)k * stride + i] = w;
  }
}

void fill_spinor(const Geom &g, Spinor &s, double re, double im, hipStream_t st) {
  const int nb = (int)((g.Vh + 255) / 256);
  if (s.prec == PREC_DOUBLE)
    hipLaunchKernelGGL(k_fill<double>, dim3(nb), dim3(256), 0, st, (int)g.Vh, s.stride, re, im, s.data, s.norm);
  else if (s.prec == PREC_SINGLE)
    hipLaunchKernelGGL(k_fill<float>, dim3(nb), dim3(256), 0, st, (int)g.Vh, s.stride, re, im, s.data, s.norm);
  else
    hipLaunchKernelGGL(k_fill<f16>, dim3(nb), dim3(256), 0, st, (int)g.Vh, s.stride, re, im, s.data, s.norm);
  QQCD_HIP(hipGetLastError());
}

}  // namespace qqcd

namespace qqcd {

// Seeded SU(3) links on the device, written into the master FP64 field.
// kind 0: Haar (Gram-Schmidt of complex Gaussian rows 0 and 1, row 2 =
// conj(row0 x row1)); kind 1: the same reunitarisation of 1 + eps*Gaussian.
__global__ void k_gauge_random(int n0, int n1, int n2, int Vh, long stride, int kind, double eps, uint64_t seed,
                               double2 *U) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int mu = blockIdx.y & 3, par = blockIdx.y >> 2;
  if (i >= Vh) return;
  const int r2 = 2 * i;
  const int x0 = r2 % n0;
  int q = r2 / n0;
  const int y = q % n1;
  q /= n1;
  const int z = q % n2, t = q / n2;
  const int x = x0 + ((par + y + z + t) & 1);
  const uint64_t s = (uint64_t)x + (uint64_t)n0 * (y + (uint64_t)n1 * (z + (uint64_t)n2 * t));
  double m[2][3][2];
  for (int r = 0; r < 2; r++)
    for (int c = 0; c < 3; c++) {
      const uint64_t a = smix(seed ^ smix(((s * 4 + mu) * 9 + r * 3 + c) * 2 + 1));
      const double u1 = ((a >> 11) + 0.5) * (1.0 / 9007199254740992.0);
      const double u2 = ((smix(a) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
      const double rr = sqrt(-2.0 * log(u1));
      double gr = rr * cos(6.283185307179586 * u2), gi = rr * sin(6.283185307179586 * u2);
      if (kind == 1) {
        gr = (r == c ? 1.0 : 0.0) + eps * gr;
        gi = eps * gi;
      }
      m[r][c][0] = gr;
      m[r][c][1] = gi;
    }
  // normalise row 0, orthogonalise and normalise row 1
  double n = 0;
  for (int c = 0; c < 3; c++) n += m[0][c][0] * m[0][c][0] + m[0][c][1] * m[0][c][1];
  n = 1.0 / sqrt(n);
  for (int c = 0; c < 3; c++) {
    m[0][c][0] *= n;
    m[0][c][1] *= n;
  }
  double pr = 0, pi = 0;  // <row0, row1>
  for (int c = 0; c < 3; c++) {
    pr += m[0][c][0] * m[1][c][0] + m[0][c][1] * m[1][c][1];
    pi += m[0][c][0] * m[1][c][1] - m[0][c][1] * m[1][c][0];
  }
  for (int c = 0; c < 3; c++) {
    m[1][c][0] -= pr * m[0][c][0] - pi * m[0][c][1];
    m[1][c][1] -= pr * m[0][c][1] + pi * m[0][c][0];
  }
  n = 0;
  for (int c = 0; c < 3; c++) n += m[1][c][0] * m[1][c][0] + m[1][c][1] * m[1][c][1];
  n = 1.0 / sqrt(n);
  for (int c = 0; c < 3; c++) {
    m[1][c][0] *= n;
    m[1][c][1] *= n;
  }
  double2 w[9];
  for (int r = 0; r < 2; r++)
    for (int c = 0; c < 3; c++) w[3 * r + c] = make_double2(m[r][c][0], m[r][c][1]);
  for (int c = 0; c < 3; c++) {
    const int a = (c + 1) % 3, b = (c + 2) % 3;
    // conj(r0[a] r1[b] - r0[b] r1[a])
    const double re = (m[0][a][0] * m[1][b][0] - m[0][a][1] * m[1][b][1]) - (m[0][b][0] * m[1][a][0] - m[0][b][1] * m[1][a][1]);
    const double im = (m[0][a][0] * m[1][b][1] + m[0][a][1] * m[1][b][0]) - (m[0][b][0] * m[1][a][1] + m[0][b][1] * m[1][a][0]);
    w[6 + c] = make_double2(re, -im);
  }
  for (int k = 0; k < 9; k++) U[((long)(par * 4 + mu) * 9 + k) * stride + i] = w[k];
}

void gauge_random(const Geom &g, Gauge &U, int kind, double eps, uint64_t seed, hipStream_t s) {
  if (U.prec != PREC_DOUBLE || U.recon != 18) throw std::runtime_error("gauge_random: master must be FP64/18");
  const int nb = (int)((g.Vh + 255) / 256);
  hipLaunchKernelGGL(k_gauge_random, dim3(nb, 8), dim3(256), 0, s, g.n[0], g.n[1], g.n[2], (int)g.Vh, U.stride, kind,
                     eps, seed, (double2 *)U.data);
  QQCD_HIP(hipGetLastError());
}

}  // namespace qqcd

namespace qqcd {

// out = a0 in + a1 (i g5) in, FP64;  i g5 v = (v2, v3, -v0, -v1) per colour
__global__ void k_local_spin(int Vh, long stride, const double2 *in, double2 *out, double a0, double a1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Vh) return;
  double2 v[12];
  for (int k = 0; k < 12; k++) v[k] = in[(long)k * stride + i];
  for (int c = 0; c < 3; c++) {
    const double2 s0 = v[c], s1 = v[3 + c], s2 = v[6 + c], s3 = v[9 + c];
    out[(long)c * stride + i] = make_double2(a0 * s0.x + a1 * s2.x, a0 * s0.y + a1 * s2.y);
    out[(long)(3 + c) * stride + i] = make_double2(a0 * s1.x + a1 * s3.x, a0 * s1.y + a1 * s3.y);
    out[(long)(6 + c) * stride + i] = make_double2(a0 * s2.x - a1 * s0.x, a0 * s2.y - a1 * s0.y);
    out[(long)(9 + c) * stride + i] = make_double2(a0 * s3.x - a1 * s1.x, a0 * s3.y - a1 * s1.y);
  }
}

void local_spin(const Geom &g, const Spinor &in, Spinor &out, double a0, double a1, hipStream_t s) {
  if (in.prec != PREC_DOUBLE || out.prec != PREC_DOUBLE) throw std::runtime_error("local_spin: FP64 only");
  const int nb = (int)((g.Vh + 255) / 256);
  hipLaunchKernelGGL(k_local_spin, dim3(nb), dim3(256), 0, s, (int)g.Vh, g.stride, (const double2 *)in.data,
                     (double2 *)out.data, a0, a1);
  QQCD_HIP(hipGetLastError());
}

}  // namespace qqcd

namespace qqcd {

// plain streaming copy, the calibration kernel for byte counters
__global__ void __launch_bounds__(256) k_copy(long n, const double2 *__restrict__ a, double2 *__restrict__ b) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) b[i] = a[i];
}

double bench_copy(size_t bytes, int iters, hipStream_t s) {
  const long n = (long)(bytes / 16);
  double2 *a = nullptr, *b = nullptr;
  QQCD_HIP(hipMalloc(&a, n * 16));
  QQCD_HIP(hipMalloc(&b, n * 16));
  QQCD_HIP(hipMemsetAsync(a, 0, n * 16, s));
  hipEvent_t e0, e1;
  QQCD_HIP(hipEventCreate(&e0));
  QQCD_HIP(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, s, n, a, b);
  QQCD_HIP(hipEventRecord(e0, s));
  for (int k = 0; k < iters; k++) hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, s, n, a, b);
  QQCD_HIP(hipEventRecord(e1, s));
  QQCD_HIP(hipEventSynchronize(e1));
  float ms = 0;
  QQCD_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return (double)ms / iters;
}

}  // namespace qqcd
