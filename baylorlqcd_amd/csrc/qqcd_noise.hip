// qqcd_noise.hip -- noise sources and local trace estimators of the
// disconnected-loop path.
//
// z4_source: Z4 noise on one checkerboard.  The reference draws one uniform
// number per (colour, spin, site) from its lagged-Fibonacci stream and maps
// it to 1 (u > 0.75), -1 (0.5 < u < 0.75), i (0.25 < u < 0.5) or -i
// (otherwise) (qqcd-test/cfgsprops/quark/disconloops.f90:4518-4569).  Here
// the uniform number of right-hand side r, global site s = x + nx(y + ny(z +
// nz t)) and component k = 3 spin + colour is the counter-based
//   u = (splitmix64(seed + G (1 + (r V + s) 12 + k)) >> 11) 2^-53,
// G = 0x9E3779B97F4A7C15, so a source does not depend on the layout or on
// how many right-hand sides are generated; the same mapping to {1,-1,i,-i}
// follows, times `scale`.  Z2 noise (z2source) maps the same uniform number
// to 1 (u > 0.5) or -1 in the real part.
//
// trace16: per timeslice t and local Dirac matrix Gamma_n, the estimator
//   T_n(t) = sum_{x in t} sum_{a,b,c} conj(eta_ac(x)) (Gamma_n)_ab x_bc(x)
// of Tr(Gamma_n M^-1) restricted to t (x = M^-1 eta).  Gamma_n, n = 0..15:
// 1; gamma_1..gamma_4; gamma_5; gamma_mu gamma_5 (mu = 1..4); sigma_munu =
// (i/2)[gamma_mu, gamma_nu] for (mu,nu) = (1,2),(1,3),(1,4),(2,3),(2,4),(3,4),
// in the reference's gamma basis (qqcd-test/cfgsprops/quark/diracops.f90:
// 61-71).  The reference's own loop operators (local scalar, local vector
// and point-split currents, disconloops.f90:7769-7910, summed per timeslice
// by spacesum :9012-9124) use the first five of these local insertions.
// Each Gamma_n has one non-zero entry per row, so a site costs 16 x 12
// complex multiply-adds and the kernel is bound by reading eta and x.
#include <complex>
#include <stdexcept>
#include <vector>

#include "qqcd_context.h"
#include "qqcd_noise.h"

namespace qqcd {

__device__ __forceinline__ uint64_t nmix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// kind 0: Z4 {1, -1, i, -i}; kind 1: Z2 {1, -1} (z2source, disconloops.f90:4583-4649: 1 for
// u > 0.5, else -1, imaginary part zero)
__global__ void k_noise(int n0, int n1, int n2, int Vh, long V, int par, long stride, uint64_t seed, long rhs,
                        double scale, int kind, double2 *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Vh) return;
  const int r2 = 2 * i;
  const int x0 = r2 % n0;
  int q = r2 / n0;
  const int y = q % n1;
  q /= n1;
  const int z = q % n2, t = q / n2;
  const int x = x0 + ((par + y + z + t) & 1);
  const long s = x + (long)n0 * (y + (long)n1 * (z + (long)n2 * t));
  for (int k = 0; k < 12; k++) {
    const uint64_t c = (uint64_t)(((uint64_t)rhs * (uint64_t)V + (uint64_t)s) * 12ull + (uint64_t)k);
    const double u = (double)(nmix(seed + 0x9E3779B97F4A7C15ull * (1ull + c)) >> 11) * (1.0 / 9007199254740992.0);
    double re = 0, im = 0;
    if (kind == 1) re = u > 0.5 ? 1 : -1;
    else if (u > 0.75) re = 1;
    else if (0.5 < u && u < 0.75) re = -1;
    else if (0.25 < u && u < 0.5) im = 1;
    else im = -1;
    out[(long)k * stride + i] = make_double2(scale * re, scale * im);
  }
}

void noise_source(const Geom &g, Spinor &out, int par, uint64_t seed, long rhs, double scale, int kind,
                  hipStream_t s) {
  if (out.prec != PREC_DOUBLE) throw std::runtime_error("noise_source: FP64 spinor expected");
  if (kind != 0 && kind != 1) throw std::runtime_error("noise_source: kind 0 (Z4) or 1 (Z2)");
  const int nb = (int)((g.Vh + 255) / 256);
  hipLaunchKernelGGL(k_noise, dim3(nb), dim3(256), 0, s, g.n[0], g.n[1], g.n[2], (int)g.Vh, g.V, par, out.stride,
                     seed, rhs, scale, kind, (double2 *)out.data);
  QQCD_HIP(hipGetLastError());
}

void z4_source(const Geom &g, Spinor &out, int par, uint64_t seed, long rhs, double scale, hipStream_t s) {
  noise_source(g, out, par, seed, rhs, scale, 0, s);
}

static constexpr int TB = 256;

__device__ __forceinline__ double tsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// grid (gx, nt): workgroup bx of timeslice t sums its share of the
// timeslice's contiguous checkerboard range; partials [(t*gx + bx)*32 + c]
__global__ void __launch_bounds__(TB) k_trace16(long ls, long stride, Gamma16 G, const double2 *eta,
                                                const double2 *x, double *part) {
  const int t = blockIdx.y, gx = gridDim.x;
  double acc[32];
#pragma unroll
  for (int c = 0; c < 32; c++) acc[c] = 0;
  for (long j = blockIdx.x * (long)TB + threadIdx.x; j < ls; j += (long)gx * TB) {
    const long i = (long)t * ls + j;
    double2 e[12], v[12];
#pragma unroll
    for (int k = 0; k < 12; k++) {
      e[k] = eta[(long)k * stride + i];
      v[k] = x[(long)k * stride + i];
    }
#pragma unroll
    for (int n = 0; n < 16; n++) {
      double sr = 0, si = 0;
#pragma unroll
      for (int a = 0; a < 4; a++) {
        const int b = G.col[n][a];
        const double pr = G.re[n][a], pi = G.im[n][a];
#pragma unroll
        for (int c = 0; c < 3; c++) {
          // conj(e_a) * ph * v_b
          const double2 ea = e[3 * a + c], vb = v[3 * b + c];
          const double wr = pr * vb.x - pi * vb.y, wi = pr * vb.y + pi * vb.x;
          sr += ea.x * wr + ea.y * wi;
          si += ea.x * wi - ea.y * wr;
        }
      }
      acc[2 * n] += sr;
      acc[2 * n + 1] += si;
    }
  }
  __shared__ double sm[TB / 64][32];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 32; c++) {
    const double s = tsum(acc[c]);
    if (lane == 0) sm[w][c] = s;
  }
  __syncthreads();
  if (threadIdx.x < 32) {
    double s = 0;
    for (int q = 0; q < TB / 64; q++) s += sm[q][threadIdx.x];
    part[((long)t * gx + blockIdx.x) * 32 + threadIdx.x] = s;
  }
}

__global__ void k_trace_final(int nt, int gx, const double *part, double *out, int accumulate) {
  const int c = threadIdx.x, t = blockIdx.x;
  if (c >= 32 || t >= nt) return;
  double s = 0;
  for (int b = 0; b < gx; b++) s += part[((long)t * gx + b) * 32 + c];
  out[(long)t * 32 + c] = accumulate ? out[(long)t * 32 + c] + s : s;
}

Gamma16 gamma16() {
  typedef std::complex<double> z;
  z g[6][4][4];
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) g[0][r][c] = r == c ? 1.0 : 0.0;
  for (int mu = 1; mu <= 5; mu++) {
    double re[4][4], im[4][4];
    gamma_dense(mu, re, im);
    for (int r = 0; r < 4; r++)
      for (int c = 0; c < 4; c++) g[mu][r][c] = z(re[r][c], im[r][c]);
  }
  auto mul = [](const z (&a)[4][4], const z (&b)[4][4], z (&o)[4][4]) {
    for (int r = 0; r < 4; r++)
      for (int c = 0; c < 4; c++) {
        z s = 0;
        for (int k = 0; k < 4; k++) s += a[r][k] * b[k][c];
        o[r][c] = s;
      }
  };
  std::vector<std::vector<z>> mats;
  auto push = [&](const z (&m)[4][4]) {
    std::vector<z> v(16);
    for (int r = 0; r < 4; r++)
      for (int c = 0; c < 4; c++) v[4 * r + c] = m[r][c];
    mats.push_back(v);
  };
  for (int n = 0; n <= 5; n++) push(g[n]);
  for (int mu = 1; mu <= 4; mu++) {
    z o[4][4];
    mul(g[mu], g[5], o);
    push(o);
  }
  const int pairs[6][2] = {{1, 2}, {1, 3}, {1, 4}, {2, 3}, {2, 4}, {3, 4}};
  for (auto &p : pairs) {
    // sigma_munu = (i/2)(g_mu g_nu - g_nu g_mu)
    z a[4][4], b[4][4], o[4][4];
    mul(g[p[0]], g[p[1]], a);
    mul(g[p[1]], g[p[0]], b);
    for (int r = 0; r < 4; r++)
      for (int c = 0; c < 4; c++) o[r][c] = z(0, 0.5) * (a[r][c] - b[r][c]);
    push(o);
  }
  Gamma16 G{};
  for (int n = 0; n < 16; n++)
    for (int r = 0; r < 4; r++) {
      int nz = 0;
      for (int c = 0; c < 4; c++) {
        const z v = mats[n][4 * r + c];
        if (std::abs(v) > 1e-12) {
          G.col[n][r] = c;
          G.re[n][r] = v.real();
          G.im[n][r] = v.imag();
          nz++;
        }
      }
      if (nz != 1) throw std::runtime_error("gamma16: a local Dirac matrix is not monomial");
    }
  return G;
}

void trace16(const Geom &g, const Spinor &eta, const Spinor &x, double *d_out, int accumulate, double *d_part,
             hipStream_t s) {
  if (eta.prec != PREC_DOUBLE || x.prec != PREC_DOUBLE) throw std::runtime_error("trace16: FP64 spinors expected");
  static const Gamma16 G = gamma16();
  const long ls = (long)g.n[0] * g.n[1] * g.n[2] / 2;
  const int nt = g.n[3];
  const int gx = trace16_gx(g);
  hipLaunchKernelGGL(k_trace16, dim3(gx, nt), dim3(TB), 0, s, ls, eta.stride, G, (const double2 *)eta.data,
                     (const double2 *)x.data, d_part);
  QQCD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_trace_final, dim3(nt), dim3(64), 0, s, nt, gx, (const double *)d_part, d_out, accumulate);
  QQCD_HIP(hipGetLastError());
}

int trace16_gx(const Geom &g) {
  const long ls = (long)g.n[0] * g.n[1] * g.n[2] / 2;
  const long want = std::max(1L, 2048L / g.n[3]);
  return (int)std::max(1L, std::min(want, (ls + TB - 1) / TB));
}

}  // namespace qqcd
