// qqcd_heatbath.hip -- quenched SU(3) gauge fields on the device.
//
// Pseudo-heatbath for the Wilson plaquette action S = beta sum_P (1 - Re Tr
// U_P / 3): Cabibbo-Marinari updates of the three SU(2) subgroups with the
// Kennedy-Pendleton SU(2) heatbath, the algorithm of the reference's sweep
// (qqcd-test/cfgsprops/cfgsprops.f90:1973-2003: staple, then phbup with
// phbsub1-3 / su2up / sprd1, qqcd-test/cfgsprops/glue/heatbath.f90:33-467).
// The random numbers come from a counter-based generator (splitmix64 of
// seed, sweep, direction, site and draw), not from the reference's
// lagged-Fibonacci stream, so trajectories are not bit-identical to the
// reference's; the equilibrium distribution is the same.
//
// All links U_mu(x) of one checkerboard and one direction are updated at
// once: their staples involve no other link of that (checkerboard,
// direction) set.  After each update the link is re-unitarised (rows 0 and
// 1 Gram-Schmidt, row 2 = conj(row0 x row1)).
#include <stdexcept>
#include <vector>

#include "qqcd_context.h"

namespace qqcd {

typedef cx<double> z;

struct M3 {
  z a[9];
};

__device__ __forceinline__ M3 ld(const double2 *U, long stride, int par, int mu, int i) {
  M3 m;
  const long base = (long)((par * 4 + mu) * 9) * stride + i;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const double2 v = U[base + (long)k * stride];
    m.a[k] = z{v.x, v.y};
  }
  return m;
}

__device__ __forceinline__ M3 mul(const M3 &A, const M3 &B) {
  M3 C;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++)
      C.a[3 * r + c] = cmul(A.a[3 * r], B.a[c]) + cmul(A.a[3 * r + 1], B.a[3 + c]) + cmul(A.a[3 * r + 2], B.a[6 + c]);
  return C;
}

// A B^dagger
__device__ __forceinline__ M3 muld(const M3 &A, const M3 &B) {
  M3 C;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      z s{0, 0};
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const z b = B.a[3 * c + k];
        s = s + cmul(A.a[3 * r + k], z{b.re, -b.im});
      }
      C.a[3 * r + c] = s;
    }
  return C;
}

// A^dagger B
__device__ __forceinline__ M3 dmul(const M3 &A, const M3 &B) {
  M3 C;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      z s{0, 0};
#pragma unroll
      for (int k = 0; k < 3; k++) s = s + cjmul(A.a[3 * k + r], B.a[3 * k + c]);
      C.a[3 * r + c] = s;
    }
  return C;
}

__device__ __forceinline__ uint64_t smix(uint64_t zz) {
  zz += 0x9E3779B97F4A7C15ull;
  zz = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ull;
  zz = (zz ^ (zz >> 27)) * 0x94D049BB133111EBull;
  return zz ^ (zz >> 31);
}

struct Rng {
  uint64_t s;
  __device__ double uni() {  // (0, 1]
    s = smix(s);
    return ((s >> 11) + 1.0) * (1.0 / 9007199254740992.0);
  }
};

struct HbArgs {
  double2 *U;
  long stride;
  int n[4];
  int Vh;
  int par, mu;
  double beta;
  uint64_t seed;
  int sweep;
};

__device__ __forceinline__ int nb_index(const int n[4], int x[4]) {
  return (x[0] + n[0] * (x[1] + n[1] * (x[2] + n[2] * x[3]))) >> 1;
}

__global__ void __launch_bounds__(128) k_heatbath(HbArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.Vh) return;
  const int *n = a.n;
  int x[4];
  {
    const int r = 2 * i;
    const int x0 = r % n[0];
    int q = r / n[0];
    x[1] = q % n[1];
    q /= n[1];
    x[2] = q % n[2];
    x[3] = q / n[2];
    x[0] = x0 + ((a.par + x[1] + x[2] + x[3]) & 1);
  }
  const int mu = a.mu, par = a.par, opar = par ^ 1;
  // staple sum A: U_mu(x) A closes the 6 plaquettes containing U_mu(x)
  M3 A;
#pragma unroll
  for (int k = 0; k < 9; k++) A.a[k] = z{0, 0};
  int xpm[4] = {x[0], x[1], x[2], x[3]};
  xpm[mu] = (xpm[mu] + 1) % n[mu];
  const int i_pm = nb_index(n, xpm);
  for (int nu = 0; nu < 4; nu++) {
    if (nu == mu) continue;
    int xpn[4] = {x[0], x[1], x[2], x[3]}, xmn[4] = {x[0], x[1], x[2], x[3]};
    int xpmmn[4] = {xpm[0], xpm[1], xpm[2], xpm[3]};
    xpn[nu] = (xpn[nu] + 1) % n[nu];
    xmn[nu] = (xmn[nu] + n[nu] - 1) % n[nu];
    xpmmn[nu] = (xpmmn[nu] + n[nu] - 1) % n[nu];
    // forward: U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
    const M3 a1 = ld(a.U, a.stride, opar, nu, i_pm);
    const M3 a2 = ld(a.U, a.stride, opar, mu, nb_index(n, xpn));
    const M3 a3 = ld(a.U, a.stride, par, nu, i);
    const M3 f = muld(muld(a1, a2), a3);
    // backward: U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu)
    const M3 b1 = ld(a.U, a.stride, par, nu, nb_index(n, xpmmn));
    const M3 b2 = ld(a.U, a.stride, opar, mu, nb_index(n, xmn));
    const M3 b3 = ld(a.U, a.stride, opar, nu, nb_index(n, xmn));
    const M3 bk = dmul(b1, dmul(b2, b3));  // b1^dag (b2^dag b3)
#pragma unroll
    for (int k = 0; k < 9; k++) A.a[k] = A.a[k] + f.a[k] + bk.a[k];
  }
  M3 U = ld(a.U, a.stride, par, mu, i);
  Rng rng{smix(a.seed ^ smix(((uint64_t)a.sweep * 8 + (uint64_t)(par * 4 + mu)) * 0x100000000ull + (uint64_t)i))};
  const int sub[3][2] = {{0, 1}, {1, 2}, {0, 2}};
  for (int s = 0; s < 3; s++) {
    const int p = sub[s][0], q = sub[s][1];
    const M3 W = mul(U, A);
    const z wpp = W.a[3 * p + p], wpq = W.a[3 * p + q], wqp = W.a[3 * q + p], wqq = W.a[3 * q + q];
    // SU(2) part of the 2x2 block: w = a0 + i a.sigma
    const double a0 = 0.5 * (wpp.re + wqq.re), a3 = 0.5 * (wpp.im - wqq.im);
    const double a1 = 0.5 * (wpq.im + wqp.im), a2 = 0.5 * (wpq.re - wqp.re);
    const double kk = sqrt(a0 * a0 + a1 * a1 + a2 * a2 + a3 * a3);
    if (kk < 1e-12) continue;
    const double alpha = 2.0 * a.beta * kk / 3.0;
    // Kennedy-Pendleton: x0 with density ~ sqrt(1 - x0^2) exp(alpha x0)
    double x0 = 1.0;
    for (int tries = 0; tries < 200; tries++) {
      const double r1 = rng.uni(), r2 = rng.uni(), r3 = rng.uni(), r4 = rng.uni();
      const double c = cos(6.283185307179586 * r2);
      const double X = -(log(r1) + c * c * log(r3)) / alpha;
      if (r4 * r4 <= 1.0 - 0.5 * X) {
        x0 = 1.0 - X;
        break;
      }
    }
    const double rr = sqrt(fmax(0.0, 1.0 - x0 * x0));
    const double ct = 2.0 * rng.uni() - 1.0, ph = 6.283185307179586 * rng.uni();
    const double st = sqrt(fmax(0.0, 1.0 - ct * ct));
    const double x1 = rr * st * cos(ph), x2 = rr * st * sin(ph), x3 = rr * ct;
    // r = x v^dag with v = w/k.  For quaternions (a0 + i a.s)(b0 + i b.s) =
    // (a0 b0 - a.b) + i (a0 b + b0 a - a x b).s; here b = (v0, -v).
    const double v0 = a0 / kk, v1 = a1 / kk, v2 = a2 / kk, v3 = a3 / kk;
    const double r0 = x0 * v0 + x1 * v1 + x2 * v2 + x3 * v3;
    const double q1 = -x0 * v1 + v0 * x1 + (x2 * v3 - x3 * v2);
    const double q2 = -x0 * v2 + v0 * x2 + (x3 * v1 - x1 * v3);
    const double q3 = -x0 * v3 + v0 * x3 + (x1 * v2 - x2 * v1);
    // 2x2 matrix of r0 + i q.sigma: [[r0 + i q3, q2 + i q1], [-q2 + i q1, r0 - i q3]]
    const z m00{r0, q3}, m01{q2, q1}, m10{-q2, q1}, m11{r0, -q3};
    // U <- R U on rows p, q
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const z up = U.a[3 * p + c], uq = U.a[3 * q + c];
      U.a[3 * p + c] = cmul(m00, up) + cmul(m01, uq);
      U.a[3 * q + c] = cmul(m10, up) + cmul(m11, uq);
    }
  }
  // re-unitarise
  double nrm = 0;
  for (int c = 0; c < 3; c++) nrm += U.a[c].re * U.a[c].re + U.a[c].im * U.a[c].im;
  nrm = 1.0 / sqrt(nrm);
  for (int c = 0; c < 3; c++) U.a[c] = nrm * U.a[c];
  z d{0, 0};
  for (int c = 0; c < 3; c++) d = d + cjmul(U.a[c], U.a[3 + c]);
  for (int c = 0; c < 3; c++) U.a[3 + c] = U.a[3 + c] - cmul(d, U.a[c]);
  nrm = 0;
  for (int c = 0; c < 3; c++) nrm += U.a[3 + c].re * U.a[3 + c].re + U.a[3 + c].im * U.a[3 + c].im;
  nrm = 1.0 / sqrt(nrm);
  for (int c = 0; c < 3; c++) U.a[3 + c] = nrm * U.a[3 + c];
  for (int c = 0; c < 3; c++) {
    const int e = (c + 1) % 3, f = (c + 2) % 3;
    const z t = cmul(U.a[e], U.a[3 + f]) - cmul(U.a[f], U.a[3 + e]);
    U.a[6 + c] = z{t.re, -t.im};
  }
  const long base = (long)((par * 4 + mu) * 9) * a.stride + i;
#pragma unroll
  for (int k = 0; k < 9; k++) a.U[base + (long)k * a.stride] = make_double2(U.a[k].re, U.a[k].im);
}

// sum over sites of parity par of Re Tr of the 6 plaquettes U_mu nu(x)
__global__ void __launch_bounds__(256) k_plaquette(const double2 *Ug, long stride, int n0, int n1, int n2, int n3,
                                                   int Vh, int par, double *part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n[4] = {n0, n1, n2, n3};
  double acc = 0;
  if (i < Vh) {
    int x[4];
    const int r = 2 * i;
    const int x0 = r % n[0];
    int q = r / n[0];
    x[1] = q % n[1];
    q /= n[1];
    x[2] = q % n[2];
    x[3] = q / n[2];
    x[0] = x0 + ((par + x[1] + x[2] + x[3]) & 1);
    for (int mu = 0; mu < 4; mu++)
      for (int nu = mu + 1; nu < 4; nu++) {
        int xm[4] = {x[0], x[1], x[2], x[3]}, xn[4] = {x[0], x[1], x[2], x[3]};
        xm[mu] = (xm[mu] + 1) % n[mu];
        xn[nu] = (xn[nu] + 1) % n[nu];
        const M3 a = ld(Ug, stride, par, mu, i), b = ld(Ug, stride, par ^ 1, nu, nb_index(n, xm));
        const M3 c = ld(Ug, stride, par ^ 1, mu, nb_index(n, xn)), d = ld(Ug, stride, par, nu, i);
        const M3 p = muld(muld(mul(a, b), c), d);
        acc += p.a[0].re + p.a[4].re + p.a[8].re;
      }
  }
  __shared__ double sm[256];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

void gauge_random(const Geom &g, Gauge &U, int kind, double eps, uint64_t seed, hipStream_t s);

double plaquette(Context &c) {
  const Geom &g = c.g;
  const int nb = (int)((g.Vh + 255) / 256);
  double *part = nullptr;
  QQCD_HIP(hipMalloc(&part, sizeof(double) * nb));
  std::vector<double> h(nb);
  double tot = 0;
  for (int par = 0; par < 2; par++) {
    hipLaunchKernelGGL(k_plaquette, dim3(nb), dim3(256), 0, c.stream, (const double2 *)c.U64.data, c.U64.stride,
                       g.n[0], g.n[1], g.n[2], g.n[3], (int)g.Vh, par, part);
    QQCD_HIP(hipGetLastError());
    QQCD_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, c.stream));
    QQCD_HIP(hipStreamSynchronize(c.stream));
    for (double v : h) tot += v;
  }
  (void)hipFree(part);
  return tot / (6.0 * 3.0 * (double)g.V);
}

void heatbath(Context &c, double beta, int nsweep, uint64_t seed, int hot_start) {
  const Geom &g = c.g;
  if (!c.U64.data) c.U64 = alloc_gauge(g, PREC_DOUBLE, 18);
  if (hot_start >= 0) {
    if (hot_start) {
      gauge_random(g, c.U64, 0, 0.0, seed ^ 0x5bd1e995ull, c.stream);
    } else {
      gauge_random(g, c.U64, 1, 0.0, seed, c.stream);  // eps = 0: the unit field
    }
  } else if (!c.have_gauge) {
    throw std::runtime_error("heatbath: no field to continue from");
  }
  HbArgs a{};
  a.U = (double2 *)c.U64.data;
  a.stride = c.U64.stride;
  for (int m = 0; m < 4; m++) a.n[m] = g.n[m];
  a.Vh = (int)g.Vh;
  a.beta = beta;
  a.seed = seed;
  const int nb = (int)((g.Vh + 127) / 128);
  // the sweep index keys the random stream: a hot or cold start begins a new
  // chain at sweep 0, a continuation carries on from the chain's last sweep
  if (hot_start >= 0) c.hb_sweep = 0;
  for (int sw = 0; sw < nsweep; sw++) {
    a.sweep = c.hb_sweep++;
    for (int par = 0; par < 2; par++)
      for (int mu = 0; mu < 4; mu++) {
        a.par = par;
        a.mu = mu;
        hipLaunchKernelGGL(k_heatbath, dim3(nb), dim3(128), 0, c.stream, a);
        QQCD_HIP(hipGetLastError());
      }
  }
  QQCD_HIP(hipStreamSynchronize(c.stream));
  c.u_host = nullptr;
  c.have_gauge = 1;
  c.gauge_gen++;
}

}  // namespace qqcd
