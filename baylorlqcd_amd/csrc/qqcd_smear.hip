// qqcd_smear.hip -- Gaussian smearing of a quark field.
//
// smear (qqcd-test/cfgsprops/quark/diracops.f90:712-856), equation (23) of
// Trottier, Phys. Rev. D55, 6844 (1997): nsmear times
//   h(x) <- b h(x) + a sum_{mu=1..3} [U_mu(x) h(x+mu) + U_mu(x-mu)^dag h(x-mu)],
// b = 1 - 6a, on both checkerboards from the previous iterate of both (the
// increments of the even and the odd sites are formed before either is
// updated), spin-diagonal, the fermion boundary condition bc(mu) on a hop
// across the spatial edge.  The increment of a direction is summed as the
// reference sums it: (acc + forward) + backward, mu = 1, 2, 3.
//
// One lane per site of the output checkerboard, each XCD on one contiguous
// range of sites; FP64 throughout.  Per site
// and iteration: the spinor at x and at its 6 spatial neighbours (7 x 192 B),
// 6 links (6 x 144 B) and the result (192 B): 2400 B.
#include <stdexcept>

#include "qqcd_context.h"

namespace qqcd {

struct SmArgs {
  int n[4];
  int bc[3];
  long Vh, stride, ustride;
  int par;
  double a, b;
};

__device__ __forceinline__ double2 sm_ld(const double2 *p, long stride, int k, long i) { return p[k * stride + i]; }

// the links and the output are touched once per launch: loaded and stored
// non-temporally, so the L2 is left to the neighbour spinors
__device__ __forceinline__ double2 nt_ld(const double2 *p) {
  const double *q = (const double *)p;
  return make_double2(__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1));
}
__device__ __forceinline__ void nt_st(double2 v, double2 *p) {
  double *q = (double *)p;
  __builtin_nontemporal_store(v.x, q);
  __builtin_nontemporal_store(v.y, q + 1);
}

// w (+)= s * M v (M row-major 3x3 at link (par, mu, i)), or M^dag v
template <bool DAG>
__device__ __forceinline__ void sm_mv(const double2 *U, long us, int par, int mu, long i, const double2 v[3],
                                      double s, double2 w[3]) {
  const long base = (long)((par * 4 + mu) * 9) * us + i;
  double2 m[9];
#pragma unroll
  for (int k = 0; k < 9; k++) m[k] = nt_ld(U + base + (long)k * us);
#pragma unroll
  for (int r = 0; r < 3; r++) {
    double re = 0, im = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double2 e = DAG ? make_double2(m[3 * c + r].x, -m[3 * c + r].y) : m[3 * r + c];
      re += e.x * v[c].x - e.y * v[c].y;
      im += e.x * v[c].y + e.y * v[c].x;
    }
    w[r] = make_double2(s * re, s * im);
  }
}

__global__ void __launch_bounds__(128) k_smear(SmArgs a, const double2 *U, const double2 *same,
                                               const double2 *other, double2 *out) {
  // workgroups are dealt round-robin to the 8 XCDs; give each XCD a
  // contiguous run of sites so the neighbours (up to n0 n1 / 2 sites away)
  // are fetched into its own L2
  const int nb = gridDim.x, per = (nb + 7) / 8;
  const int xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
  const int full = nb - 8 * (per - 1);  // XCDs that get `per` workgroups
  const int blk = xcd < full ? xcd * per + k : full * per + (xcd - full) * (per - 1) + k;
  const long i = (long)blk * blockDim.x + threadIdx.x;
  if (i >= a.Vh) return;
  const int *n = a.n;
  int x[4];
  {
    const long r = 2 * i;
    const int x0 = (int)(r % n[0]);
    long q = r / n[0];
    x[1] = (int)(q % n[1]);
    q /= n[1];
    x[2] = (int)(q % n[2]);
    x[3] = (int)(q / n[2]);
    x[0] = x0 + ((a.par + x[1] + x[2] + x[3]) & 1);
  }
  const int opar = a.par ^ 1;
  double2 acc[12];
#pragma unroll
  for (int k = 0; k < 12; k++) acc[k] = make_double2(0.0, 0.0);
  for (int mu = 0; mu < 3; mu++) {
    int xf[4] = {x[0], x[1], x[2], x[3]}, xb[4] = {x[0], x[1], x[2], x[3]};
    const bool wf = x[mu] + 1 == n[mu], wb = x[mu] == 0;
    xf[mu] = wf ? 0 : x[mu] + 1;
    xb[mu] = wb ? n[mu] - 1 : x[mu] - 1;
    const long jf = (xf[0] + (long)n[0] * (xf[1] + (long)n[1] * (xf[2] + (long)n[2] * xf[3]))) >> 1;
    const long jb = (xb[0] + (long)n[0] * (xb[1] + (long)n[1] * (xb[2] + (long)n[2] * xb[3]))) >> 1;
    const double sf = wf ? (double)a.bc[mu] : 1.0, sb = wb ? (double)a.bc[mu] : 1.0;
    for (int d = 0; d < 4; d++) {
      double2 v[3], w[3];
#pragma unroll
      for (int c = 0; c < 3; c++) v[c] = sm_ld(other, a.stride, 3 * d + c, jf);
      sm_mv<false>(U, a.ustride, a.par, mu, i, v, sf, w);
#pragma unroll
      for (int c = 0; c < 3; c++) acc[3 * d + c] = make_double2(acc[3 * d + c].x + w[c].x, acc[3 * d + c].y + w[c].y);
#pragma unroll
      for (int c = 0; c < 3; c++) v[c] = sm_ld(other, a.stride, 3 * d + c, jb);
      sm_mv<true>(U, a.ustride, opar, mu, jb, v, sb, w);
#pragma unroll
      for (int c = 0; c < 3; c++) acc[3 * d + c] = make_double2(acc[3 * d + c].x + w[c].x, acc[3 * d + c].y + w[c].y);
    }
  }
#pragma unroll
  for (int k = 0; k < 12; k++) {
    const double2 h = sm_ld(same, a.stride, k, i);
    nt_st(make_double2(a.b * h.x + a.a * acc[k].x, a.b * h.y + a.a * acc[k].y), out + (long)k * a.stride + i);
  }
}

// nsmear iterations on h[0], h[1] (FP64, both checkerboards) in place; t[0], t[1] scratch
void smear(Context &c, Spinor h[2], Spinor t[2], double asmear, int nsmear) {
  const Geom &g = c.g;
  if (nsmear < 0) throw std::runtime_error("smear: nsmear >= 0 expected");
  for (int p = 0; p < 2; p++)
    if (h[p].prec != PREC_DOUBLE || t[p].prec != PREC_DOUBLE) throw std::runtime_error("smear: FP64 fields expected");
  SmArgs a{};
  for (int m = 0; m < 4; m++) a.n[m] = g.n[m];
  for (int m = 0; m < 3; m++) a.bc[m] = g.bc[m];
  a.Vh = g.Vh;
  a.stride = h[0].stride;
  a.ustride = c.U64.stride;
  a.a = asmear;
  a.b = 1.0 - 6.0 * asmear;
  const int nb = (int)((g.Vh + 127) / 128);
  Spinor *cur = h, *nxt = t;
  for (int it = 0; it < nsmear; it++) {
    for (int p = 0; p < 2; p++) {
      a.par = p;
      hipLaunchKernelGGL(k_smear, dim3(nb), dim3(128), 0, c.stream, a, (const double2 *)c.U64.data,
                         (const double2 *)cur[p].data, (const double2 *)cur[1 - p].data, (double2 *)nxt[p].data);
      QQCD_HIP(hipGetLastError());
    }
    std::swap(cur, nxt);
  }
  if (cur != h)
    for (int p = 0; p < 2; p++)
      QQCD_HIP(hipMemcpyAsync(h[p].data, cur[p].data, h[p].bytes(), hipMemcpyDeviceToDevice, c.stream));
}

// kernel time (ms) of one smearing iteration (both checkerboards) on the
// resident links, averaged over niter iterations after nwarm
double bench_smear(Context &c, int nwarm, int niter) {
  Spinor h[2], t[2];
  for (int p = 0; p < 2; p++) {
    h[p] = alloc_spinor(c.g, PREC_DOUBLE);
    t[p] = alloc_spinor(c.g, PREC_DOUBLE);
    QQCD_HIP(hipMemsetAsync(h[p].data, 0, h[p].bytes(), c.stream));
  }
  hipEvent_t e0, e1;
  QQCD_HIP(hipEventCreate(&e0));
  QQCD_HIP(hipEventCreate(&e1));
  smear(c, h, t, 0.25, 2 * nwarm);
  QQCD_HIP(hipEventRecord(e0, c.stream));
  smear(c, h, t, 0.25, 2 * niter);  // an even count: no copy back
  QQCD_HIP(hipEventRecord(e1, c.stream));
  QQCD_HIP(hipEventSynchronize(e1));
  float ms = 0;
  QQCD_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int p = 0; p < 2; p++) {
    free_spinor(h[p]);
    free_spinor(t[p]);
  }
  return ms / (2.0 * niter);
}

}  // namespace qqcd
