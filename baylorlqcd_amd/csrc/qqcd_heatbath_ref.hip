// qqcd_heatbath_ref.hip -- the reference's pseudo-heatbath sweep, step for
// step, on the device, drawing from the reference's lagged-Fibonacci stream.
//
// sweep (qqcd-test/cfgsprops/cfgsprops.f90:1973-2003) visits the 128 link
// sets (ieo, ibl, mu) -- ieo outermost, mu innermost -- each holding the
// links U_mu(x) of the nvhalf sites x of block ibl and block-internal parity
// ieo.  No link of a set appears in the staples of another link of the same
// set, so a set is one data-parallel step; its sites are taken in the
// reference's site order i = 1..nvhalf (one lane per site) because that order
// decides which random numbers each site receives:
//   staple (qqcd-test/cfgsprops/glue/gaugetools.f90:39-345): per nu != mu the
//     plaquette paths times coact(1,mu,nu),
//       mmd(mm(U_nu(x), U_mu(x+nu)), U_nu(x+mu)) + mm(mdm(U_nu(x-nu), U_mu(x-nu)), U_nu(x-nu+mu)),
//     and for the improved actions (gaction = 2) the six 2x1 rectangles
//     (k_hb_staple lists their product chains)
//   phbup (qqcd-test/cfgsprops/glue/heatbath.f90:33-76), for the SU(2)
//   subgroups (1,2), (2,3), (1,3): phbsub1-3 project rows (p,q) of u c^dagger
//   to z (4 reals) (:80-235); su2up (:239-323) runs Kennedy-Pendleton passes:
//   a pass draws 4 numbers for each site not yet accepted, in site order
//   (4 inotup draws), until every site has accepted; then 2 numbers per site
//   (2 nvhalf draws) place (a1, a2, a3) on the sphere; sprd1 (:327-467)
//   multiplies rows p, q of u by the new SU(2) element.
// The draw of a site in a pass is its rank among the sites not yet accepted
// (each workgroup adds the counts of the workgroups before it), so the stream is consumed exactly as the reference
// consumes it and the stream state after a sweep is the reference's.
// Every sum and product is evaluated in the reference's written order
// without fused multiply-adds (mm, mmd, mdm: gaugetools.f90:741-817,
// 908-984, 1074-1151), so staples and projections are bit-identical for
// identical links; cos, sin and log come from the device math library and
// may differ from the host's in the last place.
//
// Links live on the device in the reference's own layout
// u(18, ntotal, 4, 2, 16) for the duration of the call.
#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "qqcd_context.h"
#include "qqcd_lagfib.h"

#pragma clang fp contract(off)

namespace qqcd {

static constexpr int HB_TB = 256;
// Kennedy-Pendleton passes per batch between host checks (even).  At beta = 6
// a pass accepts ~97% of the sites that draw, so nvhalf = 65536 sites need
// about four passes and a second batch is rare; passes with nothing left to
// draw still cost a launch.
static constexpr int HB_PB = 4;
// su2up's twoPi: a single-precision literal held in a double
static constexpr double HB_TWOPI = (double)6.283185307179586f;

struct HbG {
  int n[4];
  long nvh, ntot;
};

// coordinates of reference site i (0-based) of block ibl, block-internal parity ieo
__device__ __forceinline__ void hb_site(const HbG &G, long i, int ieo, int ibl, int x[4]) {
  const int ns0 = G.n[0] / 2, ns1 = G.n[1] / 2, ns2 = G.n[2] / 2;
  const long L = 2 * i;
  int s[4];
  s[0] = (int)(L % ns0);
  long q = L / ns0;
  s[1] = (int)(q % ns1);
  q /= ns1;
  s[2] = (int)(q % ns2);
  s[3] = (int)(q / ns2);
  if (((s[0] + s[1] + s[2] + s[3]) & 1) != ieo) s[0] += 1;  // ns0 is even: no carry
  for (int m = 0; m < 4; m++) x[m] = 2 * s[m] + ((ibl >> m) & 1);
}

// offset of u(1, i, mu, ieo, ibl) for the link U_mu(x)
__device__ __forceinline__ long hb_link(const HbG &G, const int x[4], int mu) {
  int s[4], b = 0;
  for (int m = 0; m < 4; m++) {
    s[m] = x[m] >> 1;
    b |= (x[m] & 1) << m;
  }
  const long L = s[0] + (long)(G.n[0] / 2) * (s[1] + (long)(G.n[1] / 2) * (s[2] + (long)(G.n[2] / 2) * s[3]));
  const int ieo = (s[0] + s[1] + s[2] + s[3]) & 1;
  return 18L * (L / 2 + G.ntot * (mu + 4L * (ieo + 2L * b)));
}

__device__ __forceinline__ void hb_load(const double *u, long off, double a[18]) {
#pragma unroll
  for (int k = 0; k < 18; k++) a[k] = u[off + k];
}

// w = y v (mm)
__device__ __forceinline__ void hb_mm(const double y[18], const double v[18], double w[18]) {
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double *a = y + 6 * r;
      w[6 * r + 2 * c] = a[0] * v[2 * c] - a[1] * v[2 * c + 1] + a[2] * v[6 + 2 * c] - a[3] * v[7 + 2 * c] +
                         a[4] * v[12 + 2 * c] - a[5] * v[13 + 2 * c];
      w[6 * r + 2 * c + 1] = a[0] * v[2 * c + 1] + a[1] * v[2 * c] + a[2] * v[7 + 2 * c] + a[3] * v[6 + 2 * c] +
                             a[4] * v[13 + 2 * c] + a[5] * v[12 + 2 * c];
    }
}

// w = y v^dagger (mmd)
__device__ __forceinline__ void hb_mmd(const double y[18], const double v[18], double w[18]) {
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double *a = y + 6 * r, *b = v + 6 * c;
      w[6 * r + 2 * c] = a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3] + a[4] * b[4] + a[5] * b[5];
      w[6 * r + 2 * c + 1] = -(a[0] * b[1]) + a[1] * b[0] - a[2] * b[3] + a[3] * b[2] - a[4] * b[5] + a[5] * b[4];
    }
}

// w = y^dagger v (mdm)
__device__ __forceinline__ void hb_mdm(const double y[18], const double v[18], double w[18]) {
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      w[6 * r + 2 * c] = y[2 * r] * v[2 * c] + y[2 * r + 1] * v[2 * c + 1] + y[6 + 2 * r] * v[6 + 2 * c] +
                         y[7 + 2 * r] * v[7 + 2 * c] + y[12 + 2 * r] * v[12 + 2 * c] + y[13 + 2 * r] * v[13 + 2 * c];
      w[6 * r + 2 * c + 1] = y[2 * r] * v[2 * c + 1] - y[2 * r + 1] * v[2 * c] + y[6 + 2 * r] * v[7 + 2 * c] -
                             y[7 + 2 * r] * v[6 + 2 * c] + y[12 + 2 * r] * v[13 + 2 * c] -
                             y[13 + 2 * r] * v[12 + 2 * c];
    }
}

// coact(1, mu, nu), coact(2, mu, nu), coact(2, nu, mu) for the link direction mu, and gaction
struct HbCoef {
  double c1[4], c2[4], c2t[4];
  int ga[4];
};

__device__ __forceinline__ void hb_at(const HbG &G, const int x[4], int d1, int k1, int d2, int k2, int y[4]) {
  for (int m = 0; m < 4; m++) y[m] = x[m];
  y[d1] = (y[d1] + k1 + 2 * G.n[d1]) % G.n[d1];
  y[d2] = (y[d2] + k2 + 2 * G.n[d2]) % G.n[d2];
}

// the link U_d at x + k1 e_d1 + k2 e_d2
__device__ __forceinline__ void hb_ld(const HbG &G, const double *u, const int x[4], int d, int d1, int k1, int d2,
                                      int k2, double a[18]) {
  int y[4];
  hb_at(G, x, d1, k1, d2, k2, y);
  hb_load(u, hb_link(G, y, d), a);
}

__device__ __forceinline__ void hb_acc(double c[18], double w, const double s[18]) {
#pragma unroll
  for (int k = 0; k < 18; k++) c[k] = c[k] + w * s[k];
}

// staple sum of the links U_mu(x), x in set (ieo, ibl): C[k*nvh + i].  Every
// term is a path from x to x+mu, in the reference's order of products:
// plaquettes 1, 2 (coact(1,mu,nu)); for gaction(mu) = 2 the rectangles 3-6,
// long in mu (coact(2,mu,nu)); for gaction(nu) = 2 the rectangles 7-8, long
// in nu (coact(2,nu,mu)).
__global__ void __launch_bounds__(HB_TB) k_hb_staple(HbG G, const double *u, int ieo, int ibl, int mu, HbCoef K,
                                                     double *C) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.nvh) return;
  int x[4];
  hb_site(G, i, ieo, ibl, x);
  double c[18];
#pragma unroll
  for (int k = 0; k < 18; k++) c[k] = 0.0;
  for (int nu = 0; nu < 4; nu++) {
    if (nu == mu) continue;
    double a[18], b[18], t[18], w[18], v[18];
    // 1: U_nu(x) U_mu(x+nu) U_nu(x+mu)^dag
    hb_ld(G, u, x, nu, nu, 0, mu, 0, a);
    hb_ld(G, u, x, mu, nu, 1, mu, 0, b);
    hb_mm(a, b, t);
    hb_ld(G, u, x, nu, mu, 1, nu, 0, a);
    hb_mmd(t, a, w);
    hb_acc(c, K.c1[nu], w);
    // 2: U_nu(x-nu)^dag U_mu(x-nu) U_nu(x-nu+mu)
    hb_ld(G, u, x, nu, nu, -1, mu, 0, a);
    hb_ld(G, u, x, mu, nu, -1, mu, 0, b);
    hb_mdm(a, b, t);
    hb_ld(G, u, x, nu, nu, -1, mu, 1, a);
    hb_mm(t, a, w);
    hb_acc(c, K.c1[nu], w);
    if (K.ga[mu] != 1) {
      // 3: [U_nu(x) U_mu(x+nu)] [U_mu(x+mu) U_nu(x+2mu) U_mu(x+mu+nu)^dag]^dag
      hb_ld(G, u, x, mu, mu, 1, nu, 0, a);
      hb_ld(G, u, x, nu, mu, 2, nu, 0, b);
      hb_mm(a, b, t);
      hb_ld(G, u, x, mu, mu, 1, nu, 1, a);
      hb_mmd(t, a, v);
      hb_ld(G, u, x, nu, nu, 0, mu, 0, a);
      hb_ld(G, u, x, mu, nu, 1, mu, 0, b);
      hb_mm(a, b, t);
      hb_mmd(t, v, w);
      hb_acc(c, K.c2[nu], w);
      // 4: [U_nu(x-nu)^dag U_mu(x-nu)] [U_mu(x-nu+mu) U_nu(x-nu+2mu) U_mu(x+mu)^dag]
      hb_ld(G, u, x, mu, nu, -1, mu, 1, a);
      hb_ld(G, u, x, nu, nu, -1, mu, 2, b);
      hb_mm(a, b, t);
      hb_ld(G, u, x, mu, mu, 1, nu, 0, a);
      hb_mmd(t, a, v);
      hb_ld(G, u, x, nu, nu, -1, mu, 0, a);
      hb_ld(G, u, x, mu, nu, -1, mu, 0, b);
      hb_mdm(a, b, t);
      hb_mm(t, v, w);
      hb_acc(c, K.c2[nu], w);
      // 5: U_mu(x-mu)^dag U_nu(x-mu) U_mu(x-mu+nu) U_mu(x+nu) U_nu(x+mu)^dag
      hb_ld(G, u, x, mu, mu, -1, nu, 0, a);
      hb_ld(G, u, x, nu, mu, -1, nu, 0, b);
      hb_mdm(a, b, t);
      hb_ld(G, u, x, mu, mu, -1, nu, 1, a);
      hb_mm(t, a, v);
      hb_ld(G, u, x, mu, nu, 1, mu, 0, a);
      hb_mm(v, a, t);
      hb_ld(G, u, x, nu, mu, 1, nu, 0, a);
      hb_mmd(t, a, w);
      hb_acc(c, K.c2[nu], w);
      // 6: [[U_mu(z)^dag U_nu(z) U_mu(z+nu)]^dag U_mu(x-nu)] U_nu(x-nu+mu), z = x-nu-mu
      hb_ld(G, u, x, mu, nu, -1, mu, -1, a);
      hb_ld(G, u, x, nu, nu, -1, mu, -1, b);
      hb_mdm(a, b, t);
      hb_ld(G, u, x, mu, mu, -1, nu, 0, a);
      hb_mm(t, a, v);
      hb_ld(G, u, x, mu, nu, -1, mu, 0, a);
      hb_mdm(v, a, t);
      hb_ld(G, u, x, nu, nu, -1, mu, 1, a);
      hb_mm(t, a, w);
      hb_acc(c, K.c2[nu], w);
    }
    if (K.ga[nu] != 1) {
      // 7: [U_nu(x) [U_nu(x+nu) U_mu(x+2nu) U_nu(x+nu+mu)^dag]] U_nu(x+mu)^dag
      hb_ld(G, u, x, nu, nu, 1, mu, 0, a);
      hb_ld(G, u, x, mu, nu, 2, mu, 0, b);
      hb_mm(a, b, t);
      hb_ld(G, u, x, nu, nu, 1, mu, 1, a);
      hb_mmd(t, a, v);
      hb_ld(G, u, x, nu, nu, 0, mu, 0, a);
      hb_mm(a, v, t);
      hb_ld(G, u, x, nu, mu, 1, nu, 0, a);
      hb_mmd(t, a, w);
      hb_acc(c, K.c2t[nu], w);
      // 8: [U_nu(x-nu)^dag [U_nu(x-2nu)^dag U_mu(x-2nu) U_nu(x-2nu+mu)]] U_nu(x-nu+mu)
      hb_ld(G, u, x, nu, nu, -2, mu, 0, a);
      hb_ld(G, u, x, mu, nu, -2, mu, 0, b);
      hb_mdm(a, b, t);
      hb_ld(G, u, x, nu, nu, -2, mu, 1, a);
      hb_mm(t, a, v);
      hb_ld(G, u, x, nu, nu, -1, mu, 0, a);
      hb_mdm(a, v, t);
      hb_ld(G, u, x, nu, nu, -1, mu, 1, a);
      hb_mm(t, a, w);
      hb_acc(c, K.c2t[nu], w);
    }
  }
#pragma unroll
  for (int k = 0; k < 18; k++) C[(long)k * G.nvh + i] = c[k];
}


// phbsub1-3: z of the link's rows (p, q) against the staple; dk = 1/|z|; the
// site starts the Kennedy-Pendleton passes not accepted
// phbsub of site i for subgroup ISU2 from its link a (18 reals) and staple C.
// Rows (p, q) of the subgroup are the offsets P, Q into the 18 reals; subgroup
// 2 flips z3
template <int ISU2>
__device__ __forceinline__ double hb_proj_site(const HbG &G, const double a[18], const double *C, long i, double *Z,
                                               double *DK, int *ID, double *A0) {
  constexpr int P = ISU2 == 2 ? 6 : 0, Q = ISU2 == 1 ? 6 : 12;
  constexpr double s3 = ISU2 == 2 ? -1.0 : 1.0;
  double c[18];
#pragma unroll
  for (int k = 0; k < 18; k++) c[k] = C[(long)k * G.nvh + i];
  double z1 = a[P] * c[P];
  for (int k = 1; k < 6; k++) z1 = z1 + a[P + k] * c[P + k];
  for (int k = 0; k < 6; k++) z1 = z1 + a[Q + k] * c[Q + k];
  double z2 = -(a[P] * c[Q + 1]);
  z2 = z2 + a[P + 1] * c[Q];
  for (int m = 1; m < 3; m++) {
    z2 = z2 - a[P + 2 * m] * c[Q + 2 * m + 1];
    z2 = z2 + a[P + 2 * m + 1] * c[Q + 2 * m];
  }
  for (int m = 0; m < 3; m++) {
    z2 = z2 - a[Q + 2 * m] * c[P + 2 * m + 1];
    z2 = z2 + a[Q + 2 * m + 1] * c[P + 2 * m];
  }
  double z3 = a[P] * c[Q];
  for (int k = 1; k < 6; k++) z3 = z3 + a[P + k] * c[Q + k];
  for (int k = 0; k < 6; k++) z3 = z3 - a[Q + k] * c[P + k];
  z3 = s3 * z3;  // phbsub2's sum is this one with every sign flipped: the same value negated
  double z4 = -(a[P] * c[P + 1]);
  z4 = z4 + a[P + 1] * c[P];
  for (int m = 1; m < 3; m++) {
    z4 = z4 - a[P + 2 * m] * c[P + 2 * m + 1];
    z4 = z4 + a[P + 2 * m + 1] * c[P + 2 * m];
  }
  for (int m = 0; m < 3; m++) {
    z4 = z4 + a[Q + 2 * m] * c[Q + 2 * m + 1];
    z4 = z4 - a[Q + 2 * m + 1] * c[Q + 2 * m];
  }
  Z[i] = z1;
  Z[G.nvh + i] = z2;
  Z[2 * G.nvh + i] = z3;
  Z[3 * G.nvh + i] = z4;
  const double dk = 1.0 / sqrt(z1 * z1 + z2 * z2 + z3 * z3 + z4 * z4);
  DK[i] = dk;
  ID[i] = 0;
  A0[i] = 0.0;
  return dk;
}



// status of a batch of passes, read back with one copy: the running stream
// position, per pass the sites that drew, and the sites that accepted in the
// batch's last pass
struct HbStatus {
  long pos;
  long tot[HB_PB];
  int acc_last;
};
static_assert(HB_PB > 1 && HB_PB % 2 == 0, "a batch starts on the counts the previous one left in cnt[0, nblk)");

__device__ __forceinline__ double hb_rn(const uint32_t *y, long j) { return (double)(int)y[j] * LAG_FNORM; }

// the first Kennedy-Pendleton pass of site i: every site draws, in site
// order, so its 4 draws start at base + 4 i; the test and a0 exactly as
// k_hb_try forms them.  1 if the site accepts.
__device__ __forceinline__ int hb_first_pass(const uint32_t *y, long base, long i, double dk, int *ID, double *A0) {
  const long j = base + 4 * i;
  const double r0 = hb_rn(y, j), r1 = hb_rn(y, j + 1), r2 = hb_rn(y, j + 2), r3 = hb_rn(y, j + 3);
  const double tt = cos(HB_TWOPI * r2);
  const double d2 = -(log(r0) + log(r1) * tt * tt) * dk * 3.0;
  if (r3 * r3 <= (1.0 - 0.5 * d2)) {
    ID[i] = 1;
    A0[i] = 1.0 - d2;
    return 1;
  }
  return 0;
}

// after the first pass: the workgroup's sites still drawing (the second
// pass's counts, cnt slot 1) and, from workgroup 0, where the second pass
// starts (dpos[1]) and the batch status; every thread of the workgroup calls it
__device__ __forceinline__ void hb_first_pass_counts(long nvh, int acc, int *cnt_next, long *dpos, long base,
                                                     HbStatus *stt) {
  __shared__ int acc_s;
  if (threadIdx.x == 0) acc_s = 0;
  __syncthreads();
  if (acc) atomicAdd(&acc_s, 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int drew = (int)min((long)HB_TB, nvh - (long)blockIdx.x * HB_TB);
    cnt_next[blockIdx.x] = drew - acc_s;
    if (blockIdx.x == 0) {
      dpos[1] = base + 4 * nvh;
      stt->tot[0] = nvh;
      stt->pos = base + 4 * nvh;
      stt->acc_last = 0;
    }
  }
}

template <int ISU2>
__global__ void __launch_bounds__(HB_TB) k_hb_proj(HbG G, const double *u, const double *C, int ieo, int ibl,
                                                   int mu, double *Z, double *DK, int *ID, double *A0,
                                                   const uint32_t *y, long base0, int *cnt_next, long *dpos,
                                                   HbStatus *stt) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int acc = 0;
  if (i < G.nvh) {
    const double *up = u + 18L * (i + G.ntot * (mu + 4L * (ieo + 2L * ibl)));
    double a[18];
#pragma unroll
    for (int k = 0; k < 18; k++) a[k] = up[k];
    const double dk = hb_proj_site<ISU2>(G, a, C, i, Z, DK, ID, A0);
    acc = hb_first_pass(y, base0, i, dk, ID, A0);
  }
  hb_first_pass_counts(G.nvh, acc, cnt_next, dpos, base0, stt);
}


// one Kennedy-Pendleton pass (su2up :270-297) over the sites not yet
// accepted.  cnt_in[b]: the sites of workgroup b that draw in this pass.
// The pass draws 4 numbers per such site, in site order, from base (the
// host's position for the first pass of a batch, where the previous pass
// ended afterwards, dpos[pass & 1]); each workgroup sums the counts of the
// workgroups before it itself (nblk reads from L2), so no scan launch and
// no wait on other workgroups.  cnt_out: the next pass's counts; workgroup
// 0 records where this pass ends in dpos[(pass + 1) & 1] and the status.
__global__ void __launch_bounds__(HB_TB) k_hb_try(long nvh, int nblk, const uint32_t *y, int *ID, const double *DK,
                                                  double *A0, const int *cnt_in, int *cnt_out, long *dpos,
                                                  long base_host, int pass, HbStatus *stt) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __shared__ int red_lo[HB_TB / 64], red_all[HB_TB / 64], ws[HB_TB / 64];
  __shared__ int acc_s;
  int lo = 0, all = 0;
  for (int b = threadIdx.x; b < nblk; b += HB_TB) {
    const int v = cnt_in[b];
    all += v;
    lo += b < (int)blockIdx.x ? v : 0;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo += __shfl_xor(lo, o);
    all += __shfl_xor(all, o);
  }
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = (i < nvh && ID[i] == 0) ? 1 : 0;
  const unsigned long long bal = __ballot(f);
  if (lane == 0) {
    red_lo[wv] = lo;
    red_all[wv] = all;
    ws[wv] = __popcll(bal);
  }
  if (threadIdx.x == 0) acc_s = 0;
  __syncthreads();
  long off = 0, total = 0;
  int before = 0;
  for (int w = 0; w < HB_TB / 64; w++) {
    off += red_lo[w];
    total += red_all[w];
    before += w < wv ? ws[w] : 0;
  }
  const long base = pass == 0 ? base_host : dpos[pass & 1];
  const unsigned long long lower = lane ? (bal & ((~0ull) >> (64 - lane))) : 0ull;
  const long rank = off + before + __popcll(lower);
  int acc = 0;
  if (f) {
    const long j = base + 4 * rank;
    const double r0 = hb_rn(y, j), r1 = hb_rn(y, j + 1), r2 = hb_rn(y, j + 2), r3 = hb_rn(y, j + 3);
    const double tt = cos(HB_TWOPI * r2);
    const double d2 = -(log(r0) + log(r1) * tt * tt) * DK[i] * 3.0;
    if (r3 * r3 <= (1.0 - 0.5 * d2)) {
      ID[i] = 1;
      A0[i] = 1.0 - d2;
      acc = 1;
    }
  }
  if (acc) atomicAdd(&acc_s, 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    int drew = 0;
    for (int w = 0; w < HB_TB / 64; w++) drew += ws[w];
    cnt_out[blockIdx.x] = drew - acc_s;  // the next pass draws for the sites still not accepted
    if (pass == HB_PB - 1 && acc_s) atomicAdd(&stt->acc_last, acc_s);
    if (blockIdx.x == 0) {
      dpos[(pass + 1) & 1] = base + 4 * total;
      stt->tot[pass] = total;
      stt->pos = base + 4 * total;
      if (pass == 0) stt->acc_last = 0;  // no workgroup of this pass adds to it
    }
  }
}

// su2up's sphere and sprd1 for site i: rows (p, q) of its link (in `up`,
// real k at up[k]) become R times those rows
template <int ISU2>
__device__ __forceinline__ void hb_hyper_site(const HbG &G, double *up, long i, const uint32_t *y, long pos,
                                              const double *Z, const double *DK, const double *A0) {
  constexpr int P = ISU2 == 2 ? 6 : 0, Q = ISU2 == 1 ? 6 : 12;
  constexpr double s3 = ISU2 == 2 ? -1.0 : 1.0;
  const double r0 = hb_rn(y, pos + 2 * i), r1 = hb_rn(y, pos + 2 * i + 1);
  const double a0 = A0[i], dk = DK[i];
  const double rad = 1.0 - a0 * a0;
  const double a3 = sqrt(rad) * (2.0 * r0 - 1.0);
  const double rad2 = sqrt(fabs(rad - a3 * a3));
  const double theta = HB_TWOPI * r1;
  const double a1 = rad2 * cos(theta);
  const double a2 = rad2 * sin(theta);
  const double z1 = Z[i], z2 = Z[G.nvh + i], z3 = Z[2 * G.nvh + i], z4 = Z[3 * G.nvh + i];
  const double n1 = dk * (a0 * z1 + a1 * z2 + a2 * z3 + a3 * z4);
  const double n2 = dk * (-(a0 * z2) + a1 * z1 + a2 * z4 - a3 * z3);
  const double n3 = dk * (-(a0 * z3) - a1 * z4 + a2 * z1 + a3 * z2);
  const double n4 = dk * (-(a0 * z4) + a1 * z3 - a2 * z2 + a3 * z1);
  // sprd1 with (z1, z2, z3, z4) = (n1, n2, s3 n3, n4): subgroup 2 uses -z3 where 1 and 3 use z3
  const double t3 = s3 * n3;
  for (int c = 0; c < 3; c++) {
    const double pr = up[P + 2 * c], pi = up[P + 2 * c + 1], qr = up[Q + 2 * c], qi = up[Q + 2 * c + 1];
    const double yp_r = n1 * pr - n4 * pi + t3 * qr - n2 * qi;
    const double yp_i = n1 * pi + n4 * pr + t3 * qi + n2 * qr;
    const double yq_r = -(t3 * pr) - n2 * pi + n1 * qr + n4 * qi;
    const double yq_i = -(t3 * pi) + n2 * pr + n1 * qi - n4 * qr;
    up[P + 2 * c] = yp_r;
    up[P + 2 * c + 1] = yp_i;
    up[Q + 2 * c] = yq_r;
    up[Q + 2 * c + 1] = yq_i;
  }
}

// su2up's sphere (:300-321) and sprd1 (:327-467): rows (p, q) of the link
// become R (rows p, q) with R = [[z1 + i z4, s3 z3 + i z2], [-s3 z3 + i z2, z1 - i z4]]
template <int ISU2>
__global__ void __launch_bounds__(HB_TB) k_hb_hyper(HbG G, double *u, int ieo, int ibl, int mu, const uint32_t *y,
                                                    long pos, const double *Z, const double *DK, const double *A0) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.nvh) return;
  hb_hyper_site<ISU2>(G, u + 18L * (i + G.ntot * (mu + 4L * (ieo + 2L * ibl))), i, y, pos, Z, DK, A0);
}

// the update of subgroup ISU2 and, on the updated link, the projection and
// first Kennedy-Pendleton pass of subgroup ISU2 + 1 (one launch instead of
// three; the same arithmetic in the same order, the link held in registers).
// The update draws 2 per site from pos, the pass 4 per site from base0.
template <int ISU2>
__global__ void __launch_bounds__(HB_TB) k_hb_hyper_proj(HbG G, double *u, const double *C, int ieo, int ibl,
                                                         int mu, const uint32_t *y, long pos, double *Z, double *DK,
                                                         int *ID, double *A0, long base0, int *cnt_next, long *dpos,
                                                         HbStatus *stt) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int acc = 0;
  if (i < G.nvh) {
    double *up = u + 18L * (i + G.ntot * (mu + 4L * (ieo + 2L * ibl)));
    double a[18];
#pragma unroll
    for (int k = 0; k < 18; k++) a[k] = up[k];
    hb_hyper_site<ISU2>(G, a, i, y, pos, Z, DK, A0);
    constexpr int P = ISU2 == 2 ? 6 : 0, Q = ISU2 == 1 ? 6 : 12;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      up[P + k] = a[P + k];
      up[Q + k] = a[Q + k];
    }
    const double dk = hb_proj_site<ISU2 + 1>(G, a, C, i, Z, DK, ID, A0);
    acc = hb_first_pass(y, base0, i, dk, ID, A0);
  }
  hb_first_pass_counts(G.nvh, acc, cnt_next, dpos, base0, stt);
}

namespace {

template <typename T> struct DBuf {
  T *p = nullptr;
  explicit DBuf(size_t n) { QQCD_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T))); }
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
};

// the draws of the stream from the start state on, kept in a device window
// [origin, gen_end) that always holds the 521 draws before the consumed count
struct HbStream {
  LagState gen;  // the generator after gen_end draws
  uint32_t start_w[521];
  int start_jr = 0;
  long origin = 0, gen_end = 0, cap = 0;
  uint32_t *buf = nullptr, *alt = nullptr;
  hipStream_t s;

  HbStream(const LagState &st, long cap_, hipStream_t s_) : gen(st), cap(cap_), s(s_) {
    for (int k = 0; k < 521; k++) start_w[k] = st.w[k];
    start_jr = st.jr;
    QQCD_HIP(hipMalloc(&buf, sizeof(uint32_t) * cap));
    QQCD_HIP(hipMalloc(&alt, sizeof(uint32_t) * cap));
  }
  ~HbStream() {
    (void)hipFree(buf);
    (void)hipFree(alt);
  }
  // make draws [pos, pos + need) resident
  void ensure(long pos, long need, long chunk) {
    if (gen_end >= pos + need) return;
    const long amount = std::max(need - (gen_end - pos), chunk);
    if (gen_end - origin + amount > cap) {
      const long keep0 = std::max(origin, pos - 521);
      const long len = gen_end - keep0;
      if (len + amount > cap) throw std::runtime_error("heatbath: stream window too small");
      if (len > 0)
        QQCD_HIP(hipMemcpyAsync(alt, buf + (keep0 - origin), sizeof(uint32_t) * len, hipMemcpyDeviceToDevice, s));
      std::swap(buf, alt);
      origin = keep0;
    }
    lagfib_generate(gen, amount, buf + (gen_end - origin), s);
    gen_end += amount;
  }
  // the rnds state after `pos` draws from the start
  LagState state_at(long pos) {
    LagState out;
    const int jr = (int)((start_jr + pos) % 521);
    std::vector<uint32_t> tail;
    const long n = std::min(pos, 521L);
    if (n > 0) {
      if (pos - n < origin) throw std::runtime_error("heatbath: draws before the window were dropped");
      tail.resize(n);
      QQCD_HIP(hipMemcpyAsync(tail.data(), buf + (pos - n - origin), sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
      QQCD_HIP(hipStreamSynchronize(s));
    }
    for (int k = 0; k < 521; k++) {
      const long t = pos - 521 + k;
      out.w[(jr + k) % 521] = t >= 0 ? tail[t - (pos - n)] : start_w[(start_jr + 521 + t) % 521];
    }
    out.jr = jr;
    out.kr = (jr + 521 - 32) % 521;
    return out;
  }
};

}  // namespace

// nsweep sweeps of the reference-layout links u_host (in place), drawing from st (advanced)
void heatbath_reference(Context &c, const int gaction[4], const double *coact, int nhit, int nsweep, double *u_host,
                        LagState &st) {
  const Geom &g = c.g;
  for (int m = 0; m < 4; m++) {
    if (gaction[m] != 1 && gaction[m] != 2)
      throw std::runtime_error("sweep: gaction(mu) is 1 (nearest neighbour) or 2 (next-nearest)");
    if (g.n[m] % 2) throw std::runtime_error("sweep: even extents expected");
  }
  if (g.n[0] % 4) throw std::runtime_error("sweep: nx must be divisible by 4");
  if (nhit < 1) throw std::runtime_error("sweep: nhit >= 1 expected");
  if (nsweep < 0) throw std::runtime_error("sweep: nsweep >= 0 expected");
  lagfib_check(st);
  HbG G{};
  for (int m = 0; m < 4; m++) G.n[m] = g.n[m];
  G.nvh = ref_nvhalf(g);
  G.ntot = ref_ntotal(g);
  const size_t ulen = 18 * (size_t)G.ntot * 128;
  const int nblk = (int)((G.nvh + HB_TB - 1) / HB_TB);
  DBuf<double> du(ulen), C(18 * G.nvh), Z(4 * G.nvh), DK(G.nvh), A0(G.nvh);
  DBuf<int> ID(G.nvh), cnt(2 * (size_t)nblk);  // the counts of a pass and of the next
  DBuf<long> dpos(2);
  DBuf<HbStatus> dst(1);
  HbStatus *hst = nullptr;  // pinned: the one status copy per batch
  QQCD_HIP(hipHostMalloc((void **)&hst, sizeof(HbStatus)));
  struct PinnedFree {
    HbStatus *p;
    ~PinnedFree() { (void)hipHostFree(p); }
  } pinned_free{hst};
  QQCD_HIP(hipMemcpyAsync(du.p, u_host, sizeof(double) * ulen, hipMemcpyHostToDevice, c.stream));
  const long need_pass = 4 * G.nvh * HB_PB, chunk = std::max(8 * need_pass, 1L << 20);
  HbStream S(st, 521 + chunk + 2 * need_pass + 2 * G.nvh, c.stream);
  long pos = 0;
  const dim3 grid(nblk), blk(HB_TB);
  for (int sw = 0; sw < nsweep; sw++)
    for (int ieo = 0; ieo < 2; ieo++)
      for (int ibl = 0; ibl < 16; ibl++)
        for (int mu = 0; mu < 4; mu++) {
          HbCoef K{};
          for (int nu = 0; nu < 4; nu++) {
            K.c1[nu] = coact[4 * mu + 16 * nu];       // coact(1, mu, nu)
            K.c2[nu] = coact[1 + 4 * mu + 16 * nu];   // coact(2, mu, nu)
            K.c2t[nu] = coact[1 + 4 * nu + 16 * mu];  // coact(2, nu, mu)
            K.ga[nu] = gaction[nu];
          }
          hipLaunchKernelGGL(k_hb_staple, grid, blk, 0, c.stream, G, du.p, ieo, ibl, mu, K, C.p);
          QQCD_HIP(hipGetLastError());
          for (int isu2 = 1; isu2 <= 3; isu2++) {
            // the projection and first Kennedy-Pendleton pass of subgroup 1 (those
            // of subgroups 2 and 3 run in the update kernel of the subgroup
            // before); the window is made to cover that pass and the first batch,
            // so it does not move before the passes that start from dpos
            if (isu2 == 1) {
              S.ensure(pos, 4 * G.nvh + need_pass, chunk);
              hipLaunchKernelGGL(k_hb_proj<1>, grid, blk, 0, c.stream, G, du.p, C.p, ieo, ibl, mu, Z.p, DK.p, ID.p,
                                 A0.p, S.buf, pos - S.origin, cnt.p + nblk, dpos.p, dst.p);
              QQCD_HIP(hipGetLastError());
              pos += 4 * G.nvh;
            }
            for (int batch = 0;; batch++) {
              // a site whose staple projects to zero never accepts (the reference would loop forever)
              if (batch == 200) throw std::runtime_error("sweep: Kennedy-Pendleton passes did not end (zero staple?)");
              S.ensure(pos, need_pass, chunk);  // already resident for the first batch
              // the first batch goes on after the first pass (its counts in
              // cnt[nblk, 2 nblk)); HB_PB is even, so every later batch starts on
              // the counts in cnt[0, nblk)
              for (int p = batch == 0 ? 1 : 0; p < HB_PB; p++) {
                hipLaunchKernelGGL(k_hb_try, grid, blk, 0, c.stream, G.nvh, nblk, S.buf, ID.p, DK.p, A0.p,
                                   cnt.p + (p & 1) * nblk, cnt.p + ((p + 1) & 1) * nblk, dpos.p, pos - S.origin,
                                   p, dst.p);
                QQCD_HIP(hipGetLastError());
              }
              QQCD_HIP(hipMemcpyAsync(hst, dst.p, sizeof(HbStatus), hipMemcpyDeviceToHost, c.stream));
              QQCD_HIP(hipStreamSynchronize(c.stream));
              pos = S.origin + hst->pos;
              const long rem = hst->tot[HB_PB - 1] > 0 ? hst->tot[HB_PB - 1] - hst->acc_last : 0;
              if (rem == 0) break;
            }
            if (isu2 < 3) {
              // the update (2 draws per site) and the next subgroup's first pass (4 per site)
              S.ensure(pos, 6 * G.nvh + need_pass, chunk);
              const long ypos = pos - S.origin;
              if (isu2 == 1)
                hipLaunchKernelGGL(k_hb_hyper_proj<1>, grid, blk, 0, c.stream, G, du.p, C.p, ieo, ibl, mu, S.buf,
                                   ypos, Z.p, DK.p, ID.p, A0.p, ypos + 2 * G.nvh, cnt.p + nblk, dpos.p, dst.p);
              else
                hipLaunchKernelGGL(k_hb_hyper_proj<2>, grid, blk, 0, c.stream, G, du.p, C.p, ieo, ibl, mu, S.buf,
                                   ypos, Z.p, DK.p, ID.p, A0.p, ypos + 2 * G.nvh, cnt.p + nblk, dpos.p, dst.p);
              QQCD_HIP(hipGetLastError());
              pos += 6 * G.nvh;
            } else {
              S.ensure(pos, 2 * G.nvh, chunk);
              hipLaunchKernelGGL(k_hb_hyper<3>, grid, blk, 0, c.stream, G, du.p, ieo, ibl, mu, S.buf, pos - S.origin,
                                 Z.p, DK.p, A0.p);
              QQCD_HIP(hipGetLastError());
              pos += 2 * G.nvh;
            }
          }
        }
  QQCD_HIP(hipMemcpyAsync(u_host, du.p, sizeof(double) * ulen, hipMemcpyDeviceToHost, c.stream));
  QQCD_HIP(hipStreamSynchronize(c.stream));
  st = S.state_at(pos);
}

}  // namespace qqcd
