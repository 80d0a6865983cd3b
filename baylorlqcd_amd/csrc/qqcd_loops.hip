// qqcd_loops.hip -- the reference's disconnected-loop operators on the device.
//
// For a source b (the noise) and a sink x (the solution of M x = b), both on
// the whole lattice, the reference forms per site nine operators
// (qqcd-test/cfgsprops/quark/disconloops.f90:7769-7910):
//   J_mu(x)   = conj( f x(x+mu)^dag (1 - g_mu) U_mu(x)^dag b(x)
//                   - f x(x)^dag (1 + g_mu) U_mu(x) b(x+mu) ),   mu = 1..4
//               (point-split conserved currents: mulfor2 :8374-8465 with vgv
//               :8829-8921, mulbac2 :8574-8675 with vv :8788-8825),
//   S(x)      = conj( x(x)^dag b(x) ) + 12 ksub          (local scalar),
//   V_mu(x)   = conj( x(x)^dag g_mu b(x) ),               mu = 1..4
//               (local vector currents, gammamultshort diracops.f90:204-257),
// f = bc[mu] for a hop across the lattice edge, and sums them per timeslice
// with the momentum factors of momfacs (:8925-9010) as spacesum does
// (:9012-9124): momentum 1 sums all nine; momentum m = 2..5 weights
//   J_1 by (F3 Im J_1, F2 Im J_1), J_2 by (F1 Im J_2, F3 Im J_2),
//   J_3 by (F2 Im J_3, F1 Im J_3), J_4 by F4, S by F5 (= F4)
// as (real, imaginary) parts, and leaves the local vector currents at zero.
// F_k of momentum m are functions of d_k = twoPi (x_k + 1 - io_k) / n_k with
// twoPi the single-precision literal 6.283185307179586 of the reference.
//
// One kernel does both parities in one launch, one site per thread:
// workgroup (bx, t, par) sums its 256 sites of timeslice t on checkerboard
// par into 58 channel partials (18 for momentum 1, 10 for each of the four
// others); a second kernel adds the partials in a fixed order
// (deterministic).  The kernel
// reads b and x at the site and its four forward neighbours and U_mu(x) --
// about 2.6 KB per site in FP64 -- and is bound by those reads.
#include <stdexcept>

#include "qqcd_context.h"
#include "qqcd_loops.h"

namespace qqcd {

static constexpr int LTB = 256;
static constexpr int NCH = 58;  // 18 + 4 * 10 channels

struct LoopArgs {
  int n[4];
  double bcf[4];
  long stride;    // spinor component run
  long ustride;   // link component run
  int ksub;
  int nmom;       // 1 or 5
  int io[3];
  int gx;
};

__device__ __forceinline__ double lsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// gamma_1..gamma_4 of the reference basis (diracops.f90:61-71) as one entry
// per row: (gamma_mu psi)_s = phase * psi_col, phase = (re, im); compile-time
// so that the spin loops index registers statically.  gamma16() is checked
// against this table before the first launch.
struct GRow {
  int col, re, im;
};
__host__ __device__ constexpr GRow grow(int mu, int s) {
  return mu == 0 ? GRow{3 - s, 1, 0}
       : mu == 1 ? GRow{3 - s, 0, (s & 1) ? 1 : -1}
       : mu == 2 ? GRow{(s + 2) & 3, (s & 1) ? -1 : 1, 0}
                 : GRow{s, s < 2 ? 1 : -1, 0};
}

// phase * v for one row of gamma_mu
template <int MU, int S> __device__ __forceinline__ double2 gph(double2 v) {
  constexpr GRow g = grow(MU, S);
  if (g.re != 0) return make_double2(g.re * v.x, g.re * v.y);
  return make_double2(-g.im * v.y, g.im * v.x);
}

// conj(a) . b summed into (sr, si)
__device__ __forceinline__ void cdot(double2 a, double2 b, double &sr, double &si) {
  sr += a.x * b.x + a.y * b.y;
  si += a.x * b.y - a.y * b.x;
}

__device__ __forceinline__ double2 cfma(double2 u, double2 v, double2 acc) {
  return make_double2(acc.x + u.x * v.x - u.y * v.y, acc.y + u.x * v.y + u.y * v.x);
}

// Spin S of the two point-split terms, with the spin projectors moved onto
// the sink (gamma_mu is hermitian):
//   x(x+mu)^dag (1 - g) U^dag b(x) = sum_S (U y_S)^dag b_S(x),   y = (1 - g) x(x+mu)
//   x(x)^dag (1 + g) U b(x+mu)     = sum_S z_S^dag (U b_S(x+mu)), z = (1 + g) x(x)
// so each spin needs only two spins of the neighbour's fields.
template <int MU, int S>
__device__ __forceinline__ void split_spin(const double2 *u, const double2 *bx, const double2 *xx, const double2 *bq,
                                           const double2 *xq, long st, long jn, double &ar, double &ai, double &br,
                                           double &bi) {
  constexpr int C = grow(MU, S).col;
  double2 y[3];
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const double2 p = gph<MU, S>(xq[(long)(3 * C + r) * st + jn]);
    const double2 v = xq[(long)(3 * S + r) * st + jn];
    y[r] = make_double2(v.x - p.x, v.y - p.y);
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double2 uy = make_double2(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < 3; r++) uy = cfma(u[3 * k + r], y[r], uy);
    cdot(uy, bx[3 * S + k], ar, ai);
  }
  double2 bn[3];
#pragma unroll
  for (int k = 0; k < 3; k++) bn[k] = bq[(long)(3 * S + k) * st + jn];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double2 ub = make_double2(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < 3; r++) ub = cfma(u[3 * k + r], bn[r], ub);
    const double2 p = gph<MU, S>(xx[3 * C + k]);
    cdot(make_double2(xx[3 * S + k].x + p.x, xx[3 * S + k].y + p.y), ub, br, bi);
  }
}

template <int MU, int S>
__device__ __forceinline__ void vec_spin(const double2 *bx, const double2 *xx, double &sr, double &si) {
  constexpr int C = grow(MU, S).col;
#pragma unroll
  for (int r = 0; r < 3; r++) cdot(xx[3 * S + r], gph<MU, S>(bx[3 * C + r]), sr, si);
}

// J_mu and V_mu of one site (conjugated as the reference does at the end)
template <int MU>
__device__ __forceinline__ void current(const LoopArgs &a, int par, long i, const int c4[4], const double2 *bx,
                                        const double2 *xx, const double2 *bq, const double2 *xq, const double2 *U,
                                        double &jr, double &ji, double &vr, double &vi) {
  const long st = a.stride;
  // forward neighbour x + mu (other checkerboard) and the edge factor
  int cn[4] = {c4[0], c4[1], c4[2], c4[3]};
  double f = 1.0;
  if (++cn[MU] == a.n[MU]) {
    cn[MU] = 0;
    f = a.bcf[MU];
  }
  const long jn = (cn[0] + (long)a.n[0] * (cn[1] + (long)a.n[1] * (cn[2] + (long)a.n[2] * cn[3]))) / 2;
  double2 u[9];
  const long ub = (long)((par * 4 + MU) * 9) * a.ustride + i;
#pragma unroll
  for (int k = 0; k < 9; k++) u[k] = U[ub + (long)k * a.ustride];
  double ar = 0, ai = 0, br = 0, bi = 0;
  split_spin<MU, 0>(u, bx, xx, bq, xq, st, jn, ar, ai, br, bi);
  split_spin<MU, 1>(u, bx, xx, bq, xq, st, jn, ar, ai, br, bi);
  split_spin<MU, 2>(u, bx, xx, bq, xq, st, jn, ar, ai, br, bi);
  split_spin<MU, 3>(u, bx, xx, bq, xq, st, jn, ar, ai, br, bi);
  jr = f * (ar - br);
  ji = -f * (ai - bi);
  double sr = 0, si = 0;
  vec_spin<MU, 0>(bx, xx, sr, si);
  vec_spin<MU, 1>(bx, xx, sr, si);
  vec_spin<MU, 2>(bx, xx, sr, si);
  vec_spin<MU, 3>(bx, xx, sr, si);
  vr = sr;
  vi = -si;
}

// grid (gx, nt, 2), one site per thread: checkerboard par, index i = t * ls + j
template <bool SITE>
__global__ void __launch_bounds__(LTB) __attribute__((amdgpu_waves_per_eu(2))) k_loopops(LoopArgs a, const double2 *b0, const double2 *b1,
                                                 const double2 *x0, const double2 *x1, const double2 *U,
                                                 double *part, double2 *jo0, double2 *jo1) {
  const int t = blockIdx.y, par = blockIdx.z;
  const long ls = (long)a.n[0] * a.n[1] * a.n[2] / 2;
  const double2 *bp = par ? b1 : b0, *xp = par ? x1 : x0;  // site parity
  const double2 *bq = par ? b0 : b1, *xq = par ? x0 : x1;  // neighbour parity
  const long st = a.stride;
  const long j = blockIdx.x * (long)LTB + threadIdx.x;
  const long i = (long)t * ls + j;
  double Jr[9], Ji[9];
  int c4[4] = {0, 0, 0, 0};
  if (j < ls) {
    const long r2 = 2 * i;
    const int x0c = (int)(r2 % a.n[0]);
    long q = r2 / a.n[0];
    c4[1] = (int)(q % a.n[1]);
    q /= a.n[1];
    c4[2] = (int)(q % a.n[2]);
    c4[3] = (int)(q / a.n[2]);
    c4[0] = x0c + ((par + c4[1] + c4[2] + c4[3]) & 1);
    double2 bx[12], xx[12];
#pragma unroll
    for (int k = 0; k < 12; k++) {
      bx[k] = bp[(long)k * st + i];
      xx[k] = xp[(long)k * st + i];
    }
    current<0>(a, par, i, c4, bx, xx, bq, xq, U, Jr[0], Ji[0], Jr[5], Ji[5]);
    current<1>(a, par, i, c4, bx, xx, bq, xq, U, Jr[1], Ji[1], Jr[6], Ji[6]);
    current<2>(a, par, i, c4, bx, xx, bq, xq, U, Jr[2], Ji[2], Jr[7], Ji[7]);
    current<3>(a, par, i, c4, bx, xx, bq, xq, U, Jr[3], Ji[3], Jr[8], Ji[8]);
    double sr = 0, si = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) cdot(xx[k], bx[k], sr, si);
    Jr[4] = sr + 12.0 * a.ksub;
    Ji[4] = -si;
    if (SITE) {
      double2 *o = par ? jo1 : jo0;
#pragma unroll
      for (int op = 0; op < 9; op++) o[(long)op * st + i] = make_double2(Jr[op], Ji[op]);
#pragma unroll
      for (int op = 9; op < 12; op++) o[(long)op * st + i] = make_double2(0.0, 0.0);
    }
  } else {
#pragma unroll
    for (int op = 0; op < 9; op++) Jr[op] = Ji[op] = 0;
  }
  double acc[NCH];
#pragma unroll
  for (int op = 0; op < 9; op++) {
    acc[2 * op] = Jr[op];
    acc[2 * op + 1] = Ji[op];
  }
#pragma unroll
  for (int c = 18; c < NCH; c++) acc[c] = 0;
  if (a.nmom > 1 && j < ls) {
    // momentum factors (momfacs): twoPi is a single-precision literal there
    const double twopi = (double)6.283185307179586f;
    double sn[3], cs[3], s2[3], c2[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double d = twopi * (double)(c4[k] + 1 - a.io[k]) / (double)a.n[k];
      sincos(d, &sn[k], &cs[k]);
      sincos(2.0 * d, &s2[k], &c2[k]);
    }
    double F[4][4];
    F[0][0] = sn[0];
    F[0][1] = sn[1];
    F[0][2] = sn[2];
    F[0][3] = (cs[0] + cs[1] + cs[2]) / 3.0;
    F[1][0] = sn[0] * (cs[1] + cs[2]) / 2.0;
    F[1][1] = sn[1] * (cs[0] + cs[2]) / 2.0;
    F[1][2] = sn[2] * (cs[0] + cs[1]) / 2.0;
    F[1][3] = (cs[0] * cs[1] + cs[0] * cs[2] + cs[1] * cs[2]) / 3.0;
    F[2][0] = sn[0] * cs[1] * cs[2];
    F[2][1] = sn[1] * cs[0] * cs[2];
    F[2][2] = sn[2] * cs[0] * cs[1];
    F[2][3] = cs[0] * cs[1] * cs[2];
    F[3][0] = s2[0];
    F[3][1] = s2[1];
    F[3][2] = s2[2];
    F[3][3] = (c2[0] + c2[1] + c2[2]) / 3.0;
#pragma unroll
    for (int m = 0; m < 4; m++) {
      double *c = acc + 18 + 10 * m;
      c[0] = F[m][2] * Ji[0];
      c[1] = F[m][1] * Ji[0];
      c[2] = F[m][0] * Ji[1];
      c[3] = F[m][2] * Ji[1];
      c[4] = F[m][1] * Ji[2];
      c[5] = F[m][0] * Ji[2];
      c[6] = F[m][3] * Jr[3];
      c[7] = F[m][3] * Ji[3];
      c[8] = F[m][3] * Jr[4];  // F5 = F4
      c[9] = F[m][3] * Ji[4];
    }
  }
  __shared__ double sm[LTB / 64][NCH];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nch = a.nmom > 1 ? NCH : 18;
#pragma unroll
  for (int c = 0; c < NCH; c++) {
    if (c >= nch) break;
    const double s = lsum(acc[c]);
    if (lane == 0) sm[wv][c] = s;
  }
  __syncthreads();
  if (threadIdx.x < NCH) {
    double s = 0;
    if (threadIdx.x < nch)
      for (int q = 0; q < LTB / 64; q++) s += sm[q][threadIdx.x];
    part[(((long)par * gridDim.y + t) * a.gx + blockIdx.x) * NCH + threadIdx.x] = s;
  }
}

// out[((t * 5 + m) * 9 + op) * 2 + ri] (+)= scale (x cur_scale for op < 4) * sum of partials
__global__ void k_loop_final(int nt, int gx, const double *part, double *out, double scale, double cur_scale,
                             int accumulate) {
  const int e = threadIdx.x, t = blockIdx.x;
  if (e >= 90 || t >= nt) return;
  const int m = e / 18, op = (e % 18) / 2, ri = e % 2;
  const int ch = m == 0 ? 2 * op + ri : (op < 5 ? 18 + 10 * (m - 1) + 2 * op + ri : -1);
  double s = 0;
  if (ch >= 0)
    for (int par = 0; par < 2; par++)
      for (int b = 0; b < gx; b++) s += part[(((long)par * nt + t) * gx + b) * NCH + ch];
  s *= op < 4 ? scale * cur_scale : scale;
  double *o = out + (long)t * 90 + e;
  *o = accumulate ? *o + s : s;
}

int loopsum_gx(const Geom &g) {
  const long ls = (long)g.n[0] * g.n[1] * g.n[2] / 2;
  return (int)((ls + LTB - 1) / LTB);
}

// the compile-time gamma table against the library's gamma matrices
static void check_gamma_table() {
  static const bool ok = [] {
    const Gamma16 G = gamma16();
    for (int mu = 0; mu < 4; mu++)
      for (int s = 0; s < 4; s++) {
        const GRow g = grow(mu, s);
        if (G.col[1 + mu][s] != g.col || G.re[1 + mu][s] != g.re || G.im[1 + mu][s] != g.im) return false;
      }
    return true;
  }();
  if (!ok) throw std::runtime_error("loop operators: gamma table disagrees with the gamma basis");
}

static LoopArgs loop_args(const Geom &g, const Gauge &U, const Spinor &b, int ksub, const int *io, int nmom) {
  LoopArgs a{};
  for (int k = 0; k < 4; k++) {
    a.n[k] = g.n[k];
    a.bcf[k] = g.bc[k];
  }
  a.stride = b.stride;
  a.ustride = U.stride;
  a.ksub = ksub;
  a.nmom = nmom;
  for (int k = 0; k < 3; k++) a.io[k] = io ? io[k] : 1;
  a.gx = loopsum_gx(g);
  return a;
}

static void check_fields(const Gauge &U, const Spinor *b, const Spinor *x) {
  if (U.prec != PREC_DOUBLE || U.recon != 18) throw std::runtime_error("loop operators: FP64 18-real links expected");
  for (int p = 0; p < 2; p++)
    if (b[p].prec != PREC_DOUBLE || x[p].prec != PREC_DOUBLE)
      throw std::runtime_error("loop operators: FP64 spinors expected");
  if (b[0].stride != b[1].stride || b[0].stride != x[0].stride || b[0].stride != x[1].stride)
    throw std::runtime_error("loop operators: spinors of one geometry expected");
}

void loopsum(const Geom &g, const Gauge &U, const Spinor b[2], const Spinor x[2], int ksub, const int *io, int nmom,
             double scale, double cur_scale, double *d_out, int accumulate, double *d_part, Spinor *J,
             hipStream_t s) {
  check_fields(U, b, x);
  check_gamma_table();
  if (nmom != 1 && nmom != 5) throw std::runtime_error("loop operators: 1 or 5 momenta");
  const LoopArgs a = loop_args(g, U, b[0], ksub, io, nmom);
  const dim3 grid(a.gx, g.n[3], 2);
  const double2 *bb0 = (const double2 *)b[0].data, *bb1 = (const double2 *)b[1].data;
  const double2 *xx0 = (const double2 *)x[0].data, *xx1 = (const double2 *)x[1].data;
  if (J) {
    if (J[0].prec != PREC_DOUBLE || J[1].prec != PREC_DOUBLE || J[0].stride != b[0].stride)
      throw std::runtime_error("loop operators: FP64 site output expected");
    hipLaunchKernelGGL(k_loopops<true>, grid, dim3(LTB), 0, s, a, bb0, bb1, xx0, xx1, (const double2 *)U.data,
                       d_part, (double2 *)J[0].data, (double2 *)J[1].data);
  } else {
    hipLaunchKernelGGL(k_loopops<false>, grid, dim3(LTB), 0, s, a, bb0, bb1, xx0, xx1, (const double2 *)U.data,
                       d_part, nullptr, nullptr);
  }
  QQCD_HIP(hipGetLastError());
  if (d_out) {
    hipLaunchKernelGGL(k_loop_final, dim3(g.n[3]), dim3(128), 0, s, g.n[3], a.gx, (const double *)d_part, d_out,
                       scale, cur_scale, accumulate);
    QQCD_HIP(hipGetLastError());
  }
}

}  // namespace qqcd
