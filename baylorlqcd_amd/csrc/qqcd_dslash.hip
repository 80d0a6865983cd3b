// qqcd_dslash.hip -- the Wilson hopping operator on one checkerboard.
//
// Computes, for every site x of checkerboard par_out,
//   D(x) = sum_mu [ (c1 - s c2 g_mu) U_mu(x) psi(x+mu)
//                 + (c1 + s c2 g_mu) U_mu(x-mu)^dag psi(x-mu) ]
// with s = +1 (idag=0) or -1 (idag=1), hops across the lattice edge in mu
// scaled by bc[mu], and then the local epilogue
//   out = a0 x + a1 (i g5 x) + b0 D + b1 (i g5 D).
// This is the reference's Hsingle (qqcd-test/cfgsprops/quark/diracops.f90:
// 554-599; mulbac :950, mulfor :1163, boundary signs :1684-1694, 1730-1740)
// and, through the epilogue, the local terms of Hdouble (:340-344) and
// Hdbletm (:484-514).
//
// MI355X mapping: one lane per output site, 256-lane workgroups, sites in
// x-fastest order so neighbour loads of a wavefront are contiguous runs.  The
// projectors (1 -/+ g_mu) have rank 2 in the reference's gamma basis
// (diracops.f90:61-71), so every hop multiplies a 2-spinor (12 reals) by the
// link.  In that basis (1 +/- g_4) touches only the upper or the lower spin
// pair, so the t hops load half of the neighbour spinor.  The link loads,
// spinor loads and the FP16 per-site norm are all coalesced 64-lane runs.
// Arithmetic is FP64 for FP64 storage and FP32 for FP32 and FP16 storage.
#include <algorithm>
#include <stdexcept>
#include <type_traits>
#include "qqcd_internal.h"

namespace qqcd {

template <typename S> struct RealOf { typedef float type; };
template <> struct RealOf<double> { typedef double type; };

// ---- storage access --------------------------------------------------
template <typename S> struct SIO;

template <> struct SIO<double> {
  typedef double R;
  typedef double2 E;
  static __device__ __forceinline__ cx<R> ld(const void *b, long stride, int k, int i, float) {
    const double2 v = ((const double2 *)b)[(long)k * stride + i];
    return cx<R>{v.x, v.y};
  }
  static __device__ __forceinline__ float norm(const float *, int) { return 1.f; }
  static __device__ __forceinline__ void st(void *b, float *, long stride, int i, const cx<R> *v) {
    double2 *p = (double2 *)b;
#pragma unroll
    for (int k = 0; k < 12; k++) p[(long)k * stride + i] = make_double2(v[k].re, v[k].im);
  }
};

template <> struct SIO<float> {
  typedef float R;
  typedef float2 E;
  static __device__ __forceinline__ cx<R> ld(const void *b, long stride, int k, int i, float) {
    const float2 v = ((const float2 *)b)[(long)k * stride + i];
    return cx<R>{v.x, v.y};
  }
  static __device__ __forceinline__ float norm(const float *, int) { return 1.f; }
  static __device__ __forceinline__ void st(void *b, float *, long stride, int i, const cx<R> *v) {
    float2 *p = (float2 *)b;
#pragma unroll
    for (int k = 0; k < 12; k++) p[(long)k * stride + i] = make_float2(v[k].re, v[k].im);
  }
};

typedef f16 h2 __attribute__((ext_vector_type(2)));

template <> struct SIO<f16> {
  typedef float R;
  typedef h2 E;
  static __device__ __forceinline__ cx<R> ld(const void *b, long stride, int k, int i, float nrm) {
    const h2 v = ((const h2 *)b)[(long)k * stride + i];
    return cx<R>{nrm * (float)v.x, nrm * (float)v.y};
  }
  static __device__ __forceinline__ float norm(const float *n, int i) { return n[i]; }
  static __device__ __forceinline__ void st(void *b, float *nrm, long stride, int i, const cx<R> *v) {
    float m = 0.f;
#pragma unroll
    for (int k = 0; k < 12; k++) m = fmaxf(m, fmaxf(fabsf(v[k].re), fabsf(v[k].im)));
    nrm[i] = m;
    const float inv = m > 0.f ? 1.f / m : 0.f;
    h2 *p = (h2 *)b;
#pragma unroll
    for (int k = 0; k < 12; k++) {
      h2 w;
      w.x = (f16)(v[k].re * inv);
      w.y = (f16)(v[k].im * inv);
      p[(long)k * stride + i] = w;
    }
  }
};

// links: component run (par, mu, k) at ((par*4 + mu)*NC + k)*stride
template <typename L, int RECON> struct LIO;

template <typename L> __device__ __forceinline__ cx<typename RealOf<L>::type> ldc(const void *b, long off);
template <> __device__ __forceinline__ cx<double> ldc<double>(const void *b, long off) {
  const double2 v = ((const double2 *)b)[off];
  return cx<double>{v.x, v.y};
}
template <> __device__ __forceinline__ cx<float> ldc<float>(const void *b, long off) {
  const float2 v = ((const float2 *)b)[off];
  return cx<float>{v.x, v.y};
}
template <> __device__ __forceinline__ cx<float> ldc<f16>(const void *b, long off) {
  const h2 v = ((const h2 *)b)[off];
  return cx<float>{(float)v.x, (float)v.y};
}

template <typename L, int RECON>
__device__ __forceinline__ void load_link(const void *b, long stride, int par, int mu, int i,
                                          cx<typename RealOf<L>::type> u[9]) {
  typedef typename RealOf<L>::type R;
  constexpr int NC = RECON == 18 ? 9 : 6;
  const long base = (long)((par * 4 + mu) * NC) * stride + i;
#pragma unroll
  for (int k = 0; k < NC; k++) u[k] = ldc<L>(b, base + (long)k * stride);
  if (RECON == 12) {
    // row 2 = conj(row 0 x row 1)
    u[6] = mk<R>(u[1].re * u[5].re - u[1].im * u[5].im - (u[2].re * u[4].re - u[2].im * u[4].im),
                 -(u[1].re * u[5].im + u[1].im * u[5].re - (u[2].re * u[4].im + u[2].im * u[4].re)));
    u[7] = mk<R>(u[2].re * u[3].re - u[2].im * u[3].im - (u[0].re * u[5].re - u[0].im * u[5].im),
                 -(u[2].re * u[3].im + u[2].im * u[3].re - (u[0].re * u[5].im + u[0].im * u[5].re)));
    u[8] = mk<R>(u[0].re * u[4].re - u[0].im * u[4].im - (u[1].re * u[3].re - u[1].im * u[3].im),
                 -(u[0].re * u[4].im + u[0].im * u[4].re - (u[1].re * u[3].im + u[1].im * u[3].re)));
  }
}

template <typename R> __device__ __forceinline__ void mv(const cx<R> u[9], const cx<R> *h, cx<R> *w) {
#pragma unroll
  for (int r = 0; r < 3; r++) w[r] = cmul(u[3 * r], h[0]) + cmul(u[3 * r + 1], h[1]) + cmul(u[3 * r + 2], h[2]);
}
template <typename R> __device__ __forceinline__ void mdv(const cx<R> u[9], const cx<R> *h, cx<R> *w) {
#pragma unroll
  for (int r = 0; r < 3; r++) w[r] = cjmul(u[r], h[0]) + cjmul(u[3 + r], h[1]) + cjmul(u[6 + r], h[2]);
}

// Spin projection (1 + sig g_mu) psi -> (h0, h1) [2 spins x 3 colours], and
// reconstruction of the 4 spins of (1 + sig g_mu) U psi from U h.
// 0-based spins; gamma basis of diracops.f90:61-71:
//   mu=0: h0 = p0 + sig p3, h1 = p1 + sig p2;   r = (h0, h1, sig h1, sig h0)
//   mu=1: h0 = p0 - i sig p3, h1 = p1 + i sig p2; r = (h0, h1, -i sig h1, i sig h0)
//   mu=2: h0 = p0 + sig p2, h1 = p1 - sig p3;   r = (h0, h1, sig h0, -sig h1)
//   mu=3: sig=+1: h = 2(p0, p1), r = (h0, h1, 0, 0); sig=-1: h = 2(p2, p3), r = (0, 0, h0, h1)
template <int MU, typename R>
__device__ __forceinline__ void project(const cx<R> p[12], R sig, cx<R> h[6]) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (MU == 0) {
      h[c] = p[c] + sig * p[9 + c];
      h[3 + c] = p[3 + c] + sig * p[6 + c];
    } else if (MU == 1) {
      h[c] = p[c] - sig * mul_i(p[9 + c]);
      h[3 + c] = p[3 + c] + sig * mul_i(p[6 + c]);
    } else if (MU == 2) {
      h[c] = p[c] + sig * p[6 + c];
      h[3 + c] = p[3 + c] - sig * p[9 + c];
    }
  }
}

template <int MU, typename R>
__device__ __forceinline__ void accum(cx<R> acc[12], const cx<R> w[6], R sig, R scl) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const cx<R> a = scl * w[c], b = scl * w[3 + c];
    if (MU == 0) {
      acc[c] = acc[c] + a;
      acc[3 + c] = acc[3 + c] + b;
      acc[6 + c] = acc[6 + c] + sig * b;
      acc[9 + c] = acc[9 + c] + sig * a;
    } else if (MU == 1) {
      acc[c] = acc[c] + a;
      acc[3 + c] = acc[3 + c] + b;
      acc[6 + c] = acc[6 + c] - sig * mul_i(b);
      acc[9 + c] = acc[9 + c] + sig * mul_i(a);
    } else if (MU == 2) {
      acc[c] = acc[c] + a;
      acc[3 + c] = acc[3 + c] + b;
      acc[6 + c] = acc[6 + c] + sig * a;
      acc[9 + c] = acc[9 + c] - sig * b;
    }
  }
}

struct DslashArgs {
  const void *in;
  const float *in_norm;
  void *out;
  float *out_norm;
  const void *x;
  const float *x_norm;
  const void *links;
  long sstride;  // spinor component stride
  long lstride;  // link component stride
  int n[4];
  int Vh;
  int par;       // output checkerboard
  float bcf[4];  // edge factor per direction
  // projector form: sig_f (forward projector sign), scl (c1)
  float sig_f[4];
  double scl[4];
  // split form
  double fw_p[4], fw_m[4], bw_p[4], bw_m[4];
  double a0, a1, b0, b1;
  int has_x;
  int xcd_remap;
  // traversal of the sites (single right-hand side): order 0 walks the
  // checkerboard index; order 1 walks blocks of bz z-planes, all t of a block
  // before the next block (z fastest inside the block), so the t and z
  // neighbours of a wavefront were read a few thousand sites earlier by the
  // same XCD and are still in its L2; order 2 does the same with blocks of
  // bz t-slices swept in z
  int order;
  int bz;
  int plane;  // sites of one (z, t) plane of a checkerboard: nx*ny/2
  // two-slot multi-RHS kernel only: per-slot scale 1/sqrt(slot_n2[slot]) of
  // the epilogue's result, and the FP16 side output side_hs x / sqrt(slot_n2)
  const double *slot_n2;
  void *side;
  float side_hs;
  // single right-hand side, FP16 spinors: the result also in FP32 (out32, no
  // norms), and skip16 leaves out / out_norm unwritten
  void *out32;
  int skip16;
};

// Pointers and strides a hop reads (scalars only, so that nothing of the
// kernel argument block is copied to private memory).
struct HopIn {
  const void *in;
  const float *in_norm;
  const void *links;
  long sstride, lstride;
};

// One hop in direction MU = 0, 1, 2 from the neighbour spinor at
// checkerboard index j (opposite parity), with link (lpar, li).
// FWD: U h ; else U^dag h.
template <typename S, typename L, int RECON, int MU, bool FWD, typename R>
__device__ __forceinline__ void hop(const HopIn &a, cx<R> acc[12], int j, int lpar, int li, R sig, R scl,
                                    R edge) {
  cx<R> u[9];
  load_link<L, RECON>(a.links, a.lstride, lpar, MU, li, u);
  const float nrm = SIO<S>::norm(a.in_norm, j);
  cx<R> p[12], h[6], w[6];
#pragma unroll
  for (int k = 0; k < 12; k++) p[k] = SIO<S>::ld(a.in, a.sstride, k, j, nrm);
  project<MU>(p, sig, h);
  if (FWD) {
    mv(u, h, w);
    mv(u, h + 3, w + 3);
  } else {
    mdv(u, h, w);
    mdv(u, h + 3, w + 3);
  }
  accum<MU>(acc, w, sig, scl * edge);
}

// One t hop.  (1 + g_4) keeps the upper spin pair, (1 - g_4) the lower one,
// each doubled; UPPER selects which, at compile time, so only that half of
// the neighbour spinor is loaded and the accumulator is indexed statically.
template <typename S, typename L, int RECON, bool FWD, bool UPPER, typename R>
__device__ __forceinline__ void hop_t(const HopIn &a, cx<R> acc[12], int j, int lpar, int li, R scl, R edge) {
  cx<R> u[9];
  load_link<L, RECON>(a.links, a.lstride, lpar, 3, li, u);
  const float nrm = SIO<S>::norm(a.in_norm, j);
  constexpr int s0 = UPPER ? 0 : 6;
  cx<R> h[6], w[6];
#pragma unroll
  for (int c = 0; c < 6; c++) h[c] = R(2) * SIO<S>::ld(a.in, a.sstride, s0 + c, j, nrm);
  if (FWD) {
    mv(u, h, w);
    mv(u, h + 3, w + 3);
  } else {
    mdv(u, h, w);
    mdv(u, h + 3, w + 3);
  }
  const R f = scl * edge;
#pragma unroll
  for (int c = 0; c < 6; c++) acc[s0 + c] = acc[s0 + c] + f * w[c];
}

// Both hops of direction MU.  Projector form: forward c1 (1 + sf g), backward
// c1 (1 - sf g); split form: fp (1+g) + fm (1-g) forward, bp (1+g) + bm (1-g)
// backward.
// Both hops of direction MU.  SIG = +1 / -1: projector form with the same
// sign in every direction, forward c1 (1 + SIG g), backward c1 (1 - SIG g),
// sc = c1; SIG = 0: split form, forward fp (1+g) + fm (1-g), backward
// bp (1+g) + bm (1-g).
// i: link index of the output site; jb: link index of the backward
// neighbour; jfs / jbs: spinor element indices of the two neighbours (equal
// to the site indices for one right-hand side, site*nrhs + r for several).
template <typename S, typename L, int RECON, int SIG, int MU, typename R>
__device__ __forceinline__ void direction(const HopIn &h, int par, cx<R> acc[12], int i, int jfs, int jb, int jbs,
                                          R ef, R eb, R sc, R fp, R fm, R bp, R bm) {
  const int opar = par ^ 1;
  const int jf = jfs;
  if (MU == 3) {
    if (SIG != 0) {
      hop_t<S, L, RECON, true, (SIG > 0)>(h, acc, jf, par, i, sc, ef);
      hop_t<S, L, RECON, false, (SIG < 0)>(h, acc, jbs, opar, jb, sc, eb);
    } else {
      if (fp != R(0)) hop_t<S, L, RECON, true, true>(h, acc, jf, par, i, fp, ef);
      if (fm != R(0)) hop_t<S, L, RECON, true, false>(h, acc, jf, par, i, fm, ef);
      if (bp != R(0)) hop_t<S, L, RECON, false, true>(h, acc, jbs, opar, jb, bp, eb);
      if (bm != R(0)) hop_t<S, L, RECON, false, false>(h, acc, jbs, opar, jb, bm, eb);
    }
    return;
  }
  if (SIG != 0) {
    hop<S, L, RECON, MU, true>(h, acc, jf, par, i, R(SIG), sc, ef);
    hop<S, L, RECON, MU, false>(h, acc, jbs, opar, jb, R(-SIG), sc, eb);
  } else {
    if (fp != R(0)) hop<S, L, RECON, MU, true>(h, acc, jf, par, i, R(1), fp, ef);
    if (fm != R(0)) hop<S, L, RECON, MU, true>(h, acc, jf, par, i, R(-1), fm, ef);
    if (bp != R(0)) hop<S, L, RECON, MU, false>(h, acc, jbs, opar, jb, R(1), bp, eb);
    if (bm != R(0)) hop<S, L, RECON, MU, false>(h, acc, jbs, opar, jb, R(-1), bm, eb);
  }
}

// Output site i of checkerboard a.par.  MULTI: the fields hold nb = 2^lognb
// right-hand sides innermost (element (site j, rhs rr) at j*nb + rr) and this
// lane computes right-hand side rr; h, out and x point at the lane's block.
// Link indices are site indices in both cases, so the lanes of one site load
// the same link words (one broadcast transaction per wavefront).
template <typename S, typename L, int RECON, int SIG, bool MULTI>
__device__ __forceinline__ void dslash_site(const DslashArgs &a, const HopIn &h, void *out, float *out_norm,
                                            const void *xin, const float *x_norm, int i, int rr, int lognb) {
  typedef typename SIO<S>::R R;
  const int nx = a.n[0], ny = a.n[1], nz = a.n[2], nt = a.n[3];
  const int r = 2 * i;
  const int x0 = r % nx;
  int q = r / nx;
  const int y = q % ny;
  q /= ny;
  const int z = q % nz;
  const int t = q / nz;
  const int x = x0 + ((a.par + y + z + t) & 1);
  const int yzt = y + ny * (z + nz * t);
  const int par = a.par;
#define SIDX(j) (MULTI ? ((j) << lognb) + rr : (j))
#define DIRARGS(M) (R) a.scl[M], (R) a.fw_p[M], (R) a.fw_m[M], (R) a.bw_p[M], (R) a.bw_m[M]

  cx<R> acc[12];
#pragma unroll
  for (int k = 0; k < 12; k++) acc[k] = mk<R>(0, 0);

  {  // x
    const int xf = x + 1 == nx ? 0 : x + 1, xb = x == 0 ? nx - 1 : x - 1;
    const R ef = x + 1 == nx ? (R)a.bcf[0] : R(1), eb = x == 0 ? (R)a.bcf[0] : R(1);
    const int jf = (xf + nx * yzt) >> 1, jb = (xb + nx * yzt) >> 1;
    direction<S, L, RECON, SIG, 0>(h, par, acc, i, SIDX(jf), jb, SIDX(jb), ef, eb, DIRARGS(0));
  }
  {  // y
    const int yf = y + 1 == ny ? 0 : y + 1, yb = y == 0 ? ny - 1 : y - 1;
    const R ef = y + 1 == ny ? (R)a.bcf[1] : R(1), eb = y == 0 ? (R)a.bcf[1] : R(1);
    const int zt = z + nz * t;
    const int jf = (x + nx * (yf + ny * zt)) >> 1, jb = (x + nx * (yb + ny * zt)) >> 1;
    direction<S, L, RECON, SIG, 1>(h, par, acc, i, SIDX(jf), jb, SIDX(jb), ef, eb, DIRARGS(1));
  }
  {  // z
    const int zf = z + 1 == nz ? 0 : z + 1, zb = z == 0 ? nz - 1 : z - 1;
    const R ef = z + 1 == nz ? (R)a.bcf[2] : R(1), eb = z == 0 ? (R)a.bcf[2] : R(1);
    const int jf = (x + nx * (y + ny * (zf + nz * t))) >> 1, jb = (x + nx * (y + ny * (zb + nz * t))) >> 1;
    direction<S, L, RECON, SIG, 2>(h, par, acc, i, SIDX(jf), jb, SIDX(jb), ef, eb, DIRARGS(2));
  }
  {  // t
    const int tf = t + 1 == nt ? 0 : t + 1, tb = t == 0 ? nt - 1 : t - 1;
    const R ef = t + 1 == nt ? (R)a.bcf[3] : R(1), eb = t == 0 ? (R)a.bcf[3] : R(1);
    const int jf = (x + nx * (y + ny * (z + nz * tf))) >> 1, jb = (x + nx * (y + ny * (z + nz * tb))) >> 1;
    direction<S, L, RECON, SIG, 3>(h, par, acc, i, SIDX(jf), jb, SIDX(jb), ef, eb, DIRARGS(3));
  }
#undef DIRARGS

  // epilogue: out = a0 x + a1 (i g5 x) + b0 D + b1 (i g5 D);  i g5 v = (v2, v3, -v0, -v1)
  const R b0 = (R)a.b0, b1 = (R)a.b1;
  cx<R> o[12];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    o[c] = b0 * acc[c] + b1 * acc[6 + c];
    o[3 + c] = b0 * acc[3 + c] + b1 * acc[9 + c];
    o[6 + c] = b0 * acc[6 + c] - b1 * acc[c];
    o[9 + c] = b0 * acc[9 + c] - b1 * acc[3 + c];
  }
  const int io = SIDX(i);
  if (a.has_x) {
    const R a0 = (R)a.a0, a1 = (R)a.a1;
    const float nrm = SIO<S>::norm(x_norm, io);
    cx<R> xv[12];
#pragma unroll
    for (int k = 0; k < 12; k++) xv[k] = SIO<S>::ld(xin, h.sstride, k, io, nrm);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      o[c] = o[c] + a0 * xv[c] + a1 * xv[6 + c];
      o[3 + c] = o[3 + c] + a0 * xv[3 + c] + a1 * xv[9 + c];
      o[6 + c] = o[6 + c] + a0 * xv[6 + c] - a1 * xv[c];
      o[9 + c] = o[9 + c] + a0 * xv[9 + c] - a1 * xv[3 + c];
    }
  }
#undef SIDX
  SIO<S>::st(out, out_norm, h.sstride, io, o);
}

template <typename S, typename L, int RECON, int SIG>
__global__ void __launch_bounds__(256) k_dslash(const DslashArgs a) {
  int blk = blockIdx.x;
  if (a.xcd_remap) {
    // workgroups are dealt round-robin to the 8 XCDs; give each XCD a
    // contiguous run of sites so neighbour reuse stays in its own L2
    const int nb = gridDim.x, per = (nb + 7) / 8;
    const int xcd = blk & 7, k = blk >> 3;
    const int full = nb - 8 * (per - 1);  // XCDs that get `per` blocks
    blk = xcd < full ? xcd * per + k : full * per + (xcd - full) * (per - 1) + k;
  }
  const int i = blk * blockDim.x + threadIdx.x;
  if (i >= a.Vh) return;
  const HopIn h{a.in, a.in_norm, a.links, a.sstride, a.lstride};
  dslash_site<S, L, RECON, SIG, false>(a, h, a.out, a.out_norm, a.x, a.x_norm, i, 0, 0);
}

// ---- pipelined hops (projector form) -----------------------------------
//
// The eight hops of a site each load one link and one (half) neighbour
// spinor.  Computed one after the other, every hop waits for its own loads,
// so a wavefront makes eight serial trips to memory.  Here the loads of hop
// k + D are issued into raw registers before hop k is computed, so D hops'
// loads are in flight per wavefront.  NT = 1 loads the links and stores the
// output non-temporally: both are touched once per launch, and streaming
// them leaves the L2 to the neighbour spinors, which are read eight times.

template <typename T> struct V2;
template <> struct V2<double> { typedef double __attribute__((ext_vector_type(2))) T; };
template <> struct V2<float> { typedef float __attribute__((ext_vector_type(2))) T; };
template <> struct V2<f16> { typedef f16 __attribute__((ext_vector_type(2))) T; };

template <typename S, typename L, int RECON> struct HopRaw {
  typename V2<L>::T u[RECON == 18 ? 9 : 6];
  typename V2<S>::T p[12];
  float nrm;
  float edge;
};

// neighbour of hop K (mu = K/2, forward for even K): checkerboard index j and
// the boundary factor of the hop
template <int K>
__device__ __forceinline__ void hop_nbr(const DslashArgs &a, int x, int y, int z, int t, int &j, float &edge) {
  constexpr int MU = K >> 1;
  constexpr bool FWD = !(K & 1);
  const int nx = a.n[0], ny = a.n[1], nz = a.n[2], nt = a.n[3];
  int c[4] = {x, y, z, t};
  const int n = a.n[MU];
  const int v = c[MU];
  edge = 1.f;
  if (FWD) {
    c[MU] = v + 1 == n ? 0 : v + 1;
    if (v + 1 == n) edge = a.bcf[MU];
  } else {
    c[MU] = v == 0 ? n - 1 : v - 1;
    if (v == 0) edge = a.bcf[MU];
  }
  (void)nt;
  j = (c[0] + nx * (c[1] + ny * (c[2] + nz * c[3]))) >> 1;
}

// Buffer access: a descriptor over a field (wavefront-uniform, in scalar
// registers), a per-lane byte offset and a per-component offset in a scalar
// register, so no load or store needs per-lane address arithmetic.  A
// descriptor covers less than 4 GiB (checked on the host).  AUX 2 = the
// non-temporal hint.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}
template <typename X, int AUX>
__device__ __forceinline__ X bld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(X) == 4) return __builtin_bit_cast(X, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, AUX));
  else if constexpr (sizeof(X) == 8)
    return __builtin_bit_cast(X, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX));
  else return __builtin_bit_cast(X, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, AUX));
}
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
template <typename X, int AUX>
__device__ __forceinline__ void bst(X v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(X) == 4) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, soff, AUX);
  else if constexpr (sizeof(X) == 8) __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, AUX);
  else __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, voff, soff, AUX);
}

template <typename S, typename L, int RECON, int SIG, bool NT, bool MULTI, int K>
__device__ __forceinline__ void hop_fetch(const DslashArgs &a, const HopIn &h, int i, int x, int y, int z, int t,
                                          int rr, int lognb, HopRaw<S, L, RECON> &r) {
  constexpr int MU = K >> 1;
  constexpr bool FWD = !(K & 1);
  constexpr int NC = RECON == 18 ? 9 : 6;
  constexpr int AUXL = NT ? 2 : 0;
  typedef typename V2<L>::T LE;
  typedef typename V2<S>::T SE;
  int j;
  hop_nbr<K>(a, x, y, z, t, j, r.edge);
  const int lpar = FWD ? a.par : (a.par ^ 1);
  const uint32_t lo = (uint32_t)(FWD ? i : j) * (uint32_t)sizeof(LE);
  const uint32_t js = MULTI ? ((uint32_t)j << lognb) + rr : (uint32_t)j;
  const uint32_t so = js * (uint32_t)sizeof(SE);
  const uint32_t lrun = (uint32_t)h.lstride * (uint32_t)sizeof(LE), srun = (uint32_t)h.sstride * (uint32_t)sizeof(SE);
  // the link block of (lpar, MU): NC runs
  const __amdgpu_buffer_rsrc_t lr =
      make_rsrc((const char *)h.links + (size_t)((lpar * 4 + MU) * NC) * lrun, NC * lrun);
#pragma unroll
  for (int k = 0; k < NC; k++) r.u[k] = bld<LE, AUXL>(lr, lo, k * lrun);
  // (1 +/- g_4) keeps the upper or the lower spin pair of the neighbour
  constexpr int s0 = MU == 3 ? (((FWD ? SIG : -SIG) > 0) ? 0 : 6) : 0;
  constexpr int ns = MU == 3 ? 6 : 12;
  const __amdgpu_buffer_rsrc_t sr = make_rsrc(h.in, 12 * srun);
#pragma unroll
  for (int k = 0; k < ns; k++) r.p[k] = bld<SE, 0>(sr, so, (s0 + k) * srun);
  if constexpr (std::is_same<S, f16>::value) r.nrm = bld<float, 0>(make_rsrc(h.in_norm, 4u * (uint32_t)h.sstride), js * 4u, 0);
  else r.nrm = 1.f;
}

template <typename R> __device__ __forceinline__ void recon_row2(cx<R> u[9]) {
  u[6] = mk<R>(u[1].re * u[5].re - u[1].im * u[5].im - (u[2].re * u[4].re - u[2].im * u[4].im),
               -(u[1].re * u[5].im + u[1].im * u[5].re - (u[2].re * u[4].im + u[2].im * u[4].re)));
  u[7] = mk<R>(u[2].re * u[3].re - u[2].im * u[3].im - (u[0].re * u[5].re - u[0].im * u[5].im),
               -(u[2].re * u[3].im + u[2].im * u[3].re - (u[0].re * u[5].im + u[0].im * u[5].re)));
  u[8] = mk<R>(u[0].re * u[4].re - u[0].im * u[4].im - (u[1].re * u[3].re - u[1].im * u[3].im),
               -(u[0].re * u[4].im + u[0].im * u[4].re - (u[1].re * u[3].im + u[1].im * u[3].re)));
}

// acc += f w and acc += f (i w) for a real f, as two fused multiply-adds each
template <typename R> __device__ __forceinline__ void fma_acc(cx<R> &acc, R f, cx<R> w) {
  acc.re = fma(f, w.re, acc.re);
  acc.im = fma(f, w.im, acc.im);
}
template <typename R> __device__ __forceinline__ void fma_acc_i(cx<R> &acc, R f, cx<R> w) {
  acc.re = fma(-f, w.im, acc.re);
  acc.im = fma(f, w.re, acc.im);
}

// The hop is linear in the neighbour spinor, so its scale factors (c1, the
// boundary factor, the FP16 per-site norm, the 2 of the t projector) are
// one scalar f applied in the accumulation, not to the loaded values.
// hop_math: acc += f (1 +/- g_mu) U p for the link u of hop K and the loaded
// neighbour components p (for a t hop, the kept spin pair in p[0..5]).
template <int SIG, int K, typename R>
__device__ __forceinline__ void hop_math(const cx<R> u[9], const cx<R> p[12], cx<R> acc[12], R f) {
  constexpr int MU = K >> 1;
  constexpr bool FWD = !(K & 1);
  cx<R> w[6];
  if constexpr (MU == 3) {
    constexpr int s0 = ((FWD ? SIG : -SIG) > 0) ? 0 : 6;
    cx<R> hh[6];
#pragma unroll
    for (int c = 0; c < 6; c++) hh[c] = p[c];
    if (FWD) {
      mv(u, hh, w);
      mv(u, hh + 3, w + 3);
    } else {
      mdv(u, hh, w);
      mdv(u, hh + 3, w + 3);
    }
    const R f2 = R(2) * f;
#pragma unroll
    for (int c = 0; c < 6; c++) fma_acc(acc[s0 + c], f2, w[c]);
  } else {
    constexpr int sg = FWD ? SIG : -SIG;
    const R sig = R(sg);
    cx<R> hh[6];
    project<MU>(p, sig, hh);
    if (FWD) {
      mv(u, hh, w);
      mv(u, hh + 3, w + 3);
    } else {
      mdv(u, hh, w);
      mdv(u, hh + 3, w + 3);
    }
    const R fs = sg > 0 ? f : -f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      fma_acc(acc[c], f, w[c]);
      fma_acc(acc[3 + c], f, w[3 + c]);
      if constexpr (MU == 0) {
        fma_acc(acc[6 + c], fs, w[3 + c]);
        fma_acc(acc[9 + c], fs, w[c]);
      } else if constexpr (MU == 1) {
        fma_acc_i(acc[6 + c], -fs, w[3 + c]);
        fma_acc_i(acc[9 + c], fs, w[c]);
      } else {
        fma_acc(acc[6 + c], fs, w[c]);
        fma_acc(acc[9 + c], -fs, w[3 + c]);
      }
    }
  }
}

template <typename S, typename L, int RECON, int SIG, int K, typename R>
__device__ __forceinline__ void hop_consume(const HopRaw<S, L, RECON> &r, cx<R> acc[12], R sc) {
  constexpr int NC = RECON == 18 ? 9 : 6;
  constexpr int NP = (K >> 1) == 3 ? 6 : 12;
  cx<R> u[9];
#pragma unroll
  for (int k = 0; k < NC; k++) u[k] = mk<R>((R)r.u[k].x, (R)r.u[k].y);
  if constexpr (RECON == 12) recon_row2(u);
  R f = sc * (R)r.edge;
  if constexpr (std::is_same<S, f16>::value) f *= (R)r.nrm;
  cx<R> p[12];
#pragma unroll
  for (int k = 0; k < NP; k++) p[k] = mk<R>((R)r.p[k].x, (R)r.p[k].y);
  hop_math<SIG, K>(u, p, acc, f);
}

// Stage boundary of the pipeline: the accumulators pass through an empty
// asm statement (no instruction is emitted), so the arithmetic of the hops
// before it is finished there, and the memory clobber keeps the loads of
// the stages after it from being hoisted above it.  Without it the compiler
// issues every load first and keeps all raw buffers live at once.
template <typename R> __device__ __forceinline__ void stage_fence(cx<R> acc[12]) {
#pragma unroll
  for (int k = 0; k < 12; k++) asm volatile("" : "+v"(acc[k].re), "+v"(acc[k].im));
  asm volatile("" ::: "memory");
}

// The other side of a stage: hop K's raw registers pass through an empty
// asm statement right before hop K is computed, so its arithmetic cannot be
// scheduled above the loads issued for later hops (the wait for hop K's
// loads lands here, with the later hops' loads already in flight).
template <typename S, typename L, int RECON> __device__ __forceinline__ void stage_pin(HopRaw<S, L, RECON> &r) {
  constexpr int NC = RECON == 18 ? 9 : 6;
  asm volatile("" ::: "memory");
#pragma unroll
  for (int k = 0; k < NC; k++) asm volatile("" : "+v"(r.u[k]));
#pragma unroll
  for (int k = 0; k < 12; k++) asm volatile("" : "+v"(r.p[k]));
  asm volatile("" : "+v"(r.nrm), "+v"(r.edge));
}

// hops K .. 7: compute hop K from its buffer, refill the buffer with hop
// K + D; once no hop is left to issue, the epilogue's local spinor is loaded
// in its place
template <typename S, typename L, int RECON, int SIG, bool NT, bool MULTI, int D, int K, typename R>
__device__ __forceinline__ void hop_pipe(const DslashArgs &a, const HopIn &h, int i, int x, int y, int z, int t,
                                         int rr, int lognb, HopRaw<S, L, RECON> *buf, cx<R> acc[12],
                                         const void *xin, const float *x_norm, int io, typename V2<S>::T *xr,
                                         float &xn) {
  if constexpr (K < 8) {
    stage_pin(buf[K % D]);
    hop_consume<S, L, RECON, SIG, K>(buf[K % D], acc, (R)a.scl[K >> 1]);
    stage_fence(acc);
    if constexpr (K + D < 8) {
      hop_fetch<S, L, RECON, SIG, NT, MULTI, K + D>(a, h, i, x, y, z, t, rr, lognb, buf[K % D]);
    } else if constexpr (K + D == 8) {
      if (a.has_x) {
        typedef typename V2<S>::T SE;
        const uint32_t srun = (uint32_t)h.sstride * (uint32_t)sizeof(SE);
        const __amdgpu_buffer_rsrc_t xrs = make_rsrc(xin, 12 * srun);
#pragma unroll
        for (int k = 0; k < 12; k++) xr[k] = bld<SE, 0>(xrs, (uint32_t)io * (uint32_t)sizeof(SE), k * srun);
        if constexpr (std::is_same<S, f16>::value)
          xn = bld<float, 0>(make_rsrc(x_norm, 4u * (uint32_t)h.sstride), (uint32_t)io * 4u, 0);
      }
    }
    hop_pipe<S, L, RECON, SIG, NT, MULTI, D, K + 1>(a, h, i, x, y, z, t, rr, lognb, buf, acc, xin, x_norm, io, xr,
                                                     xn);
  }
}

template <typename S, typename L, int RECON, int SIG, bool NT, bool MULTI, int D, int K>
__device__ __forceinline__ void hop_prefetch(const DslashArgs &a, const HopIn &h, int i, int x, int y, int z, int t,
                                             int rr, int lognb, HopRaw<S, L, RECON> *buf) {
  if constexpr (K < D) {
    hop_fetch<S, L, RECON, SIG, NT, MULTI, K>(a, h, i, x, y, z, t, rr, lognb, buf[K]);
    asm volatile("" ::: "memory");
    hop_prefetch<S, L, RECON, SIG, NT, MULTI, D, K + 1>(a, h, i, x, y, z, t, rr, lognb, buf);
  }
}

template <typename S, typename L, int RECON, int SIG, bool NT, bool MULTI, int D>
__device__ __forceinline__ void dslash_site_pipe(const DslashArgs &a, const HopIn &h, void *out, float *out_norm,
                                                 const void *xin, const float *x_norm, int i, int rr, int lognb) {
  typedef typename SIO<S>::R R;
  typedef typename V2<S>::T SE;
  const int nx = a.n[0], ny = a.n[1], nz = a.n[2];
  const int r2 = 2 * i;
  const int x0 = r2 % nx;
  int q = r2 / nx;
  const int y = q % ny;
  q /= ny;
  const int z = q % nz;
  const int t = q / nz;
  const int x = x0 + ((a.par + y + z + t) & 1);

  HopRaw<S, L, RECON> buf[D];
  hop_prefetch<S, L, RECON, SIG, NT, MULTI, D, 0>(a, h, i, x, y, z, t, rr, lognb, buf);
  const int io = MULTI ? (i << lognb) + rr : i;
  SE xr[12];
  float xn = 1.f;
  cx<R> acc[12];
#pragma unroll
  for (int k = 0; k < 12; k++) acc[k] = mk<R>(0, 0);
  hop_pipe<S, L, RECON, SIG, NT, MULTI, D, 0>(a, h, i, x, y, z, t, rr, lognb, buf, acc, xin, x_norm, io, xr, xn);

  // epilogue: out = a0 x + a1 (i g5 x) + b0 D + b1 (i g5 D);  i g5 v = (v2, v3, -v0, -v1)
  const R b0 = (R)a.b0, b1 = (R)a.b1;
  cx<R> o[12];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    o[c] = b0 * acc[c] + b1 * acc[6 + c];
    o[3 + c] = b0 * acc[3 + c] + b1 * acc[9 + c];
    o[6 + c] = b0 * acc[6 + c] - b1 * acc[c];
    o[9 + c] = b0 * acc[9 + c] - b1 * acc[3 + c];
  }
  if (a.has_x) {
    const R a0 = (R)a.a0, a1 = (R)a.a1;
    cx<R> xv[12];
#pragma unroll
    for (int k = 0; k < 12; k++) xv[k] = mk<R>((R)xn * (R)xr[k].x, (R)xn * (R)xr[k].y);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      o[c] = o[c] + a0 * xv[c] + a1 * xv[6 + c];
      o[3 + c] = o[3 + c] + a0 * xv[3 + c] + a1 * xv[9 + c];
      o[6 + c] = o[6 + c] + a0 * xv[6 + c] - a1 * xv[c];
      o[9 + c] = o[9 + c] + a0 * xv[9 + c] - a1 * xv[3 + c];
    }
  }
  constexpr int AUXS = NT ? 2 : 0;
  const uint32_t srun = (uint32_t)h.sstride * (uint32_t)sizeof(SE);
  const uint32_t oo = (uint32_t)io * (uint32_t)sizeof(SE);
  const __amdgpu_buffer_rsrc_t ors = make_rsrc(out, 12 * srun);
  if constexpr (std::is_same<S, f16>::value) {
    if (MULTI || !a.skip16) {
      float m = 0.f;
#pragma unroll
      for (int k = 0; k < 12; k++) m = fmaxf(m, fmaxf(fabsf(o[k].re), fabsf(o[k].im)));
      bst<float, AUXS>(m, make_rsrc(out_norm, 4u * (uint32_t)h.sstride), (uint32_t)io * 4u, 0);
      const float inv = m > 0.f ? 1.f / m : 0.f;
#pragma unroll
      for (int k = 0; k < 12; k++) {
        SE w;
        w.x = (f16)(o[k].re * inv);
        w.y = (f16)(o[k].im * inv);
        bst<SE, AUXS>(w, ors, oo, k * srun);
      }
    }
    if constexpr (!MULTI) {
      if (a.out32) {
        typedef typename V2<float>::T FE;
        const uint32_t frun = (uint32_t)h.sstride * (uint32_t)sizeof(FE);
        const __amdgpu_buffer_rsrc_t frs = make_rsrc(a.out32, 12 * frun);
#pragma unroll
        for (int k = 0; k < 12; k++) {
          FE w;
          w.x = o[k].re;
          w.y = o[k].im;
          bst<FE, AUXS>(w, frs, (uint32_t)io * (uint32_t)sizeof(FE), k * frun);
        }
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < 12; k++) {
      SE w;
      w.x = o[k].re;
      w.y = o[k].im;
      bst<SE, AUXS>(w, ors, oo, k * srun);
    }
  }
}

// logical lane index -> checkerboard site (see DslashArgs::order)
__device__ __forceinline__ int site_of(const DslashArgs &a, int lidx) {
  if (a.order == 0) return lidx;
  const int p = lidx / a.plane, o = lidx - p * a.plane;
  if (a.order == 2) {  // blocks of bz t-slices, swept in z
    const int per = a.bz * a.n[2];
    const int tb = p / per, r = p - tb * per;
    const int z = r / a.bz, t = tb * a.bz + (r - z * a.bz);
    return (z + a.n[2] * t) * a.plane + o;
  }
  const int per = a.bz * a.n[3];
  const int zb = p / per, r = p - zb * per;
  const int t = r / a.bz, z = zb * a.bz + (r - t * a.bz);
  return (z + a.n[2] * t) * a.plane + o;
}

template <typename S, typename L, int RECON, int SIG, bool NT, int D>
__global__ void __launch_bounds__(256) k_dslash_pipe(const DslashArgs a) {
  int blk = blockIdx.x;
  if (a.xcd_remap) {
    const int nb = gridDim.x, per = (nb + 7) / 8;
    const int xcd = blk & 7, k = blk >> 3;
    const int full = nb - 8 * (per - 1);
    blk = xcd < full ? xcd * per + k : full * per + (xcd - full) * (per - 1) + k;
  }
  const int l = blk * blockDim.x + threadIdx.x;
  if (l >= a.Vh) return;
  const int i = site_of(a, l);
  const HopIn h{a.in, a.in_norm, a.links, a.sstride, a.lstride};
  dslash_site_pipe<S, L, RECON, SIG, NT, false, D>(a, h, a.out, a.out_norm, a.x, a.x_norm, i, 0, 0);
}

// Several right-hand sides, RHS-innermost in blocks of nb = 2^lognb
// (element (rhs r, component k, site j) of a field at
// ((r/nb * 12 + k) * stride + j) * nb + r%nb).  A work tile is 256 lanes of
// one RHS block: nb lanes per site, 256/nb consecutive sites.  The kernel is
// persistent: a fixed grid (a multiple of 8 workgroups) walks the tiles of
// the *active* blocks, and the active RHS count is read from device memory
// at launch, so a launch captured once in a graph keeps working while
// converged right-hand sides are retired (moved behind the active ones).
// Tiles are dealt so that each XCD works on one contiguous range of tiles
// per sweep.  order 0: RHS block outermost; 1: RHS block innermost (the
// blocks of one site range run back to back and share the link loads in L2).
struct MultiArgs {
  const int *nact;
  int lognb;
  int order;
  int site_tiles;
  long blk_elems;  // elements of one RHS block (12 * stride * nb)
  long blk_norms;  // norms of one RHS block (stride * nb)
  int per_tile;    // k_dslash_multi_pipe: one workgroup per tile instead of a persistent grid
};

template <typename S, typename L, int RECON, int SIG>
__global__ void __launch_bounds__(256) k_dslash_multi(const DslashArgs a, const MultiArgs m) {
  typedef typename SIO<S>::E E;
  const int nact = *m.nact;
  const int nb = 1 << m.lognb;
  const int nblk = (nact + nb - 1) >> m.lognb;
  const long ntiles = (long)m.site_tiles * nblk;
  const int G = gridDim.x;
  const int vw = (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);
  for (long t = vw; t < ntiles; t += G) {
    int blk, st;
    if (m.order == 0) {
      blk = (int)(t / m.site_tiles);
      st = (int)(t - (long)blk * m.site_tiles);
    } else {
      st = (int)(t / nblk);
      blk = (int)(t - (long)st * nblk);
    }
    const int lane = st * 256 + threadIdx.x;
    const int i = lane >> m.lognb, rr = lane & (nb - 1);
    if (i >= a.Vh || (blk << m.lognb) + rr >= nact) continue;
    const long eo = (long)blk * m.blk_elems, no = (long)blk * m.blk_norms;
    const HopIn h{(const E *)a.in + eo, a.in_norm ? a.in_norm + no : nullptr, a.links, a.sstride, a.lstride};
    dslash_site<S, L, RECON, SIG, true>(a, h, (E *)a.out + eo, a.out_norm ? a.out_norm + no : nullptr,
                                        a.x ? (const E *)a.x + eo : nullptr, a.x_norm ? a.x_norm + no : nullptr, i,
                                        rr, m.lognb);
  }
}

// Pipelined form of k_dslash_multi.  The tiles (RHS block, 256/nb sites in
// the z-blocked t-sweep order of site_of) are split into eight contiguous
// ranges, one per XCD for the whole launch, which its workgroups walk in
// order: with the RHS block outermost (order 0) and nb * 8 right-hand sides,
// every XCD sweeps all sites of one block, so the t and z neighbours of its
// wavefronts were read a few slabs earlier into its own L2 and the links,
// read by all XCDs at about the same pace, are shared in the last-level
// cache.  Persistent like k_dslash_multi (the active count is read at
// launch), or (per_tile) one workgroup per tile of the full capacity, whose
// workgroups beyond the active tiles return at once: the hardware then hands
// each XCD its range in order, so the sites in flight on an XCD stay within a
// few slabs, where the workgroups of a persistent grid drift apart over the
// sweep and spread its reads beyond the L2.
template <typename S, typename L, int RECON, int SIG, bool NT, int D>
__global__ void __launch_bounds__(256) k_dslash_multi_pipe(const DslashArgs a, const MultiArgs m) {
  typedef typename V2<S>::T SE;
  const int nact = *m.nact;
  const int nb = 1 << m.lognb;
  const int nblk = (nact + nb - 1) >> m.lognb;
  const long ntiles = (long)m.site_tiles * nblk;
  const int g8 = gridDim.x >> 3;  // workgroups per XCD (the grid is a multiple of 8)
  const int xcd = blockIdx.x & 7, w = blockIdx.x >> 3;
  const long per = (ntiles + 7) >> 3;
  const long r0 = xcd * per, r1 = r0 + per < ntiles ? r0 + per : ntiles;
  // per_tile: the one tile w of the XCD's range (none when w is past it)
  const long t0 = r0 + w, t1 = m.per_tile ? (t0 + 1 < r1 ? t0 + 1 : r1) : r1, step = m.per_tile ? 1 : g8;
  const int spt = 256 >> m.lognb;  // sites per tile
  for (long t = t0; t < t1; t += step) {
    int blk, st;
    if (m.order == 0) {
      blk = (int)(t / m.site_tiles);
      st = (int)(t - (long)blk * m.site_tiles);
    } else {
      st = (int)(t / nblk);
      blk = (int)(t - (long)st * nblk);
    }
    const int lidx = st * spt + ((int)threadIdx.x >> m.lognb);
    const int rr = (int)threadIdx.x & (nb - 1);
    if (lidx >= a.Vh || (blk << m.lognb) + rr >= nact) continue;
    const int i = site_of(a, lidx);
    const long eo = (long)blk * m.blk_elems, no = (long)blk * m.blk_norms;
    const HopIn h{(const SE *)a.in + eo, a.in_norm ? a.in_norm + no : nullptr, a.links, a.sstride, a.lstride};
    dslash_site_pipe<S, L, RECON, SIG, NT, true, D>(a, h, (SE *)a.out + eo, a.out_norm ? a.out_norm + no : nullptr,
                                                    a.x ? (const SE *)a.x + eo : nullptr,
                                                    a.x_norm ? a.x_norm + no : nullptr, i, rr, m.lognb);
  }
}

// ---- two right-hand sides per lane (FP32 spinors) -------------------------
// Slots 2q and 2q+1 of an RHS block are adjacent in memory (element
// j*nb + r), so one 16-byte load fetches a component of both; the lane loads
// and rebuilds the link of each hop once for the two and keeps two
// accumulators.  Per right-hand side this halves the load instructions and
// the link and index arithmetic of the one-lane-per-slot kernel, whose
// issue rate, not its HBM traffic, bounds it.
// FP16 input or output (IN16 / OUT16): FP16 fields with one norm per (site,
// slot), a component of both slots in one 8-byte load, both norms in another.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef f16 f16x4 __attribute__((ext_vector_type(4)));

template <typename L, int RECON, bool IN16> struct HopRaw2 {
  typename V2<L>::T u[RECON == 18 ? 9 : 6];
  typename std::conditional<IN16, f16x4, f32x4>::type p[12];  // (re, im) of slot 2q, then of slot 2q+1
  typename V2<float>::T nrm;  // IN16: the norms of the two slots at the neighbour
  float edge;
};

template <typename L, int RECON, int SIG, bool NT, int K, bool IN16>
__device__ __forceinline__ void hop_fetch2(const DslashArgs &a, const HopIn &h, int i, int x, int y, int z, int t,
                                           int q2, int lognb, HopRaw2<L, RECON, IN16> &r) {
  constexpr int MU = K >> 1;
  constexpr bool FWD = !(K & 1);
  constexpr int NC = RECON == 18 ? 9 : 6;
  constexpr int AUXL = NT ? 2 : 0;
  typedef typename V2<L>::T LE;
  int j;
  hop_nbr<K>(a, x, y, z, t, j, r.edge);
  const int lpar = FWD ? a.par : (a.par ^ 1);
  constexpr uint32_t ES = IN16 ? 4u : 8u;  // bytes of one complex element
  typedef typename std::conditional<IN16, f16x4, f32x4>::type PE;
  const uint32_t lo = (uint32_t)(FWD ? i : j) * (uint32_t)sizeof(LE);
  const uint32_t je = ((uint32_t)j << lognb) + (uint32_t)q2;
  const uint32_t so = je * ES;
  const uint32_t lrun = (uint32_t)h.lstride * (uint32_t)sizeof(LE), srun = (uint32_t)h.sstride * ES;
  const __amdgpu_buffer_rsrc_t lr =
      make_rsrc((const char *)h.links + (size_t)((lpar * 4 + MU) * NC) * lrun, NC * lrun);
#pragma unroll
  for (int k = 0; k < NC; k++) r.u[k] = bld<LE, AUXL>(lr, lo, k * lrun);
  constexpr int s0 = MU == 3 ? (((FWD ? SIG : -SIG) > 0) ? 0 : 6) : 0;
  constexpr int ns = MU == 3 ? 6 : 12;
  const __amdgpu_buffer_rsrc_t sr = make_rsrc(h.in, 12 * srun);
#pragma unroll
  for (int k = 0; k < ns; k++) r.p[k] = bld<PE, 0>(sr, so, (s0 + k) * srun);
  if constexpr (IN16)
    r.nrm = bld<typename V2<float>::T, 0>(make_rsrc(h.in_norm, 4u * (uint32_t)h.sstride), je * 4u, 0);
}

template <typename L, int RECON, int SIG, int K, bool IN16>
__device__ __forceinline__ void hop_consume2(const HopRaw2<L, RECON, IN16> &r, cx<float> acc0[12],
                                             cx<float> acc1[12], float sc) {
  constexpr int NC = RECON == 18 ? 9 : 6;
  constexpr int NP = (K >> 1) == 3 ? 6 : 12;
  cx<float> u[9];
#pragma unroll
  for (int k = 0; k < NC; k++) u[k] = mk<float>((float)r.u[k].x, (float)r.u[k].y);
  if constexpr (RECON == 12) recon_row2(u);
  const float f = sc * r.edge;
  float f0 = f, f1 = f;
  if constexpr (IN16) {
    f0 *= r.nrm.x;
    f1 *= r.nrm.y;
  }
  cx<float> p[12];
#pragma unroll
  for (int k = 0; k < NP; k++) p[k] = mk<float>((float)r.p[k].x, (float)r.p[k].y);
  hop_math<SIG, K>(u, p, acc0, f0);
#pragma unroll
  for (int k = 0; k < NP; k++) p[k] = mk<float>((float)r.p[k].z, (float)r.p[k].w);
  hop_math<SIG, K>(u, p, acc1, f1);
}

template <typename L, int RECON, bool IN16> __device__ __forceinline__ void stage_pin2(HopRaw2<L, RECON, IN16> &r) {
  constexpr int NC = RECON == 18 ? 9 : 6;
  asm volatile("" ::: "memory");
#pragma unroll
  for (int k = 0; k < NC; k++) asm volatile("" : "+v"(r.u[k]));
#pragma unroll
  for (int k = 0; k < 12; k++) asm volatile("" : "+v"(r.p[k]));
  if constexpr (IN16) asm volatile("" : "+v"(r.nrm));
  asm volatile("" : "+v"(r.edge));
}

template <typename L, int RECON, int SIG, bool NT, int D, int K, bool IN16>
__device__ __forceinline__ void hop_prefetch2(const DslashArgs &a, const HopIn &h, int i, int x, int y, int z, int t,
                                              int q2, int lognb, HopRaw2<L, RECON, IN16> *buf) {
  if constexpr (K < D) {
    hop_fetch2<L, RECON, SIG, NT, K, IN16>(a, h, i, x, y, z, t, q2, lognb, buf[K]);
    asm volatile("" ::: "memory");
    hop_prefetch2<L, RECON, SIG, NT, D, K + 1, IN16>(a, h, i, x, y, z, t, q2, lognb, buf);
  }
}

// hops K .. 7 of the pair, pipelined as hop_pipe: hop K is computed from its
// buffer, which is refilled with hop K + D, or with the local pair of the
// epilogue once no hop is left
template <typename L, int RECON, int SIG, bool NT, int D, int K, bool IN16>
__device__ __forceinline__ void hop_pipe2(const DslashArgs &a, const HopIn &h, int i, int x, int y, int z, int t,
                                          int q2, int lognb, HopRaw2<L, RECON, IN16> *buf, cx<float> acc0[12],
                                          cx<float> acc1[12], const void *xin, int io, f32x4 *xr) {
  if constexpr (K < 8) {
    stage_pin2(buf[K % D]);
    hop_consume2<L, RECON, SIG, K, IN16>(buf[K % D], acc0, acc1, (float)a.scl[K >> 1]);
    stage_fence(acc0);
    stage_fence(acc1);
    if constexpr (K + D < 8) {
      hop_fetch2<L, RECON, SIG, NT, K + D, IN16>(a, h, i, x, y, z, t, q2, lognb, buf[K % D]);
    } else if constexpr (K + D == 8) {
      if (a.has_x) {
        const uint32_t srun = (uint32_t)h.sstride * 8u;
        const __amdgpu_buffer_rsrc_t xrs = make_rsrc(xin, 12 * srun);
#pragma unroll
        for (int k = 0; k < 12; k++) xr[k] = bld<f32x4, 0>(xrs, (uint32_t)io * 8u, k * srun);
      }
    }
    hop_pipe2<L, RECON, SIG, NT, D, K + 1, IN16>(a, h, i, x, y, z, t, q2, lognb, buf, acc0, acc1, xin, io, xr);
  }
}

// out = a0 x + a1 (i g5 x) + b0 D + b1 (i g5 D) for one right-hand side
__device__ __forceinline__ void epilogue_f(const DslashArgs &a, const cx<float> acc[12], const cx<float> *xv,
                                           cx<float> o[12]) {
  const float b0 = (float)a.b0, b1 = (float)a.b1;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    o[c] = b0 * acc[c] + b1 * acc[6 + c];
    o[3 + c] = b0 * acc[3 + c] + b1 * acc[9 + c];
    o[6 + c] = b0 * acc[6 + c] - b1 * acc[c];
    o[9 + c] = b0 * acc[9 + c] - b1 * acc[3 + c];
  }
  if (xv) {
    const float a0 = (float)a.a0, a1 = (float)a.a1;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      o[c] = o[c] + a0 * xv[c] + a1 * xv[6 + c];
      o[3 + c] = o[3 + c] + a0 * xv[3 + c] + a1 * xv[9 + c];
      o[6 + c] = o[6 + c] + a0 * xv[6 + c] - a1 * xv[c];
      o[9 + c] = o[9 + c] + a0 * xv[9 + c] - a1 * xv[3 + c];
    }
  }
}

// Output site i, slots q2 and q2 + 1 of the lane's RHS block (both: whether
// slot q2 + 1 is active; a retired slot is computed but not written).
template <typename L, int RECON, int SIG, bool NT, int D, bool IN16, bool OUT16>
__device__ __forceinline__ void dslash_site_pair(const DslashArgs &a, const HopIn &h, void *out, float *out_norm,
                                                 const void *xin, void *side, int i, int q2, int lognb, int slot0,
                                                 bool both) {
  const int nx = a.n[0], ny = a.n[1], nz = a.n[2];
  const int r2 = 2 * i;
  const int x0 = r2 % nx;
  int q = r2 / nx;
  const int y = q % ny;
  q /= ny;
  const int z = q % nz;
  const int t = q / nz;
  const int x = x0 + ((a.par + y + z + t) & 1);

  HopRaw2<L, RECON, IN16> buf[D];
  hop_prefetch2<L, RECON, SIG, NT, D, 0, IN16>(a, h, i, x, y, z, t, q2, lognb, buf);
  const int io = (i << lognb) + q2;
  f32x4 xr[12];
  cx<float> acc0[12], acc1[12];
#pragma unroll
  for (int k = 0; k < 12; k++) acc0[k] = acc1[k] = mk<float>(0, 0);
  hop_pipe2<L, RECON, SIG, NT, D, 0, IN16>(a, h, i, x, y, z, t, q2, lognb, buf, acc0, acc1, xin, io, xr);

  cx<float> o0[12], o1[12], xv[12];
#pragma unroll
  for (int k = 0; k < 12; k++) xv[k] = mk<float>(xr[k].x, xr[k].y);
  epilogue_f(a, acc0, a.has_x ? xv : nullptr, o0);
#pragma unroll
  for (int k = 0; k < 12; k++) xv[k] = mk<float>(xr[k].z, xr[k].w);
  epilogue_f(a, acc1, a.has_x ? xv : nullptr, o1);

  constexpr int AUXS = NT ? 2 : 0;
  const uint32_t srun = (uint32_t)h.sstride * 8u;
  const uint32_t oo = (uint32_t)io * 8u;
  if (a.slot_n2) {
    // the result is A x / ||x|| per slot; side = hs x / ||x|| in FP16
    const double n0 = a.slot_n2[slot0], n1 = both ? a.slot_n2[slot0 + 1] : 0.0;
    const float s0 = n0 > 0.0 ? (float)(1.0 / sqrt(n0)) : 0.f, s1 = n1 > 0.0 ? (float)(1.0 / sqrt(n1)) : 0.f;
#pragma unroll
    for (int k = 0; k < 12; k++) {
      o0[k] = s0 * o0[k];
      o1[k] = s1 * o1[k];
    }
    if (side) {
      typedef f16 h4 __attribute__((ext_vector_type(4)));
      const float f0 = a.side_hs * s0, f1 = a.side_hs * s1;
      const uint32_t hrun = (uint32_t)h.sstride * 4u;
      const __amdgpu_buffer_rsrc_t hrs = make_rsrc(side, 12 * hrun);
#pragma unroll
      for (int k = 0; k < 12; k++) {
        if (both) {
          h4 v;
          v.x = (f16)(f0 * xr[k].x);
          v.y = (f16)(f0 * xr[k].y);
          v.z = (f16)(f1 * xr[k].z);
          v.w = (f16)(f1 * xr[k].w);
          bst<h4, AUXS>(v, hrs, (uint32_t)io * 4u, k * hrun);
        } else {
          h2 v;
          v.x = (f16)(f0 * xr[k].x);
          v.y = (f16)(f0 * xr[k].y);
          bst<h2, AUXS>(v, hrs, (uint32_t)io * 4u, k * hrun);
        }
      }
    }
  }
  if constexpr (OUT16) {
    // per slot: norm = the largest |re|, |im| of its 24 reals, stored values / norm
    float m0 = 0.f, m1 = 0.f;
#pragma unroll
    for (int k = 0; k < 12; k++) {
      m0 = fmaxf(m0, fmaxf(fabsf(o0[k].re), fabsf(o0[k].im)));
      m1 = fmaxf(m1, fmaxf(fabsf(o1[k].re), fabsf(o1[k].im)));
    }
    const float i0 = m0 > 0.f ? 1.f / m0 : 0.f, i1 = m1 > 0.f ? 1.f / m1 : 0.f;
    const uint32_t hrun = (uint32_t)h.sstride * 4u;
    const __amdgpu_buffer_rsrc_t ors = make_rsrc(out, 12 * hrun);
    const __amdgpu_buffer_rsrc_t nrs = make_rsrc(out_norm, 4u * (uint32_t)h.sstride);
    if (both) {
      typename V2<float>::T nn;
      nn.x = m0;
      nn.y = m1;
      bst<typename V2<float>::T, AUXS>(nn, nrs, (uint32_t)io * 4u, 0);
#pragma unroll
      for (int k = 0; k < 12; k++) {
        f16x4 w;
        w.x = (f16)(o0[k].re * i0);
        w.y = (f16)(o0[k].im * i0);
        w.z = (f16)(o1[k].re * i1);
        w.w = (f16)(o1[k].im * i1);
        bst<f16x4, AUXS>(w, ors, (uint32_t)io * 4u, k * hrun);
      }
    } else {
      bst<float, AUXS>(m0, nrs, (uint32_t)io * 4u, 0);
#pragma unroll
      for (int k = 0; k < 12; k++) {
        h2 w;
        w.x = (f16)(o0[k].re * i0);
        w.y = (f16)(o0[k].im * i0);
        bst<h2, AUXS>(w, ors, (uint32_t)io * 4u, k * hrun);
      }
    }
    return;
  }
  const __amdgpu_buffer_rsrc_t ors = make_rsrc(out, 12 * srun);
  if (both) {
#pragma unroll
    for (int k = 0; k < 12; k++) {
      f32x4 w;
      w.x = o0[k].re;
      w.y = o0[k].im;
      w.z = o1[k].re;
      w.w = o1[k].im;
      bst<f32x4, AUXS>(w, ors, oo, k * srun);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 12; k++) {
      typename V2<float>::T w;
      w.x = o0[k].re;
      w.y = o0[k].im;
      bst<typename V2<float>::T, AUXS>(w, ors, oo, k * srun);
    }
  }
}

// One workgroup per tile of 256 lanes = 512/nb sites x nb/2 slot pairs of one
// RHS block (nb >= 2), tiles split into eight contiguous ranges, one per XCD,
// as k_dslash_multi_pipe with per_tile; m.site_tiles counts tiles of pairs.
template <typename L, int RECON, int SIG, bool NT, int D, bool IN16, bool OUT16>
__global__ void __launch_bounds__(256) k_dslash_multi_pair(const DslashArgs a, const MultiArgs m) {
  const int nact = *m.nact;
  const int nb = 1 << m.lognb;
  const int nblk = (nact + nb - 1) >> m.lognb;
  const long ntiles = (long)m.site_tiles * nblk;
  const int xcd = blockIdx.x & 7, w = blockIdx.x >> 3;
  const long per = (ntiles + 7) >> 3;
  const long tt = xcd * per + w;
  if (w >= per || tt >= ntiles) return;
  int blk, st;
  if (m.order == 0) {
    blk = (int)(tt / m.site_tiles);
    st = (int)(tt - (long)blk * m.site_tiles);
  } else {
    st = (int)(tt / nblk);
    blk = (int)(tt - (long)st * nblk);
  }
  const int lp = m.lognb - 1;  // log2 of the pair lanes of a site
  const int lidx = st * (256 >> lp) + ((int)threadIdx.x >> lp);
  const int q2 = ((int)threadIdx.x & ((1 << lp) - 1)) << 1;
  const int r0 = (blk << m.lognb) + q2;
  if (lidx >= a.Vh || r0 >= nact) return;
  const int i = site_of(a, lidx);
  typedef typename V2<float>::T SE;
  const long eo = (long)blk * m.blk_elems, no = (long)blk * m.blk_norms;
  const void *in = IN16 ? (const void *)((const h2 *)a.in + eo) : (const void *)((const SE *)a.in + eo);
  void *out = OUT16 ? (void *)((h2 *)a.out + eo) : (void *)((SE *)a.out + eo);
  const HopIn h{in, IN16 ? a.in_norm + no : nullptr, a.links, a.sstride, a.lstride};
  dslash_site_pair<L, RECON, SIG, NT, D, IN16, OUT16>(a, h, out, OUT16 ? a.out_norm + no : nullptr,
                                                      a.x ? (const SE *)a.x + eo : nullptr,
                                                      a.side ? (h2 *)a.side + eo : nullptr, i, q2, m.lognb, r0,
                                                      r0 + 1 < nact);
}

HopCoef make_hopcoef(const Geom &g, int idag) {
  HopCoef h{};
  const double s = idag ? -1.0 : 1.0;
  h.proj = 1;
  for (int mu = 0; mu < 4; mu++) {
    const double c1 = g.c1[mu], c2 = s * g.c2[mu];
    h.fw_p[mu] = 0.5 * (c1 - c2);
    h.fw_m[mu] = 0.5 * (c1 + c2);
    h.bw_p[mu] = 0.5 * (c1 + c2);
    h.bw_m[mu] = 0.5 * (c1 - c2);
    if (h.fw_p[mu] != 0.0 && h.fw_m[mu] != 0.0) h.proj = 0;
  }
  return h;
}

static int g_xcd_remap = 1;
void set_xcd_remap(int v) { g_xcd_remap = v; }

static DslashTune g_tune;
void set_dslash_tune(const DslashTune &t) {
  if (t.order < 0 || t.order > 2 || t.bz < 0 || (t.depth != 0 && t.depth != 1 && t.depth != 2 && t.depth != 4) ||
      t.nt < 0 || t.nt > 1)
    throw std::runtime_error("dslash tuning: order 0/1/2, bz >= 0, depth 0/1/2/4, nt 0/1");
  g_tune = t;
}
DslashTune get_dslash_tune() { return g_tune; }

template <typename S, typename L, int RECON, int SIG, bool NT>
static void launch_pipe_nt(const DslashArgs &a, int depth, int nb, hipStream_t s) {
  switch (depth) {
    case 1: hipLaunchKernelGGL((k_dslash_pipe<S, L, RECON, SIG, NT, 1>), dim3(nb), dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_dslash_pipe<S, L, RECON, SIG, NT, 2>), dim3(nb), dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL((k_dslash_pipe<S, L, RECON, SIG, NT, 4>), dim3(nb), dim3(256), 0, s, a); break;
  }
}

template <typename S, typename L, int RECON, int SIG>
static void launch_pipe(const DslashArgs &a, int depth, int nt, int nb, hipStream_t s) {
  if (nt) launch_pipe_nt<S, L, RECON, SIG, true>(a, depth, nb, s);
  else launch_pipe_nt<S, L, RECON, SIG, false>(a, depth, nb, s);
}

template <typename S, typename L, int RECON>
static bool launch_t(const DslashArgs &a0, int sig, hipStream_t s) {
  const int nb = (a0.Vh + 255) / 256;
  const DslashTune &t = g_tune;
  // the buffer descriptors of the pipelined kernel cover < 4 GiB each: a
  // spinor field (12 runs) and one (parity, mu) block of links
  const bool fits = 12.0 * a0.sstride * sizeof(typename V2<S>::T) < 4294967295.0 &&
                    (RECON == 18 ? 9.0 : 6.0) * a0.lstride * sizeof(typename V2<L>::T) < 4294967295.0;
  if (sig != 0 && t.depth > 0 && fits) {
    DslashArgs a = a0;
    a.plane = a.n[0] * a.n[1] / 2;
    int bz = t.bz;
    // order 1: blocks of bz z-planes swept in t; order 2: blocks of bz
    // t-slices swept in z (bz 0: the largest of 16, 8, 4, 2 dividing the
    // blocked extent)
    const int ext = t.order == 2 ? a.n[3] : a.n[2];
    if (bz == 0)
      for (bz = 16; bz > 1 && ext % bz; bz /= 2) {
      }
    a.order = ((t.order == 1 || t.order == 2) && bz > 1 && ext % bz == 0) ? t.order : 0;
    a.bz = a.order ? bz : 1;
    if (sig > 0) launch_pipe<S, L, RECON, 1>(a, t.depth, t.nt, nb, s);
    else launch_pipe<S, L, RECON, -1>(a, t.depth, t.nt, nb, s);
    QQCD_HIP(hipGetLastError());
    return true;
  }
  if (sig > 0)
    hipLaunchKernelGGL((k_dslash<S, L, RECON, 1>), dim3(nb), dim3(256), 0, s, a0);
  else if (sig < 0)
    hipLaunchKernelGGL((k_dslash<S, L, RECON, -1>), dim3(nb), dim3(256), 0, s, a0);
  else
    hipLaunchKernelGGL((k_dslash<S, L, RECON, 0>), dim3(nb), dim3(256), 0, s, a0);
  QQCD_HIP(hipGetLastError());
  return false;
}

// resident workgroups of a persistent kernel: occupancy x CUs, a multiple of 8
template <typename K> static int persistent_grid(K kern) {
  int dev = 0, cus = 0, per = 0;
  QQCD_HIP(hipGetDevice(&dev));
  QQCD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  QQCD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 256, 0));
  const int g = std::max(8, (cus * std::max(per, 1)) / 8 * 8);
  return g;
}

template <typename S, typename L, int RECON, int SIG>
static void launch_multi_sig(const DslashArgs &a, const MultiArgs &m, int grid, hipStream_t s) {
  static int g0 = 0;
  if (!g0) g0 = persistent_grid(k_dslash_multi<S, L, RECON, SIG>);
  const int g = grid > 0 ? (grid + 7) / 8 * 8 : g0;
  hipLaunchKernelGGL((k_dslash_multi<S, L, RECON, SIG>), dim3(g), dim3(256), 0, s, a, m);
  QQCD_HIP(hipGetLastError());
}

template <typename S, typename L, int RECON, int SIG, bool NT>
static void launch_multi_pipe(const DslashArgs &a, const MultiArgs &m, int grid, long max_tiles, int lds,
                              hipStream_t s) {
  static int g0 = 0;
  if (!g0) g0 = persistent_grid(k_dslash_multi_pipe<S, L, RECON, SIG, NT, 2>);
  // per tile: eight ranges of the capacity's tiles, each range's workgroups
  // spaced eight apart
  const long gt = 8 * ((max_tiles + 7) / 8);
  if (m.per_tile && gt > 0x7fffffffL) throw std::runtime_error("dslash_multi: too many tiles for one launch");
  const int g = m.per_tile ? (int)gt : grid > 0 ? (grid + 7) / 8 * 8 : g0;
  // lds (per tile only): unused dynamic LDS that caps the workgroups resident
  // on a CU, and with them the sites in flight on an XCD
  hipLaunchKernelGGL((k_dslash_multi_pipe<S, L, RECON, SIG, NT, 2>), dim3(g), dim3(256), m.per_tile ? lds : 0, s, a,
                     m);
  QQCD_HIP(hipGetLastError());
}

template <typename L, int RECON, int SIG, bool NT>
static void launch_multi_pair(const DslashArgs &a, const MultiArgs &m0, int depth, long max_blk, int io16,
                              hipStream_t s) {
  MultiArgs m = m0;
  m.site_tiles = (int)(((long)a.Vh * (1L << (m.lognb - 1)) + 255) / 256);
  const long gt = 8 * (((long)m.site_tiles * max_blk + 7) / 8);
  if (gt > 0x7fffffffL) throw std::runtime_error("dslash_multi: too many tiles for one launch");
  // io16: 1 FP32 in, FP16 out; 2 FP16 in, FP32 out (two hops in flight)
  if (io16 == 1)
    hipLaunchKernelGGL((k_dslash_multi_pair<L, RECON, SIG, NT, 2, false, true>), dim3((int)gt), dim3(256), 0, s, a,
                       m);
  else if (io16 == 2)
    hipLaunchKernelGGL((k_dslash_multi_pair<L, RECON, SIG, NT, 2, true, false>), dim3((int)gt), dim3(256), 0, s, a,
                       m);
  else if (depth == 1)
    hipLaunchKernelGGL((k_dslash_multi_pair<L, RECON, SIG, NT, 1, false, false>), dim3((int)gt), dim3(256), 0, s, a,
                       m);
  else
    hipLaunchKernelGGL((k_dslash_multi_pair<L, RECON, SIG, NT, 2, false, false>), dim3((int)gt), dim3(256), 0, s, a,
                       m);
  QQCD_HIP(hipGetLastError());
}

template <typename S, typename L, int RECON>
static void launch_multi(const DslashArgs &a0, const MultiArgs &m0, int sig, const MultiTune &t, long max_blk,
                         int io16, hipStream_t s) {
  const long max_tiles = (long)m0.site_tiles * max_blk;
  // the pipelined kernel: projector form, buffer descriptors of < 4 GiB (one
  // RHS block of a field, one (parity, mu) block of links)
  const bool fits = 12.0 * a0.sstride * sizeof(typename V2<S>::T) < 4294967295.0 &&
                    (RECON == 18 ? 9.0 : 6.0) * a0.lstride * sizeof(typename V2<L>::T) < 4294967295.0;
  if (t.kernel >= 1 && t.kernel <= 4 && sig != 0 && fits) {
    MultiArgs m = m0;
    // kernels 3 / 4 with FP64 spinors: the persistent form, the faster one-slot
    // form there
    m.per_tile = t.kernel == 2 || (t.kernel >= 3 && std::is_same<S, float>::value);
    DslashArgs a = a0;
    a.plane = a.n[0] * a.n[1] / 2;
    int bz = t.bz;
    if (bz <= 0 || a.n[2] % bz)  // the largest of 8, 4, 2 that divides nz
      for (bz = 8; bz > 1 && a.n[2] % bz; bz /= 2) {
      }
    a.order = bz > 1 ? 1 : 0;
    a.bz = bz > 1 ? bz : 1;
    // kernels 3 / 4: two right-hand sides per lane, FP32 spinors, nb >= 2
    if constexpr (std::is_same<S, float>::value) {
      if (t.kernel >= 3 && m.lognb >= 1) {
        const int depth = t.kernel == 3 ? 2 : 1;
        if (sig > 0) {
          if (t.nt) launch_multi_pair<L, RECON, 1, true>(a, m, depth, max_blk, io16, s);
          else launch_multi_pair<L, RECON, 1, false>(a, m, depth, max_blk, io16, s);
        } else {
          if (t.nt) launch_multi_pair<L, RECON, -1, true>(a, m, depth, max_blk, io16, s);
          else launch_multi_pair<L, RECON, -1, false>(a, m, depth, max_blk, io16, s);
        }
        return;
      }
    }
    if (a0.slot_n2 || a0.side || io16)
      throw std::runtime_error("dslash_multi: the scaled epilogue and FP16 in/out need the two-slot FP32 kernel");
    if (sig > 0) {
      if (t.nt) launch_multi_pipe<S, L, RECON, 1, true>(a, m, t.grid, max_tiles, t.lds, s);
      else launch_multi_pipe<S, L, RECON, 1, false>(a, m, t.grid, max_tiles, t.lds, s);
    } else {
      if (t.nt) launch_multi_pipe<S, L, RECON, -1, true>(a, m, t.grid, max_tiles, t.lds, s);
      else launch_multi_pipe<S, L, RECON, -1, false>(a, m, t.grid, max_tiles, t.lds, s);
    }
    return;
  }
  if (a0.slot_n2 || a0.side || io16)
    throw std::runtime_error("dslash_multi: the scaled epilogue and FP16 in/out need the two-slot FP32 kernel");
  if (sig > 0) launch_multi_sig<S, L, RECON, 1>(a0, m0, t.grid, s);
  else if (sig < 0) launch_multi_sig<S, L, RECON, -1>(a0, m0, t.grid, s);
  else launch_multi_sig<S, L, RECON, 0>(a0, m0, t.grid, s);
}

// Kernel arguments shared by the single- and multi-RHS launches; returns
// the projector sign (+1 / -1 in every direction, or 0 for the split form).
static int make_args(const Geom &g, const Gauge &U, int par_out, int idag, const Epilogue &ep, DslashArgs &a) {
  a.links = U.data;
  a.sstride = g.stride;
  a.lstride = U.stride;
  for (int m = 0; m < 4; m++) {
    a.n[m] = g.n[m];
    a.bcf[m] = (float)g.bc[m];
  }
  a.Vh = (int)g.Vh;
  a.par = par_out;
  HopCoef h = make_hopcoef(g, idag);
  for (int m = 0; m < 4; m++) {
    a.fw_p[m] = h.fw_p[m];
    a.fw_m[m] = h.fw_m[m];
    a.bw_p[m] = h.bw_p[m];
    a.bw_m[m] = h.bw_m[m];
    // projector form: forward is c1 (1 + sig g) with sig = +1 when fw_m == 0
    a.sig_f[m] = h.fw_m[m] == 0.0 ? 1.f : -1.f;
    a.scl[m] = h.fw_p[m] + h.fw_m[m];
  }
  a.a0 = ep.a0;
  a.a1 = ep.a1;
  a.b0 = ep.b0;
  a.b1 = ep.b1;
  a.xcd_remap = g_xcd_remap;
  int proj = 0;
  if (h.proj) {
    proj = a.sig_f[0] > 0 ? 1 : -1;
    for (int m = 1; m < 4; m++)
      if ((a.sig_f[m] > 0 ? 1 : -1) != proj) proj = 0;
  }
  return proj;
}

// Storage combinations: FP64 spinors with FP64 links; FP32 spinors with FP32
// or FP16 links; FP16 spinors with FP16 or FP32 links.
template <typename F>
static void dispatch_prec(Prec sp, const Gauge &U, const char *who, F &&f) {
  auto bad = [&](const char *m) { throw std::runtime_error(std::string(who) + ": " + m); };
  const bool r18 = U.recon == 18;
  if (sp == PREC_DOUBLE) {
    if (U.prec != PREC_DOUBLE) bad("FP64 spinors need FP64 links");
    if (r18) f((double *)0, (double *)0, std::integral_constant<int, 18>());
    else f((double *)0, (double *)0, std::integral_constant<int, 12>());
  } else if (sp == PREC_SINGLE) {
    if (U.prec == PREC_SINGLE) {
      if (r18) f((float *)0, (float *)0, std::integral_constant<int, 18>());
      else f((float *)0, (float *)0, std::integral_constant<int, 12>());
    } else if (U.prec == PREC_HALF) {
      if (r18) f((float *)0, (f16 *)0, std::integral_constant<int, 18>());
      else f((float *)0, (f16 *)0, std::integral_constant<int, 12>());
    } else {
      bad("FP32 spinors need FP32 or FP16 links");
    }
  } else {
    if (U.prec == PREC_HALF) {
      if (r18) f((f16 *)0, (f16 *)0, std::integral_constant<int, 18>());
      else f((f16 *)0, (f16 *)0, std::integral_constant<int, 12>());
    } else if (U.prec == PREC_SINGLE) {
      if (r18) f((f16 *)0, (float *)0, std::integral_constant<int, 18>());
      else f((f16 *)0, (float *)0, std::integral_constant<int, 12>());
    } else {
      bad("FP16 spinors need FP32 or FP16 links");
    }
  }
}

void dslash(const Geom &g, const Gauge &U, const Spinor &in, Spinor &out, int par_out, int idag,
            const Epilogue &ep, const Spinor *x, hipStream_t s, Spinor *out32, bool only32) {
  if (in.prec != out.prec || (x && x->prec != out.prec))
    throw std::runtime_error("dslash: spinor precisions differ");
  if (in.stride != g.stride || out.stride != g.stride || U.stride != g.stride)
    throw std::runtime_error("dslash: field stride does not match the geometry");
  if (in.data == out.data) throw std::runtime_error("dslash: in and out alias");
  if (out32 && (out.prec != PREC_HALF || out32->prec != PREC_SINGLE || out32->stride != g.stride ||
                out32->data == in.data || (x && out32->data == x->data)))
    throw std::runtime_error("dslash: the FP32 copy of the result is for FP16 spinors, a separate FP32 field");
  DslashArgs a{};
  a.in = in.data;
  a.in_norm = in.norm;
  a.out = out.data;
  a.out_norm = out.norm;
  a.x = x ? x->data : nullptr;
  a.x_norm = x ? x->norm : nullptr;
  a.has_x = x ? 1 : 0;
  a.out32 = out32 ? out32->data : nullptr;
  a.skip16 = out32 && only32 ? 1 : 0;
  const int proj = make_args(g, U, par_out, idag, ep, a);
  bool piped = false;
  dispatch_prec(in.prec, U, "dslash", [&](auto *sp, auto *lp, auto rc) {
    typedef typename std::remove_pointer<decltype(sp)>::type S;
    typedef typename std::remove_pointer<decltype(lp)>::type L;
    piped = launch_t<S, L, decltype(rc)::value>(a, proj, s);
  });
  // the other kernels write the FP16 result only
  if (out32 && !piped) spinor_copy_convert(g, out, *out32, s);
}

size_t MultiSpinor::bytes() const {
  const size_t e = prec == PREC_DOUBLE ? 16 : prec == PREC_SINGLE ? 8 : 4;
  return (size_t)elems() * e + (prec == PREC_HALF ? (size_t)stride * nb * nblk() * sizeof(float) : 0);
}

static int g_multi_kernel = -1, g_multi_bz = -1, g_multi_nt = -1, g_multi_lds = -1;
void set_multi_override(int kernel, int bz, int nt) {
  if (kernel < -1 || kernel > 4 || bz < -1 || nt < -1 || nt > 1)
    throw std::runtime_error("multi-RHS tuning: kernel -1..4, bz >= -1, nt -1/0/1 (-1: the caller's choice)");
  g_multi_kernel = kernel;
  g_multi_bz = bz;
  g_multi_nt = nt;
}
void set_multi_lds(int bytes) {
  int dev = 0, most = 0;
  QQCD_HIP(hipGetDevice(&dev));
  QQCD_HIP(hipDeviceGetAttribute(&most, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  if (bytes < -1 || bytes > most)
    throw std::runtime_error("multi-RHS tuning: LDS bytes -1 (the default) or 0.." + std::to_string(most));
  g_multi_lds = bytes;
}

static MultiTune multi_tune_of(const MultiTune &t0) {
  MultiTune t = t0;
  if (g_multi_kernel >= 0) t.kernel = g_multi_kernel;
  if (g_multi_bz >= 0) t.bz = g_multi_bz;
  if (g_multi_nt >= 0) t.nt = g_multi_nt;
  if (g_multi_lds >= 0) t.lds = g_multi_lds;
  return t;
}

bool dslash_multi_can_scale(const Geom &g, Prec p, int nb, const MultiTune &t0) {
  const MultiTune t = multi_tune_of(t0);
  // the launch conditions of launch_multi's two-slot branch: projector form,
  // FP32 spinors, nb >= 2, buffer descriptors of < 4 GiB (links: at most 9
  // FP32 complex runs per (parity, mu) block)
  const HopCoef h = make_hopcoef(g, 0);
  bool proj = h.proj != 0;
  if (proj) {
    const bool s0 = h.fw_m[0] == 0.0;
    for (int mu = 1; mu < 4; mu++)
      if ((h.fw_m[mu] == 0.0) != s0) proj = false;
  }
  const double sst = (double)g.stride * nb;
  return proj && p == PREC_SINGLE && nb >= 2 && (t.kernel == 3 || t.kernel == 4) && 12.0 * sst * 8 < 4294967295.0 &&
         9.0 * g.stride * 8 < 4294967295.0;
}

void dslash_multi(const Geom &g, const Gauge &U, const MultiSpinor &in, MultiSpinor &out, int par_out, int idag,
                  const Epilogue &ep, const MultiSpinor *x, const int *d_nact, const MultiTune &t0, hipStream_t s,
                  const double *slot_n2, MultiSpinor *side, double side_hs) {
  const MultiTune t = multi_tune_of(t0);
  // FP16 (with norms) on one side of an FP32 hop: the two-slot kernel's
  // FP16 intermediate of the Schur operator (io16 1: FP32 -> FP16, 2: FP16 -> FP32)
  int io16 = 0;
  if (in.prec == PREC_SINGLE && out.prec == PREC_HALF && out.norm && !x) io16 = 1;
  if (in.prec == PREC_HALF && in.norm && out.prec == PREC_SINGLE) io16 = 2;
  if ((in.prec != out.prec && !io16) || (x && x->prec != out.prec))
    throw std::runtime_error("dslash_multi: spinor precisions differ");
  if (io16 && !dslash_multi_can_scale(g, PREC_SINGLE, in.nb, t0))
    throw std::runtime_error("dslash_multi: FP16 in or out needs the two-slot FP32 kernel");
  if (in.stride != g.stride || out.stride != g.stride || U.stride != g.stride)
    throw std::runtime_error("dslash_multi: field stride does not match the geometry");
  if (in.nb != out.nb || in.cap != out.cap || (x && (x->nb != in.nb || x->cap != in.cap)))
    throw std::runtime_error("dslash_multi: right-hand-side layouts differ");
  if (in.nb < 1 || in.nb > 256 || (in.nb & (in.nb - 1)) || (1 << in.lognb) != in.nb)
    throw std::runtime_error("dslash_multi: RHS block must be a power of two <= 256");
  if (in.data == out.data) throw std::runtime_error("dslash_multi: in and out alias");
  if (!d_nact) throw std::runtime_error("dslash_multi: no active count");
  DslashArgs a{};
  a.in = in.data;
  a.in_norm = in.norm;
  a.out = out.data;
  a.out_norm = out.norm;
  a.x = x ? x->data : nullptr;
  a.x_norm = x ? x->norm : nullptr;
  a.has_x = x ? 1 : 0;
  const int proj = make_args(g, U, par_out, idag, ep, a);
  a.sstride = g.stride * in.nb;
  a.slot_n2 = slot_n2;
  a.side = nullptr;
  a.side_hs = 0.f;
  if (side) {
    if (!slot_n2 || !x) throw std::runtime_error("dslash_multi: the FP16 side output needs slot norms and x");
    if (side->prec != PREC_HALF || side->norm || side->nb != in.nb || side->cap != in.cap ||
        side->stride != in.stride)
      throw std::runtime_error("dslash_multi: the side output is an FP16 field without norms in the same layout");
    a.side = side->data;
    a.side_hs = (float)side_hs;
  }
  MultiArgs m{};
  m.nact = d_nact;
  m.lognb = in.lognb;
  m.order = t.order;
  m.site_tiles = (int)((g.Vh * in.nb + 255) / 256);
  m.blk_elems = in.blk_elems();
  m.blk_norms = g.stride * in.nb;
  if (in.prec == PREC_HALF && !io16) throw std::runtime_error("dslash_multi: FP64 or FP32 right-hand sides");
  dispatch_prec(io16 ? PREC_SINGLE : in.prec, U, "dslash_multi", [&](auto *sp, auto *lp, auto rc) {
    typedef typename std::remove_pointer<decltype(sp)>::type S;
    typedef typename std::remove_pointer<decltype(lp)>::type L;
    if constexpr (!std::is_same<S, f16>::value)
      launch_multi<S, L, decltype(rc)::value>(a, m, proj, t, (long)in.nblk(), io16, s);
  });
}

}  // namespace qqcd
