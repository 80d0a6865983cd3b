// qqcd_nn.hip -- parallel-transport convolution layers of the preconditioner
// network: forward / adjoint layer and weight gradient, BF16 MFMA with FP32
// accumulation (CDNA4 v_mfma_f32_32x32x16_bf16).
//
// A layer is a GEMM per site after the gauge transports:
//   out[(x, colour)][(a, s)] = sum_{(p, b, s')} T[(x, colour)][(p, b, s')] W[a][b][p][s][s'],
// rows = (site, colour), K = (path, input channel, spin), N = (output channel,
// spin); complex numbers are real-ified as [re | im] blocks:
//   [out_re | out_im] = [T_re | T_im] [[W_re^T, W_im^T], [-W_im^T, W_re^T]].
// The transports U_p psi(x+p) (3x3 colour matrix times 4 spins, FP32 on the
// vector units, straight from the link field) are written to LDS as BF16
// in 32-column K chunks and multiplied on the matrix cores with the weight
// matrix, which sits in LDS for the life of the workgroup.  A wavefront owns
// 32 consecutive sites (one 32x32 output block per colour and per 32 output
// columns) and walks the K chunks; the workgroup (8 wavefronts) is
// persistent over the lattice.
//
// K chunk ch (32 real columns):
//   C_in = 8: path ch/2, input channels 4 (ch%2) .. 4 (ch%2) + 3, column
//             (b%4)*4 + s' (re) and 16 + (b%4)*4 + s' (im);
//   C_in = 1: paths 4 ch .. 4 ch + 3 (paths >= 9 are zero padding), column
//             8 pp + s' (re) and 8 pp + 4 + s' (im).
// The weight gradient uses the same chunks with the roles swapped:
//   P[ch] = X^T Y_ch, X = [g_re | g_im] (rows (colour, site), output columns),
// accumulated per workgroup and summed over workgroups in a fixed order.
#include <algorithm>
#include <cstring>
#include <random>
#include <stdexcept>

#include "qqcd_nn.h"

namespace qqcd {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

static constexpr int NNW = 8;  // wavefronts per workgroup

__device__ __forceinline__ float bf_lo(uint32_t v) { return __uint_as_float(v << 16); }
__device__ __forceinline__ float bf_hi(uint32_t v) { return __uint_as_float(v & 0xFFFF0000u); }
// (re, im) -> packed BF16 pair, hardware round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t pack_bf(float re, float im) {
  bf16x2 v;
  v.x = (__bf16)re;
  v.y = (__bf16)im;
  return __builtin_bit_cast(uint32_t, v);
}
// a.x * b.x + a.y * b.y + c with BF16 pairs, FP32 accumulation (v_dot2c_f32_bf16)
__device__ __forceinline__ float dot2(uint32_t a, uint32_t b, float c) {
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, a), __builtin_bit_cast(bf16x2, b), c, false);
}
__device__ __forceinline__ uint32_t swap16(uint32_t v) { return (v >> 16) | (v << 16); }

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct NNGeo {
  int n[4];
  int Vh;
  long stride;   // network-field site stride per parity
  long lstride;  // link component stride
  float bcf[4];
  int src_even;  // the input field is zero on the odd sites: transports from them are skipped
};

template <int CIN> struct ChunkMap {
  static constexpr int NCH = CIN == 8 ? 18 : 3;
  static constexpr int KP = NCH * 32;
};

// Neighbour and link of path pe at site (par, cb): spinor site (npar, ncb),
// link site (lpar, lcb) and direction mu, backward flag, edge factor f.
struct Hop {
  int npar, ncb, lpar, lcb, mu;
  bool back;
  float f;
};
__device__ __forceinline__ Hop hop_of(const NNGeo &g, int par, int cb, const int x[4], int pe) {
  Hop h{par, cb, par, cb, 0, false, 1.f};
  if (pe > 0) {
    h.mu = (pe - 1) >> 1;
    h.back = (pe - 1) & 1;
    int y[4] = {x[0], x[1], x[2], x[3]};
    const int mu = h.mu;
    if (!h.back) {
      if (x[mu] + 1 == g.n[mu]) {
        y[mu] = 0;
        h.f = g.bcf[mu];
      } else {
        y[mu] = x[mu] + 1;
      }
    } else {
      if (x[mu] == 0) {
        y[mu] = g.n[mu] - 1;
        h.f = g.bcf[mu];
      } else {
        y[mu] = x[mu] - 1;
      }
    }
    h.npar = par ^ 1;
    h.ncb = (y[0] + g.n[0] * (y[1] + g.n[1] * (y[2] + g.n[2] * y[3]))) >> 1;
    if (h.back) {
      h.lpar = h.npar;
      h.lcb = h.ncb;
    }
  }
  return h;
}

// Raw operands of one transport: the neighbour spinor, the link, the edge
// factor and the kind (0: zero, 1: identity, 2: forward hop, 3: backward hop).
struct TRaw {
  uint32_t p[12];
  uint32_t u[9];
  float f;
  int kind;
};

__device__ __forceinline__ void transport_load(const NNGeo &g, const uint32_t *in, int C, const uint32_t *links,
                                               int valid, int par, int cb, const int x[4], int pe, int b, TRaw &r) {
  // every load is issued on every path (an unused path reads valid
  // addresses and is masked), so the count of loads in flight is static
  const bool ok0 = valid && pe < NN_PATHS;
  const int pe2 = ok0 ? pe : 0;
  const Hop h = hop_of(g, par, cb, x, pe2);
  const bool ok = ok0 && !(g.src_even && h.npar == 1);
  const uint4 *src = (const uint4 *)(in + ((long)(h.npar * g.stride + h.ncb) * C + b) * 12);
  const uint4 w0 = src[0], w1 = src[1], w2 = src[2];
  const uint32_t q[12] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w, w2.x, w2.y, w2.z, w2.w};
  const uint32_t keep = ok ? 0xFFFFFFFFu : 0u;
#pragma unroll
  for (int k = 0; k < 12; k++) r.p[k] = q[k] & keep;
  const long lb = (long)((h.lpar * 4 + h.mu) * 9) * g.lstride + h.lcb;
#pragma unroll
  for (int k = 0; k < 9; k++) r.u[k] = links[lb + (long)k * g.lstride];
  r.f = h.f;
  r.kind = !ok ? 0 : (pe2 == 0 ? 1 : (h.back ? 3 : 2));
}

__device__ __forceinline__ void transport_compute(const TRaw &r, uint32_t t[12]) {
  if (r.kind <= 1) {
#pragma unroll
    for (int k = 0; k < 12; k++) t[k] = r.p[k];  // zero or identity (the zero case loaded zeros)
    return;
  }
  const bool back = r.kind == 3;
  uint32_t ore[9], oim[9];
#pragma unroll
  for (int rr = 0; rr < 3; rr++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const uint32_t e = back ? r.u[3 * c + rr] : r.u[3 * rr + c];
      ore[3 * rr + c] = back ? e : (e ^ 0x80000000u);
      oim[3 * rr + c] = back ? (swap16(e) ^ 0x00008000u) : swap16(e);
    }
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int rr = 0; rr < 3; rr++) {
      float re = 0.f, im = 0.f;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        re = dot2(ore[3 * rr + c], r.p[3 * s + c], re);
        im = dot2(oim[3 * rr + c], r.p[3 * s + c], im);
      }
      t[3 * s + rr] = pack_bf(r.f * re, r.f * im);
    }
}

// The transported channel b, T_pe psi_b at site (par, cb), as 12 packed BF16
// complex numbers (spin-major).  Links are BF16 complex pairs at
// ((par*4 + mu)*9 + 3 r + c)*lstride + site; the colour products run as
// BF16 dot products with FP32 accumulation.  pe >= 9 or invalid sites give zero.
__device__ __forceinline__ void transport(const NNGeo &g, const uint32_t *in, int C, const uint32_t *links,
                                          int valid, int par, int cb, const int x[4], int pe, int b, uint32_t t[12]) {
  if (!valid || pe >= NN_PATHS) {
#pragma unroll
    for (int k = 0; k < 12; k++) t[k] = 0u;
    return;
  }
  const Hop h = hop_of(g, par, cb, x, pe);
  const uint4 *src = (const uint4 *)(in + ((long)(h.npar * g.stride + h.ncb) * C + b) * 12);
  const uint4 w0 = src[0], w1 = src[1], w2 = src[2];
  const uint32_t p[12] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w, w2.x, w2.y, w2.z, w2.w};
  if (pe == 0) {
#pragma unroll
    for (int k = 0; k < 12; k++) t[k] = p[k];
    return;
  }
  uint32_t u[9];
  const long lb = (long)((h.lpar * 4 + h.mu) * 9) * g.lstride + h.lcb;
#pragma unroll
  for (int k = 0; k < 9; k++) u[k] = links[lb + (long)k * g.lstride];
  // operands: forward m = U[r][c]: re = (ur, -ui).v, im = (ui, ur).v
  //           backward m = conj(U[c][r]): re = (ur, ui).v, im = (-ui, ur).v
  uint32_t ore[9], oim[9];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const uint32_t e = h.back ? u[3 * c + r] : u[3 * r + c];
      ore[3 * r + c] = h.back ? e : (e ^ 0x80000000u);
      oim[3 * r + c] = h.back ? (swap16(e) ^ 0x00008000u) : swap16(e);
    }
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int r = 0; r < 3; r++) {
      float re = 0.f, im = 0.f;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        re = dot2(ore[3 * r + c], p[3 * s + c], re);
        im = dot2(oim[3 * r + c], p[3 * s + c], im);
      }
      t[3 * s + r] = pack_bf(h.f * re, h.f * im);
    }
}

__device__ __forceinline__ void site_of(const NNGeo &g, long q, int &valid, int &par, int &cb, int x[4]) {
  valid = q < 2L * g.Vh;
  const long qq = valid ? q : 0;
  par = qq >= g.Vh ? 1 : 0;
  cb = (int)(qq - (long)par * g.Vh);
  const int r = 2 * cb;
  const int x0 = r % g.n[0];
  int t = r / g.n[0];
  x[1] = t % g.n[1];
  t /= g.n[1];
  x[2] = t % g.n[2];
  x[3] = t / g.n[2];
  x[0] = x0 + ((par + x[1] + x[2] + x[3]) & 1);
}

// item j (0, 1) of this lane in chunk ch: (path, channel, first complex K
// column kc0; spin s' sits at complex column kc0 + s', i.e. real columns
// 2 (kc0 + s') (re) and 2 (kc0 + s') + 1 (im))
template <int CIN>
__device__ __forceinline__ void chunk_item(int ch, int lane, int j, int &pg, int &b, int &kc0) {
  if (CIN == 8) {
    pg = ch >> 1;
    b = 4 * (ch & 1) + ((lane >> 5) << 1) + j;
    kc0 = (b & 3) * 4;
  } else {
    const int pp = ((lane >> 5) << 1) + j;
    pg = 4 * ch + pp;
    b = 0;
    kc0 = 4 * pp;
  }
}

struct PtcArgs {
  const uint32_t *in;
  uint32_t *out;
  const uint32_t *links;  // BF16 complex
  const uint16_t *B;  // [NP][KP] bf16
  NNGeo g;
  int adjoint;
};

template <int CIN, int COUT>
__global__ void __launch_bounds__(64 * NNW) k_ptc(const PtcArgs a) {
  constexpr int KP = ChunkMap<CIN>::KP, NCH = ChunkMap<CIN>::NCH, KPP = KP + 8;
  constexpr int NC = 4 * COUT, NP = COUT == 8 ? 64 : 32, NB = NP / 32;
  constexpr int CKP = 40;  // padded row of an A chunk (32 + 8)
  __shared__ __attribute__((aligned(16))) uint16_t sB[NP * KPP];
  __shared__ __attribute__((aligned(16))) uint16_t sA[NNW][3][32][CKP];
  for (int i = threadIdx.x; i < NP * KP / 8; i += 64 * NNW) {
    const int row = (i * 8) / KP, col = (i * 8) % KP;
    *(uint4 *)&sB[row * KPP + col] = ((const uint4 *)a.B)[i];
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const NNGeo &g = a.g;
  const long nsites = 2L * g.Vh;
  const long ntile = (nsites + 32 * NNW - 1) / (32 * NNW);
  const int row = lane & 31;
  for (long wt = blockIdx.x; wt < ntile; wt += gridDim.x) {
    const long q0 = wt * 32 * NNW + w * 32;
    int valid, par, cb, x[4];
    site_of(g, q0 + row, valid, par, cb, x);
    f32x16 acc[3][NB];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int nb = 0; nb < NB; nb++)
#pragma unroll
        for (int i = 0; i < 16; i++) acc[c][nb][i] = 0.f;
    // software pipeline: the gathers of chunk ch+1 are issued before the
    // transports and matrix products of chunk ch
    auto load_chunk = [&](int ch, TRaw (&r)[2]) {
#pragma unroll
      for (int j = 0; j < 2; j++) {
        int pg, b, kc0;
        chunk_item<CIN>(ch, lane, j, pg, b, kc0);
        int pe = pg;
        if (a.adjoint && pg > 0 && pg < NN_PATHS) pe = (pg & 1) ? pg + 1 : pg - 1;
        transport_load(g, a.in, CIN, a.links, valid, par, cb, x, pe, b, r[j]);
      }
    };
    TRaw cur[2], nxt[2];
    load_chunk(0, nxt);
    for (int ch = 0; ch < NCH; ch++) {
#pragma unroll
      for (int j = 0; j < 2; j++) cur[j] = nxt[j];
      if (ch + 1 < NCH) load_chunk(ch + 1, nxt);
#pragma unroll
      for (int j = 0; j < 2; j++) {
        int pg, b, kc0;
        chunk_item<CIN>(ch, lane, j, pg, b, kc0);
        uint32_t t[12];
        transport_compute(cur[j], t);
        // the four spins of one colour are four consecutive complex K columns: one 16-byte store
#pragma unroll
        for (int c = 0; c < 3; c++)
          *(uint4 *)&sA[w][c][row][2 * kc0] = make_uint4(t[c], t[3 + c], t[6 + c], t[9 + c]);
      }
      wave_sync();
#pragma unroll
      for (int ks = 0; ks < 2; ks++) {
        const int kk = ks * 16 + (lane >> 5) * 8;
        bf16x8 bv[NB];
#pragma unroll
        for (int nb = 0; nb < NB; nb++) bv[nb] = *(const bf16x8 *)&sB[(nb * 32 + row) * KPP + ch * 32 + kk];
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const bf16x8 av = *(const bf16x8 *)&sA[w][c][row][kk];
#pragma unroll
          for (int nb = 0; nb < NB; nb++)
            acc[c][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv[nb], acc[c][nb], 0, 0, 0);
        }
      }
      wave_sync();
    }
    // output: acc[c][nb][i] is D[r][n], r = 8 (i/4) + 4 (lane/32) + i%4, n = nb*32 + lane%32
    const int n = lane & 31;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int r = 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
      int v2, p2, cb2, x2[4];
      site_of(g, q0 + r, v2, p2, cb2, x2);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        float re, im;
        bool st;
        if (COUT == 8) {
          re = acc[c][0][i];
          im = acc[c][NB - 1][i];
          st = true;
        } else {
          re = acc[c][0][i];
          im = __shfl(acc[c][0][i], (lane & 32) + ((n + 4) & 31), 64);
          st = n < NC;
        }
        if (st && v2) {
          const int ch_out = n >> 2, s = n & 3;
          a.out[((long)(p2 * g.stride + cb2) * COUT + ch_out) * 12 + 3 * s + c] = pack_bf(re, im);
        }
      }
    }
  }
}

// ---- C_in = 8: both items of a lane share site and path ----------------------
// In chunk ch every lane transports channels 4 (ch%2) + 2 (lane/32) + {0, 1}
// of its site along path ch/2 (reversed for the adjoint layer), so the path
// is the same for the whole wavefront: the hop direction is a scalar branch
// and one set of link operands serves both channels.
struct Raw8 {
  uint32_t p0[12], p1[12], u[9];
  float f;
};

__device__ __forceinline__ int chunk_path8(int ch, int adjoint) {
  const int pg = ch >> 1;
  return (adjoint && pg > 0) ? ((pg & 1) ? pg + 1 : pg - 1) : pg;
}

__device__ __forceinline__ void load8(const NNGeo &g, const uint32_t *in, const uint32_t *links, int valid, int par,
                                      int cb, const int x[4], int ch, int adjoint, int lane, Raw8 &r) {
  const int pe = chunk_path8(ch, adjoint);
  const int b0 = 4 * (ch & 1) + ((lane >> 5) << 1);
  const Hop h = hop_of(g, par, cb, x, pe);
  const uint4 *src = (const uint4 *)(in + ((long)(h.npar * g.stride + h.ncb) * 8 + b0) * 12);
  uint4 w[6];
#pragma unroll
  for (int k = 0; k < 6; k++) w[k] = src[k];  // channels b0 and b0 + 1 are adjacent
  const uint32_t q[24] = {w[0].x, w[0].y, w[0].z, w[0].w, w[1].x, w[1].y, w[1].z, w[1].w,
                          w[2].x, w[2].y, w[2].z, w[2].w, w[3].x, w[3].y, w[3].z, w[3].w,
                          w[4].x, w[4].y, w[4].z, w[4].w, w[5].x, w[5].y, w[5].z, w[5].w};
  // invalid sites and fixed boundaries (f = 0) give zero, antiperiodic
  // edges (f = -1) flip both halves' signs: folded into the loaded spinors
  const uint32_t keep = (valid && h.f != 0.f) ? 0xFFFFFFFFu : 0u;
  const uint32_t flip = h.f < 0.f ? 0x80008000u : 0u;
#pragma unroll
  for (int k = 0; k < 12; k++) {
    r.p0[k] = (q[k] & keep) ^ (flip & keep);
    r.p1[k] = (q[12 + k] & keep) ^ (flip & keep);
  }
  r.f = 1.f;
  // loaded unconditionally (path 0 reads the site's own mu = 0 link and does
  // not use it) so that the number of loads in flight is the same on every
  // path and the wait before the next chunk covers only this chunk's loads
  const long lb = (long)((h.lpar * 4 + h.mu) * 9) * g.lstride + h.lcb;
#pragma unroll
  for (int k = 0; k < 9; k++) r.u[k] = links[lb + (long)k * g.lstride];
}

template <bool BACK>
__device__ __forceinline__ void hop8(const Raw8 &r, uint32_t t0[12], uint32_t t1[12]) {
  uint32_t ore[9], oim[9];
#pragma unroll
  for (int rr = 0; rr < 3; rr++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const uint32_t e = BACK ? r.u[3 * c + rr] : r.u[3 * rr + c];
      ore[3 * rr + c] = BACK ? e : (e ^ 0x80000000u);
      oim[3 * rr + c] = BACK ? (swap16(e) ^ 0x00008000u) : swap16(e);
    }
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int rr = 0; rr < 3; rr++) {
      float a = dot2(ore[3 * rr], r.p0[3 * s], 0.f), b = dot2(oim[3 * rr], r.p0[3 * s], 0.f);
      float c2 = dot2(ore[3 * rr], r.p1[3 * s], 0.f), d = dot2(oim[3 * rr], r.p1[3 * s], 0.f);
#pragma unroll
      for (int c = 1; c < 3; c++) {
        a = dot2(ore[3 * rr + c], r.p0[3 * s + c], a);
        b = dot2(oim[3 * rr + c], r.p0[3 * s + c], b);
        c2 = dot2(ore[3 * rr + c], r.p1[3 * s + c], c2);
        d = dot2(oim[3 * rr + c], r.p1[3 * s + c], d);
      }
      t0[3 * s + rr] = pack_bf(a, b);
      t1[3 * s + rr] = pack_bf(c2, d);
    }
}

template <int COUT>
__global__ void __launch_bounds__(64 * NNW) k_ptc8(const PtcArgs a) {
  constexpr int KP = ChunkMap<8>::KP, NCH = ChunkMap<8>::NCH, KPP = KP + 8;
  constexpr int NP = COUT == 8 ? 64 : 32, NB = NP / 32, NC = 4 * COUT;
  constexpr int CKP = 40;
  __shared__ __attribute__((aligned(16))) uint16_t sB[NP * KPP];
  __shared__ __attribute__((aligned(16))) uint16_t sA[NNW][3][32][CKP];
  for (int i = threadIdx.x; i < NP * KP / 8; i += 64 * NNW) {
    const int row = (i * 8) / KP, col = (i * 8) % KP;
    *(uint4 *)&sB[row * KPP + col] = ((const uint4 *)a.B)[i];
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, row = lane & 31;
  const NNGeo &g = a.g;
  const long nsites = 2L * g.Vh;
  const long ntile = (nsites + 32 * NNW - 1) / (32 * NNW);
  for (long wt = blockIdx.x; wt < ntile; wt += gridDim.x) {
    const long q0 = wt * 32 * NNW + w * 32;
    int valid, par, cb, x[4];
    site_of(g, q0 + row, valid, par, cb, x);
    f32x16 acc[3][NB];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int nb = 0; nb < NB; nb++)
#pragma unroll
        for (int i = 0; i < 16; i++) acc[c][nb][i] = 0.f;
    auto process = [&](int ch, const Raw8 &r) {
      const int kc0 = ((lane >> 5) << 1) * 4;  // channel b0 % 4 = 2 (lane/32)
      uint32_t t0[12], t1[12];
      const int pe = __builtin_amdgcn_readfirstlane(chunk_path8(ch, a.adjoint));
      if (pe == 0) {
#pragma unroll
        for (int k = 0; k < 12; k++) {
          t0[k] = r.p0[k];
          t1[k] = r.p1[k];
        }
      } else if ((pe - 1) & 1) {
        hop8<true>(r, t0, t1);
      } else {
        hop8<false>(r, t0, t1);
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
        *(uint4 *)&sA[w][c][row][2 * kc0] = make_uint4(t0[c], t0[3 + c], t0[6 + c], t0[9 + c]);
        *(uint4 *)&sA[w][c][row][2 * kc0 + 8] = make_uint4(t1[c], t1[3 + c], t1[6 + c], t1[9 + c]);
      }
      wave_sync();
#pragma unroll
      for (int ks = 0; ks < 2; ks++) {
        const int kk = ks * 16 + (lane >> 5) * 8;
        bf16x8 bv[NB];
#pragma unroll
        for (int nb = 0; nb < NB; nb++) bv[nb] = *(const bf16x8 *)&sB[(nb * 32 + row) * KPP + ch * 32 + kk];
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const bf16x8 av = *(const bf16x8 *)&sA[w][c][row][kk];
#pragma unroll
          for (int nb = 0; nb < NB; nb++)
            acc[c][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv[nb], acc[c][nb], 0, 0, 0);
        }
      }
      wave_sync();
    };
    // ping-pong software pipeline: the gathers of the next chunk are in
    // flight while the current one is transported and multiplied
    Raw8 ra, rb;
    load8(g, a.in, a.links, valid, par, cb, x, 0, a.adjoint, lane, ra);
#pragma unroll 1
    for (int ch = 0; ch < NCH; ch += 2) {
      load8(g, a.in, a.links, valid, par, cb, x, ch + 1, a.adjoint, lane, rb);
      process(ch, ra);
      if (ch + 2 < NCH) load8(g, a.in, a.links, valid, par, cb, x, ch + 2, a.adjoint, lane, ra);
      process(ch + 1, rb);
    }
    const int n = lane & 31;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int r = 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
      int v2, p2, cb2, x2[4];
      site_of(g, q0 + r, v2, p2, cb2, x2);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        float re, im;
        bool st;
        if (COUT == 8) {
          re = acc[c][0][i];
          im = acc[c][NB - 1][i];
          st = true;
        } else {
          re = acc[c][0][i];
          im = __shfl(acc[c][0][i], (lane & 32) + ((n + 4) & 31), 64);
          st = n < NC;
        }
        if (st && v2)
          a.out[((long)(p2 * g.stride + cb2) * COUT + (n >> 2)) * 12 + 3 * (n & 3) + c] = pack_bf(re, im);
      }
    }
  }
}

// ---- 8 -> 1 layers: mix first, then gather ----------------------------------
// out(x) = sum_p T_p phi_p(x), phi_p(y) = sum_b W[0][b][p] psi_b(y): the
// matrix cores form the nine path fields phi_p of every site (K = 8 channels
// x 4 spins, N = 9 paths x 4 spins, complex, interleaved re / im), then one
// gauge transport per path gathers them; 8x fewer transports than the
// transport-first order for a single output channel.
struct MixArgs {
  const uint32_t *in;  // 8 channels
  uint32_t *phi;       // 9 path fields
  const uint16_t *B;   // [96][64] bf16
  NNGeo g;
  int even_out;        // only the even sites' output is wanted: phi_0 on even, phi_1..8 on odd sites
};

__global__ void __launch_bounds__(64 * NNW) k_ptc_mix(const MixArgs a) {
  constexpr int KP = 64, KPP = KP + 8, NP = 96, NB = 3;
  __shared__ __attribute__((aligned(16))) uint16_t sB[NP * KPP];
  __shared__ __attribute__((aligned(16))) uint16_t sA[NNW][3][32][KPP];
  for (int i = threadIdx.x; i < NP * KP / 8; i += 64 * NNW) {
    const int row = (i * 8) / KP, col = (i * 8) % KP;
    *(uint4 *)&sB[row * KPP + col] = ((const uint4 *)a.B)[i];
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, row = lane & 31;
  const NNGeo &g = a.g;
  const long nsites = 2L * g.Vh;
  const long ntile = (nsites + 32 * NNW - 1) / (32 * NNW);
  for (long wt = blockIdx.x; wt < ntile; wt += gridDim.x) {
    const long q0 = wt * 32 * NNW + w * 32;
    int valid, par, cb, x[4];
    site_of(g, q0 + row, valid, par, cb, x);
    // stage psi: lane (site, channel half) -> sA[c][site][2 (b*4 + s') + im]
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int b = (lane >> 5) * 4 + j;
      uint32_t p[12];
      if (valid) {
        const uint4 *src = (const uint4 *)(a.in + ((long)(par * g.stride + cb) * 8 + b) * 12);
        const uint4 w0 = src[0], w1 = src[1], w2 = src[2];
        const uint32_t q[12] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w, w2.x, w2.y, w2.z, w2.w};
#pragma unroll
        for (int k = 0; k < 12; k++) p[k] = q[k];
      } else {
#pragma unroll
        for (int k = 0; k < 12; k++) p[k] = 0u;
      }
#pragma unroll
      for (int c = 0; c < 3; c++)
        *(uint4 *)&sA[w][c][row][8 * b] = make_uint4(p[c], p[3 + c], p[6 + c], p[9 + c]);
    }
    wave_sync();
    f32x16 acc[3][NB];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int nb = 0; nb < NB; nb++)
#pragma unroll
        for (int i = 0; i < 16; i++) acc[c][nb][i] = 0.f;
#pragma unroll
    for (int ks = 0; ks < KP / 16; ks++) {
      const int kk = ks * 16 + (lane >> 5) * 8;
      bf16x8 bv[NB];
#pragma unroll
      for (int nb = 0; nb < NB; nb++) bv[nb] = *(const bf16x8 *)&sB[(nb * 32 + row) * KPP + kk];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const bf16x8 av = *(const bf16x8 *)&sA[w][c][row][kk];
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
          acc[c][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv[nb], acc[c][nb], 0, 0, 0);
      }
    }
    wave_sync();
    // D[r][n], n = 2 (p*4 + s) + im: the odd lane hands its imaginary part to the even one
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int r = 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
      int v2, p2, cb2, x2[4];
      site_of(g, q0 + r, v2, p2, cb2, x2);
#pragma unroll
      for (int nb = 0; nb < NB; nb++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const float re = acc[c][nb][i];
          const float im = __shfl_xor(re, 1, 64);
          const int n = nb * 32 + (lane & 31), jj = n >> 1;
          const int pth = jj >> 2, sp = jj & 3;
          if (!(lane & 1) && jj < 4 * NN_PATHS && v2 && (!a.even_out || (p2 == 0) == (pth == 0))) {
            a.phi[((long)(p2 * g.stride + cb2) * NN_PATHS + pth) * 12 + 3 * sp + c] = pack_bf(re, im);
          }
        }
    }
  }
}

// out(x) = sum_p T_pe phi_p(x), pe = p (or its reverse for the adjoint layer)
// sites q < nsite (2 Vh: all; Vh: the even checkerboard only)
__global__ void __launch_bounds__(256) k_ptc_gather(const uint32_t *phi, uint32_t *out, const uint32_t *links,
                                                    NNGeo g, int adjoint, long nsite) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= nsite) return;
  int valid, par, cb, x[4];
  site_of(g, q, valid, par, cb, x);
  if (!valid) return;
  float re[12], im[12];
#pragma unroll
  for (int k = 0; k < 12; k++) re[k] = im[k] = 0.f;
  // gathers of three paths in flight at a time
  for (int p0 = 0; p0 < NN_PATHS; p0 += 3) {
    TRaw r[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const int p = p0 + j;
      int pe = p;
      if (adjoint && p > 0) pe = (p & 1) ? p + 1 : p - 1;
      transport_load(g, phi, NN_PATHS, links, 1, par, cb, x, pe, p, r[j]);
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
      uint32_t t[12];
      transport_compute(r[j], t);
#pragma unroll
      for (int k = 0; k < 12; k++) {
        re[k] += bf_lo(t[k]);
        im[k] += bf_hi(t[k]);
      }
    }
  }
  uint32_t *o = out + (long)(par * g.stride + cb) * 12;
#pragma unroll
  for (int k = 0; k < 12; k++) o[k] = pack_bf(re[k], im[k]);
}

// ---- weight gradient -----------------------------------------------------
struct WgArgs {
  const uint32_t *in;    // layer input (C_in channels)
  const uint32_t *gout;  // output gradient (C_out channels)
  const uint32_t *links; // BF16 complex
  float *part;           // [wg][NCH][NP][32]
  NNGeo g;
};

template <int CIN, int COUT>
__global__ void __launch_bounds__(64 * NNW) k_ptc_wgrad(const WgArgs a) {
  constexpr int NCH = ChunkMap<CIN>::NCH, NC = 4 * COUT, NP = COUT == 8 ? 64 : 32, MB = NP / 32;
  constexpr int KR = 96, KRP = KR + 8;  // rows (colour, site) of a tile, padded
  constexpr int CPW = (NCH + NNW - 1) / NNW;  // chunks per wavefront
  __shared__ __attribute__((aligned(16))) uint16_t sX[NP][KRP];
  __shared__ __attribute__((aligned(16))) uint16_t sY[NNW][32][KRP];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const NNGeo &g = a.g;
  for (int i = threadIdx.x; i < NP * KRP; i += 64 * NNW) (&sX[0][0])[i] = 0;
  f32x16 acc[CPW][MB];
#pragma unroll
  for (int q = 0; q < CPW; q++)
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
#pragma unroll
      for (int i = 0; i < 16; i++) acc[q][mb][i] = 0.f;
  const long nsites = 2L * g.Vh;
  const long ntile = (nsites + 31) / 32;
  for (long tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const long q0 = tile * 32;
    __syncthreads();
    // stage X^T: sX[m][c*32 + site], m = a*4 + s (re) or NC + a*4 + s (im)
    for (int e = threadIdx.x; e < 32 * COUT * 12; e += 64 * NNW) {
      const int site = e / (COUT * 12), rem = e % (COUT * 12), ao = rem / 12, k = rem % 12;
      const int s = k / 3, c = k % 3;
      int valid, par, cb, x[4];
      site_of(g, q0 + site, valid, par, cb, x);
      uint32_t v = 0;
      if (valid) v = a.gout[((long)(par * g.stride + cb) * COUT + ao) * 12 + k];
      sX[ao * 4 + s][c * 32 + site] = (uint16_t)(v & 0xFFFF);
      sX[NC + ao * 4 + s][c * 32 + site] = (uint16_t)(v >> 16);
    }
    __syncthreads();
    const int row = lane & 31;
    int valid, par, cb, x[4];
    site_of(g, q0 + row, valid, par, cb, x);
#pragma unroll
    for (int q = 0; q < CPW; q++) {
      const int ch = w + q * NNW;
      if (ch >= NCH) break;
#pragma unroll
      for (int j = 0; j < 2; j++) {
        int pg, b, kc0;
        chunk_item<CIN>(ch, lane, j, pg, b, kc0);
        uint32_t t[12];
        transport(g, a.in, CIN, a.links, valid, par, cb, x, pg, b, t);
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
          for (int c = 0; c < 3; c++) {
            sY[w][2 * (kc0 + s)][c * 32 + row] = (uint16_t)(t[3 * s + c] & 0xFFFFu);
            sY[w][2 * (kc0 + s) + 1][c * 32 + row] = (uint16_t)(t[3 * s + c] >> 16);
          }
      }
      wave_sync();
#pragma unroll
      for (int ks = 0; ks < KR / 16; ks++) {
        const int kk = ks * 16 + (lane >> 5) * 8;
        const bf16x8 bv = *(const bf16x8 *)&sY[w][row][kk];
#pragma unroll
        for (int mb = 0; mb < MB; mb++) {
          const bf16x8 av = *(const bf16x8 *)&sX[mb * 32 + row][kk];
          acc[q][mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[q][mb], 0, 0, 0);
        }
      }
      wave_sync();
    }
  }
  // partials: part[((wg * NCH + ch) * NP + m) * 32 + n]
#pragma unroll
  for (int q = 0; q < CPW; q++) {
    const int ch = w + q * NNW;
    if (ch >= NCH) break;
#pragma unroll
    for (int mb = 0; mb < MB; mb++)
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int m = mb * 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        a.part[(((long)blockIdx.x * NCH + ch) * NP + m) * 32 + (lane & 31)] = acc[q][mb][i];
      }
  }
}

// sum of the workgroup partials in a fixed order: res[e] = sum_wg part[wg*ne + e]
__global__ void k_wsum(int nwg, long ne, const float *part, double *res) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double s = 0;
  for (int w = 0; w < nwg; w++) s += part[(long)w * ne + e];
  res[e] = s;
}

// ---- conversions ---------------------------------------------------------
// FP32 18-real links -> BF16 complex pairs, same component-run layout
__global__ void k_nn_links(long n, const float2 *U, uint32_t *out) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 v = U[i];
  out[i] = pack_bf(v.x, v.y);
}

template <typename T>
__global__ void k_nn_from_spinor(int Vh, long sstride, long fstride, const T *sp, uint32_t *f, float scale) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= 2L * Vh) return;
  const int par = q >= Vh, cb = (int)(q - (long)par * Vh);
  uint32_t *o = f + (long)(par * fstride + cb) * 12;
  for (int k = 0; k < 12; k++) {
    uint32_t v = 0;
    if (par == 0) {
      const T re = sp[2 * ((long)k * sstride + cb)], im = sp[2 * ((long)k * sstride + cb) + 1];
      v = pack_bf(scale * (float)re, scale * (float)im);
    }
    o[k] = v;
  }
}

template <typename T>
__global__ void k_nn_to_spinor(int Vh, long sstride, long fstride, int C, const uint32_t *f, T *sp, float scale) {
  const int cb = blockIdx.x * blockDim.x + threadIdx.x;
  if (cb >= Vh) return;
  const uint32_t *o = f + (long)cb * C * 12;
  for (int k = 0; k < 12; k++) {
    sp[2 * ((long)k * sstride + cb)] = (T)(scale * bf_lo(o[k]));
    sp[2 * ((long)k * sstride + cb) + 1] = (T)(scale * bf_hi(o[k]));
  }
}

__global__ void k_nn_from_canon(NNGeo g, int C, int ch, const double2 *cn, uint32_t *f) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  int valid, par, cb, x[4];
  site_of(g, q, valid, par, cb, x);
  if (!valid) return;
  const long s = x[0] + (long)g.n[0] * (x[1] + (long)g.n[1] * (x[2] + (long)g.n[2] * x[3]));
  for (int k = 0; k < 12; k++) {
    const double2 v = cn[s * 12 + k];
    f[((long)(par * g.stride + cb) * C + ch) * 12 + k] = pack_bf((float)v.x, (float)v.y);
  }
}

__global__ void k_nn_to_canon(NNGeo g, int C, int ch, const uint32_t *f, double2 *cn) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  int valid, par, cb, x[4];
  site_of(g, q, valid, par, cb, x);
  if (!valid) return;
  const long s = x[0] + (long)g.n[0] * (x[1] + (long)g.n[1] * (x[2] + (long)g.n[2] * x[3]));
  for (int k = 0; k < 12; k++) {
    const uint32_t v = f[((long)(par * g.stride + cb) * C + ch) * 12 + k];
    cn[s * 12 + k] = make_double2(bf_lo(v), bf_hi(v));
  }
}

// g = so * out(ch 0, even) - st * target on the even sites, 0 on the odd;
// per-block partial sums of |g|^2
template <typename T>
__global__ void __launch_bounds__(256) k_nn_diff(int Vh, long sstride, long fstride, int C, const uint32_t *out,
                                                 const T *tg, float so, float st, uint32_t *gout, double *part) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double acc = 0;
  if (q < 2L * Vh) {
    const int par = q >= Vh, cb = (int)(q - (long)par * Vh);
    uint32_t *o = gout + (long)(par * fstride + cb) * 12;
    for (int k = 0; k < 12; k++) {
      uint32_t v = 0;
      if (par == 0) {
        const uint32_t y = out[(long)cb * C * 12 + k];
        const float re = so * bf_lo(y) - st * (float)tg[2 * ((long)k * sstride + cb)];
        const float im = so * bf_hi(y) - st * (float)tg[2 * ((long)k * sstride + cb) + 1];
        v = pack_bf(re, im);
        acc += (double)re * re + (double)im * im;
      }
      o[k] = v;
    }
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  __shared__ double sm[4];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
}

// ===================================================================== //

static NNGeo nngeo(const Geom &g, const Gauge *U) {
  NNGeo n{};
  for (int m = 0; m < 4; m++) {
    n.n[m] = g.n[m];
    n.bcf[m] = (float)g.bc[m];
  }
  n.Vh = (int)g.Vh;
  n.stride = g.stride;
  n.lstride = U ? U->stride : g.stride;
  return n;
}

NNField alloc_nn(const Geom &g, int C) {
  NNField f;
  f.C = C;
  f.stride = g.stride;
  QQCD_HIP(hipMalloc(&f.data, f.bytes()));
  QQCD_HIP(hipMemset(f.data, 0, f.bytes()));
  return f;
}

void free_nn(NNField &f) {
  if (f.data) (void)hipFree(f.data);
  f.data = nullptr;
}

static int nn_grid() {
  static int g = 0;
  if (!g) {
    int dev = 0, cus = 0;
    QQCD_HIP(hipGetDevice(&dev));
    QQCD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    g = std::max(1, cus);
  }
  return g;
}

static void check_layer(int cin, int cout) {
  if ((cin != 1 && cin != 8) || (cout != 1 && cout != 8))
    throw std::runtime_error("network layers have 1 or 8 channels");
}

static uint16_t to_bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// The weight seen by a layer kernel: forward W[a][b][p][s][s'], or for the
// adjoint layer (which maps forward-output channels a' to forward-input
// channels b') W'[b'][a'][p][s][s'] = conj(W[a'][b'][p][s'][s]); the path
// reversal of the adjoint is done by the kernels' transports.
static std::complex<double> wview(const NNLayer &L, bool adjoint, int a, int b, int p, int s, int sp) {
  if (!adjoint) return L.W[(((size_t)a * L.cin + b) * NN_PATHS + p) * 16 + 4 * s + sp];
  return std::conj(L.W[(((size_t)b * L.cin + a) * NN_PATHS + p) * 16 + 4 * sp + s]);
}

// B[n][kk] of the transport-first kernel: kk = 32 ch + 2 kc + im_in (the
// chunk's complex column kc of (path, input channel, spin s')), n = nc
// (output re) or NC + nc (output im), nc = 4 a + s:
//   [out_re | out_im] = [T_re, T_im](interleaved) x [[W_re^T, W_im^T], [-W_im^T, W_re^T]]
static std::vector<uint16_t> build_B(const NNLayer &L, bool adjoint) {
  const int ki = adjoint ? L.cout : L.cin, ko = adjoint ? L.cin : L.cout;
  const int NCH = ki == 8 ? 18 : 3, KP = NCH * 32, NC = 4 * ko, NP = ko == 8 ? 64 : 32;
  std::vector<uint16_t> B((size_t)NP * KP, 0);
  for (int ch = 0; ch < NCH; ch++)
    for (int kc = 0; kc < 16; kc++) {
      int p, b;
      if (ki == 8) {
        p = ch / 2;
        b = 4 * (ch % 2) + kc / 4;
      } else {
        p = 4 * ch + kc / 4;
        b = 0;
      }
      if (p >= NN_PATHS) continue;
      const int sp = kc % 4, kre = ch * 32 + 2 * kc, kim = kre + 1;
      for (int nc = 0; nc < NC; nc++) {
        const std::complex<double> w = wview(L, adjoint, nc / 4, b, p, nc % 4, sp);
        B[(size_t)nc * KP + kre] = to_bf((float)w.real());
        B[(size_t)nc * KP + kim] = to_bf((float)-w.imag());
        B[(size_t)(NC + nc) * KP + kre] = to_bf((float)w.imag());
        B[(size_t)(NC + nc) * KP + kim] = to_bf((float)w.real());
      }
    }
  return B;
}

// Bm[n][k] of the mix kernel (8 input channels -> 9 path fields of output
// channel 0): n = 2 (p*4 + s) + im_out, k = 2 (b*4 + s') + im_in
static std::vector<uint16_t> build_Bmix(const NNLayer &L, bool adjoint) {
  const int KP = 64, NP = 96;
  std::vector<uint16_t> B((size_t)NP * KP, 0);
  for (int p = 0; p < NN_PATHS; p++)
    for (int s = 0; s < 4; s++)
      for (int b = 0; b < 8; b++)
        for (int sp = 0; sp < 4; sp++) {
          const std::complex<double> w = wview(L, adjoint, 0, b, p, s, sp);
          const int n = 2 * (p * 4 + s), k = 2 * (b * 4 + sp);
          B[(size_t)n * KP + k] = to_bf((float)w.real());
          B[(size_t)n * KP + k + 1] = to_bf((float)-w.imag());
          B[(size_t)(n + 1) * KP + k] = to_bf((float)w.imag());
          B[(size_t)(n + 1) * KP + k + 1] = to_bf((float)w.real());
        }
  return B;
}

static void upload_vec(const std::vector<uint16_t> &h, uint16_t *&d, size_t &n, hipStream_t s) {
  if (d && n != h.size()) {
    (void)hipFree(d);
    d = nullptr;
  }
  if (!d) QQCD_HIP(hipMalloc(&d, h.size() * 2));
  n = h.size();
  QQCD_HIP(hipMemcpyAsync(d, h.data(), h.size() * 2, hipMemcpyHostToDevice, s));
}

void nn_upload(const NNLayer &L, NNLayerDev &D, hipStream_t s) {
  check_layer(L.cin, L.cout);
  if (L.W.size() != (size_t)L.cout * L.cin * NN_PATHS * 16) throw std::runtime_error("nn_upload: weight size");
  D.cin = L.cin;
  D.cout = L.cout;
  upload_vec(build_B(L, false), D.B, D.nB, s);
  upload_vec(build_B(L, true), D.Badj, D.nBadj, s);
  if (L.cin == 8 && L.cout == 1) upload_vec(build_Bmix(L, false), D.Bmix, D.nBmix, s);
  if (L.cin == 1 && L.cout == 8) upload_vec(build_Bmix(L, true), D.Bmix, D.nBmix, s);
  QQCD_HIP(hipStreamSynchronize(s));
}

void nn_free(NNLayerDev &D) {
  if (D.B) (void)hipFree(D.B);
  if (D.Badj) (void)hipFree(D.Badj);
  if (D.Bmix) (void)hipFree(D.Bmix);
  D.B = D.Badj = D.Bmix = nullptr;
}

NNLinks nn_links(const Geom &g, const Gauge &U32, hipStream_t s) {
  if (U32.prec != PREC_SINGLE || U32.recon != 18) throw std::runtime_error("nn_links: FP32 18-real links expected");
  NNLinks L;
  L.stride = U32.stride;
  const long n = 2L * 4 * 9 * U32.stride;
  QQCD_HIP(hipMalloc(&L.data, n * sizeof(uint32_t)));
  hipLaunchKernelGGL(k_nn_links, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, (const float2 *)U32.data,
                     L.data);
  QQCD_HIP(hipGetLastError());
  (void)g;
  return L;
}

void nn_free_links(NNLinks &L) {
  if (L.data) (void)hipFree(L.data);
  L.data = nullptr;
}

void nn_layer(const Geom &g, const NNLinks &U, const NNLayerDev &D, int adjoint, const NNField &in, NNField &out,
              hipStream_t s, int sparse) {
  const int ci = adjoint ? D.cout : D.cin, co = adjoint ? D.cin : D.cout;
  if (in.C != ci || out.C != co) throw std::runtime_error("nn_layer: channel counts do not match the layer");
  if (in.data == out.data) throw std::runtime_error("nn_layer: in and out alias");
  NNGeo geo = nngeo(g, nullptr);
  geo.lstride = U.stride;
  geo.src_even = (sparse & NN_SRC_EVEN) ? 1 : 0;
  const bool even_out = (sparse & NN_OUT_EVEN) != 0;
  if (ci == 8 && co == 1) {
    // mix, then gather: the nine path fields live in a scratch field kept between calls
    static NNField phi;
    if (!phi.data || phi.stride != g.stride) {
      free_nn(phi);
      phi = alloc_nn(g, NN_PATHS);
    }
    if (!D.Bmix) throw std::runtime_error("nn_layer: mix weights missing");
    MixArgs m{};
    m.in = in.data;
    m.phi = phi.data;
    m.B = D.Bmix;
    m.g = geo;
    m.g.src_even = 0;  // the mix reads its own site only; zero odd input gives zero there
    m.even_out = even_out;
    hipLaunchKernelGGL(k_ptc_mix, dim3(nn_grid()), dim3(64 * NNW), 0, s, m);
    QQCD_HIP(hipGetLastError());
    const long n = even_out ? g.Vh : 2 * g.Vh;
    hipLaunchKernelGGL(k_ptc_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (const uint32_t *)phi.data,
                       out.data, (const uint32_t *)U.data, geo, adjoint, n);
    QQCD_HIP(hipGetLastError());
    return;
  }
  PtcArgs a{};
  a.in = in.data;
  a.out = out.data;
  a.links = U.data;
  a.B = adjoint ? D.Badj : D.B;
  a.g = geo;
  a.adjoint = adjoint;
  const dim3 grid(nn_grid()), blk(64 * NNW);
  if (ci == 8 && co == 8) hipLaunchKernelGGL((k_ptc8<8>), grid, blk, 0, s, a);
  else if (ci == 1 && co == 8) hipLaunchKernelGGL((k_ptc<1, 8>), grid, blk, 0, s, a);
  else hipLaunchKernelGGL((k_ptc<1, 1>), grid, blk, 0, s, a);
  QQCD_HIP(hipGetLastError());
}

void nn_wgrad(const Geom &g, const NNLinks &U, int cin, int cout, const NNField &psi, const NNField &gout,
              std::complex<double> *dW, hipStream_t s) {
  check_layer(cin, cout);
  if (psi.C != cin || gout.C != cout) throw std::runtime_error("nn_wgrad: channel counts");
  const int NCH = cin == 8 ? 18 : 3, NP = cout == 8 ? 64 : 32, NC = 4 * cout;
  const int nwg = nn_grid();
  const long ne = (long)NCH * NP * 32;
  // scratch kept between calls (the trainer calls this for every layer and pair)
  static float *part = nullptr;
  static double *res = nullptr;
  static long cap = 0;
  if (ne * nwg > cap) {
    if (part) (void)hipFree(part);
    if (res) (void)hipFree(res);
    part = nullptr;
    res = nullptr;
    cap = 0;
    QQCD_HIP(hipMalloc(&part, sizeof(float) * ne * nwg));
    QQCD_HIP(hipMalloc(&res, sizeof(double) * ne * nwg));
    cap = ne * nwg;
  }
  std::vector<double> P(ne);
  {
    WgArgs a{};
    a.in = psi.data;
    a.gout = gout.data;
    a.links = U.data;
    a.part = part;
    a.g = nngeo(g, nullptr);
    a.g.lstride = U.stride;
    const dim3 grid(nwg), blk(64 * NNW);
    if (cin == 8 && cout == 8) hipLaunchKernelGGL((k_ptc_wgrad<8, 8>), grid, blk, 0, s, a);
    else if (cin == 1 && cout == 8) hipLaunchKernelGGL((k_ptc_wgrad<1, 8>), grid, blk, 0, s, a);
    else if (cin == 8 && cout == 1) hipLaunchKernelGGL((k_ptc_wgrad<8, 1>), grid, blk, 0, s, a);
    else hipLaunchKernelGGL((k_ptc_wgrad<1, 1>), grid, blk, 0, s, a);
    QQCD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_wsum, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, nwg, ne, (const float *)part, res);
    QQCD_HIP(hipGetLastError());
    QQCD_HIP(hipMemcpyAsync(P.data(), res, sizeof(double) * ne, hipMemcpyDeviceToHost, s));
    QQCD_HIP(hipStreamSynchronize(s));
  }
  // P[ch][m][n]: m = a*4 + s (g re) / NC + a*4 + s (g im); n = 2 kc + im (T)
  // dW = sum g conj(T): re = P[gre,Tre] + P[gim,Tim], im = P[gim,Tre] - P[gre,Tim]
  std::fill(dW, dW + (size_t)cout * cin * NN_PATHS * 16, std::complex<double>(0.0));
  for (int ch = 0; ch < NCH; ch++)
    for (int kc = 0; kc < 16; kc++) {
      int p, b;
      if (cin == 8) {
        p = ch / 2;
        b = 4 * (ch % 2) + kc / 4;
      } else {
        p = 4 * ch + kc / 4;
        b = 0;
      }
      if (p >= NN_PATHS) continue;
      const int sp = kc % 4, cre = 2 * kc, cim = cre + 1;
      for (int nc = 0; nc < NC; nc++) {
        const int a = nc / 4, sidx = nc % 4;
        auto Pv = [&](int m, int n) { return P[((size_t)ch * NP + m) * 32 + n]; };
        const double re = Pv(nc, cre) + Pv(NC + nc, cim);
        const double imv = Pv(NC + nc, cre) - Pv(nc, cim);
        dW[(((size_t)a * cin + b) * NN_PATHS + p) * 16 + 4 * sidx + sp] = std::complex<double>(re, imv);
      }
    }
}

void nn_from_spinor(const Geom &g, const Spinor &sp, NNField &f, float scale, hipStream_t s) {
  if (f.C != 1) throw std::runtime_error("nn_from_spinor: one-channel field expected");
  const long n = 2 * g.Vh;
  const dim3 grid((unsigned)((n + 255) / 256));
  if (sp.prec == PREC_DOUBLE)
    hipLaunchKernelGGL(k_nn_from_spinor<double>, grid, dim3(256), 0, s, (int)g.Vh, sp.stride, f.stride,
                       (const double *)sp.data, f.data, scale);
  else if (sp.prec == PREC_SINGLE)
    hipLaunchKernelGGL(k_nn_from_spinor<float>, grid, dim3(256), 0, s, (int)g.Vh, sp.stride, f.stride,
                       (const float *)sp.data, f.data, scale);
  else
    throw std::runtime_error("nn_from_spinor: FP64 or FP32 spinor expected");
  QQCD_HIP(hipGetLastError());
}

void nn_to_spinor(const Geom &g, const NNField &f, Spinor &sp, float scale, hipStream_t s) {
  const dim3 grid((unsigned)((g.Vh + 255) / 256));
  if (sp.prec == PREC_DOUBLE)
    hipLaunchKernelGGL(k_nn_to_spinor<double>, grid, dim3(256), 0, s, (int)g.Vh, sp.stride, f.stride, f.C, f.data,
                       (double *)sp.data, scale);
  else if (sp.prec == PREC_SINGLE)
    hipLaunchKernelGGL(k_nn_to_spinor<float>, grid, dim3(256), 0, s, (int)g.Vh, sp.stride, f.stride, f.C, f.data,
                       (float *)sp.data, scale);
  else
    throw std::runtime_error("nn_to_spinor: FP64 or FP32 spinor expected");
  QQCD_HIP(hipGetLastError());
}

void nn_from_canon(const Geom &g, const double *dcanon, NNField &f, int ch, hipStream_t s) {
  const long n = 2 * g.Vh;
  hipLaunchKernelGGL(k_nn_from_canon, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nngeo(g, nullptr), f.C, ch,
                     (const double2 *)dcanon, f.data);
  QQCD_HIP(hipGetLastError());
}

void nn_to_canon(const Geom &g, const NNField &f, int ch, double *dcanon, hipStream_t s) {
  const long n = 2 * g.Vh;
  hipLaunchKernelGGL(k_nn_to_canon, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nngeo(g, nullptr), f.C, ch,
                     (const uint32_t *)f.data, (double2 *)dcanon);
  QQCD_HIP(hipGetLastError());
}

double nn_diff(const Geom &g, const NNField &out, const Spinor &target, float so, float st, NNField &gout,
               double *d_scratch, hipStream_t s) {
  if (gout.C != 1) throw std::runtime_error("nn_diff: one-channel gradient expected");
  const long n = 2 * g.Vh;
  const int nb = (int)((n + 255) / 256);
  if (target.prec == PREC_SINGLE)
    hipLaunchKernelGGL(k_nn_diff<float>, dim3(nb), dim3(256), 0, s, (int)g.Vh, target.stride, gout.stride, out.C,
                       (const uint32_t *)out.data, (const float *)target.data, so, st, gout.data, d_scratch);
  else
    hipLaunchKernelGGL(k_nn_diff<double>, dim3(nb), dim3(256), 0, s, (int)g.Vh, target.stride, gout.stride, out.C,
                       (const uint32_t *)out.data, (const double *)target.data, so, st, gout.data, d_scratch);
  QQCD_HIP(hipGetLastError());
  std::vector<double> h(nb);
  QQCD_HIP(hipMemcpyAsync(h.data(), d_scratch, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  QQCD_HIP(hipStreamSynchronize(s));
  double t = 0;
  for (double v : h) t += v;
  return t;
}

// ---- model ---------------------------------------------------------------
void NNModel::init(const std::vector<int> &ch, uint64_t seed) {
  if (ch.size() < 2 || ch.front() != 1 || ch.back() != 1) throw std::runtime_error("network: channels 1, ..., 1");
  layers.clear();
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (size_t l = 0; l + 1 < ch.size(); l++) {
    NNLayer L;
    L.cin = ch[l];
    L.cout = ch[l + 1];
    check_layer(L.cin, L.cout);
    L.W.resize((size_t)L.cout * L.cin * NN_PATHS * 16);
    const double sc = 1.0 / std::sqrt(2.0 * L.cin * NN_PATHS * 4);
    for (auto &w : L.W) w = std::complex<double>(sc * nd(rng), sc * nd(rng));
    layers.push_back(L);
  }
  ready = false;
}

void NNModel::upload(hipStream_t s) {
  dev.resize(layers.size());
  for (size_t l = 0; l < layers.size(); l++) nn_upload(layers[l], dev[l], s);
  ready = true;
}

void NNModel::alloc(const Geom &g) {
  if (!acts.empty() && acts[0].stride == g.stride && acts.size() == layers.size() + 1) return;
  for (auto &f : acts) free_nn(f);
  for (auto &f : grads) free_nn(f);
  acts.clear();
  grads.clear();
  acts.push_back(alloc_nn(g, 1));
  for (auto &L : layers) acts.push_back(alloc_nn(g, L.cout));
}

void NNModel::alloc_grads(const Geom &g) {
  alloc(g);
  if (grads.size() == acts.size() && grads[0].stride == g.stride) return;
  for (auto &f : grads) free_nn(f);
  grads.clear();
  for (auto &f : acts) grads.push_back(alloc_nn(g, f.C));
}

void NNModel::release() {
  nn_free_links(links);
  for (auto &f : acts) free_nn(f);
  for (auto &f : grads) free_nn(f);
  acts.clear();
  grads.clear();
  for (auto &d : dev) nn_free(d);
  dev.clear();
  ready = false;
}

const NNLinks &NNModel::links_for(Context &c) {
  if (!links.data || links_gen != c.gauge_gen || links.stride != c.g.stride) {
    nn_free_links(links);
    links = nn_links(c.g, c.gauge(PREC_SINGLE, 18), c.stream);
    links_gen = c.gauge_gen;
  }
  return links;
}

void nn_apply(Context &c, NNModel &M, const Spinor &r, Spinor &z) {
  if (!M.ready) throw std::runtime_error("nn_apply: network weights not uploaded");
  M.alloc(c.g);
  const NNLinks &U = M.links_for(c);
  nn_from_spinor(c.g, r, M.acts[0], 1.f, c.stream);
  // the input lives on the even sites and only the even sites of the output
  // are read: the first layer skips the transports out of odd sites, the
  // last one forms the even sites only
  const size_t L = M.layers.size();
  for (size_t l = 0; l < L; l++)
    nn_layer(c.g, U, M.dev[l], 0, M.acts[l], M.acts[l + 1], c.stream,
             (l == 0 ? NN_SRC_EVEN : 0) | (l + 1 == L ? NN_OUT_EVEN : 0));
  nn_to_spinor(c.g, M.acts.back(), z, 1.f, c.stream);
}

}  // namespace qqcd
