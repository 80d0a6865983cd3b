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
typedef float f32x16 __attribute__((ext_vector_type(16)));

static constexpr int NNW = 8;  // wavefronts per workgroup

__device__ __forceinline__ float bf_lo(uint32_t v) { return __uint_as_float(v << 16); }
__device__ __forceinline__ float bf_hi(uint32_t v) { return __uint_as_float(v & 0xFFFF0000u); }
// round to nearest even (finite values)
__device__ __forceinline__ uint32_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u >> 16;
}
__device__ __forceinline__ uint32_t pack_bf(float re, float im) { return f2bf(re) | (f2bf(im) << 16); }

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
};

template <int CIN> struct ChunkMap {
  static constexpr int NCH = CIN == 8 ? 18 : 3;
  static constexpr int KP = NCH * 32;
};

// The transported input T_p psi_b at site (par, cb) into t[12] (spin-major).
// pe: path (0..8); pe >= 9 or invalid sites give zero.
template <int CIN>
__device__ __forceinline__ void transport(const NNGeo &g, const uint32_t *in, const float2 *links, int valid,
                                          int par, int cb, const int x[4], int pe, int b, float2 t[12]) {
  if (!valid || pe >= NN_PATHS) {
#pragma unroll
    for (int k = 0; k < 12; k++) t[k] = make_float2(0.f, 0.f);
    return;
  }
  int npar = par, ncb = cb, mu = 0, lpar = par, lcb = cb;
  bool back = false;
  float f = 1.f;
  if (pe > 0) {
    mu = (pe - 1) >> 1;
    back = (pe - 1) & 1;
    int y[4] = {x[0], x[1], x[2], x[3]};
    if (!back) {
      if (x[mu] + 1 == g.n[mu]) {
        y[mu] = 0;
        f = g.bcf[mu];
      } else {
        y[mu] = x[mu] + 1;
      }
    } else {
      if (x[mu] == 0) {
        y[mu] = g.n[mu] - 1;
        f = g.bcf[mu];
      } else {
        y[mu] = x[mu] - 1;
      }
    }
    npar = par ^ 1;
    ncb = (y[0] + g.n[0] * (y[1] + g.n[1] * (y[2] + g.n[2] * y[3]))) >> 1;
    if (back) {
      lpar = npar;
      lcb = ncb;
    }
  }
  const uint4 *src = (const uint4 *)(in + ((long)(npar * g.stride + ncb) * CIN + b) * 12);
  const uint4 w0 = src[0], w1 = src[1], w2 = src[2];
  const uint32_t raw[12] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w, w2.x, w2.y, w2.z, w2.w};
  float2 p[12];
#pragma unroll
  for (int k = 0; k < 12; k++) p[k] = make_float2(bf_lo(raw[k]), bf_hi(raw[k]));
  if (pe == 0) {
#pragma unroll
    for (int k = 0; k < 12; k++) t[k] = p[k];
    return;
  }
  float2 u[9];
  const long lb = (long)((lpar * 4 + mu) * 9) * g.lstride + lcb;
#pragma unroll
  for (int k = 0; k < 9; k++) u[k] = links[lb + (long)k * g.lstride];
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int r = 0; r < 3; r++) {
      float re = 0.f, im = 0.f;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        // U[r][c] forward, conj(U[c][r]) backward
        const float2 m = back ? make_float2(u[3 * c + r].x, -u[3 * c + r].y) : u[3 * r + c];
        const float2 v = p[3 * s + c];
        re += m.x * v.x - m.y * v.y;
        im += m.x * v.y + m.y * v.x;
      }
      t[3 * s + r] = make_float2(f * re, f * im);
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

// item j (0, 1) of this lane in chunk ch: (path, channel, K column base)
template <int CIN>
__device__ __forceinline__ void chunk_item(int ch, int lane, int j, int &pg, int &b, int &col) {
  if (CIN == 8) {
    pg = ch >> 1;
    b = 4 * (ch & 1) + ((lane >> 5) << 1) + j;
    col = (b & 3) * 4;  // re at col + s, im at 16 + col + s
  } else {
    const int pp = ((lane >> 5) << 1) + j;
    pg = 4 * ch + pp;
    b = 0;
    col = 8 * pp;  // re at col + s, im at col + 4 + s
  }
}

template <int CIN> __device__ __forceinline__ int im_off() { return CIN == 8 ? 16 : 4; }

struct PtcArgs {
  const uint32_t *in;
  uint32_t *out;
  const float2 *links;
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
    for (int ch = 0; ch < NCH; ch++) {
#pragma unroll
      for (int j = 0; j < 2; j++) {
        int pg, b, col;
        chunk_item<CIN>(ch, lane, j, pg, b, col);
        int pe = pg;
        if (a.adjoint && pg > 0 && pg < NN_PATHS) pe = (pg & 1) ? pg + 1 : pg - 1;
        float2 t[12];
        transport<CIN>(g, a.in, a.links, valid, par, cb, x, pe, b, t);
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
          for (int c = 0; c < 3; c++) {
            sA[w][c][row][col + s] = (uint16_t)f2bf(t[3 * s + c].x);
            sA[w][c][row][col + im_off<CIN>() + s] = (uint16_t)f2bf(t[3 * s + c].y);
          }
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

// ---- weight gradient -----------------------------------------------------
struct WgArgs {
  const uint32_t *in;    // layer input (C_in channels)
  const uint32_t *gout;  // output gradient (C_out channels)
  const float2 *links;
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
        int pg, b, col;
        chunk_item<CIN>(ch, lane, j, pg, b, col);
        float2 t[12];
        transport<CIN>(g, a.in, a.links, valid, par, cb, x, pg, b, t);
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
          for (int c = 0; c < 3; c++) {
            sY[w][col + s][c * 32 + row] = (uint16_t)f2bf(t[3 * s + c].x);
            sY[w][col + im_off<CIN>() + s][c * 32 + row] = (uint16_t)f2bf(t[3 * s + c].y);
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

// B[n][kk] of the forward layer (W) or of the adjoint layer (W^dag, reversed
// paths; its K runs over the output channels of the forward layer)
static std::vector<uint16_t> build_B(const NNLayer &L, bool adjoint) {
  const int ki = adjoint ? L.cout : L.cin, ko = adjoint ? L.cin : L.cout;
  const int NCH = ki == 8 ? 18 : 3, KP = NCH * 32, NC = 4 * ko, NP = ko == 8 ? 64 : 32;
  std::vector<float> B((size_t)NP * KP, 0.f);
  auto Wf = [&](int a, int b, int p, int s, int sp) {
    // forward: W[a][b][p][s][sp]; adjoint: W'[b'=a_fwd...]: (W[a][b][rev p])^dag
    if (!adjoint) return L.W[(((size_t)a * L.cin + b) * NN_PATHS + p) * 16 + 4 * s + sp];
    // adjoint layer maps forward-output channel a' to forward-input channel b':
    // W'[b'][a'][p][s][sp] = conj(W[a'][b'][p][sp][s]); the path reversal is in the kernel
    return std::conj(L.W[(((size_t)b * L.cin + a) * NN_PATHS + p) * 16 + 4 * sp + s]);
  };
  for (int ch = 0; ch < NCH; ch++)
    for (int col = 0; col < 32; col++) {
      int p, b, sp, im;
      if (ki == 8) {
        p = ch / 2;
        im = col >= 16;
        b = 4 * (ch % 2) + (col % 16) / 4;
        sp = col % 4;
      } else {
        const int pp = col / 8;
        p = 4 * ch + pp;
        im = (col % 8) >= 4;
        b = 0;
        sp = col % 4;
      }
      if (p >= NN_PATHS) continue;
      const int kk = ch * 32 + col;
      for (int nc = 0; nc < NC; nc++) {
        const int a = nc / 4, s = nc % 4;
        const std::complex<double> w = Wf(a, b, p, s, sp);
        // [out_re | out_im] = [T_re | T_im] [[W_re^T, W_im^T], [-W_im^T, W_re^T]]
        B[(size_t)nc * KP + kk] = (float)(im ? -w.imag() : w.real());
        B[(size_t)(NC + nc) * KP + kk] = (float)(im ? w.real() : w.imag());
      }
    }
  std::vector<uint16_t> out(B.size());
  for (size_t i = 0; i < B.size(); i++) {
    uint32_t u;
    std::memcpy(&u, &B[i], 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    out[i] = (uint16_t)(u >> 16);
  }
  return out;
}

void nn_upload(const NNLayer &L, NNLayerDev &D, hipStream_t s) {
  check_layer(L.cin, L.cout);
  if (L.W.size() != (size_t)L.cout * L.cin * NN_PATHS * 16) throw std::runtime_error("nn_upload: weight size");
  std::vector<uint16_t> B = build_B(L, false), Ba = build_B(L, true);
  if (D.B && (D.nB != B.size() || D.nBadj != Ba.size())) nn_free(D);
  if (!D.B) {
    QQCD_HIP(hipMalloc(&D.B, B.size() * 2));
    QQCD_HIP(hipMalloc(&D.Badj, Ba.size() * 2));
  }
  D.nB = B.size();
  D.nBadj = Ba.size();
  D.cin = L.cin;
  D.cout = L.cout;
  QQCD_HIP(hipMemcpyAsync(D.B, B.data(), B.size() * 2, hipMemcpyHostToDevice, s));
  QQCD_HIP(hipMemcpyAsync(D.Badj, Ba.data(), Ba.size() * 2, hipMemcpyHostToDevice, s));
  QQCD_HIP(hipStreamSynchronize(s));
}

void nn_free(NNLayerDev &D) {
  if (D.B) (void)hipFree(D.B);
  if (D.Badj) (void)hipFree(D.Badj);
  D.B = D.Badj = nullptr;
}

void nn_layer(const Geom &g, const Gauge &U32, const NNLayerDev &D, int adjoint, const NNField &in, NNField &out,
              hipStream_t s) {
  if (U32.prec != PREC_SINGLE || U32.recon != 18) throw std::runtime_error("nn_layer: FP32 18-real links expected");
  const int ci = adjoint ? D.cout : D.cin, co = adjoint ? D.cin : D.cout;
  if (in.C != ci || out.C != co) throw std::runtime_error("nn_layer: channel counts do not match the layer");
  if (in.data == out.data) throw std::runtime_error("nn_layer: in and out alias");
  PtcArgs a{};
  a.in = in.data;
  a.out = out.data;
  a.links = (const float2 *)U32.data;
  a.B = adjoint ? D.Badj : D.B;
  a.g = nngeo(g, &U32);
  a.adjoint = adjoint;
  const dim3 grid(nn_grid()), blk(64 * NNW);
  if (ci == 8 && co == 8) hipLaunchKernelGGL((k_ptc<8, 8>), grid, blk, 0, s, a);
  else if (ci == 1 && co == 8) hipLaunchKernelGGL((k_ptc<1, 8>), grid, blk, 0, s, a);
  else if (ci == 8 && co == 1) hipLaunchKernelGGL((k_ptc<8, 1>), grid, blk, 0, s, a);
  else hipLaunchKernelGGL((k_ptc<1, 1>), grid, blk, 0, s, a);
  QQCD_HIP(hipGetLastError());
}

void nn_wgrad(const Geom &g, const Gauge &U32, int cin, int cout, const NNField &psi, const NNField &gout,
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
    a.links = (const float2 *)U32.data;
    a.part = part;
    a.g = nngeo(g, &U32);
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
  // P[ch][m][n]: m = a*4 + s (g re) / NC + a*4 + s (g im); n = chunk column (T re / im)
  // dW = sum g conj(T): re = P[gre,Tre] + P[gim,Tim], im = P[gim,Tre] - P[gre,Tim]
  std::fill(dW, dW + (size_t)cout * cin * NN_PATHS * 16, std::complex<double>(0.0));
  for (int ch = 0; ch < NCH; ch++)
    for (int col = 0; col < 32; col++) {
      int p, b, sp, im;
      if (cin == 8) {
        p = ch / 2;
        im = col >= 16;
        b = 4 * (ch % 2) + (col % 16) / 4;
        sp = col % 4;
      } else {
        const int pp = col / 8;
        p = 4 * ch + pp;
        im = (col % 8) >= 4;
        b = 0;
        sp = col % 4;
      }
      if (p >= NN_PATHS || im) continue;
      const int col_im = cin == 8 ? col + 16 : col + 4;
      for (int nc = 0; nc < NC; nc++) {
        const int a = nc / 4, sidx = nc % 4;
        auto Pv = [&](int m, int n) { return P[((size_t)ch * NP + m) * 32 + n]; };
        const double re = Pv(nc, col) + Pv(NC + nc, col_im);
        const double imv = Pv(NC + nc, col) - Pv(nc, col_im);
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
  for (auto &f : acts) free_nn(f);
  for (auto &f : grads) free_nn(f);
  acts.clear();
  grads.clear();
  for (auto &d : dev) nn_free(d);
  dev.clear();
  ready = false;
}

void nn_apply(Context &c, NNModel &M, const Spinor &r, Spinor &z) {
  if (!M.ready) throw std::runtime_error("nn_apply: network weights not uploaded");
  M.alloc(c.g);
  const Gauge &U = c.gauge(PREC_SINGLE, 18);
  nn_from_spinor(c.g, r, M.acts[0], 1.f, c.stream);
  for (size_t l = 0; l < M.layers.size(); l++) nn_layer(c.g, U, M.dev[l], 0, M.acts[l], M.acts[l + 1], c.stream);
  nn_to_spinor(c.g, M.acts.back(), z, 1.f, c.stream);
}

}  // namespace qqcd
