// qqcd_lagfib.hip -- the reference's lagged-Fibonacci random numbers on the device.
//
// The reference's generator (qqcd-test/cfgsprops/glue/lagfib.f90) keeps 521
// words w and two cursors jr, kr = jr - 32 (mod 521).  A draw sets
// w[jr] ^= w[kr] -- again while the result is zero -- returns w[jr] times
// fnorm and advances both cursors.  As a sequence of draws,
//   y_t = y_{t-521} ^ y_{t-32},   y_{-521+k} = w[(jr + k) mod 521],
// bitwise linear over GF(2) with characteristic polynomial
// P(x) = x^521 + x^489 + 1: y_t = XOR_k a_k y_{k-521} with
// sum_k a_k x^k = x^(t+521) mod P.
//
// Device generation of m draws: chunks of L draws (8192 2^k, at most 2048
// per launch).  Because every shift of the sequence obeys the recurrence,
// x^s mod P = sum_k a_k x^k gives y_{s+t} = XOR_k a_k y_{k+t} for every t, so
// the 521 words before chunk c (s = cL) are one polynomial (17 words, built
// once per L on the host) applied to the first 1041 terms of the stream (the
// base window and the next 520 terms, extended in LDS); then one wave per
// chunk runs the recurrence through an LDS ring: the first 1600 draws 64 a
// step (lanes 32-63 use y_t = y_{t-521} ^ y_{t-553} ^ y_{t-64}, so every
// operand comes from an earlier step), the rest 128 a step by the
// recurrence of P(x)^4 = P(x^4), y_t = y_{t-128} ^ y_{t-2084}, which holds
// for the same sequence (it is P(x) times P(x)^3).  A zero draw (probability 2^-31) is
// the one non-linear step: the reference XORs again, so y_t becomes
// y_{t-32}.  The first zero is found with an atomic minimum, corrected, and
// the stream after it generated again from the corrected window.
#include <algorithm>
#include <array>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "qqcd_lagfib.h"

namespace qqcd {

static constexpr int LP = 521, LQ = 32, LW = 17, RING = 4096;
// draws of a chunk made 64 a step before the ring holds the 2084 draws the
// fourth-power recurrence reaches back to (521 + B1 >= 2084, B1 % 64 == 0)
static constexpr int B1 = 1600;
// ring slot of draw r of a chunk (r >= -521): (r + S0) mod RING, S0 a
// multiple of 64 at least 521
static constexpr int S0 = 576;
static_assert(LP + B1 >= 4 * LP && B1 % 64 == 0, "bootstrap too short");
static_assert(S0 >= LP && S0 % 64 == 0 && (B1 + S0) % 64 == 0, "slot blocks must not wrap");
static_assert(4 * LP + 128 <= RING && LP - 1 + 575 < RING, "ring too small");
static constexpr int KMAX = 2048;  // chunks per launch
typedef std::array<uint32_t, LW> Poly;

// ---- initlagfib (lagfib.f90:80-216) ----------------------------------------

void lagfib_init(int myid, int iseed, LagState &st) {
  // initbirep: one bit per word from a multiplicative congruential sequence,
  // then 32 steps of the bit recurrence per word of the common start
  int ib[LP];
  uint32_t ix = 1;
  for (int i = 0; i < LP; i++) {
    ix *= 69069u;
    ib[i] = (int)(ix >> 31);
  }
  int jr = 0, kr = LP - LQ;
  uint32_t w0[LP];
  for (int j = 0; j < LP; j++) {
    uint32_t work = 0;
    for (int i = 0; i < 32; i++) {
      work = 2u * work + (uint32_t)ib[jr];
      ib[jr] ^= ib[kr];
      jr = jr + 1 == LP ? 0 : jr + 1;
      kr = kr + 1 == LP ? 0 : kr + 1;
    }
    w0[j] = work >> 1;
  }
  // bigdelay(myid + iseed): the common start advanced by a polynomial in x
  // built from the binary digits of ndelay (40 squarings first)
  const int mu = 40;
  int m = myid + iseed;
  if (m <= 0) throw std::runtime_error("lagfib_init: myid + iseed must be positive");
  std::vector<uint32_t> wk(2 * LP - 1);
  for (int i = 0; i < LP; i++) wk[i] = w0[i];
  for (int i = LP; i < 2 * LP - 1; i++) wk[i] = wk[i - LP] ^ wk[i - LQ];
  int bits[LP + 32] = {0};
  int nb = mu - 1;
  while (m > LP - 1) {
    bits[++nb] = m % 2;
    m /= 2;
  }
  std::vector<int> c(2 * LP, 0);
  c[m] = 1;
  for (int j = nb; j >= 0; j--) {
    for (int i = LP - 1; i >= 0; i--) {
      c[2 * i + bits[j]] = c[i];
      c[2 * i + 1 - bits[j]] = 0;
    }
    for (int i = 2 * LP - 1; i >= LP; i--) {
      c[i - LP] ^= c[i];
      c[i - LQ] ^= c[i];
    }
  }
  for (int j = 0; j < LP; j++) {
    uint32_t v = 0;
    for (int i = 0; i < LP; i++)
      if (c[i]) v ^= wk[j + i];
    st.w[j] = v;
  }
  st.jr = jr;
  st.kr = kr;
}

void lagfib_check(const LagState &st) {
  if (st.jr < 0 || st.jr >= LP || st.kr != (st.jr + LP - LQ) % LP)
    throw std::runtime_error("lagfib: cursors must satisfy 0 <= jr < 521 and kr = jr - 32 mod 521 (got jr=" +
                             std::to_string(st.jr) + ", kr=" + std::to_string(st.kr) + ")");
}

// ---- jump-ahead tables -----------------------------------------------------

static void mulx(Poly &a) {
  uint32_t carry = 0;
  for (int w = 0; w < LW; w++) {
    const uint32_t nc = a[w] >> 31;
    a[w] = (a[w] << 1) | carry;
    carry = nc;
  }
  if (a[LP / 32] & (1u << (LP % 32))) {  // x^521 = x^489 + 1
    a[LP / 32] &= ~(1u << (LP % 32));
    a[0] ^= 1u;
    a[(LP - LQ) / 32] ^= 1u << ((LP - LQ) % 32);
  }
}

static Poly mulmod(const Poly &a, const Poly &b) {
  Poly r{};
  for (int i = LP - 1; i >= 0; i--) {
    mulx(r);
    if ((a[i / 32] >> (i % 32)) & 1u)
      for (int w = 0; w < LW; w++) r[w] ^= b[w];
  }
  return r;
}

static Poly xpow(long e) {
  Poly r{}, b{};
  r[0] = 1;
  b[0] = 2;  // x
  while (e) {
    if (e & 1) r = mulmod(r, b);
    b = mulmod(b, b);
    e >>= 1;
  }
  return r;
}

struct LagTable {
  long L = 0;
  int K = 0;
  uint32_t *d = nullptr;
  std::vector<uint32_t> h;
};
static std::vector<LagTable> g_tables;

// x^(cL) mod P for c < K, 17 words each, on the device
static const uint32_t *table_for(long L, int K, hipStream_t s) {
  LagTable *t = nullptr;
  for (auto &e : g_tables)
    if (e.L == L) t = &e;
  if (!t) {
    g_tables.push_back(LagTable{});
    t = &g_tables.back();
    t->L = L;
  }
  if (t->K >= K) return t->d;
  const int K0 = t->K;
  t->h.resize((size_t)K * LW);
  const Poly xL = xpow(L);
  Poly p = xpow((long)K0 * L);
  for (int c = K0; c < K; c++) {
    std::memcpy(&t->h[(size_t)c * LW], p.data(), sizeof(uint32_t) * LW);
    p = mulmod(p, xL);
  }
  if (t->d) {
    QQCD_HIP(hipStreamSynchronize(s));
    QQCD_HIP(hipFree(t->d));
    t->d = nullptr;
  }
  QQCD_HIP(hipMalloc(&t->d, sizeof(uint32_t) * t->h.size()));
  QQCD_HIP(hipMemcpyAsync(t->d, t->h.data(), sizeof(uint32_t) * t->h.size(), hipMemcpyHostToDevice, s));
  QQCD_HIP(hipStreamSynchronize(s));
  t->K = K;
  return t->d;
}

void lagfib_release_device() {
  for (auto &e : g_tables)
    if (e.d) (void)hipFree(e.d);
  g_tables.clear();
}

// ---- device generation -----------------------------------------------------

// Workgroup of one wave per chunk c = blockIdx.x.  Slots 55..575 of the
// ring hold the window y_{cL-521} .. y_{cL-1} at the start.  LDS ordering
// between steps needs only wavefront scope: the wave's LDS operations
// complete in order.
__device__ __forceinline__ int lag_slot(int r) { return (r + S0) & (RING - 1); }
__device__ __forceinline__ void lag_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// one step of the fourth-power recurrence: draws r0 + lane and r0 + 64 + lane
// (r0 + S0 a multiple of 64, so the blocks of 64 slots at r0 - 128, r0 - 64,
// r0 and r0 + 64 never wrap; the taps 2084 back are wrapped per lane)
template <bool TAIL>
__device__ __forceinline__ void lag_step4(uint32_t *ring, uint32_t *o, int r0, int lane, int len, uint32_t &vmin) {
  const int s1 = (r0 + S0 - 4 * LQ) & (RING - 1), s2 = (r0 + S0 - 2 * LQ) & (RING - 1);
  const int sw = (r0 + S0) & (RING - 1), sw2 = (r0 + S0 + 64) & (RING - 1);
  const int sf = r0 + S0 - 4 * LP;  // >= 0 for r0 >= B1
  const uint32_t ya = ring[s1 + lane] ^ ring[(sf + lane) & (RING - 1)];
  const uint32_t yb = ring[s2 + lane] ^ ring[(sf + 64 + lane) & (RING - 1)];
  if (!TAIL || r0 + lane < len) {
    ring[sw + lane] = ya;
    o[r0 + lane] = ya;
    vmin = min(vmin, ya);
  }
  if (!TAIL || r0 + 64 + lane < len) {
    ring[sw2 + lane] = yb;
    o[r0 + 64 + lane] = yb;
    vmin = min(vmin, yb);
  }
}

__global__ void __launch_bounds__(64) k_lag_gen(const uint32_t *base, const uint32_t *polys, long L, long m, int K,
                                                uint32_t *out, unsigned long long *first_zero) {
  // ring[k] = y_{k-521}, k = 0..1040, while the window is formed: the base
  // window, then 520 terms by the recurrence; afterwards the ring of draws
  __shared__ uint32_t ring[RING];
  const int lane = threadIdx.x;
  const int c = blockIdx.x;
  // the chunk's polynomial x^(cL) mod P, word w in lane w, and the base
  // window, all loads in flight together
  const uint32_t pw = lane < LW ? polys[(long)c * LW + lane] : 0u;
  uint32_t v[9];
#pragma unroll
  for (int q = 0; q < 9; q++) v[q] = lane + 64 * q < LP ? base[lane + 64 * q] : 0u;
#pragma unroll
  for (int q = 0; q < 9; q++)
    if (lane + 64 * q < LP) ring[lane + 64 * q] = v[q];
  lag_wave_sync();
  for (int k0 = LP; k0 < 2 * LP - 1; k0 += 32) {
    const int k = k0 + lane;
    if (lane < 32 && k < 2 * LP - 1) ring[k] = ring[k - LP] ^ ring[k - LQ];
    lag_wave_sync();
  }
  // window: w_j = XOR over the set bits k of the polynomial of y[k + j], two
  // bits a round, every read unconditional (k + j < 1096 stays in the ring;
  // lanes with j >= 521 read words they never keep)
#pragma unroll
  for (int q = 0; q < 9; q++) v[q] = 0u;
  for (int w = 0; w < LW; w++) {
    uint32_t bits = (uint32_t)__builtin_amdgcn_readlane((int)pw, w);
    while (bits) {
      const int k1 = 32 * w + __builtin_ctz(bits);
      bits &= bits - 1;
      int k2 = k1;
      uint32_t m2 = 0u;
      if (bits) {
        k2 = 32 * w + __builtin_ctz(bits);
        bits &= bits - 1;
        m2 = ~0u;
      }
#pragma unroll
      for (int q = 0; q < 9; q++) v[q] ^= ring[k1 + lane + 64 * q] ^ (ring[k2 + lane + 64 * q] & m2);
    }
  }
  __syncthreads();  // every read of the start terms is done before the window replaces them
#pragma unroll
  for (int q = 0; q < 9; q++)
    if (lane + 64 * q < LP) ring[lag_slot(lane + 64 * q - LP)] = v[q];
  __syncthreads();
  const long t0 = (long)c * L;
  const int len = (int)std::min(L, m - t0);
  uint32_t *o = out + t0;
  uint32_t vmin = ~0u;  // a zero draw shows as a zero minimum
  // draws 0 .. B1-1 of the chunk, 64 a step
  const int n1 = std::min(len, B1);
  for (int r0 = 0; r0 < n1; r0 += 64) {
    const int r = r0 + lane;
    uint32_t yy = ring[lag_slot(r - LP)];
    if (lane < 32) yy ^= ring[lag_slot(r - LQ)];
    else yy ^= ring[lag_slot(r - LP - LQ)] ^ ring[lag_slot(r - 2 * LQ)];
    if (r < n1) {
      ring[lag_slot(r)] = yy;
      o[r] = yy;
      vmin = min(vmin, yy);
    }
    lag_wave_sync();
  }
  // the rest 128 a step by y_t = y_{t-128} ^ y_{t-2084} (P(x)^4 = P(x^4)):
  // every tap lies at least 64 draws back, in an earlier step
  int r0 = B1;
  for (; r0 + 128 <= len; r0 += 128) {
    lag_step4<false>(ring, o, r0, lane, len, vmin);
    lag_wave_sync();
  }
  if (r0 < len) lag_step4<true>(ring, o, r0, lane, len, vmin);
  // the first zero of the chunk (probability about L 2^-32): found again
  // from the chunk's own output, each lane reading back what it wrote
  if (__ballot(vmin == 0u)) {
    for (int q0 = 0; q0 < len; q0 += 64) {
      const unsigned long long z = __ballot(q0 + lane < len && o[q0 + lane] == 0u);
      if (z) {
        if (lane == 0) atomicMin(first_zero, (unsigned long long)(t0 + q0 + __ffsll((long long)z) - 1));
        break;
      }
    }
  }
}

// nb[k] = y_{tend-521+k} (y_t = out[t] for t >= 0, base[521+t] below); with
// fix >= 0 the draw at t = fix (== tend-1) becomes y_{fix-32}, in out too,
// as the reference's zero skip makes it; bad[0] = 1 if that is zero as well.
__global__ void k_lag_rebase(const uint32_t *base, uint32_t *out, long tend, long fix, uint32_t *nb, int *bad) {
  __shared__ uint32_t fixed;
  if (threadIdx.x == 0) {
    if (fix >= 0) {
      const long t = fix - LQ;
      const uint32_t v = t >= 0 ? out[t] : base[LP + t];
      fixed = v;
      out[fix] = v;
      if (v == 0u) bad[0] = 1;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < LP; k += blockDim.x) {
    const long t = tend - LP + k;
    uint32_t v = t >= 0 ? out[t] : base[LP + t];
    if (t == fix) v = fixed;
    nb[k] = v;
  }
}

// chunk length for n draws: 8192 2^k, the least that needs at most KMAX chunks
static long chunk_len(long n) {
  long L = 8192;
  while ((n + L - 1) / L > KMAX && L < (1L << 21)) L *= 2;
  return L;
}

__global__ void k_lag_reset(unsigned long long *first_zero, int *bad, int clear_bad) {
  if (threadIdx.x == 0) {
    first_zero[0] = ~0ull;
    if (clear_bad) bad[0] = 0;
  }
}

// Device buffers of one generation: the two base windows, the first-zero
// word and the degenerate-state flag, and the jump table of chunk polynomials
// (table(L, K) returns x^(cL) mod P for c < K).
struct LagWork {
  uint32_t *base = nullptr;  // 2 * 521 words
  unsigned long long *dz = nullptr;
  int *bad = nullptr;
  const uint32_t *(*table)(void *ctx, long L, int K, hipStream_t s) = nullptr;
  void *table_ctx = nullptr;
};

static void lag_generate_with(LagState &st, long m, uint32_t *d_out, hipStream_t s, const LagWork &w) {
  uint32_t hb[LP];
  for (int k = 0; k < LP; k++) hb[k] = st.w[(st.jr + k) % LP];
  QQCD_HIP(hipMemcpyAsync(w.base, hb, sizeof(hb), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_lag_reset, dim3(1), dim3(64), 0, s, w.dz, w.bad, 1);
  QQCD_HIP(hipGetLastError());
  uint32_t *cur = w.base, *nxt = w.base + LP;
  uint32_t *o = d_out;
  long left = m;
  while (left > 0) {
    const long L = chunk_len(left);
    const long piece = std::min(left, (long)KMAX * L);
    const int K = (int)((piece + L - 1) / L);
    const uint32_t *rows = w.table(w.table_ctx, L, K, s);
    if (o != d_out) {
      hipLaunchKernelGGL(k_lag_reset, dim3(1), dim3(64), 0, s, w.dz, w.bad, 0);
      QQCD_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_lag_gen, dim3(K), dim3(64), 0, s, cur, rows, L, piece, K, o, w.dz);
    QQCD_HIP(hipGetLastError());
    unsigned long long z = 0;
    QQCD_HIP(hipMemcpyAsync(&z, w.dz, sizeof(z), hipMemcpyDeviceToHost, s));
    QQCD_HIP(hipStreamSynchronize(s));
    long used = piece;
    long fix = -1;
    if (z < (unsigned long long)piece) {
      used = (long)z + 1;
      fix = (long)z;
    }
    if (used < left) {
      hipLaunchKernelGGL(k_lag_rebase, dim3(1), dim3(256), 0, s, cur, o, used, fix, nxt, w.bad);
      QQCD_HIP(hipGetLastError());
      std::swap(cur, nxt);
    } else if (fix >= 0) {  // the last draw was the zero: correct it in place
      hipLaunchKernelGGL(k_lag_rebase, dim3(1), dim3(256), 0, s, cur, o, used, fix, nxt, w.bad);
      QQCD_HIP(hipGetLastError());
    }
    // a rebase reads the old base, so keep it until the kernel has run
    int bad = 0;
    QQCD_HIP(hipMemcpyAsync(&bad, w.bad, sizeof(int), hipMemcpyDeviceToHost, s));
    QQCD_HIP(hipStreamSynchronize(s));
    if (bad) throw std::runtime_error("lagfib: degenerate state (the zero skip of rnds would not end)");
    o += used;
    left -= used;
  }
  // the new state: w'[(jr + m + k) mod 521] = y_{m-521+k}
  const long nlast = std::min(m, (long)LP);
  uint32_t tail[LP];
  if (nlast) {
    QQCD_HIP(hipMemcpyAsync(tail, d_out + (m - nlast), sizeof(uint32_t) * nlast, hipMemcpyDeviceToHost, s));
    QQCD_HIP(hipStreamSynchronize(s));
  }
  LagState ns;
  const int njr = (int)((st.jr + m) % LP);
  for (int k = 0; k < LP; k++) {
    const long t = m - LP + k;
    ns.w[(njr + k) % LP] = t >= 0 ? tail[t - (m - nlast)] : hb[LP + t];
  }
  ns.jr = njr;
  ns.kr = (njr + LP - LQ) % LP;
  st = ns;
}

void lagfib_generate(LagState &st, long m, uint32_t *d_out, hipStream_t s) {
  lagfib_check(st);
  if (m < 0) throw std::runtime_error("lagfib_generate: negative count");
  LagWork w;
  w.table = [](void *, long L, int K, hipStream_t st) { return table_for(L, K, st); };
  QQCD_HIP(hipMalloc(&w.base, sizeof(uint32_t) * 2 * LP));
  QQCD_HIP(hipMalloc(&w.dz, sizeof(unsigned long long)));
  QQCD_HIP(hipMalloc(&w.bad, sizeof(int)));
  try {
    lag_generate_with(st, m, d_out, s, w);
  } catch (...) {
    (void)hipFree(w.base);
    (void)hipFree(w.dz);
    (void)hipFree(w.bad);
    throw;
  }
  QQCD_HIP(hipFree(w.base));
  QQCD_HIP(hipFree(w.dz));
  QQCD_HIP(hipFree(w.bad));
}

// ---- rnds without allocation -------------------------------------------------
//
// qqcd_rnds_noalloc works in one static device array per GPU instead of
// memory it allocates: the raw draws, their doubles, the two base windows,
// the first-zero word and flag, and the jump table (LAG_NOALLOC_MAX draws
// need chunks of 8192 draws, at most 128 of them per launch).  Its address on
// each device is taken once, at qqcd_init with that device current; host
// code passes that address, never the array's name, to kernels and copies.
// Every call on one GPU uses the same array, so calls that overlap on one GPU
// race; the array keeps the last call's contents (the table is reused when
// the next call needs no more of it).
static constexpr long NA_Y = 0;                                   // LAG_NOALLOC_MAX raw draws
static constexpr long NA_D = NA_Y + 4 * LAG_NOALLOC_MAX;           // LAG_NOALLOC_MAX doubles
static constexpr long NA_BASE = NA_D + 8 * LAG_NOALLOC_MAX;        // 2 * 521 words
static constexpr long NA_FLAGS = NA_BASE + 4352;                  // first zero (8 B), flag (4 B)
static constexpr long NA_TAB = NA_FLAGS + 256;                    // NA_CHUNKS * 17 words
static constexpr int NA_CHUNKS = (int)((LAG_NOALLOC_MAX + 8191) / 8192);
static constexpr long NA_BYTES = NA_TAB + 4L * LW * NA_CHUNKS;
static_assert(8 * LP <= 4352 && NA_BYTES % 256 == 0, "workspace layout");
static_assert(NA_CHUNKS <= KMAX, "one launch covers the maximum");

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

struct NoAllocDevice {
  unsigned char *p = nullptr;  // device address of scratch on this GPU
  long tab_L = 0;              // the jump table scratch holds: chunk length, chunks
  int tab_K = 0;
};
static constexpr int NA_MAXDEV = 64;
static NoAllocDevice g_noalloc[NA_MAXDEV];

void lagfib_noalloc_setup(int dev) {
  if (dev < 0 || dev >= NA_MAXDEV) throw std::runtime_error("lagfib: device index beyond 63");
  if (g_noalloc[dev].p) return;
  void *ptr = nullptr;
  QQCD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) throw std::runtime_error("lagfib: no device address for the static workspace");
  g_noalloc[dev].p = (unsigned char *)ptr;
}

// the jump table in the workspace, rebuilt when the call needs a different
// chunk length or more chunks than it holds
static const uint32_t *noalloc_table(void *ctx, long L, int K, hipStream_t s) {
  NoAllocDevice &d = *(NoAllocDevice *)ctx;
  uint32_t *tab = (uint32_t *)(d.p + NA_TAB);
  if (K > NA_CHUNKS) throw std::runtime_error("qqcd_rnds_noalloc: jump table beyond its workspace");
  if (d.tab_L == L && d.tab_K >= K) return tab;
  static thread_local uint32_t h[NA_CHUNKS * LW];
  const Poly xL = xpow(L);
  Poly p = xpow(0);
  for (int c = 0; c < K; c++) {
    std::memcpy(&h[(size_t)c * LW], p.data(), sizeof(uint32_t) * LW);
    p = mulmod(p, xL);
  }
  d.tab_L = 0;  // invalid until the copy has landed
  QQCD_HIP(hipMemcpyAsync(tab, h, sizeof(uint32_t) * LW * K, hipMemcpyHostToDevice, s));
  QQCD_HIP(hipStreamSynchronize(s));
  d.tab_L = L;
  d.tab_K = K;
  return tab;
}

void rnds_noalloc(LagState &st, long m, double *rn, int dev, hipStream_t s) {
  lagfib_check(st);
  if (m < 0 || m > LAG_NOALLOC_MAX)
    throw std::runtime_error("qqcd_rnds_noalloc: m = " + std::to_string(m) + " outside 0 .. " +
                             std::to_string(LAG_NOALLOC_MAX) + " (the size of its static workspace; qqcd_rnds has no limit)");
  if (m > 0 && !rn) throw std::runtime_error("qqcd_rnds_noalloc: no output array");
  if (dev < 0 || dev >= NA_MAXDEV || !g_noalloc[dev].p)
    throw std::runtime_error("qqcd_rnds_noalloc: workspace not set up on this device (qqcd_init)");
  NoAllocDevice &d = g_noalloc[dev];
  LagWork w;
  w.base = (uint32_t *)(d.p + NA_BASE);
  w.dz = (unsigned long long *)(d.p + NA_FLAGS);
  w.bad = (int *)(d.p + NA_FLAGS + 8);
  w.table = noalloc_table;
  w.table_ctx = &d;
  uint32_t *y = (uint32_t *)(d.p + NA_Y);
  double *dd = (double *)(d.p + NA_D);
  lag_generate_with(st, m, y, s, w);
  if (m > 0) {
    lagfib_to_double(y, m, dd, s);
    QQCD_HIP(hipMemcpyAsync(rn, dd, sizeof(double) * m, hipMemcpyDeviceToHost, s));
    QQCD_HIP(hipStreamSynchronize(s));
  }
}

double bench_lagfib(long m, int niter, long chunk, hipStream_t s) {
  if (m <= 0 || niter <= 0) throw std::runtime_error("bench_lagfib: m, niter > 0");
  if (chunk < 0 || chunk > (1L << 21)) throw std::runtime_error("bench_lagfib: chunk in 0 (default) .. 2^21");
  LagState st;
  lagfib_init(0, 1, st);
  uint32_t hb[LP];
  for (int k = 0; k < LP; k++) hb[k] = st.w[(st.jr + k) % LP];
  const long L = chunk ? chunk : chunk_len(m);
  m = std::min(m, (long)KMAX * L);
  const int K = (int)((m + L - 1) / L);
  const uint32_t *rows = table_for(L, K, s);
  uint32_t *dbase = nullptr, *out = nullptr;
  unsigned long long *dz = nullptr;
  QQCD_HIP(hipMalloc(&dbase, sizeof(hb)));
  QQCD_HIP(hipMalloc(&dz, sizeof(unsigned long long)));
  QQCD_HIP(hipMalloc(&out, sizeof(uint32_t) * m));
  QQCD_HIP(hipMemcpyAsync(dbase, hb, sizeof(hb), hipMemcpyHostToDevice, s));
  QQCD_HIP(hipMemsetAsync(dz, 0xff, sizeof(unsigned long long), s));
  hipEvent_t e0, e1;
  QQCD_HIP(hipEventCreate(&e0));
  QQCD_HIP(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_lag_gen, dim3(K), dim3(64), 0, s, dbase, rows, L, m, K, out, dz);
  QQCD_HIP(hipEventRecord(e0, s));
  for (int it = 0; it < niter; it++)
    hipLaunchKernelGGL(k_lag_gen, dim3(K), dim3(64), 0, s, dbase, rows, L, m, K, out, dz);
  QQCD_HIP(hipEventRecord(e1, s));
  QQCD_HIP(hipEventSynchronize(e1));
  float ms = 0;
  QQCD_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(dbase);
  (void)hipFree(dz);
  (void)hipFree(out);
  return ms / niter;
}

__global__ void k_lag_double(const uint32_t *y, long m, double *rn) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) rn[t] = (double)(int)y[t] * LAG_FNORM;
}

void lagfib_to_double(const uint32_t *d_y, long m, double *d_rn, hipStream_t s) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_lag_double, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, d_y, m, d_rn);
  QQCD_HIP(hipGetLastError());
}

// ---- z4source / z2source ---------------------------------------------------

// Site i of checkerboard par; the reference slot (isite, ieo, ibleo) of the
// site gives the draw of (colour ic, spin id): z4source fills each
// checkerboard in the order ic, isite, id, ieo, ibleo (ibleo fastest),
// disconloops.f90:4518-4569 (Z4: 1 for u > 0.75, -1 for 0.5 < u < 0.75, i for
// 0.25 < u < 0.5, else -i) and :4583-4649 (Z2: 1 for u > 0.5, else -1).
__global__ void k_lag_noise(int n0, int n1, int n2, int n3, int Vh, long nvh, int par, long stride,
                            const uint32_t *y, int kind, double scale, double2 *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Vh) return;
  int x[4];
  {
    const int r = 2 * i;
    const int x0 = r % n0;
    int q = r / n0;
    x[1] = q % n1;
    q /= n1;
    x[2] = q % n2;
    x[3] = q / n2;
    x[0] = x0 + ((par + x[1] + x[2] + x[3]) & 1);
  }
  const int ns[4] = {n0 / 2, n1 / 2, n2 / 2, n3 / 2};
  int b[4], sv[4];
  for (int m = 0; m < 4; m++) {
    b[m] = x[m] & 1;
    sv[m] = x[m] >> 1;
  }
  const long L = sv[0] + (long)ns[0] * (sv[1] + (long)ns[1] * (sv[2] + (long)ns[2] * sv[3]));
  const long isite = L / 2;
  const int ieo = (sv[0] + sv[1] + sv[2] + sv[3]) & 1;
  const int ibleo = (b[0] + 2 * b[1] + 4 * b[2] + 8 * b[3]) >> 1;
  for (int d = 0; d < 4; d++)
    for (int ic = 0; ic < 3; ic++) {
      const long e = (long)ic * 64 * nvh + isite * 64 + d * 16 + ieo * 8 + ibleo;
      const double u = (double)(int)y[e] * LAG_FNORM;
      double re = 0, im = 0;
      if (kind == 1) re = u > 0.5 ? 1.0 : -1.0;
      else if (u > 0.75) re = 1;
      else if (0.5 < u && u < 0.75) re = -1;
      else if (0.25 < u && u < 0.5) im = 1;
      else im = -1;
      out[(long)(3 * d + ic) * stride + i] = make_double2(scale * re, scale * im);
    }
}

void lag_noise_source(const Geom &g, Spinor &out, int par, const uint32_t *d_draws, int kind, double scale,
                      hipStream_t s) {
  if (out.prec != PREC_DOUBLE) throw std::runtime_error("lag_noise_source: FP64 spinor expected");
  if (kind != 0 && kind != 1) throw std::runtime_error("lag_noise_source: kind 0 (Z4) or 1 (Z2)");
  for (int m = 0; m < 4; m++)
    if (g.n[m] % 2) throw std::runtime_error("lag_noise_source: even extents expected");
  const int nb = (int)((g.Vh + 255) / 256);
  hipLaunchKernelGGL(k_lag_noise, dim3(nb), dim3(256), 0, s, g.n[0], g.n[1], g.n[2], g.n[3], (int)g.Vh, g.V / 32,
                     par, out.stride, d_draws, kind, scale, (double2 *)out.data);
  QQCD_HIP(hipGetLastError());
}

}  // namespace qqcd
