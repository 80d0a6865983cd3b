// qqcd_batch.cpp -- batched GMRES-Proj with reliable updates, and the
// autotuner of the multi-RHS hopping kernel.
#include "qqcd_batch.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <stdexcept>
#include <tuple>

namespace qqcd {

using dense::zd;

MultiSchur::MultiSchur(Context &c_, Prec p, int recon, const double k[2], int cap, int nb, const MultiTune &t,
                       int idag_)
    : c(c_), U(c_.gauge(p, recon)), idag(idag_), tmp(c_.g, p, cap, nb), tune(t) {
  kappa[0] = k[0];
  kappa[1] = k[1];
}

void MultiSchur::operator()(const MultiSpinor &in, MultiSpinor &out, const int *nact) {
  const double k = kappa[0], mu = kappa[1];
  Epilogue e1, e2;
  if (mu < 1e-8) {
    e2.a0 = 1.0 / (k * k);
    e2.b0 = -1.0;
  } else {
    const double a = 2.0 * k * mu;
    e1.b1 = -a;
    e2.a0 = 1.0 / (k * k);
    e2.a1 = 2.0 * mu / k;
    e2.b0 = -1.0 / (1.0 + a * a);
  }
  dslash_multi(c.g, U, in, mid(), 1, idag, e1, nullptr, nact, tune, c.stream);
  dslash_multi(c.g, U, mid(), out, 0, idag, e2, &in, nact, tune, c.stream);
}

bool MultiSchur::can_scale() const { return dslash_multi_can_scale(c.g, tmp.m.prec, tmp.m.nb, tune); }

void MultiSchur::set_half_mid(bool on) {
  if (on && tmp.m.prec == PREC_SINGLE && can_scale()) {
    if (!tmp16) tmp16.reset(new MultiBuf(c.g, PREC_HALF, tmp.m.cap, tmp.m.nb, true));
  } else {
    tmp16.reset();
  }
}

void MultiSchur::scaled(const MultiSpinor &in, MultiSpinor &out, const int *nact, const double *n2,
                        MultiSpinor *side, double hs) {
  const double k = kappa[0], mu = kappa[1];
  Epilogue e1, e2;
  if (mu < 1e-8) {
    e2.a0 = 1.0 / (k * k);
    e2.b0 = -1.0;
  } else {
    const double a = 2.0 * k * mu;
    e1.b1 = -a;
    e2.a0 = 1.0 / (k * k);
    e2.a1 = 2.0 * mu / k;
    e2.b0 = -1.0 / (1.0 + a * a);
  }
  // the operator is linear: the first hop runs on the unscaled input, the
  // second hop's epilogue scales the whole result
  dslash_multi(c.g, U, in, mid(), 1, idag, e1, nullptr, nact, tune, c.stream);
  dslash_multi(c.g, U, mid(), out, 0, idag, e2, &in, nact, tune, c.stream, n2, side, hs);
}

namespace {

// pinned host mirror of a device array
template <typename T> struct HostBuf {
  T *p = nullptr;
  size_t n = 0;
  explicit HostBuf(size_t n_) : n(n_) { QQCD_HIP(hipHostMalloc((void **)&p, n * sizeof(T), hipHostMallocDefault)); }
  ~HostBuf() {
    if (p) (void)hipHostFree(p);
  }
  T &operator[](size_t i) { return p[i]; }
};

template <typename T> struct DevBuf {
  T *p = nullptr;
  explicit DevBuf(size_t n) {
    QQCD_HIP(hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T)));
    QQCD_HIP(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

struct Graph {
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
  ~Graph() {
    if (e) (void)hipGraphExecDestroy(e);
    if (g) (void)hipGraphDestroy(g);
  }
};

class Batch {
 public:
  Batch(Context &c, const double kappa[2], MultiSpinor &B, MultiSpinor &X, const Deflation &D, const BatchParams &p,
        std::vector<int> &id);
  BatchStats run(int n);

 private:
  Context &c;
  const BatchParams &p;
  MultiSpinor &B, &X;
  const Deflation &D;
  std::vector<int> &id;
  const int cap, nb, m, k;
  MBlas mb;
  MultiSchur A64, A32;
  MultiBuf R64, Bin, E, R;
  std::vector<MultiBuf *> basis_bufs;
  std::vector<MultiSpinor> V;
  // FP16 basis only: the FP32 vector being orthogonalised and the newest
  // basis vector in FP32 (the operator's input)
  std::unique_ptr<MultiBuf> Wf, Qf;
  bool half() const { return p.basis == PREC_HALF; }
  // FP16 basis with the two-slot operator: each step's normalisation is
  // done by the next operator application (MultiSchur::scaled); hs: the
  // basis scale
  bool fused = false;
  double hs = 0.0;
  // two-step blocks (BatchParams::sstep): A^2 v_j in U2f, the four scalars
  // of the update pass and the per-slot coefficients of the next vector
  bool paired = false;
  std::unique_ptr<MultiBuf> U2f;
  // device state: ints [0] n_outer, [1] n_inner, [2] any, [3, 3+cap) flags
  DevBuf<int> di;
  DevBuf<double> h1, h2, n2a, n2b, nf, H, coef, coef2, n2, ssc, sbc, sself;
  HostBuf<double> hH, hc, hcoef, hcoef2, hn2;
  std::vector<double> bn, rn, tgt;  // per slot: ||b||, current ||r||, inner target
  BatchStats st;
  Graph graph;
  bool have_graph = false;
  int n_out = 0, n_in = 0;

  int *d_nout() { return di.p; }
  int *d_nin() { return di.p + 1; }
  int *d_any() { return di.p + 2; }
  int *d_flag() { return di.p + 3; }
  void set_counts() {
    int v[2] = {n_out, n_in};
    QQCD_HIP(hipMemcpyAsync(di.p, v, sizeof(v), hipMemcpyHostToDevice, c.stream));
    QQCD_HIP(hipStreamSynchronize(c.stream));
  }
  void download(double *h, const double *d, size_t n) {
    QQCD_HIP(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    QQCD_HIP(hipStreamSynchronize(c.stream));
  }
  void upload(double *d, const double *h, size_t n) {
    QQCD_HIP(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  }
  void swap_slots(int a, int b, bool inner_fields);
  void arnoldi_step(int j);
  void arnoldi_block(int j);
  void arnoldi_cycle();
  void inner();
  void outer_residual();
};

Batch::Batch(Context &c_, const double kappa[2], MultiSpinor &B_, MultiSpinor &X_, const Deflation &D_,
             const BatchParams &p_, std::vector<int> &id_)
    : c(c_),
      p(p_),
      B(B_),
      X(X_),
      D(D_),
      id(id_),
      cap(B_.cap),
      nb(B_.nb),
      m(p_.m),
      k(D_.k),
      mb(c_.stream, B_.cap),
      A64(c_, PREC_DOUBLE, 18, kappa, B_.cap, B_.nb, p_.tune),
      A32(c_, p_.op_prec, p_.op_recon, kappa, B_.cap, B_.nb, p_.tune),
      R64(c_.g, PREC_DOUBLE, B_.cap, B_.nb),
      Bin(c_.g, PREC_SINGLE, B_.cap, B_.nb),
      E(c_.g, PREC_SINGLE, B_.cap, B_.nb),
      R(c_.g, PREC_SINGLE, B_.cap, B_.nb),
      di(3 + B_.cap),
      h1((size_t)2 * (p_.m + 2) * B_.cap),
      h2((size_t)2 * (p_.m + 2) * B_.cap),
      n2a(B_.cap),
      n2b(B_.cap),
      nf(B_.cap),
      H((size_t)2 * p_.m * (p_.m + 1) * B_.cap),
      coef((size_t)2 * (p_.m + 1) * B_.cap),
      coef2((size_t)2 * (p_.m + 1) * B_.cap),
      n2(B_.cap),
      ssc((size_t)4 * B_.cap),
      sbc((size_t)4 * B_.cap),
      sself((size_t)4 * B_.cap),
      hH((size_t)2 * p_.m * (p_.m + 1) * B_.cap),
      hc((size_t)2 * (p_.m + 1) * B_.cap),
      hcoef((size_t)2 * (p_.m + 1) * B_.cap),
      hcoef2((size_t)2 * (p_.m + 1) * B_.cap),
      hn2(B_.cap) {
  if (X.prec != PREC_DOUBLE || B.prec != PREC_DOUBLE || X.cap != cap || X.nb != nb)
    throw std::runtime_error("batched_solve: B and X must be FP64 fields of one layout");
  if (p.op_prec == PREC_DOUBLE) throw std::runtime_error("batched_solve: the inner operator is FP32 or FP16");
  if (m < 1) throw std::runtime_error("batched_solve: m must be >= 1");
  if (k > 0) {
    if ((int)D.V.size() != k + 1 || D.prec != PREC_SINGLE)
      throw std::runtime_error("batched_solve: the deflation space must hold k+1 FP32 vectors");
    if (k > m) throw std::runtime_error("batched_solve: k > m");
  }
  if (p.basis != PREC_SINGLE && p.basis != PREC_HALF)
    throw std::runtime_error("batched_solve: the Krylov basis is FP32 or FP16");
  basis_bufs.reserve(m + 1);
  for (int i = 0; i <= m; i++) {
    basis_bufs.push_back(new MultiBuf(c.g, p.basis, cap, nb, false));
    V.push_back(basis_bufs.back()->m);
  }
  if (half()) {
    Wf.reset(new MultiBuf(c.g, PREC_SINGLE, cap, nb));
    Qf.reset(new MultiBuf(c.g, PREC_SINGLE, cap, nb));
    fused = p.fuse_norm && A32.can_scale();
    hs = std::sqrt(12.0 * (double)c.g.stride);
    paired = fused && p.sstep == 2 && p.reorth == 0 && m % 2 == 0;
    if (paired) U2f.reset(new MultiBuf(c.g, PREC_SINGLE, cap, nb));
  }
  A32.set_half_mid(p.half_mid != 0);
  mb.reserve(mb.need());
  clear_flags(d_flag(), d_any(), cap, c.stream);
  st.iters.assign(cap, 0);
  st.rel_res.assign(cap, 0.0);
}

void Batch::swap_slots(int a, int b, bool inner_fields) {
  if (a == b) return;
  mb.swap(B, a, b);
  mb.swap(X, a, b);
  mb.swap(R64.m, a, b);
  if (inner_fields) {
    mb.swap(Bin.m, a, b);
    mb.swap(E.m, a, b);
    mb.swap(R.m, a, b);
  }
  std::swap(id[a], id[b]);
  std::swap(bn[a], bn[b]);
  std::swap(rn[a], rn[b]);
  std::swap(tgt[a], tgt[b]);
}

// one Arnoldi step of every active inner slot; device-only (capturable)
void Batch::arnoldi_step(int j) {
  int *nin = d_nin();
  if (fused) {
    // FP16 basis, normalisation folded into the operator: step j's input
    // is the orthogonalised but unnormalised w of step j - 1 (r / ||r|| in
    // Qf for step 0), its squared norm nf left by that step's column; the
    // operator writes V[j] = hs w / ||w|| and A w / ||w||; the two FP32
    // fields take turns as input and output
    MultiSpinor &in = (j & 1) ? Wf->m : Qf->m;
    MultiSpinor &w = (j & 1) ? Qf->m : Wf->m;
    if (j == 0) A32(in, w, nin);
    else A32.scaled(in, w, nin, nf.p, &V[j], hs);
    mb.dot(j + 1, V.data(), w, nin, h1.p, nullptr, true);  // <V_q, w> and <w, w> in one pass
    mb.axpy(j + 1, V.data(), h1.p, 0, -1.0, w, nin, n2a.p);
    if (p.reorth > 0) {
      arnoldi_flags(n2a.p, h1.p + (size_t)2 * (j + 1) * cap, d_flag(), d_any(), nin, cap, p.reorth == 2, c.stream);
      mb.dot(j + 1, V.data(), w, nin, h2.p, d_any());
      mb.axpy(j + 1, V.data(), h2.p, 0, -1.0, w, nin, n2b.p, d_flag(), d_any());
    }
    arnoldi_column(H.p, m + 1, j, h1.p, h2.p, d_flag(), n2a.p, n2b.p, nf.p, nin, cap, c.stream);
    return;
  }
  // w: the new vector, orthogonalised in place (FP32; FP16 basis: in Wf)
  MultiSpinor &w = half() ? Wf->m : V[j + 1];
  A32(half() ? Qf->m : V[j], w, nin);
  if (half()) {
    mb.dot(j + 1, V.data(), w, nin, h1.p, nullptr, true);  // <V_q, w> and <w, w> in one pass
  } else {
    mb.dot(j + 2, V.data(), w, nin, h1.p);
  }
  mb.axpy(j + 1, V.data(), h1.p, 0, -1.0, w, nin, n2a.p);
  if (p.reorth > 0) {
    arnoldi_flags(n2a.p, h1.p + (size_t)2 * (j + 1) * cap, d_flag(), d_any(), nin, cap, p.reorth == 2, c.stream);
    mb.dot(j + 1, V.data(), w, nin, h2.p, d_any());
    mb.axpy(j + 1, V.data(), h2.p, 0, -1.0, w, nin, n2b.p, d_flag(), d_any());
  }
  arnoldi_column(H.p, m + 1, j, h1.p, h2.p, d_flag(), n2a.p, n2b.p, nf.p, nin, cap, c.stream);
  if (half()) mb.scale_inv_norm2(w, Qf->m, V[j + 1], nf.p, nin);
  else mb.scale_inv_norm(w, w, nf.p, nin);
}

// Arnoldi steps j and j + 1 (j even) of every active inner slot in one pair:
// u1 = A v_j (the operator also normalises v_j and writes it to V[j]), u2 =
// A u1, the dots of both against V_0..V_j in one pass, both updates
// u1' = u1 - V g1, u2' = u2 - V g2 in another; then v_{j+1} = u1' / nu and
// the part of A v_{j+1} orthogonal to v_0..v_{j+1}, (u2' - alpha u1') / nu,
// which the next pair's operator normalises.  Half the passes over the basis
// of two single steps; device-only (capturable).
void Batch::arnoldi_block(int j) {
  int *nin = d_nin();
  MultiSpinor &rin = Qf->m, &u1 = Wf->m, &u2 = U2f->m;
  if (j == 0) A32(rin, u1, nin);
  else A32.scaled(rin, u1, nin, nf.p, &V[j], hs);
  A32(u1, u2, nin);
  if (p.pair_norms) {
    mb.dot2(j + 1, V.data(), u1, u2, nin, h1.p, h2.p, sself.p);
    pythag_scalars(j + 1, h1.p, h2.p, sself.p, ssc.p, nin, cap, c.stream);
    sstep_columns(H.p, m + 1, j, h1.p, h2.p, ssc.p, sbc.p, nin, cap, hs, c.stream);
    mb.axpy2_next(j + 1, V.data(), h1.p, h2.p, u1, u2, sbc.p, rin, V[j + 1], nin, nf.p);
  } else {
    mb.dot2(j + 1, V.data(), u1, u2, nin, h1.p, h2.p);
    mb.axpy2(j + 1, V.data(), h1.p, h2.p, u1, u2, nin, ssc.p);
    sstep_columns(H.p, m + 1, j, h1.p, h2.p, ssc.p, sbc.p, nin, cap, hs, c.stream);
    mb.next2(u1, u2, rin, V[j + 1], sbc.p, nin, nf.p);
  }
  set_subdiag(H.p, m + 1, j + 1, nf.p, nin, cap, c.stream);
}

void Batch::arnoldi_cycle() {
  auto run_steps = [&]() {
    if (paired)
      for (int j = 0; j < m; j += 2) arnoldi_block(j);
    else
      for (int j = 0; j < m; j++) arnoldi_step(j);
  };
  if (p.use_graph && have_graph) {
    QQCD_HIP(hipGraphLaunch(graph.e, c.stream));
    st.graph_replays++;
  } else {
    run_steps();
    if (p.use_graph) {
      // every kernel has run once (static launch shapes are set), capture now
      QQCD_HIP(hipStreamSynchronize(c.stream));
      QQCD_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
      try {
        run_steps();
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      QQCD_HIP(hipStreamEndCapture(c.stream, &graph.g));
      QQCD_HIP(hipGraphInstantiate(&graph.e, graph.g, nullptr, nullptr, 0));
      have_graph = true;
    }
  }
  st.batched_matvecs += m;
}

// ||b - A x|| of the outer slots in FP64 (into rn)
void Batch::outer_residual() {
  A64(X, R64.m, d_nout());
  st.batched_matvecs++;
  mb.axpby(1.0, B, -1.0, R64.m, d_nout(), n2.p);
  download(hn2.p, n2.p, cap);
  for (int s = 0; s < n_out; s++) rn[s] = std::sqrt(hn2[s]);
  st.slot_matvecs += n_out;
}

void Batch::inner() {
  const int ldh = m + 1;
  // inner system: A e = Bin (FP32 copy of the outer residual), e0 = 0
  mb.axpby(1.0, R64.m, 0.0, Bin.m, d_nout(), nullptr);
  mb.axpby(0.0, Bin.m, 0.0, E.m, d_nout(), nullptr);
  mb.axpby(1.0, Bin.m, 0.0, R.m, d_nout(), n2.p);
  download(hn2.p, n2.p, cap);
  std::vector<double> rin(cap, 0.0);
  for (int s = 0; s < n_in; s++) rin[s] = std::sqrt(hn2[s]);
  std::vector<zd> Hs((size_t)ldh * m), cs(ldh), y(m), dd(std::max(k, 1)), cd(k + 1);
  for (int cyc = 0; cyc < p.inner_cycles && n_in > 0; cyc++) {
    if (k > 0) {
      // projection: min || V_{k+1}^H r - Hbar_k d ||, e += V_k d, r -= V_{k+1} Hbar_k d
      mb.dot_shared(k + 1, D.V.data(), R.m, d_nin(), coef.p);
      download(hc.p, coef.p, (size_t)2 * (k + 1) * cap);
      for (int s = 0; s < n_in; s++) {
        for (int q = 0; q <= k; q++) cd[q] = zd(hc[2 * ((size_t)q * cap + s)], hc[2 * ((size_t)q * cap + s) + 1]);
        dense::lsq(k + 1, k, D.H.data(), k + 1, cd.data(), dd.data());
        for (int q = 0; q < k; q++) {
          hcoef[2 * ((size_t)q * cap + s)] = dd[q].real();
          hcoef[2 * ((size_t)q * cap + s) + 1] = dd[q].imag();
        }
        for (int q = 0; q <= k; q++) {
          zd g = 0;
          for (int l = 0; l < k; l++) g += D.H[q + (size_t)(k + 1) * l] * dd[l];
          hcoef2[2 * ((size_t)q * cap + s)] = g.real();
          hcoef2[2 * ((size_t)q * cap + s) + 1] = g.imag();
        }
      }
      upload(coef.p, hcoef.p, (size_t)2 * k * cap);
      upload(coef2.p, hcoef2.p, (size_t)2 * (k + 1) * cap);
      mb.axpy_shared(k, D.V.data(), coef.p, 0, 1.0, E.m, d_nin(), nullptr);
      mb.axpy_shared(k + 1, D.V.data(), coef2.p, 0, -1.0, R.m, d_nin(), n2.p);
    } else {
      mb.norm2(R.m, d_nin(), n2.p);
    }
    // GMRES(m) from r: V_0 = r / ||r||
    if (half()) mb.scale_inv_norm2(R.m, Qf->m, V[0], n2.p, d_nin());
    else mb.scale_inv_norm(R.m, V[0], n2.p, d_nin());
    download(hn2.p, n2.p, cap);
    std::vector<double> beta(n_in);
    for (int s = 0; s < n_in; s++) beta[s] = std::sqrt(hn2[s]);
    arnoldi_cycle();
    download(hH.p, H.p, (size_t)2 * m * ldh * cap);
    for (int s = 0; s < n_in; s++) {
      // the slot's (m+1) x m Hessenberg matrix, cut at the first breakdown
      double hmax = 0;
      for (int j = 0; j < m; j++)
        for (int q = 0; q <= j + 1; q++) {
          const size_t o = 2 * (((size_t)j * ldh + q) * cap + s);
          Hs[q + (size_t)ldh * j] = zd(hH[o], hH[o + 1]);
          hmax = std::max(hmax, std::abs(Hs[q + (size_t)ldh * j]));
        }
      int jj = m;
      for (int j = 0; j < m; j++)
        if (std::abs(Hs[j + 1 + (size_t)ldh * j]) <= 1e-14 * hmax) {
          jj = j + 1;
          break;
        }
      std::fill(cs.begin(), cs.end(), zd(0.0));
      cs[0] = beta[s];
      std::fill(y.begin(), y.end(), zd(0.0));
      if (beta[s] > 0 && hmax > 0) dense::lsq(jj + 1, jj, Hs.data(), ldh, cs.data(), y.data());
      for (int q = 0; q < m; q++) {
        hcoef[2 * ((size_t)q * cap + s)] = q < jj ? y[q].real() : 0.0;
        hcoef[2 * ((size_t)q * cap + s) + 1] = q < jj ? y[q].imag() : 0.0;
      }
      st.iters[id[s]] += m;
    }
    st.slot_matvecs += (long)m * n_in;
    st.cycles++;
    upload(coef.p, hcoef.p, (size_t)2 * m * cap);
    mb.axpy(m, V.data(), coef.p, 0, 1.0, E.m, d_nin(), nullptr);
    // true residual of the inner system
    A32(E.m, R.m, d_nin());
    st.batched_matvecs++;
    st.slot_matvecs += n_in;
    mb.axpby(1.0, Bin.m, -1.0, R.m, d_nin(), n2.p);
    download(hn2.p, n2.p, cap);
    for (int s = 0; s < n_in; s++) rin[s] = std::sqrt(hn2[s]);
    // retire the slots that met their inner target
    for (int s = n_in - 1; s >= 0; s--)
      if (rin[s] < tgt[s] || !(rin[s] == rin[s])) {
        const int last = n_in - 1;
        swap_slots(s, last, true);
        std::swap(rin[s], rin[last]);
        n_in--;
      }
    set_counts();
  }
}

BatchStats Batch::run(int n) {
  if (n < 0 || n > cap) throw std::runtime_error("batched_solve: bad right-hand-side count");
  bn.assign(cap, 0.0);
  rn.assign(cap, 0.0);
  tgt.assign(cap, 0.0);
  n_out = n;
  n_in = 0;
  set_counts();
  mb.norm2(B, d_nout(), n2.p);
  download(hn2.p, n2.p, cap);
  for (int s = 0; s < n_out; s++) bn[s] = std::sqrt(hn2[s]);
  outer_residual();
  auto retire_outer = [&]() {
    for (int s = n_out - 1; s >= 0; s--)
      if (rn[s] < p.tol * bn[s] || bn[s] == 0.0) {
        const int last = n_out - 1;
        swap_slots(s, last, false);
        st.rel_res[id[last]] = bn[last] > 0 ? rn[last] / bn[last] : 0.0;
        n_out--;
      }
    set_counts();
  };
  retire_outer();
  while (n_out > 0 && st.outer < p.max_outer) {
    for (int s = 0; s < n_out; s++) tgt[s] = std::max(p.inner_tol, 0.5 * p.tol * bn[s] / rn[s]) * rn[s];
    n_in = n_out;
    set_counts();
    inner();
    // x += e over every slot of this outer step
    n_in = 0;
    set_counts();
    const FaultPlan &fault = fault_plan();
    if (st.outer < fault.corrupt_outer && fault.batch_rhs >= 0) {
      for (int s = 0; s < n_out; s++)
        if (id[s] == fault.batch_rhs) {
          SpinorBuf bad(c.g, PREC_SINGLE);
          fill_spinor(c.g, bad.s, std::nan(""), std::nan(""), c.stream);
          mb.put(bad.s, E.m, s);
          QQCD_HIP(hipStreamSynchronize(c.stream));
        }
    }
    // a slot whose correction holds a NaN or infinity keeps its x (the next
    // outer step, or the single-RHS solver, works from its true residual)
    mb.norm2(E.m, d_nout(), n2.p);
    download(hn2.p, n2.p, cap);
    for (int s = 0; s < n_out; s++)
      if (!std::isfinite(hn2[s])) {
        SpinorBuf zero(c.g, PREC_SINGLE);
        mb.put(zero.s, E.m, s);
        QQCD_HIP(hipStreamSynchronize(c.stream));
        st.rejected++;
      }
    mb.axpby(1.0, E.m, 1.0, X, d_nout(), nullptr);
    outer_residual();
    st.outer++;
    retire_outer();
  }
  // stragglers: the single-RHS mixed solver (plain FP64 GMRES-DR as its last resort)
  if (n_out > 0) {
    SpinorBuf b1(c.g, PREC_DOUBLE), x1(c.g, PREC_DOUBLE);
    Deflation local;
    for (int s = 0; s < n_out; s++) {
      mb.get(B, s, b1.s);
      mb.get(X, s, x1.s);
      MixedParams mp;
      mp.m = std::max(m, 20);
      mp.k = std::min(16, mp.m / 2);
      mp.tol = p.tol;
      mp.op_recon = p.op_recon;
      MixedStats ms = solve_mixed(c, A64.kappa, b1.s, x1.s, mp, &local);
      st.iters[id[s]] += ms.inner.iters;
      st.slot_matvecs += ms.inner.matvecs + ms.fp64_matvecs;
      mb.put(x1.s, X, s);
      st.fallback++;
    }
    outer_residual();
    for (int s = 0; s < n_out; s++) st.rel_res[id[s]] = bn[s] > 0 ? rn[s] / bn[s] : 0.0;
  }
  QQCD_HIP(hipStreamSynchronize(c.stream));
  st.converged = 1;
  for (int s = 0; s < n; s++)
    if (!(st.rel_res[id[s]] < p.tol)) st.converged = 0;
  return st;
}

}  // namespace

BatchStats batched_solve(Context &c, const double kappa[2], MultiSpinor &B, MultiSpinor &X, int n,
                         const Deflation &D, const BatchParams &p, std::vector<int> &id) {
  if ((int)id.size() != B.cap) {
    id.resize(B.cap);
    for (int s = 0; s < B.cap; s++) id[s] = s;
  }
  std::unique_ptr<Batch> b(new Batch(c, kappa, B, X, D, p, id));
  BatchStats st = b->run(n);
  // per-id arrays cover the ids [0, cap)
  return st;
}

TuneResult autotune_multi(Context &c, Prec p, int recon, int nrhs) {
  static std::map<std::tuple<long, int, int, int, uint64_t>, TuneResult> cache;
  const auto key = std::make_tuple(c.g.V, (int)p, recon, nrhs, (uint64_t)c.g.n[0] * 1000003ull + c.g.n[3]);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  const Gauge &U = c.gauge(p == PREC_DOUBLE ? PREC_DOUBLE : p, p == PREC_DOUBLE ? 18 : recon);
  int top = 1;
  while (top < nrhs && top < 64) top <<= 1;
  TuneResult best;
  best.ms = 1e30;
  int *d_n = nullptr;
  QQCD_HIP(hipMalloc(&d_n, sizeof(int)));
  QQCD_HIP(hipMemcpy(d_n, &nrhs, sizeof(int), hipMemcpyHostToDevice));
  hipEvent_t e0, e1;
  QQCD_HIP(hipEventCreate(&e0));
  QQCD_HIP(hipEventCreate(&e1));
  try {
    for (int nb = std::min(8, top); nb <= top; nb *= 2) {
      MultiBuf in(c.g, p, nrhs, nb), out(c.g, p, nrhs, nb);
      for (int order = 0; order < 2; order++) {
        if (order == 1 && in.m.nblk() == 1) continue;  // same tiles as order 0
        MultiTune t;
        t.order = order;
        Epilogue ep;
        for (int w = 0; w < 2; w++) dslash_multi(c.g, U, in.m, out.m, 0, 0, ep, nullptr, d_n, t, c.stream);
        QQCD_HIP(hipEventRecord(e0, c.stream));
        const int reps = 5;
        for (int r = 0; r < reps; r++) dslash_multi(c.g, U, in.m, out.m, r & 1, 0, ep, nullptr, d_n, t, c.stream);
        QQCD_HIP(hipEventRecord(e1, c.stream));
        QQCD_HIP(hipEventSynchronize(e1));
        float ms = 0;
        QQCD_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double per = ms / reps;
        // a larger RHS block must beat the best smaller one by 5%: the hop is
        // a third of a batched step, and the passes over the slots' bases run
        // faster in smaller blocks (63 sources at 24^3x48: 10.5 s with block
        // 16 against 11.0 s with block 32, whose hop was 3% faster)
        const double need = best.nb > 0 && nb > best.nb ? 0.95 * best.ms : best.ms;
        if (per < need) {
          best.ms = per;
          best.nb = nb;
          best.t = t;
        }
      }
    }
  } catch (...) {
    (void)hipFree(d_n);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipFree(d_n);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  cache[key] = best;
  return best;
}

}  // namespace qqcd
