// qqcd_solvers.cpp -- GMRES-DR, GMRES-Proj and BiCGStab on the device.
//
// Algorithms (the reference's, restated):
//  * GMRES-DR(m,k) (R. B. Morgan, SIAM J. Sci. Comput. 24 (2002) 20;
//    qqcd-test/cfgsprops/quark/inverters.f90:2871-3572, Matlab
//    Matlab/gmresdr5.m): Arnoldi to m vectors, least squares, true residual
//    at the cycle end, then restart with the k harmonic Ritz vectors of
//    smallest modulus (Schur vectors of H_m + |h|^2 H_m^{-H} e_m e_m^T)
//    plus the residual vector of the small problem.
//  * GMRES-Proj (Morgan & Wilcox; inverters.f90:7183-8304,
//    Matlab/gmresproj5.m): a minimal-residual projection over the
//    deflation space (V_{k+1}, Hbar_k) of an earlier GMRES-DR solve,
//    alternated with GMRES(m) cycles.
//  * BiCGStab (inverters.f90:302-487).
//
// Differences of mechanism (same mathematics): the basis is orthogonalised
// by classical Gram-Schmidt with one reorthogonalisation (two fused
// multi-dot / multi-axpy passes per step) instead of modified Gram-Schmidt
// (j dependent reductions per step, :3120-3140); the solution is formed
// once per cycle instead of at every Arnoldi step (:3242-3255); the small
// least-squares problem is solved by Householder QR on the host; a cycle ends
// early when its least-squares residual already meets the tolerance.
#include "qqcd_solvers.h"

#include <cmath>
#include <memory>
#include <stdexcept>

namespace qqcd {

using dense::zd;

void Deflation::clear() {
  for (auto &s : V) free_spinor(s);
  V.clear();
  H.clear();
  k = 0;
}

static Vec vv(const Spinor &s) { return as_vec(s); }

static std::vector<Vec> vecs(const std::vector<Spinor> &S, int n) {
  std::vector<Vec> v(n);
  for (int i = 0; i < n; i++) v[i] = as_vec(S[i]);
  return v;
}

// w = b - A x ; returns ||w||
static double residual(Context &c, LinOp &A, const Spinor &b, const Spinor &x, Spinor &w, SolveStats &st) {
  A.apply(x, w);
  st.matvecs++;
  Vec wv = vv(w);
  return std::sqrt(c.blas->axpby(1.0, vv(b), -1.0, wv, true));
}

// One Arnoldi step: V[j+1] = A V[j] orthonormalised against V[0..j] by
// classical Gram-Schmidt with selective reorthogonalisation (the DGKS
// criterion: a second pass only when the first removed more than half of
// ||w||^2, i.e. when cancellation may have spoiled orthogonality).  The
// first pass computes <V_i, w> and <w, w> in one sweep; the Hessenberg
// column goes to h[0..j+1].  Returns the new norm.
// With a separate work vector (a basis kept in a lower storage precision),
// w = A v_j is formed and orthogonalised there and V[j+1] = w / ||w|| is
// written once, by the site-wise conversion.
static double arnoldi_step(Context &c, LinOp &A, std::vector<Spinor> &V, int j, zd *h, int reorth,
                           LinOp *M = nullptr, std::vector<Spinor> *Z = nullptr, Spinor *work = nullptr) {
  Spinor &ws = work ? *work : V[j + 1];
  if (M) {
    // flexible right preconditioning: z_j = M v_j, w = A z_j
    M->apply(V[j], (*Z)[j]);
    A.apply((*Z)[j], ws);
  } else {
    A.apply(V[j], ws);
  }
  Vec w = vv(ws);
  std::vector<Vec> Vb = vecs(V, j + 1);
  std::vector<zd> h1(j + 2), h2(j + 1);
  double w2;
  if (!work) {
    Vb.push_back(w);  // <w, w> in the same pass
    c.blas->multi_dot(j + 2, Vb.data(), w, h1.data());
    Vb.pop_back();
    w2 = h1[j + 1].real();
  } else {
    c.blas->multi_dot(j + 1, Vb.data(), w, h1.data());
    w2 = reorth ? c.blas->norm2(w) : 0.0;
  }
  double n2 = c.blas->multi_axpy(j + 1, h1.data(), Vb.data(), w, true);
  for (int i = 0; i <= j; i++) h[i] = h1[i];
  if (reorth && n2 < 0.5 * w2) {
    c.blas->multi_dot(j + 1, Vb.data(), w, h2.data());
    n2 = c.blas->multi_axpy(j + 1, h2.data(), Vb.data(), w, true);
    for (int i = 0; i <= j; i++) h[i] += h2[i];
  }
  const double hn = std::sqrt(n2);
  h[j + 1] = hn;
  if (work) spinor_copy_convert(c.g, ws, V[j + 1], c.stream, hn > 0 ? 1.0 / hn : 0.0);
  else if (hn > 0) c.blas->axpby(1.0 / hn, w, 0.0, w, false);
  return hn;
}

// ---- synchronisation-free Arnoldi steps ----------------------------------
// Without the DGKS test (one Gram-Schmidt pass), an Arnoldi step needs no
// host round trip: the multi-dot leaves <V_i, w> in device memory, the
// multi-axpy reads its coefficients from there and leaves ||w||^2, and the
// normalisation reads that.  The host collects the Hessenberg columns every
// few steps (and at the end of a cycle) for the least-squares estimate, so
// the device runs ahead of the host by up to kCheck steps; a cycle may take
// up to kCheck - 1 steps more than it needed.
static constexpr int kCheck = 4;

struct DevHess {
  int m, ld;              // ld = 2 (m + 1) doubles per column
  double *d = nullptr;    // m columns of complex <V_i, w>, then m squared norms
  double *h = nullptr;    // pinned host copy
  double *x = nullptr;    // paired steps: g2 (ld), the four scalars, the three coefficients
  explicit DevHess(int m_) : m(m_), ld(2 * (m_ + 1)) {
    const size_t n = (size_t)(ld + 1) * m;
    QQCD_HIP(hipMalloc(&d, sizeof(double) * n));
    QQCD_HIP(hipMalloc(&x, sizeof(double) * (ld + 8)));
    QQCD_HIP(hipHostMalloc(&h, sizeof(double) * n, hipHostMallocDefault));
  }
  ~DevHess() {
    if (d) (void)hipFree(d);
    if (x) (void)hipFree(x);
    if (h) (void)hipHostFree(h);
  }
  double *g2() { return x; }
  double *sc() { return x + ld; }
  double *cb() { return x + ld + 4; }
  DevHess(const DevHess &) = delete;
  DevHess &operator=(const DevHess &) = delete;
  double *col(int j) { return d + (size_t)ld * j; }
  double *n2(int j) { return d + (size_t)ld * m + j; }
  // Hessenberg columns [j0, j1) into H; returns the first column whose new
  // vector vanished (breakdown), or -1
  int fetch(hipStream_t s, int j0, int j1, std::vector<zd> &H, int ldh) {
    const size_t n = (size_t)(ld + 1) * m;
    QQCD_HIP(hipMemcpyAsync(h, d, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    QQCD_HIP(hipStreamSynchronize(s));
    int jb = -1;
    for (int j = j0; j < j1; j++) {
      for (int i = 0; i <= j; i++) H[i + (size_t)ldh * j] = zd(h[(size_t)ld * j + 2 * i], h[(size_t)ld * j + 2 * i + 1]);
      const double q = h[(size_t)ld * m + j];
      H[j + 1 + (size_t)ldh * j] = q > 0 ? std::sqrt(q) : 0.0;
      if (!(q > 0) && jb < 0) jb = j;
    }
    return jb;
  }
};

static void arnoldi_step_dev(Context &c, LinOp &A, std::vector<Spinor> &V, int j, DevHess &dh, LinOp *M,
                             std::vector<Spinor> *Z, Spinor *work) {
  Spinor &ws = work ? *work : V[j + 1];
  if (M) {
    M->apply(V[j], (*Z)[j]);
    A.apply((*Z)[j], ws);
  } else {
    A.apply(V[j], ws);
  }
  Vec w = vv(ws);
  std::vector<Vec> Vb = vecs(V, j + 1);
  c.blas->multi_dot_dev(j + 1, Vb.data(), w, dh.col(j));
  c.blas->multi_axpy_dev(j + 1, dh.col(j), Vb.data(), w, dh.n2(j));
  if (work) spinor_convert_rnorm(c.g, ws, V[j + 1], dh.n2(j), c.stream);
  else c.blas->rscale_dev(w, dh.n2(j));
}

// Arnoldi steps j and j + 1 in one pair (FP16 basis with per-site norms, FP32
// work vectors w1, w2): u1 = A v_j, u2 = A u1, the dots of both against
// V_0..V_j in one pass over the basis, both updates in another, column j + 1
// of H from the Arnoldi relation (pair_columns); then v_{j+1} = u1' / ||u1'||
// and v_{j+2} from the part of A v_{j+1} orthogonal to v_0..v_{j+1}.  Half
// the passes over the basis of two single steps.  The derivation of column
// j + 1 uses the orthonormality of V_j, which the FP16 basis holds to its
// rounding, as the single-step passes do.
static int g_arnoldi_steps = 2;
void set_arnoldi_steps(int steps) {
  if (steps != 1 && steps != 2) throw std::runtime_error("Arnoldi steps per Gram-Schmidt pass: 1 or 2");
  g_arnoldi_steps = steps;
}
int arnoldi_steps() { return g_arnoldi_steps; }

static void arnoldi_pair_dev(Context &c, LinOp &A, std::vector<Spinor> &V, int j, DevHess &dh, Spinor &w1,
                             Spinor &w2) {
  if (A.apply_sq) {
    A.apply_sq(V[j], w1, w2);
  } else {
    A.apply(V[j], w1);
    A.apply(w1, w2);
  }
  Vec a1 = vv(w1), a2 = vv(w2);
  std::vector<Vec> Vb = vecs(V, j + 1);
  c.blas->pair_dot_dev(j + 1, Vb.data(), a1, a2, dh.col(j), dh.g2());
  c.blas->pair_axpy_dev(j + 1, dh.col(j), dh.g2(), Vb.data(), a1, a2, dh.sc());
  pair_columns(dh.d, dh.ld, dh.m, j, dh.g2(), dh.sc(), dh.cb(), c.stream);
  if (V[j + 1].prec == PREC_HALF && V[j + 1].norm && V[j + 2].prec == PREC_HALF && V[j + 2].norm &&
      w1.prec == PREC_SINGLE && w2.prec == PREC_SINGLE) {
    c.blas->pair_comb_basis(c.g.Vh, dh.cb(), w1, w2, dh.n2(j), V[j + 1], V[j + 2], dh.n2(j + 1));
    return;
  }
  c.blas->pair_comb_dev(dh.cb(), a1, a2, dh.n2(j + 1));
  spinor_convert_rnorm(c.g, w1, V[j + 1], dh.n2(j), c.stream);
  spinor_convert_rnorm(c.g, w2, V[j + 2], dh.n2(j + 1), c.stream);
}

// Groups of synchronisation-free steps replayed from HIP graphs: a group
// [j0, j0 + n) that already ran on the same vectors is launched as one
// graph, which removes the launch gaps between its ~13 kernels per step.
// Only for operators whose buffers stay put between applications (every
// LinOp here except a preconditioner's).  A graph captured within one call
// on a fresh basis is replayed about three times, and then the captures
// cost more than the replays save (32^3x64 mixed FP16 solve: 1.32-1.35 s
// with graphs, 1.25 s without); so by default (mode 2) graphs are kept only
// with a ProjWork, whose basis outlives the call and whose graphs serve
// every later outer step of the mixed solver.
static int g_step_graphs = 2;
void set_step_graphs(int mode) {
  if (mode < 0 || mode > 2) throw std::runtime_error("step graphs: mode 0 (never), 1 (always) or 2 (kept buffers)");
  g_step_graphs = mode;
}

struct StepGraphs {
  struct Entry {
    std::vector<const void *> key;
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
  };
  std::vector<Entry> list;
  int replays = 0;
  StepGraphs() = default;
  StepGraphs(const StepGraphs &) = delete;
  StepGraphs &operator=(const StepGraphs &) = delete;
  ~StepGraphs() {
    for (auto &x : list) {
      if (x.e) (void)hipGraphExecDestroy(x.e);
      if (x.g) (void)hipGraphDestroy(x.g);
    }
  }
};

static void arnoldi_steps_dev(Context &c, LinOp &A, std::vector<Spinor> &V, int j0, int n, DevHess &dh,
                              Spinor *work, StepGraphs *sg, Spinor *work2 = nullptr) {
  // in pairs when a second work vector is given and the group is whole pairs
  auto run = [&]() {
    if (work && work2 && j0 % 2 == 0 && n % 2 == 0)
      for (int j = j0; j < j0 + n; j += 2) arnoldi_pair_dev(c, A, V, j, dh, *work, *work2);
    else
      for (int j = j0; j < j0 + n; j++) arnoldi_step_dev(c, A, V, j, dh, nullptr, nullptr, work);
  };
  if (!sg) {
    run();
    return;
  }
  std::vector<const void *> key{(const void *)(intptr_t)j0, (const void *)(intptr_t)n, work ? work->data : nullptr,
                                work2 ? work2->data : nullptr, dh.d};
  for (int q = 0; q <= j0 + n; q++) {
    key.push_back(V[q].data);
    key.push_back(V[q].norm);
  }
  for (auto &x : sg->list)
    if (x.key == key) {
      QQCD_HIP(hipGraphLaunch(x.e, c.stream));
      sg->replays++;
      return;
    }
  // first time: run the steps, then record them for the next time
  run();
  QQCD_HIP(hipStreamSynchronize(c.stream));
  StepGraphs::Entry e;
  e.key = key;
  QQCD_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  try {
    run();
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(c.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  QQCD_HIP(hipStreamEndCapture(c.stream, &e.g));
  QQCD_HIP(hipGraphInstantiate(&e.e, e.g, nullptr, nullptr, 0));
  sg->list.push_back(std::move(e));
  e.g = nullptr;
  e.e = nullptr;
}

// dst = s * src across precisions (an FP16 destination through the conversion)
static void scaled_copy(Context &c, const Spinor &src, Spinor &dst, double s) {
  if (dst.prec == PREC_HALF || src.prec == PREC_HALF) {
    spinor_copy_convert(c.g, src, dst, c.stream, s);
  } else {
    Vec d = vv(dst);
    c.blas->axpby(s, vv(src), 0.0, d, false);
  }
}

static void alloc_basis(Context &c, std::vector<Spinor> &V, int n, Prec p) {
  V.resize(n);
  for (auto &s : V) s = alloc_spinor(c.g, p);
}
static void free_basis(std::vector<Spinor> &V) {
  for (auto &s : V) free_spinor(s);
  V.clear();
}

struct BasisGuard {
  std::vector<Spinor> &V;
  ~BasisGuard() { free_basis(V); }
};

ProjWork::~ProjWork() { reset(); }
void ProjWork::reset() {
  sg.reset();  // the graphs name the buffers: they go first
  dh.reset();
  work.reset();
  work2.reset();
  free_basis(V);
  m = 0;
}

// Deflated restart of GMRES-DR from a cycle of size j (> k):
// builds Z = [[P_k; 0], srv_perp], replaces V[0..k] by V[0..j] Z, returns the
// new (k+1) x k matrix and right-hand side.
static bool deflate(Context &c, std::vector<Spinor> &V, std::vector<Spinor> &W, int j, int k,
                    const std::vector<zd> &H, int ldh, const std::vector<zd> &cv, const std::vector<zd> &y,
                    std::vector<zd> &Hnew, std::vector<zd> &cnew, std::vector<zd> *Zout = nullptr) {
  const int rows = j + 1;
  std::vector<zd> srv(rows), P((size_t)j * k), theta(j);
  for (int i = 0; i < rows; i++) {
    zd s = cv[i];
    for (int q = 0; q < j; q++) s -= H[i + (size_t)ldh * q] * y[q];
    srv[i] = s;
  }
  if (!dense::harmonic_ritz(j, H.data(), ldh, k, P.data(), theta.data())) return false;
  std::vector<zd> Z((size_t)rows * (k + 1), 0.0);
  for (int q = 0; q < k; q++)
    for (int i = 0; i < j; i++) Z[i + (size_t)rows * q] = P[i + (size_t)j * q];
  // last column: srv orthogonalised against the first k (twice), normalised
  for (int i = 0; i < rows; i++) Z[i + (size_t)rows * k] = srv[i];
  for (int pass = 0; pass < 2; pass++)
    for (int q = 0; q < k; q++) {
      zd d = 0;
      for (int i = 0; i < rows; i++) d += std::conj(Z[i + (size_t)rows * q]) * Z[i + (size_t)rows * k];
      for (int i = 0; i < rows; i++) Z[i + (size_t)rows * k] -= d * Z[i + (size_t)rows * q];
    }
  double nz = 0;
  for (int i = 0; i < rows; i++) nz += std::norm(Z[i + (size_t)rows * k]);
  nz = std::sqrt(nz);
  if (nz == 0.0) return false;
  for (int i = 0; i < rows; i++) Z[i + (size_t)rows * k] /= nz;
  // Hnew = Z^H (H P), cnew = Z^H srv
  std::vector<zd> HP((size_t)rows * k, 0.0);
  for (int q = 0; q < k; q++)
    for (int i = 0; i < rows; i++) {
      zd s = 0;
      for (int l = 0; l < j; l++) s += H[i + (size_t)ldh * l] * P[l + (size_t)j * q];
      HP[i + (size_t)rows * q] = s;
    }
  Hnew.assign((size_t)(k + 1) * k, 0.0);
  cnew.assign(k + 1, 0.0);
  for (int q = 0; q < k; q++)
    for (int p = 0; p <= k; p++) {
      zd s = 0;
      for (int i = 0; i < rows; i++) s += std::conj(Z[i + (size_t)rows * p]) * HP[i + (size_t)rows * q];
      Hnew[p + (size_t)(k + 1) * q] = s;
    }
  for (int p = 0; p <= k; p++) {
    zd s = 0;
    for (int i = 0; i < rows; i++) s += std::conj(Z[i + (size_t)rows * p]) * srv[i];
    cnew[p] = s;
  }
  if (Zout) *Zout = Z;
  // V[0..k] <- V[0..j] Z  (formed in W)
  std::vector<Vec> Vin = vecs(V, rows), Wout = vecs(W, k + 1);
  c.blas->combine(rows, Vin.data(), k + 1, Z.data(), rows, Wout.data(), false);
  // re-orthogonalise the last vector against the first k (the reference
  // does the same, inverters.f90:3491-3508)
  std::vector<Vec> Wk = vecs(W, k);
  Vec wk = vv(W[k]);
  std::vector<zd> h(k);
  c.blas->multi_dot(k, Wk.data(), wk, h.data());
  const double n2 = c.blas->multi_axpy(k, h.data(), Wk.data(), wk, true);
  const double s = 1.0 / std::sqrt(n2);
  if (V[0].prec == W[0].prec) {
    for (int q = 0; q <= k; q++) std::swap(V[q], W[q]);
    c.blas->axpby(s, wk, 0.0, wk, false);  // wk now lives in V[k]
  } else {
    // basis kept in a lower storage precision: convert the new vectors in
    for (int q = 0; q < k; q++) spinor_copy_convert(c.g, W[q], V[q], c.stream);
    spinor_copy_convert(c.g, W[k], V[k], c.stream, s);
  }
  return true;
}

static void store_deflation(Context &c, Deflation *defl, const std::vector<Spinor> &V, int k,
                            const std::vector<zd> &Hnew, Prec p) {
  if (!defl) return;
  defl->clear();
  defl->k = k;
  defl->prec = p;
  defl->H = Hnew;
  defl->V.resize(k + 1);
  for (int q = 0; q <= k; q++) {
    defl->V[q] = alloc_spinor(c.g, p);
    Vec d = vv(defl->V[q]);
    c.blas->copy(vv(V[q]), d);
  }
}

SolveStats gmresdr(Context &c, LinOp &A, const Spinor &b, Spinor &x, int m, int k, double tol, int itermin,
                   int maxiter, Deflation *defl, int reorth, LinOp *M, const CycleHook *hook,
                   const StepHook *step_hook) {
  if (m < 2 || k < 0 || k >= m) throw std::runtime_error("gmresdr: need 0 <= k < m and m >= 2");
  if (M && defl) throw std::runtime_error("gmresdr: no deflation space is kept with a preconditioner");
  SolveStats st;
  const Prec p = A.prec, vp = A.basis();
  if (M && vp != p) throw std::runtime_error("gmresdr: the flexible variant keeps its basis in the work precision");
  std::vector<Spinor> V, W, Z, ZW;
  BasisGuard g1{V}, g2{W}, g3{Z}, g4{ZW};
  alloc_basis(c, V, m + 1, vp);
  alloc_basis(c, W, k + 1, p);
  if (M) {
    alloc_basis(c, Z, m, p);
    alloc_basis(c, ZW, k, p);
  }
  std::unique_ptr<SpinorBuf> work, work2;
  if (vp != p) work.reset(new SpinorBuf(c.g, p));
  // paired Arnoldi steps: one Gram-Schmidt pass, FP32 work over an FP16 basis
  if (g_arnoldi_steps == 2 && !reorth && !M && p == PREC_SINGLE && vp == PREC_HALF) work2.reset(new SpinorBuf(c.g, p));
  std::unique_ptr<DevHess> dh;
  if (!reorth) dh.reset(new DevHess(m));
  std::unique_ptr<StepGraphs> sg;
  if (dh && !M && g_step_graphs == 1) sg.reset(new StepGraphs);
  std::vector<zd> Zmat;
  SpinorBuf r(c.g, p);
  double rn = residual(c, A, b, x, r.s, st);
  const double r0 = rn;
  st.rel_res = 1.0;
  if (r0 == 0.0) {
    st.converged = 1;
    st.rel_res = 0.0;
    return st;
  }
  const int ldh = m + 1;
  std::vector<zd> H((size_t)ldh * m, 0.0), cv(m + 1, 0.0), y(m), col(m + 2);
  scaled_copy(c, r.s, V[0], 1.0 / rn);
  cv[0] = rn;
  int j = 0;
  bool have_defl = false;
  std::vector<zd> Hnew, cnew;
  for (;;) {
    // Arnoldi
    int jchk = j;
    while (j < m) {
      if (dh) {
        if (sg || work2) {
          const int n = std::min(std::min(kCheck, m - j), std::max(1, maxiter - st.iters));
          arnoldi_steps_dev(c, A, V, j, n, *dh, work ? &work->s : nullptr, sg.get(), work2 ? &work2->s : nullptr);
          st.matvecs += n;
          st.iters += n;
          j += n;
        } else {
          arnoldi_step_dev(c, A, V, j, *dh, M, M ? &Z : nullptr, work ? &work->s : nullptr);
          st.matvecs++;
          st.iters++;
          j++;
          if (j < m && st.iters < maxiter && j - jchk < kCheck) continue;
        }
        const int jb = dh->fetch(c.stream, jchk, j, H, ldh);
        jchk = j;
        if (jb >= 0) {
          j = jb + 1;  // the basis ended at column jb
          break;
        }
        const double est = dense::lsq(j + 1, j, H.data(), ldh, cv.data(), y.data());
        if ((est < tol * r0 && st.iters >= itermin) || st.iters >= maxiter) break;
        continue;
      }
      const double hn = arnoldi_step(c, A, V, j, col.data(), reorth, M, M ? &Z : nullptr, work ? &work->s : nullptr);
      st.matvecs++;
      st.iters++;
      for (int i = 0; i <= j + 1; i++) H[i + (size_t)ldh * j] = col[i];
      j++;
      const double est = dense::lsq(j + 1, j, H.data(), ldh, cv.data(), y.data());
      if (step_hook && *step_hook && !M) (*step_hook)(x, V, j, y.data());
      if ((est < tol * r0 && st.iters >= itermin) || hn == 0.0 || st.iters >= maxiter) break;
    }
    // solution update (x += V y, or Z y when preconditioned) and true residual
    dense::lsq(j + 1, j, H.data(), ldh, cv.data(), y.data());
    {
      std::vector<Vec> Vb = vecs(M ? Z : V, j);
      Vec xv = vv(x);
      c.blas->combine(j, Vb.data(), 1, y.data(), j, &xv, true);
    }
    rn = residual(c, A, b, x, r.s, st);
    st.cycles++;
    st.rel_res = rn / r0;
    if (!std::isfinite(rn)) {
      // a NaN or infinity reached x: no restart can remove it
      st.nonfinite = 1;
      if (defl) defl->clear();
      return st;
    }
    if (hook && *hook) (*hook)(x, r.s, st);
    const bool done = (rn < tol * r0 && st.iters >= itermin) || st.iters >= maxiter;
    if (done) {
      st.converged = rn < tol * r0;
      // keep the harmonic Ritz space of the final cycle for later right-hand sides
      if (defl && j > k && k > 0 && deflate(c, V, W, j, k, H, ldh, cv, y, Hnew, cnew)) {
        store_deflation(c, defl, V, k, Hnew, vp);
        have_defl = true;
      }
      if (defl && !have_defl) defl->clear();
      return st;
    }
    if (j > k && k > 0 && deflate(c, V, W, j, k, H, ldh, cv, y, Hnew, cnew, M ? &Zmat : nullptr)) {
      if (M) {
        // Z_k <- Z_j P_k (the first k columns of the restart combination; A Z_k = V_{k+1} Hbar_k)
        std::vector<Vec> Zin = vecs(Z, j), Zo = vecs(ZW, k);
        c.blas->combine(j, Zin.data(), k, Zmat.data(), j + 1, Zo.data(), false);
        for (int q = 0; q < k; q++) std::swap(Z[q], ZW[q]);
      }
      std::fill(H.begin(), H.end(), zd(0.0));
      std::fill(cv.begin(), cv.end(), zd(0.0));
      for (int q = 0; q < k; q++)
        for (int i = 0; i <= k; i++) H[i + (size_t)ldh * q] = Hnew[i + (size_t)(k + 1) * q];
      for (int i = 0; i <= k; i++) cv[i] = cnew[i];
      j = k;
      if (defl) {
        store_deflation(c, defl, V, k, Hnew, vp);
        have_defl = true;
      }
    } else {
      // plain restart from the true residual
      scaled_copy(c, r.s, V[0], 1.0 / rn);
      std::fill(H.begin(), H.end(), zd(0.0));
      std::fill(cv.begin(), cv.end(), zd(0.0));
      cv[0] = rn;
      j = 0;
    }
  }
}

SolveStats gmresproj(Context &c, LinOp &A, const Spinor &b, Spinor &x, int m, const Deflation &defl, double tol,
                     int itermin, int maxiter, int reorth, ResHistory *hist, ProjWork *ws) {
  SolveStats st;
  const Prec p = A.prec, vp = A.basis();
  const int k = defl.k;
  if (m < 1) throw std::runtime_error("gmresproj: m must be >= 1");
  // the buffers: this call's own, or the workspace's (rebuilt when the shape
  // or the precisions change)
  std::vector<Spinor> Vown;
  BasisGuard g1{Vown};
  std::unique_ptr<SpinorBuf> work_own, work2_own;
  std::unique_ptr<DevHess> dh_own;
  std::unique_ptr<StepGraphs> sg_own;
  std::vector<Spinor> *Vp = &Vown;
  SpinorBuf *work = nullptr, *work2 = nullptr;
  DevHess *dh = nullptr;
  StepGraphs *sg = nullptr;
  // paired Arnoldi steps: one Gram-Schmidt pass, FP32 work over an FP16 basis
  const bool pairs = g_arnoldi_steps == 2 && !reorth && p == PREC_SINGLE && vp == PREC_HALF;
  if (ws) {
    if (ws->m != m || ws->p != p || ws->vp != vp || (!reorth) != (ws->dh != nullptr) ||
        pairs != (ws->work2 != nullptr)) {
      ws->reset();
      alloc_basis(c, ws->V, m + 1, vp);
      if (vp != p) ws->work.reset(new SpinorBuf(c.g, p));
      if (pairs) ws->work2.reset(new SpinorBuf(c.g, p));
      if (!reorth) ws->dh.reset(new DevHess(m));
      ws->m = m;
      ws->p = p;
      ws->vp = vp;
    }
    if (ws->dh && g_step_graphs && !ws->sg) ws->sg.reset(new StepGraphs);
    if (!g_step_graphs) ws->sg.reset();
    Vp = &ws->V;
    work = ws->work.get();
    work2 = ws->work2.get();
    dh = ws->dh.get();
    sg = ws->sg.get();
  } else {
    alloc_basis(c, Vown, m + 1, vp);
    if (vp != p) work_own.reset(new SpinorBuf(c.g, p));
    if (pairs) work2_own.reset(new SpinorBuf(c.g, p));
    if (!reorth) dh_own.reset(new DevHess(m));
    if (dh_own && g_step_graphs == 1) sg_own.reset(new StepGraphs);
    work = work_own.get();
    work2 = work2_own.get();
    dh = dh_own.get();
    sg = sg_own.get();
  }
  std::vector<Spinor> &V = *Vp;
  SpinorBuf r(c.g, p);
  double rn = residual(c, A, b, x, r.s, st);
  const double r0 = rn;
  if (r0 == 0.0) {
    st.converged = 1;
    return st;
  }
  const int ldh = m + 1;
  std::vector<zd> H((size_t)ldh * m), cv(m + 1), y(m), col(m + 2);
  for (;;) {
    // projection over the deflation space: min || V_{k+1}^H r - Hbar_k d ||
    if (k > 0 && (int)defl.V.size() == k + 1) {
      std::vector<Vec> Dv = vecs(defl.V, k + 1);
      std::vector<zd> cd(k + 1), d(k);
      c.blas->multi_dot(k + 1, Dv.data(), vv(r.s), cd.data());
      dense::lsq(k + 1, k, defl.H.data(), k + 1, cd.data(), d.data());
      Vec xv = vv(x);
      c.blas->combine(k, Dv.data(), 1, d.data(), k, &xv, true);
      rn = residual(c, A, b, x, r.s, st);
      st.rel_res = rn / r0;
      if (!std::isfinite(rn)) {
        st.nonfinite = 1;
        break;
      }
      if ((rn < tol * r0 && st.iters >= itermin) || st.iters >= maxiter) break;
    }
    // GMRES(m) cycle from r
    std::fill(H.begin(), H.end(), zd(0.0));
    std::fill(cv.begin(), cv.end(), zd(0.0));
    scaled_copy(c, r.s, V[0], 1.0 / rn);
    cv[0] = rn;
    int j = 0, jchk = 0;
    while (j < m) {
      if (dh) {
        if (sg || work2) {
          const int n = std::min(std::min(kCheck, m - j), std::max(1, maxiter - st.iters));
          arnoldi_steps_dev(c, A, V, j, n, *dh, work ? &work->s : nullptr, sg, work2 ? &work2->s : nullptr);
          st.matvecs += n;
          st.iters += n;
          j += n;
        } else {
          arnoldi_step_dev(c, A, V, j, *dh, nullptr, nullptr, work ? &work->s : nullptr);
          st.matvecs++;
          st.iters++;
          j++;
          if (j < m && st.iters < maxiter && j - jchk < kCheck) continue;
        }
        const int jb = dh->fetch(c.stream, jchk, j, H, ldh);
        const int jlast = jb >= 0 ? jb + 1 : j;
        if (hist)  // one entry per Arnoldi step of this fetch
          for (int jj = jchk + 1; jj <= jlast; jj++)
            hist->emplace_back(st.matvecs - (j - jj), dense::lsq(jj + 1, jj, H.data(), ldh, cv.data(), y.data()));
        jchk = j;
        if (jb >= 0) {
          j = jb + 1;
          break;
        }
        const double est = dense::lsq(j + 1, j, H.data(), ldh, cv.data(), y.data());
        if ((est < tol * r0 && st.iters >= itermin) || st.iters >= maxiter) break;
        continue;
      }
      const double hn = arnoldi_step(c, A, V, j, col.data(), reorth, nullptr, nullptr, work ? &work->s : nullptr);
      st.matvecs++;
      st.iters++;
      for (int i = 0; i <= j + 1; i++) H[i + (size_t)ldh * j] = col[i];
      j++;
      const double est = dense::lsq(j + 1, j, H.data(), ldh, cv.data(), y.data());
      if (hist) hist->emplace_back(st.matvecs, est);
      if ((est < tol * r0 && st.iters >= itermin) || hn == 0.0 || st.iters >= maxiter) break;
    }
    dense::lsq(j + 1, j, H.data(), ldh, cv.data(), y.data());
    {
      std::vector<Vec> Vb = vecs(V, j);
      Vec xv = vv(x);
      c.blas->combine(j, Vb.data(), 1, y.data(), j, &xv, true);
    }
    rn = residual(c, A, b, x, r.s, st);
    st.cycles++;
    st.rel_res = rn / r0;
    if (!std::isfinite(rn)) {
      st.nonfinite = 1;
      break;
    }
    if ((rn < tol * r0 && st.iters >= itermin) || st.iters >= maxiter) break;
  }
  st.converged = st.rel_res < tol;
  return st;
}

SolveStats bicgstab(Context &c, LinOp &A, const Spinor &b, Spinor &x, double tol, int itermin, int maxiter,
                    const IterHook *hook) {
  SolveStats st;
  const Prec p = A.prec;
  SpinorBuf r(c.g, p), rhat(c.g, p), pp(c.g, p), v(c.g, p), t(c.g, p);
  Vec R = vv(r.s), Rh = vv(rhat.s), P = vv(pp.s), Vv = vv(v.s), T = vv(t.s), X = vv(x);
  double rn = residual(c, A, b, x, r.s, st);
  const double r0 = rn;
  if (r0 == 0.0) {
    st.converged = 1;
    return st;
  }
  for (;;) {
    c.blas->copy(R, Rh);
    c.blas->zero(P);
    c.blas->zero(Vv);
    zd rho = 1.0, alpha = 1.0, omega = 1.0;
    double rr = rn;
    while (!((rr < tol * r0 && st.iters >= itermin) || st.iters >= maxiter)) {
      const zd oldrho = rho;
      rho = c.blas->dot(Rh, R);
      const zd beta = (rho / oldrho) * (alpha / omega);
      // p = r + beta (p - omega v)
      c.blas->axpby(-beta * omega, Vv, beta, P, false);
      c.blas->axpby(1.0, R, 1.0, P, false);
      A.apply(pp.s, v.s);
      st.matvecs++;
      alpha = rho / c.blas->dot(Rh, Vv);
      c.blas->axpby(-alpha, Vv, 1.0, R, false);
      A.apply(r.s, t.s);
      st.matvecs++;
      const zd top = c.blas->dot(T, R);
      const double bot = c.blas->norm2(T);
      omega = top / bot;
      c.blas->axpby(omega, R, 1.0, X, false);
      c.blas->axpby(alpha, P, 1.0, X, false);
      rr = std::sqrt(c.blas->axpby(-omega, T, 1.0, R, true));
      st.iters++;
      if (hook && *hook) (*hook)(st.iters, rr, r0);
    }
    // confirm on the true residual
    rn = residual(c, A, b, x, r.s, st);
    st.cycles++;
    st.rel_res = rn / r0;
    if (!std::isfinite(rn)) {
      st.nonfinite = 1;
      break;
    }
    if ((rn < tol * r0 && st.iters >= itermin) || st.iters >= maxiter) break;
  }
  st.converged = st.rel_res < tol;
  return st;
}

}  // namespace qqcd

namespace qqcd {

FaultPlan &fault_plan() {
  static FaultPlan f;
  return f;
}

MixedStats solve_mixed(Context &c, const double kappa[2], const Spinor &b, Spinor &x, const MixedParams &p0,
                       Deflation *defl) {
  MixedStats ms;
  const FaultPlan fault = fault_plan();
  MixedParams p = p0;
  if (fault.maxiter > 0) p.maxiter = std::min(p.maxiter, fault.maxiter);
  // the preconditioner as the inner solves see it (its output replaced when a fault is planned)
  LinOp pre_f;
  if (p.precond && fault.precond) {
    LinOp *orig = p.precond;
    pre_f.prec = orig->prec;
    pre_f.store = orig->store;
    const int mode = fault.precond;
    uint64_t seed = 977;
    pre_f.apply = [&c, orig, mode, seed](const Spinor &a, Spinor &z) mutable {
      orig->apply(a, z);
      if (mode == 1) fill_spinor(c.g, z, std::nan(""), std::nan(""), c.stream);
      else if (mode == 2) fill_spinor(c.g, z, 0.0, 0.0, c.stream);
      else random_spinor(c.g, z, seed++, c.stream);
    };
    p.precond = &pre_f;
  }
  SchurOp A64(c, PREC_DOUBLE, 18, kappa);
  LinOp op64 = A64.op();
  // operator levels: the requested storage, then FP32, then FP64 (fallback
  // when an outer step stalls)
  Prec op_prec = p.op_prec;
  std::unique_ptr<SchurOp> Alo(new SchurOp(c, op_prec, p.op_recon, kappa));
  std::unique_ptr<SpinorBuf> in_lo, out_lo;
  // inner solves work in FP32 (or FP64); an FP16 basis is kept in storage
  // only (inputs of the operator may then be FP16 basis vectors)
  const Prec work = p.basis == PREC_HALF ? PREC_SINGLE : p.basis;
  auto make_inner = [&]() {
    LinOp o;
    o.prec = work;
    if (p.basis == PREC_HALF) o.store = PREC_HALF;
    in_lo.reset(new SpinorBuf(c.g, op_prec));
    out_lo.reset(new SpinorBuf(c.g, op_prec));
    // an FP16 operator writes an FP32 result itself (the hop's epilogue)
    const bool direct32 = op_prec == PREC_HALF && work == PREC_SINGLE;
    o.apply = [&, direct32](const Spinor &a, Spinor &r) {
      const Spinor *in = &a;
      if (a.prec != op_prec) {
        spinor_copy_convert(c.g, a, in_lo->s, c.stream);
        in = &in_lo->s;
      }
      if (r.prec == op_prec) {
        (*Alo)(*in, r);
      } else if (direct32 && r.prec == PREC_SINGLE) {
        (*Alo)(*in, out_lo->s, &r, true);
      } else {
        (*Alo)(*in, out_lo->s);
        spinor_copy_convert(c.g, out_lo->s, r, c.stream);
      }
    };
    if (direct32) {
      // r1 = A a in FP32 and, rounded to FP16 by the same launch, the input
      // of r2 = A r1: the operator reads its input in FP16, so r2 is A
      // applied to r1 as the operator sees it
      o.apply_sq = [&](const Spinor &a, Spinor &r1, Spinor &r2) {
        if (r1.prec != PREC_SINGLE || r2.prec != PREC_SINGLE) throw std::runtime_error("apply_sq: FP32 results");
        const Spinor *in = &a;
        if (a.prec != op_prec) {
          spinor_copy_convert(c.g, a, in_lo->s, c.stream);
          in = &in_lo->s;
        }
        (*Alo)(*in, out_lo->s, &r1, false);
        (*Alo)(out_lo->s, in_lo->s, &r2, true);
      };
    }
    return o;
  };
  LinOp inner = make_inner();
  double inner_tol = p.inner_tol;
  SpinorBuf r(c.g, PREC_DOUBLE), rl(c.g, work), el(c.g, work);
  SolveStats dummy;
  const double bn = std::sqrt(c.blas->norm2(as_vec(b)));
  double rn = residual(c, op64, b, x, r.s, dummy);
  ms.fp64_matvecs++;
  if (!std::isfinite(rn) && std::isfinite(bn)) {
    // a starting guess with a NaN or infinite entry: start from zero
    Vec xv = as_vec(x);
    c.blas->zero(xv);
    rn = residual(c, op64, b, x, r.s, dummy);
    ms.fp64_matvecs++;
  }
  Deflation local;
  Deflation &D = defl ? *defl : local;
  // the later outer steps' GMRES-Proj keeps its basis and step graphs here
  ProjWork pw;
  bool use_pre = p.precond != nullptr;
  if (D.k > 0 && (D.prec != p.basis || D.k != p.k)) D.clear();
  while (rn >= p.tol * bn && ms.outer < p.max_outer && ms.inner.iters < p.maxiter) {
    if (op_prec == PREC_DOUBLE && p.basis == PREC_DOUBLE && ms.outer > 0 && ms.stalled &&
        p.maxiter > ms.inner.iters) {
      // last resort: plain FP64 GMRES-DR on the true residual equation
      SolveStats st = gmresdr(c, op64, b, x, p.m, p.k, p.tol * bn / rn * 0.5, 0, p.maxiter - ms.inner.iters, nullptr);
      ms.inner.iters += st.iters;
      ms.inner.matvecs += st.matvecs;
      ms.fallback = 2;
      rn = residual(c, op64, b, x, r.s, dummy);
      ms.outer++;
      break;
    }
    Vec rv = as_vec(r.s), rlv = as_vec(rl.s), elv = as_vec(el.s), xv = as_vec(x);
    c.blas->copy(rv, rlv);
    c.blas->zero(elv);
    // never ask the inner solve for more than the outer target needs
    const double itol = std::max(inner_tol, 0.5 * p.tol * bn / rn);
    SolveStats st;
    const int budget = p.maxiter - ms.inner.iters;
    if (use_pre) {
      st = gmresdr(c, inner, rl.s, el.s, p.m, p.k, itol, 0, std::min(budget, p.first_cycles * p.m), nullptr,
                   p.reorth, p.precond);
      ms.precond_steps += st.iters;
    } else if (D.k == 0)
      st = gmresdr(c, inner, rl.s, el.s, p.m, p.k, itol, 0, std::min(budget, p.first_cycles * p.m), &D, p.reorth);
    else
      st = gmresproj(c, inner, rl.s, el.s, p.m - p.k, D, itol, 0, std::min(budget, p.later_cycles * (p.m - p.k)),
                     p.reorth, nullptr, &pw);
    ms.inner.iters += st.iters;
    ms.inner.matvecs += st.matvecs;
    ms.inner.cycles += st.cycles;
    if (fault.batch_rhs < 0 && ms.outer < fault.corrupt_outer) fill_spinor(c.g, el.s, std::nan(""), std::nan(""), c.stream);
    // the correction is added only when every entry is finite: a NaN from the
    // preconditioner or a low-precision overflow must not reach x (x + e with
    // finite x and e cannot overflow the FP64 residual, ||e|| < 1e154)
    const double en2 = c.blas->norm2(elv);
    const double rn_old = rn;
    if (std::isfinite(en2)) {
      c.blas->axpby(1.0, elv, 1.0, xv, false);
      rn = residual(c, op64, b, x, r.s, dummy);
      ms.fp64_matvecs++;
    } else {
      ms.rejected++;
    }
    ms.outer++;
    // written so that a NaN residual counts as a stall
    ms.stalled = !(rn <= p.stall * rn_old);
    if (ms.stalled && use_pre) {
      // the unpreconditioned solver takes over
      use_pre = false;
      ms.precond_dropped = 1;
      ms.stalled = 0;
    } else if (ms.stalled) {
      // the low-precision operator no longer reduces the true residual:
      // climb to the next operator precision
      if (op_prec == PREC_HALF) {
        op_prec = PREC_SINGLE;
        ms.fallback = 1;
      } else {
        op_prec = PREC_DOUBLE;
        ms.fallback = 2;
      }
      if (op_prec == PREC_DOUBLE && p.basis != PREC_DOUBLE) {
        // the FP32 basis cannot carry an FP64 operator: finish with plain FP64 GMRES-DR
        if (p.maxiter <= ms.inner.iters) break;
        SolveStats st2 = gmresdr(c, op64, b, x, p.m, p.k, 0.5 * p.tol * bn / rn, 0, p.maxiter - ms.inner.iters,
                                 nullptr);
        ms.inner.iters += st2.iters;
        ms.inner.matvecs += st2.matvecs;
        rn = residual(c, op64, b, x, r.s, dummy);
        ms.outer++;
        break;
      }
      pw.reset();  // its graphs name the old operator's kernels and buffers
      Alo.reset(new SchurOp(c, op_prec, op_prec == PREC_DOUBLE ? 18 : p.op_recon, kappa));
      inner = make_inner();
      inner_tol = op_prec == PREC_SINGLE ? std::min(inner_tol, 1e-6) : inner_tol;
    }
  }
  ms.rel_res = bn > 0 ? rn / bn : 0.0;
  ms.converged = rn < p.tol * bn;
  last_mixed_stats() = ms;
  return ms;
}

MixedStats &last_mixed_stats() {
  static MixedStats m;
  return m;
}

void prep_source(Context &c, const double kappa[2], const Spinor &b_e, const Spinor &b_o, Spinor &src) {
  const double k = kappa[0], a = 2.0 * kappa[0] * kappa[1];
  SpinorBuf t(c.g, PREC_DOUBLE);
  local_spin(c.g, b_o, t.s, 1.0, -a, c.stream);
  Epilogue es;
  es.a0 = 1.0 / (k * k);
  es.b0 = 1.0 / (k * (1.0 + a * a));
  dslash(c.g, c.U64, t.s, src, 0, 0, es, &b_e, c.stream);
}

void reconstruct_odd(Context &c, const double kappa[2], const Spinor &b_o, const Spinor &x_e, Spinor &x_o) {
  const double k = kappa[0], a = 2.0 * kappa[0] * kappa[1];
  const double f = 1.0 / (1.0 + a * a);
  Epilogue er;
  er.a0 = f;
  er.a1 = -a * f;
  er.b0 = k * f;
  er.b1 = -a * k * f;
  dslash(c.g, c.U64, x_e, x_o, 1, 0, er, &b_o, c.stream);
}

// ||b - M x|| / ||b|| for M = (1 + a i g5) - kappa H on the whole lattice
double full_residual(Context &c, const double kappa[2], const Spinor &b_e, const Spinor &b_o, const Spinor &x_e,
                     const Spinor &x_o) {
  const double k = kappa[0], a = 2.0 * kappa[0] * kappa[1];
  Epilogue e;
  e.a0 = -1.0;
  e.a1 = -a;
  e.b0 = k;
  SpinorBuf t(c.g, PREC_DOUBLE);
  double num = 0, den = 0;
  const Spinor *xs[2] = {&x_e, &x_o}, *bs[2] = {&b_e, &b_o};
  for (int par = 0; par < 2; par++) {
    dslash(c.g, c.U64, *xs[1 - par], t.s, par, 0, e, xs[par], c.stream);
    Vec tv = as_vec(t.s);
    num += c.blas->axpby(1.0, as_vec(*bs[par]), 1.0, tv, true);
    den += c.blas->norm2(as_vec(*bs[par]));
  }
  return den > 0 ? std::sqrt(num / den) : 0.0;
}

}  // namespace qqcd
