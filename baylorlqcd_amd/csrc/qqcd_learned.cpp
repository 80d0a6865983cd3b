// qqcd_learned.cpp -- the learned preconditioner: (residual, error) pairs
// harvested from the plain solver, training with the device forward and
// backward kernels and Adam, and the mixed-precision solve with the network
// as the right preconditioner of flexible GMRES-DR.  None of this has a
// counterpart in the reference (qqcd has no preconditioner on this path).
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <vector>

#include "../../include/qqcd_hip.h"
#include "qqcd_nn.h"
#include "qqcd_solvers.h"

using namespace qqcd;
using dense::zd;

#define API_BEGIN try {
#define API_END                     \
  }                                 \
  catch (const std::exception &e) { \
    set_error(e.what());            \
    return -1;                      \
  }                                 \
  return 0;

namespace qqcd {

// The library's network.  The host weights outlive a context, so a network
// trained on one lattice is applied on another; the device copy belongs to
// the context it was made for.
struct NNState {
  std::vector<NNLayer> layers;
  std::unique_ptr<NNModel> model;
  const Context *owner = nullptr;
};
static NNState g_nn;

void nn_release_device() {
  g_nn.model.reset();
  g_nn.owner = nullptr;
}

static NNModel &nn_model(Context &c) {
  if (g_nn.layers.empty()) throw std::runtime_error("no preconditioner network: train or load one first");
  if (!g_nn.model || g_nn.owner != &c) {
    g_nn.model.reset(new NNModel());
    g_nn.owner = &c;
  }
  NNModel &M = *g_nn.model;
  if (!M.ready) {
    M.layers = g_nn.layers;
    M.upload(c.stream);
    M.alloc(c.g);
  }
  return M;
}

static void nn_weights_changed(Context &c) {
  if (g_nn.model && g_nn.owner == &c) {
    g_nn.model->layers = g_nn.layers;
    g_nn.model->upload(c.stream);
  }
}

// Network initialised near the point-Jacobi preconditioner kappa^2 * 1
// (path 0, channel 0 carries the input through every layer) plus small
// random weights elsewhere.
static void nn_init_identity(const std::vector<int> &ch, double scale, double noise, uint64_t seed) {
  NNModel tmp;
  tmp.init(ch, seed);
  const int L = (int)tmp.layers.size();
  const double per = std::pow(scale, 1.0 / L);
  for (auto &layer : tmp.layers) {
    for (auto &w : layer.W) w *= noise / (1.0 / std::sqrt(2.0 * layer.cin * NN_PATHS * 4));
    for (int s = 0; s < 4; s++) layer.W[(size_t)4 * s + s] += per;  // W[0][0][0] = per * 1
  }
  g_nn.layers = tmp.layers;
}

struct Pair {
  Spinor r, e;  // even checkerboard, FP16 with per-site norms, ||r|| = 1
  double e2;    // ||e||^2
  int conf;
};

struct TrainSet {
  std::vector<NNLinks> U;  // the BF16 links of every configuration
  std::vector<Pair> pairs;
  std::unique_ptr<SpinorBuf> r32, e32;  // a pair in FP32, as the network reads it
  ~TrainSet() {
    for (auto &u : U) nn_free_links(u);
    for (auto &p : pairs) {
      free_spinor(p.r);
      free_spinor(p.e);
    }
  }
};

// Pairs (r_j, e_j) from every Arnoldi step j of plain FP64 GMRES-DR solves
// of Gaussian sources: the step's iterate x_j = x + V y_j (the reference
// forms it at every step, inverters.f90:3242-3255), e_j = x* - x_j and
// r_j = b - A x_j, both divided by ||r_j||, kept in FP16 with per-site norms
// (6.8 MB per pair at 16^3x32); the starting guess x_0 = 0 gives the first.
// Sources are solved until max_pairs pairs are held or nsrc sources are done.
static void harvest(Context &c, const double kappa[2], int nsrc, int m, int k, double tol, uint64_t seed, int conf,
                    int max_pairs, TrainSet &T) {
  SchurOp A(c, PREC_DOUBLE, 18, kappa);
  LinOp op = A.op();
  for (int src = 0; src < nsrc && (int)T.pairs.size() < max_pairs; src++) {
    SpinorBuf b(c.g, PREC_DOUBLE), x(c.g, PREC_DOUBLE);
    random_spinor(c.g, b.s, seed * 1000003ull + src, c.stream);
    const int room = max_pairs - (int)T.pairs.size() - 1;  // the x_0 = 0 pair comes first
    std::vector<Spinor> xs;
    StepHook hook = [&](const Spinor &x0, const std::vector<Spinor> &V, int j, const dense::zd *y) {
      if ((int)xs.size() >= room) return;
      xs.push_back(alloc_spinor(c.g, PREC_DOUBLE));
      Vec av = as_vec(xs.back());
      std::vector<Vec> Vb(j);
      for (int q = 0; q < j; q++) Vb[q] = as_vec(V[q]);
      c.blas->combine(j, Vb.data(), 1, y, j, &av, false);
      c.blas->axpby(1.0, as_vec(x0), 1.0, av, false);
    };
    try {
      gmresdr(c, op, b.s, x.s, m, k, tol, 0, 5000 * m, nullptr, 1, nullptr, nullptr, &hook);
      SpinorBuf e64(c.g, PREC_DOUBLE), r64(c.g, PREC_DOUBLE);
      // iterates whose error is at the rounding level of x* (the last step's
      // iterate is x* itself) carry no target
      const double xn2 = c.blas->norm2(as_vec(x.s));
      for (int i = -1; i < (int)xs.size() && (int)T.pairs.size() < max_pairs; i++) {
        Vec ev = as_vec(e64.s), rv = as_vec(r64.s);
        c.blas->copy(as_vec(x.s), ev);
        if (i < 0) {
          c.blas->copy(as_vec(b.s), rv);
        } else {
          c.blas->axpby(-1.0, as_vec(xs[i]), 1.0, ev, false);
          op.apply(xs[i], r64.s);
          c.blas->axpby(1.0, as_vec(b.s), -1.0, rv, false);
        }
        const double rn = std::sqrt(c.blas->norm2(rv));
        const double en2 = c.blas->norm2(ev);
        if (!(rn > 0) || !(en2 > 1e-24 * xn2)) continue;
        Pair p;
        p.conf = conf;
        p.e2 = en2 / (rn * rn);
        p.r = alloc_spinor(c.g, PREC_HALF);
        p.e = alloc_spinor(c.g, PREC_HALF);
        T.pairs.push_back(p);
        spinor_copy_convert(c.g, r64.s, T.pairs.back().r, c.stream, 1.0 / rn);
        spinor_copy_convert(c.g, e64.s, T.pairs.back().e, c.stream, 1.0 / rn);
      }
    } catch (...) {
      for (auto &s : xs) free_spinor(s);
      throw;
    }
    for (auto &s : xs) free_spinor(s);
  }
}

// the pair's r and e in FP32 scratch spinors
static void pair_fp32(Context &c, TrainSet &T, const Pair &p) {
  if (!T.r32) T.r32.reset(new SpinorBuf(c.g, PREC_SINGLE));
  if (!T.e32) T.e32.reset(new SpinorBuf(c.g, PREC_SINGLE));
  spinor_copy_convert(c.g, p.r, T.r32->s, c.stream);
  spinor_copy_convert(c.g, p.e, T.e32->s, c.stream);
}

// One pass over a pair: forward, loss ||N(r) - e||^2 / ||e||^2, backward
// into dW (accumulated).  Returns the loss.
static double pair_grad(Context &c, NNModel &M, TrainSet &T, const Pair &p, double *d_scr,
                        std::vector<std::vector<std::complex<double>>> &dW) {
  const int L = (int)M.layers.size();
  const NNLinks &U = T.U[p.conf];
  pair_fp32(c, T, p);
  nn_from_spinor(c.g, T.r32->s, M.acts[0], 1.f, c.stream);
  for (int l = 0; l < L; l++) nn_layer(c.g, U, M.dev[l], 0, M.acts[l], M.acts[l + 1], c.stream);
  const float f = (float)(1.0 / p.e2);
  // g = (N(r) - e) / ||e||^2 = dL/dconj(N(r)); loss = ||g||^2 ||e||^2
  const double g2 = nn_diff(c.g, M.acts[L], T.e32->s, f, f, M.grads[L], d_scr, c.stream);
  std::vector<std::complex<double>> tmp;
  for (int l = L - 1; l >= 0; l--) {
    tmp.resize(dW[l].size());
    nn_wgrad(c.g, U, M.layers[l].cin, M.layers[l].cout, M.acts[l], M.grads[l + 1], tmp.data(), c.stream);
    for (size_t i = 0; i < tmp.size(); i++) dW[l][i] += tmp[i];
    if (l > 0) nn_layer(c.g, U, M.dev[l], 1, M.grads[l + 1], M.grads[l], c.stream);
  }
  return g2 * p.e2;
}

LinOp nn_precond(Context &c) {
  NNModel &M = nn_model(c);
  LinOp pre;
  pre.prec = PREC_SINGLE;
  pre.apply = [&c, &M](const Spinor &a, Spinor &z) { nn_apply(c, M, a, z); };
  return pre;
}

// loss only (evaluation)
static double pair_loss(Context &c, NNModel &M, TrainSet &T, const Pair &p, double *d_scr) {
  const int L = (int)M.layers.size();
  pair_fp32(c, T, p);
  nn_from_spinor(c.g, T.r32->s, M.acts[0], 1.f, c.stream);
  for (int l = 0; l < L; l++) nn_layer(c.g, T.U[p.conf], M.dev[l], 0, M.acts[l], M.acts[l + 1], c.stream);
  const float f = (float)(1.0 / std::sqrt(p.e2));
  return nn_diff(c.g, M.acts[L], T.e32->s, f, f, M.grads[L], d_scr, c.stream);
}

}  // namespace qqcd

// ---------------------------------------------------------------------------

extern "C" int qqcd_nn_init(int nlayers, const int *channels, double scale, double noise, unsigned long long seed) {
  API_BEGIN
  if (nlayers < 1 || nlayers > 16) throw std::runtime_error("nn_init: 1..16 layers");
  std::vector<int> ch(channels, channels + nlayers + 1);
  nn_init_identity(ch, scale, noise, seed);
  if (g_ctx) nn_weights_changed(*g_ctx);
  API_END
}

extern "C" int qqcd_nn_set_weights(int nlayers, const int *channels, const double *W) {
  API_BEGIN
  if (nlayers < 1 || nlayers > 16) throw std::runtime_error("nn_set_weights: 1..16 layers");
  std::vector<NNLayer> layers;
  size_t off = 0;
  for (int l = 0; l < nlayers; l++) {
    NNLayer L;
    L.cin = channels[l];
    L.cout = channels[l + 1];
    if ((L.cin != 1 && L.cin != 8) || (L.cout != 1 && L.cout != 8))
      throw std::runtime_error("network layers have 1 or 8 channels");
    const size_t n = (size_t)L.cout * L.cin * NN_PATHS * 16;
    L.W.resize(n);
    for (size_t i = 0; i < n; i++) L.W[i] = std::complex<double>(W[2 * (off + i)], W[2 * (off + i) + 1]);
    off += n;
    layers.push_back(L);
  }
  if (channels[0] != 1 || channels[nlayers] != 1) throw std::runtime_error("network maps 1 channel to 1 channel");
  g_nn.layers = layers;
  if (g_ctx) nn_weights_changed(*g_ctx);
  API_END
}

extern "C" int qqcd_nn_get_weights(int *nlayers, int *channels, double *W) {
  API_BEGIN
  if (g_nn.layers.empty()) throw std::runtime_error("no preconditioner network");
  *nlayers = (int)g_nn.layers.size();
  channels[0] = g_nn.layers[0].cin;
  size_t off = 0;
  for (size_t l = 0; l < g_nn.layers.size(); l++) {
    channels[l + 1] = g_nn.layers[l].cout;
    if (W)
      for (size_t i = 0; i < g_nn.layers[l].W.size(); i++) {
        W[2 * (off + i)] = g_nn.layers[l].W[i].real();
        W[2 * (off + i) + 1] = g_nn.layers[l].W[i].imag();
      }
    off += g_nn.layers[l].W.size();
  }
  API_END
}

// layer l (or its adjoint) on canonical arrays: in[C_in][V][12], out[C_out][V][12] complex
extern "C" int qqcd_nn_layer_canonical(int layer, int adjoint, const double *in, double *out) {
  API_BEGIN
  Context &c = ctx();
  if (!c.have_gauge) throw std::runtime_error("no gauge field loaded");
  NNModel &M = nn_model(c);
  if (layer < 0 || layer >= (int)M.layers.size()) throw std::runtime_error("no such layer");
  const int ci = adjoint ? M.layers[layer].cout : M.layers[layer].cin;
  const int co = adjoint ? M.layers[layer].cin : M.layers[layer].cout;
  const size_t n = (size_t)24 * c.g.V;
  NNField fi = alloc_nn(c.g, ci), fo = alloc_nn(c.g, co);
  double *d = nullptr;
  try {
    QQCD_HIP(hipMalloc(&d, n * sizeof(double)));
    for (int ch = 0; ch < ci; ch++) {
      QQCD_HIP(hipMemcpyAsync(d, in + ch * n, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
      nn_from_canon(c.g, d, fi, ch, c.stream);
    }
    nn_layer(c.g, M.links_for(c), M.dev[layer], adjoint, fi, fo, c.stream);
    for (int ch = 0; ch < co; ch++) {
      nn_to_canon(c.g, fo, ch, d, c.stream);
      QQCD_HIP(hipMemcpyAsync(out + ch * n, d, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    }
    QQCD_HIP(hipStreamSynchronize(c.stream));
  } catch (...) {
    free_nn(fi);
    free_nn(fo);
    if (d) (void)hipFree(d);
    throw;
  }
  free_nn(fi);
  free_nn(fo);
  (void)hipFree(d);
  API_END
}

// dW (complex, layer weight order) = sum_x,c g (T_p psi)^dag for layer l
extern "C" int qqcd_nn_wgrad_canonical(int layer, const double *psi, const double *g, double *dW) {
  API_BEGIN
  Context &c = ctx();
  if (!c.have_gauge) throw std::runtime_error("no gauge field loaded");
  NNModel &M = nn_model(c);
  if (layer < 0 || layer >= (int)M.layers.size()) throw std::runtime_error("no such layer");
  const int ci = M.layers[layer].cin, co = M.layers[layer].cout;
  const size_t n = (size_t)24 * c.g.V;
  NNField fi = alloc_nn(c.g, ci), fg = alloc_nn(c.g, co);
  double *d = nullptr;
  try {
    QQCD_HIP(hipMalloc(&d, n * sizeof(double)));
    for (int ch = 0; ch < ci; ch++) {
      QQCD_HIP(hipMemcpyAsync(d, psi + ch * n, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
      nn_from_canon(c.g, d, fi, ch, c.stream);
    }
    for (int ch = 0; ch < co; ch++) {
      QQCD_HIP(hipMemcpyAsync(d, g + ch * n, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
      nn_from_canon(c.g, d, fg, ch, c.stream);
    }
    nn_wgrad(c.g, M.links_for(c), ci, co, fi, fg, (std::complex<double> *)dW, c.stream);
  } catch (...) {
    free_nn(fi);
    free_nn(fg);
    if (d) (void)hipFree(d);
    throw;
  }
  free_nn(fi);
  free_nn(fg);
  (void)hipFree(d);
  API_END
}

// Train the network on nconf seeded heatbath configurations of the current
// lattice (beta, nsweep sweeps from a hot start, kappa uniform in
// [kappa_lo, kappa_hi], Wilson), pairs from plain FP64 GMRES-DR(m, k) solves
// of nsrc Gaussian sources each (to harvest_tol), then `steps` Adam steps
// (learning rate lr) on minibatches of `batch` pairs.  The network is
// (re)initialised near kappa^2 * 1 with the given channels unless keep = 1.
// Pairs come from every Arnoldi step (max_pairs in all, spread evenly over the
// configurations).  loss[steps]: mean minibatch loss per step; stats = {pairs,
// harvest seconds, training seconds, mean loss over a fixed sample of at most
// 512 pairs before and after (for the kept weights), validation steps of the
// starting and of the kept weights}.
extern "C" int qqcd_nn_train(int nconf, double beta, int nsweep, double kappa_lo, double kappa_hi, int nsrc, int m,
                             int k, double harvest_tol, int max_pairs, int nlayers, const int *channels, int keep,
                             int steps, int batch, double lr, unsigned long long seed, double *loss, double *stats) {
  API_BEGIN
  Context &c = ctx();
  if (nconf < 1 || nsrc < 1 || steps < 0 || batch < 1) throw std::runtime_error("nn_train: bad sizes");
  if (!keep || g_nn.layers.empty()) {
    std::vector<int> ch(channels, channels + nlayers + 1);
    const double kap = 0.5 * (kappa_lo + kappa_hi);
    nn_init_identity(ch, kap * kap, 0.02, seed);
  }
  const auto t0 = std::chrono::steady_clock::now();
  TrainSet T;
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> uni(0.0, 1.0);
  double coact[64] = {0};
  for (int mu = 0; mu < 4; mu++) {
    coact[2 + 4 * (0 + 4 * mu)] = 1.0;
    coact[2 + 4 * (1 + 4 * mu)] = 1.0;
  }
  const int per_conf = (max_pairs + nconf - 1) / nconf;
  double last_kappa = kappa_hi;
  for (int conf = 0; conf < nconf; conf++) {
    if (qqcd_gauge_heatbath(beta, nsweep, seed * 7919ull + conf, 1, coact) != 0)
      throw std::runtime_error(std::string("nn_train: heatbath: ") + qqcd_last_error());
    const double kap[2] = {kappa_lo + (kappa_hi - kappa_lo) * uni(rng), 0.0};
    last_kappa = kap[0];
    const int before = (int)T.pairs.size();
    harvest(c, kap, nsrc, m, k, harvest_tol, seed + 31ull * conf, conf, before + per_conf, T);
    // keep this configuration's links
    T.U.push_back(nn_links(c.g, c.gauge(PREC_SINGLE, 18), c.stream));
  }
  QQCD_HIP(hipStreamSynchronize(c.stream));
  const auto t1 = std::chrono::steady_clock::now();
  if (T.pairs.empty()) throw std::runtime_error("nn_train: no pairs harvested");
  NNModel &M = nn_model(c);
  M.layers = g_nn.layers;
  M.upload(c.stream);
  M.alloc_grads(c.g);
  double *d_scr = nullptr;
  QQCD_HIP(hipMalloc(&d_scr, sizeof(double) * ((2 * c.g.Vh + 255) / 256 + 16)));
  // the mean loss over a fixed sample of at most 512 pairs, evenly spaced
  // over the harvest (every configuration and every stage of the solves)
  const size_t nsample = std::min<size_t>(512, T.pairs.size());
  auto mean_loss = [&]() {
    double s = 0;
    for (size_t q = 0; q < nsample; q++) {
      const Pair &p = T.pairs[q * T.pairs.size() / nsample];
      s += pair_loss(c, M, T, p, d_scr);
    }
    return s / nsample;
  };
  // Validation by the solver itself: Krylov steps of flexible GMRES-DR(40,16)
  // with the network as right preconditioner to 1e-6 on a point source (the
  // origin, spin and colour 0) of the last training configuration, FP32
  // operator; the weights that need the fewest steps are kept (the starting
  // weights, near kappa^2 * 1, included), checked every val_every Adam steps.
  // A lower loss alone does not make a better preconditioner.
  const int val_every = 150;
  auto validate = [&]() -> int {
    const double kv[2] = {last_kappa, 0.0};
    SchurOp A32(c, PREC_SINGLE, 12, kv);
    LinOp op = A32.op();
    LinOp pre = nn_precond(c);
    SpinorBuf b(c.g, PREC_SINGLE), x(c.g, PREC_SINGLE);
    Vec bv = as_vec(b.s), xv = as_vec(x.s);
    c.blas->zero(bv);
    c.blas->zero(xv);
    const float one[2] = {1.f, 0.f};  // component 0 of checkerboard site 0, the origin
    QQCD_HIP(hipMemcpyAsync(b.s.data, one, sizeof(one), hipMemcpyHostToDevice, c.stream));
    SolveStats st = gmresdr(c, op, b.s, x.s, 40, 16, 1e-6, 0, 2000, nullptr, 0, &pre);
    return st.converged ? st.iters : 1000000 + st.iters;
  };
  try {
    const double before = mean_loss();
    int val_best = validate();
    const int val_init = val_best;
    std::vector<NNLayer> best = M.layers;
    // Adam on the real and imaginary parts of every weight
    const int L = (int)M.layers.size();
    std::vector<std::vector<double>> m1(L), m2(L);
    for (int l = 0; l < L; l++) {
      m1[l].assign(2 * M.layers[l].W.size(), 0.0);
      m2[l].assign(2 * M.layers[l].W.size(), 0.0);
    }
    const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
    std::uniform_int_distribution<int> pick(0, (int)T.pairs.size() - 1);
    std::vector<std::vector<std::complex<double>>> dW(L);
    for (int step = 0; step < steps; step++) {
      for (int l = 0; l < L; l++) dW[l].assign(M.layers[l].W.size(), std::complex<double>(0.0));
      double ls = 0;
      for (int bi = 0; bi < batch; bi++) {
        const Pair &p = T.pairs[pick(rng)];
        ls += pair_grad(c, M, T, p, d_scr, dW);
      }
      loss[step] = ls / batch;
      const double bc1 = 1.0 - std::pow(b1, step + 1), bc2 = 1.0 - std::pow(b2, step + 1);
      for (int l = 0; l < L; l++) {
        auto &W = M.layers[l].W;
        for (size_t i = 0; i < W.size(); i++) {
          // dL/dRe W = 2 Re dW, dL/dIm W = 2 Im dW (Wirtinger), averaged over the batch
          const double gr[2] = {2.0 * dW[l][i].real() / batch, 2.0 * dW[l][i].imag() / batch};
          double upd[2];
          for (int q = 0; q < 2; q++) {
            double &a = m1[l][2 * i + q], &v = m2[l][2 * i + q];
            a = b1 * a + (1 - b1) * gr[q];
            v = b2 * v + (1 - b2) * gr[q] * gr[q];
            upd[q] = lr * (a / bc1) / (std::sqrt(v / bc2) + eps);
          }
          W[i] -= std::complex<double>(upd[0], upd[1]);
        }
      }
      M.upload(c.stream);
      if ((step + 1) % val_every == 0 || step + 1 == steps) {
        const int v = validate();
        if (v < val_best) {
          val_best = v;
          best = M.layers;
        }
      }
    }
    M.layers = best;
    M.upload(c.stream);
    const double after = mean_loss();
    g_nn.layers = M.layers;
    const auto t2 = std::chrono::steady_clock::now();
    stats[0] = (double)T.pairs.size();
    stats[1] = std::chrono::duration<double>(t1 - t0).count();
    stats[2] = std::chrono::duration<double>(t2 - t1).count();
    stats[3] = before;
    stats[4] = after;
    stats[5] = val_init;
    stats[6] = val_best;
  } catch (...) {
    (void)hipFree(d_scr);
    throw;
  }
  (void)hipFree(d_scr);
  API_END
}

// z = N(r) for a canonical array r (its even sites); z receives the even sites
extern "C" int qqcd_nn_apply_canonical(const double *r, double *z) {
  API_BEGIN
  Context &c = ctx();
  if (!c.have_gauge) throw std::runtime_error("no gauge field loaded");
  NNModel &M = nn_model(c);
  const size_t n = (size_t)24 * c.g.V;
  double *d = nullptr;
  QQCD_HIP(hipMalloc(&d, n * sizeof(double)));
  try {
    SpinorBuf a(c.g, PREC_SINGLE), b(c.g, PREC_SINGLE);
    QQCD_HIP(hipMemcpyAsync(d, r, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    canon_to_spinor(c.g, d, a.s, 0, c.stream);
    nn_apply(c, M, a.s, b.s);
    QQCD_HIP(hipMemcpyAsync(d, z, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    spinor_to_canon(c.g, b.s, d, 0, c.stream);
    QQCD_HIP(hipMemcpyAsync(z, d, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    QQCD_HIP(hipStreamSynchronize(c.stream));
  } catch (...) {
    (void)hipFree(d);
    throw;
  }
  (void)hipFree(d);
  API_END
}

// Time a solve of M x = point source on the resident field: method 0 plain
// FP64 GMRES-DR, 1 mixed precision, 2 mixed precision with the network as the
// flexible right preconditioner of the inner solves.

extern "C" int qqcd_bench_learned(int method, const int src[4], int spin, int colour, const double kappa[2],
                                  const int GMRES[2], double tol, int op_prec, int op_recon, double *seconds,
                                  int *iters, int *outer, double *true_res, int *nn_steps, int *dropped) {
  API_BEGIN
  Context &c = ctx();
  if (!c.have_gauge) throw std::runtime_error("no gauge field loaded");
  SpinorBuf b_e(c.g, PREC_DOUBLE), b_o(c.g, PREC_DOUBLE), src_e(c.g, PREC_DOUBLE);
  SpinorBuf x_e(c.g, PREC_DOUBLE), x_o(c.g, PREC_DOUBLE);
  point_source_dev(c, b_e.s, b_o.s, src, spin, colour);
  NNModel *M = method == 2 ? &nn_model(c) : nullptr;
  LinOp pre;
  pre.prec = PREC_SINGLE;
  if (M) pre.apply = [&](const Spinor &a, Spinor &z) { nn_apply(c, *M, a, z); };
  QQCD_HIP(hipDeviceSynchronize());
  const auto t0 = std::chrono::steady_clock::now();
  prep_source(c, kappa, b_e.s, b_o.s, src_e.s);
  *nn_steps = 0;
  *dropped = 0;
  if (method == 0) {
    SchurOp A(c, PREC_DOUBLE, 18, kappa);
    LinOp op = A.op();
    SolveStats st = gmresdr(c, op, src_e.s, x_e.s, GMRES[0], GMRES[1], tol, 0, 5000 * GMRES[0], nullptr);
    *iters = st.iters;
    *outer = st.cycles;
  } else if (method == 1 || method == 2) {
    MixedParams p = mixed_params(GMRES, tol, 1, op_prec, op_recon, 0.0);
    if (M) p.precond = &pre;
    Deflation D;
    MixedStats ms = solve_mixed(c, kappa, src_e.s, x_e.s, p, &D);
    *iters = ms.inner.iters;
    *outer = ms.outer;
    *nn_steps = ms.precond_steps;
    *dropped = ms.precond_dropped;
  } else {
    throw std::runtime_error("bench_learned: method 0, 1 or 2");
  }
  reconstruct_odd(c, kappa, b_o.s, x_e.s, x_o.s);
  QQCD_HIP(hipDeviceSynchronize());
  const auto t1 = std::chrono::steady_clock::now();
  *seconds = std::chrono::duration<double>(t1 - t0).count();
  *true_res = full_residual(c, kappa, b_e.s, b_o.s, x_e.s, x_o.s);
  API_END
}

// Average time of one network application z = N(r) (device-resident, seeded
// Gaussian r on the even sites) over niter launches after nwarm.
extern "C" int qqcd_bench_nn(int nwarm, int niter, double *ms) {
  API_BEGIN
  Context &c = ctx();
  if (!c.have_gauge) throw std::runtime_error("no gauge field loaded");
  NNModel &M = nn_model(c);
  SpinorBuf r(c.g, PREC_SINGLE), z(c.g, PREC_SINGLE);
  random_spinor(c.g, r.s, 99, c.stream);
  for (int i = 0; i < nwarm; i++) nn_apply(c, M, r.s, z.s);
  hipEvent_t e0, e1;
  QQCD_HIP(hipEventCreate(&e0));
  QQCD_HIP(hipEventCreate(&e1));
  QQCD_HIP(hipEventRecord(e0, c.stream));
  for (int i = 0; i < niter; i++) nn_apply(c, M, r.s, z.s);
  QQCD_HIP(hipEventRecord(e1, c.stream));
  QQCD_HIP(hipEventSynchronize(e1));
  float f = 0;
  QQCD_HIP(hipEventElapsedTime(&f, e0, e1));
  *ms = f / std::max(niter, 1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  API_END
}
