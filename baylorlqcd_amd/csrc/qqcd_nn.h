// qqcd_nn.h -- the gauge-equivariant preconditioner network on the device.
//
// A layer maps C_in channels to C_out channels of full-lattice spin-colour
// fields:
//   out_a(x) = sum_{b, p} W[a][b][p] (T_p psi_b)(x),
// p over the paths {0, +x, -x, +y, -y, +z, -z, +t, -t} (index 0, 1+2mu,
// 2+2mu), T_0 = 1, (T_+mu psi)(x) = U_mu(x) psi(x+mu), (T_-mu psi)(x) =
// U_mu(x-mu)^dag psi(x-mu), hops across the lattice edge scaled by bc[mu],
// and W[a][b][p] a complex 4x4 spin matrix (identity on colour, so the layer
// commutes with gauge transformations).
//
// Device layout of a network field (BF16 complex, 4 bytes per component),
// site-major: component k = 3 spin + colour of channel ch at checkerboard
// site (par, cb) is element ((par * stride + cb) * C + ch) * 12 + k.
#pragma once
#include <complex>
#include <vector>

#include "qqcd_context.h"

namespace qqcd {

constexpr int NN_PATHS = 9;

struct NNField {
  int C = 0;
  long stride = 0;
  uint32_t *data = nullptr;  // BF16 (re, im) pairs
  size_t bytes() const { return (size_t)2 * stride * C * 12 * 4; }
};
NNField alloc_nn(const Geom &g, int C);
void free_nn(NNField &f);

// Weights of one layer, W[((a * cin + b) * 9 + p) * 16 + 4 s + s'] complex.
struct NNLayer {
  int cin = 0, cout = 0;
  std::vector<std::complex<double>> W;
};

// Device operand of a layer: the real-ified weight matrix in the order the
// kernel consumes (BF16), for the forward layer and for its adjoint.
struct NNLayerDev {
  int cin = 0, cout = 0;
  uint16_t *B = nullptr;     // forward (transport first)
  uint16_t *Badj = nullptr;  // adjoint (reversed paths, W^dag)
  uint16_t *Bmix = nullptr;  // the 8 -> 1 direction as mix-then-gather (forward of 8->1, adjoint of 1->8)
  size_t nB = 0, nBadj = 0, nBmix = 0;
};
void nn_upload(const NNLayer &L, NNLayerDev &D, hipStream_t s);
void nn_free(NNLayerDev &D);

// Links as BF16 complex pairs in the 18-real component-run layout.
struct NNLinks {
  uint32_t *data = nullptr;
  long stride = 0;
};
NNLinks nn_links(const Geom &g, const Gauge &U32, hipStream_t s);
void nn_free_links(NNLinks &L);

// out = layer(in) (adjoint = 0) or layer^dag(in) (adjoint = 1).  sparse:
// NN_SRC_EVEN promises that `in` is zero on the odd sites (their transports
// are skipped); NN_OUT_EVEN asks for the even sites of `out` only (the odd
// ones are left undefined; honoured by the 8 -> 1 layer).
constexpr int NN_SRC_EVEN = 1, NN_OUT_EVEN = 2;
void nn_layer(const Geom &g, const NNLinks &U, const NNLayerDev &D, int adjoint, const NNField &in, NNField &out,
              hipStream_t s, int sparse = 0);
// Weight gradient dL/dconj(W) of one layer, summed over the lattice:
// dW[a][b][p][s][s'] = sum_x,c g_a(x)[s,c] conj((T_p psi_b)(x)[s',c]); written
// as complex doubles in the weight order into host array dW (cout*cin*9*16).
void nn_wgrad(const Geom &g, const NNLinks &U, int cin, int cout, const NNField &psi, const NNField &gout,
              std::complex<double> *dW, hipStream_t s);

// even-checkerboard FP32/FP64 spinor <-> channel 0 of a network field (the
// other checkerboard set to zero on the way in); scale multiplies.
void nn_from_spinor(const Geom &g, const Spinor &sp, NNField &f, float scale, hipStream_t s);
void nn_to_spinor(const Geom &g, const NNField &f, Spinor &sp, float scale, hipStream_t s);
// canonical complex array psi[((site*4 + spin)*3 + colour)] (both parities) <-> channel ch
void nn_from_canon(const Geom &g, const double *dcanon, NNField &f, int ch, hipStream_t s);
void nn_to_canon(const Geom &g, const NNField &f, int ch, double *dcanon, hipStream_t s);
// g = (out ch 0 on the even sites) * so - target * st, zero on the odd sites; returns
// ||g||^2 (FP64) -- the loss term and output gradient of one training pair
double nn_diff(const Geom &g, const NNField &out, const Spinor &target, float so, float st, NNField &gout,
               double *d_scratch, hipStream_t s);

// The network: a stack of layers with channel counts ch[0] = 1, ..., ch[L] = 1.
struct NNModel {
  std::vector<NNLayer> layers;
  std::vector<NNLayerDev> dev;
  std::vector<NNField> acts;  // activations of the last forward pass (L + 1 fields)
  std::vector<NNField> grads; // gradients with respect to the activations (training only)
  bool ready = false;
  NNLinks links;  // BF16 links of the context's gauge field
  uint64_t links_gen = ~0ull;
  const NNLinks &links_for(Context &c);
  void init(const std::vector<int> &channels, uint64_t seed);
  void upload(hipStream_t s);
  void alloc(const Geom &g);
  void alloc_grads(const Geom &g);
  void release();
  ~NNModel() { release(); }
};

// z_e = N(r_e) for even-checkerboard spinors (FP32 or FP64)
void nn_apply(Context &c, NNModel &M, const Spinor &r, Spinor &z);

}  // namespace qqcd
