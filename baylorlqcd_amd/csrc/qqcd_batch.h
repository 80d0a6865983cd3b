// qqcd_batch.h -- batched solves of many right-hand sides on one gauge field.
#pragma once
#include <vector>

#include "qqcd_multi.h"
#include "qqcd_solvers.h"

namespace qqcd {

// The even-odd Schur operator (schur_apply) on the active slots of multi-RHS fields.
struct MultiSchur {
  Context &c;
  const Gauge &U;
  double kappa[2];
  int idag;
  MultiBuf tmp;
  MultiTune tune;
  // the odd-checkerboard intermediate in FP16 with per-(site, slot) norms
  // (set_half_mid, where the two-slot FP32 kernel runs): the first hop writes
  // it, the second reads it, half the bytes of FP32
  std::unique_ptr<MultiBuf> tmp16;
  MultiSchur(Context &c_, Prec p, int recon, const double k[2], int cap, int nb, const MultiTune &t,
             int idag_ = 0);
  void set_half_mid(bool on);
  MultiSpinor &mid() { return tmp16 ? tmp16->m : tmp.m; }
  void operator()(const MultiSpinor &in, MultiSpinor &out, const int *nact);
  // out = A in / ||in|| per slot (n2: ||in||^2 per slot in device memory, 0
  // where it is not positive) and side = hs in / ||in|| as an FP16 field
  // without norms, both in the second hop's epilogue: the normalisation of
  // an Arnoldi step folded into the next operator application.  Needs
  // can_scale().
  void scaled(const MultiSpinor &in, MultiSpinor &out, const int *nact, const double *n2, MultiSpinor *side,
              double hs);
  bool can_scale() const;
};

// Batched GMRES-Proj with reliable updates.  Every right-hand side b_r is
// solved to ||b_r - A x_r|| < tol ||b_r|| in FP64: an outer loop recomputes
// the true residuals with the FP64 operator and adds corrections from an
// inner batched GMRES-Proj(m) over the FP32 deflation space of an earlier
// GMRES-DR solve (projection over (V_{k+1}, Hbar_k), then a GMRES(m) cycle;
// the reference's gmresproject, qqcd-test/cfgsprops/quark/inverters.f90:
// 7183-8304, with psc's split m-k / k, :4439-4458).  All right-hand sides
// advance in lockstep: one multi-RHS operator application serves every
// active slot, the Gram-Schmidt coefficients stay on the device, and a slot
// whose residual meets its target is retired (swapped behind the active
// slots, so later launches skip it).  Right-hand sides that do not
// converge within max_outer outer steps are finished one by one with the
// single-RHS mixed solver (whose last resort is plain FP64 GMRES-DR).
struct BatchParams {
  int m = 40;             // GMRES cycle length
  double tol = 1e-10;     // per right-hand side, FP64 true residual
  double inner_tol = 1e-5;
  int inner_cycles = 4;   // GMRES-Proj cycles per inner solve
  int max_outer = 30;
  int reorth = 0;         // 0: one Gram-Schmidt pass; 1: second pass per slot when needed; 2: always
  int use_graph = 1;      // replay the Arnoldi cycle from a captured graph
  Prec op_prec = PREC_SINGLE;
  int op_recon = 12;
  // Krylov basis storage of the inner GMRES: FP32, or FP16 scaled unit
  // vectors (the operator then reads an FP32 copy of the newest one)
  Prec basis = PREC_HALF;
  // FP16 basis: fold each Arnoldi step's normalisation into the next
  // operator application where the launch form allows it
  int fuse_norm = 1;
  // Arnoldi steps per Gram-Schmidt pass over the basis: 2 runs the steps in
  // pairs (u1 = A v_j, u2 = A u1, one pass over V for both sets of dots and
  // one for both updates, column j + 1 from the Arnoldi relation) where the
  // FP16 basis, the normalising operator, one Gram-Schmidt pass and an even
  // m allow it; 1 one step at a time
  int sstep = 2;
  // paired steps: 1 takes the norm and inner product of the updated vectors
  // from those of u1, u2 (formed in the pass of dots) less the projections,
  // so the update pass also writes the next basis vector and operator input;
  // 0 computes them from the updated vectors, with a third pass for those two
  int pair_norms = 1;
  // the inner operator's intermediate (odd-checkerboard) field in FP16 with
  // norms where the two-slot kernel runs (1), or in FP32 (0)
  int half_mid = 1;
  MultiTune tune;
};

struct BatchStats {
  int outer = 0;
  int cycles = 0;
  long slot_matvecs = 0;     // operator applications summed over right-hand sides
  long batched_matvecs = 0;  // multi-RHS operator launches (pairs of hopping kernels)
  int graph_replays = 0;
  int fallback = 0;          // right-hand sides finished by the single-RHS solver
  int rejected = 0;          // per-slot inner corrections with a NaN or infinite entry, not added to x
  std::vector<int> iters;    // per right-hand side: Arnoldi steps while active
  std::vector<double> rel_res;
  int converged = 0;
};

// Solves A x_r = b_r for the right-hand sides in slots [0, n) of B (FP64);
// X (FP64, same layout) holds the initial guesses and gets the solutions.
// Slots are permuted while right-hand sides retire; on return slot s holds
// right-hand side id[s] (id is the identity on entry unless given).
BatchStats batched_solve(Context &c, const double kappa[2], MultiSpinor &B, MultiSpinor &X, int n,
                         const Deflation &D, const BatchParams &p, std::vector<int> &id);

// Autotuner of the multi-RHS hopping kernel: times every RHS block nb in
// {8, 16, 32, 64} (<= the power of two above nrhs) and tile order for the
// given storage and reports the fastest.  Results are cached per shape.
struct TuneResult {
  int nb = 16;
  MultiTune t;
  double ms = 0;
};
TuneResult autotune_multi(Context &c, Prec p, int recon, int nrhs);

}  // namespace qqcd
