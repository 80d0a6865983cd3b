// qqcd_solvers.h -- Krylov solvers on device-resident spinors.
#pragma once
#include <functional>
#include <memory>
#include <utility>
#include <vector>

#include "qqcd_context.h"
#include "qqcd_dense.h"

namespace qqcd {

// A linear operator on checkerboard-0 spinors of precision prec.  store:
// the precision in which the Krylov solvers keep their basis (-1: prec);
// PREC_HALF with prec PREC_SINGLE keeps the basis and deflation vectors as
// FP16 fields (half the bytes of every orthogonalisation pass) while every
// new vector is formed and orthogonalised in FP32.  apply then also takes
// FP16 input.
struct LinOp {
  Prec prec = PREC_DOUBLE;
  int store = -1;
  std::function<void(const Spinor &, Spinor &)> apply;
  // optional: r1 = A a, r2 = A r1 in one call (an operator that rounds its
  // input to a lower storage precision can then pass its own rounded r1 on)
  std::function<void(const Spinor &, Spinor &, Spinor &)> apply_sq;
  Prec basis() const { return store < 0 ? prec : (Prec)store; }
};

// The even-odd Schur operator (Hdbletm semantics) with storage (prec, recon).
struct SchurOp {
  Context &c;
  const Gauge &U;
  double kappa[2];
  int idag;
  SpinorBuf tmp;
  SchurOp(Context &c_, Prec p, int recon, const double k[2], int idag_ = 0)
      : c(c_), U(c_.gauge(p, recon)), idag(idag_), tmp(c_.g, p) {
    kappa[0] = k[0];
    kappa[1] = k[1];
  }
  void operator()(const Spinor &in, Spinor &out, Spinor *out32 = nullptr, bool only32 = false) {
    schur_apply(c, U, kappa, idag, in, out, tmp.s, out32, only32);
  }
  LinOp op() {
    LinOp o;
    o.prec = tmp.s.prec;
    o.apply = [this](const Spinor &a, Spinor &b) { (*this)(a, b); };
    return o;
  }
};

struct SolveStats {
  int iters = 0;       // Arnoldi / BiCGStab steps (operator applications in the Krylov iterations)
  int matvecs = 0;     // all operator applications
  int cycles = 0;
  double rel_res = 0;  // final ||b - A x|| / ||r0|| with the true residual
  int converged = 0;
  int nonfinite = 0;   // the true residual became NaN or infinite: the solve stopped there
};

// Deflation space carried between solves: A V_k = V_{k+1} Hbar_k.
struct Deflation {
  std::vector<Spinor> V;  // k+1 vectors
  std::vector<dense::zd> H;  // (k+1) x k, column-major
  int k = 0;
  Prec prec = PREC_DOUBLE;
  void clear();
  ~Deflation() { clear(); }
};

// GMRES-DR(m, k): x holds x0 on entry.  Stops when ||b - A x|| < tol ||b - A x0||.
// When defl is given, the deflation space of the last restart (or of the
// final cycle) is stored there.
// Called after every cycle with the current iterate, its true residual and
// the statistics so far.
typedef std::function<void(const Spinor &x, const Spinor &r, const SolveStats &st)> CycleHook;

// reorth = 1: a second classical Gram-Schmidt pass whenever the first
// removed more than half of ||w||^2 (the DGKS criterion); 0: one pass.
// M: right preconditioner, applied flexibly (FGMRES-DR: z_j = M v_j is kept,
// A Z_m = V_{m+1} Hbar_m, x += Z y, and the deflated restart combines Z with
// the same harmonic Ritz coefficients as V); no deflation space is kept then.
// Called after every Arnoldi step of the synchronising path (reorth != 0,
// no preconditioner) with the iterate x of the cycle start, the basis and
// the step's least-squares solution y[0..j): the step's iterate is
// x + V[0..j) y (the reference forms it at every step, inverters.f90:3242-3255).
typedef std::function<void(const Spinor &x, const std::vector<Spinor> &V, int j, const dense::zd *y)> StepHook;

SolveStats gmresdr(Context &c, LinOp &A, const Spinor &b, Spinor &x, int m, int k, double tol, int itermin,
                   int maxiter, Deflation *defl, int reorth = 1, LinOp *M = nullptr,
                   const CycleHook *hook = nullptr, const StepHook *step_hook = nullptr);

// Residual history of a solve: (number of operator applications so far, the
// least-squares residual norm ||c - Hbar y|| after that Arnoldi step).
typedef std::vector<std::pair<int, double>> ResHistory;

struct DevHess;
struct StepGraphs;

// Buffers of GMRES-Proj kept from one call to the next (the basis, the
// device Hessenberg, the operator's work vector) together with the HIP
// graphs of its groups of Arnoldi steps, which then replay in every later
// call.  Owned by a caller that solves one system after another with the
// same operator; reset() whenever that operator is replaced.
struct ProjWork {
  std::vector<Spinor> V;
  std::unique_ptr<SpinorBuf> work, work2;
  std::unique_ptr<DevHess> dh;
  std::unique_ptr<StepGraphs> sg;
  int m = 0;
  Prec p = PREC_DOUBLE, vp = PREC_DOUBLE;
  ProjWork() = default;
  ProjWork(const ProjWork &) = delete;
  ProjWork &operator=(const ProjWork &) = delete;
  ~ProjWork();
  void reset();
};

// GMRES-Proj(m, k) with deflation space defl: alternates a projection over
// (V_{k+1}, Hbar_k) with a GMRES(m) cycle.  x holds x0 on entry.  hist, if
// given, receives one entry per Arnoldi step.  ws, if given, holds the
// buffers between calls.
SolveStats gmresproj(Context &c, LinOp &A, const Spinor &b, Spinor &x, int m, const Deflation &defl, double tol,
                     int itermin, int maxiter, int reorth = 1, ResHistory *hist = nullptr, ProjWork *ws = nullptr);

// BiCGStab from x (Frommer et al.; the reference's bicgstab).  Convergence
// on the recursive residual, then confirmed on the true residual (restarting
// from it if the two disagree).
// Called after every BiCGStab iteration with the iteration count, the
// recursive residual norm and the initial residual norm.
typedef std::function<void(int iters, double rnorm, double r0)> IterHook;
SolveStats bicgstab(Context &c, LinOp &A, const Spinor &b, Spinor &x, double tol, int itermin, int maxiter,
                    const IterHook *hook = nullptr);

// out = a0 in + a1 (i g5) in, FP64, on one checkerboard
void local_spin(const Geom &g, const Spinor &in, Spinor &out, double a0, double a1, hipStream_t s);

}  // namespace qqcd

namespace qqcd {

// Mixed-precision solve of the even-odd system (defect correction with
// reliable updates): the solution, the true residual and every outer update
// are FP64; each correction is solved with an inner GMRES-DR(m,k) (first
// time) or GMRES-Proj(m-k,k) over the kept deflation space (afterwards) on
// vectors of precision `basis`, with the operator stored in (op_prec,
// op_recon).  Stops when ||b - A x|| < tol ||b||.  An outer step that does
// not halve the true residual raises the operator precision (FP16 -> FP32
// -> plain FP64 GMRES-DR), so the FP64 solver is the fallback and the FP64
// true residual the judge of every result.
// replay repeated groups of synchronisation-free Arnoldi steps from HIP
// graphs: 0 never, 1 in every solve, 2 (the default) only where the basis is
// kept between calls (ProjWork), so that a capture serves many solves
void set_step_graphs(int mode);
// Arnoldi steps per Gram-Schmidt pass over an FP16 basis (single-RHS
// solvers): 2 (the default) in pairs, 1 one at a time
void set_arnoldi_steps(int steps);
int arnoldi_steps();

struct MixedParams {
  int m = 40, k = 16;
  Prec basis = PREC_SINGLE;
  Prec op_prec = PREC_SINGLE;
  int op_recon = 12;
  double tol = 1e-10;
  double inner_tol = 1e-6;
  int max_outer = 60;
  int maxiter = 200000;
  int first_cycles = 30;  // GMRES-DR cycles allowed in the first inner solve
  int later_cycles = 3;   // GMRES-Proj cycles per later inner solve
  double stall = 0.5;     // an outer step that shrinks ||r|| by less stalls
  int reorth = 0;         // Gram-Schmidt of the inner (low-precision) solves: 0 one pass, 1 DGKS
  // right preconditioner of the inner solves (flexible GMRES-DR, restarted
  // fresh in every outer step); dropped for good at the first stalled outer step
  LinOp *precond = nullptr;
};

struct MixedStats {
  SolveStats inner;  // summed over outer iterations
  int outer = 0;
  int fp64_matvecs = 0;
  double rel_res = 0;  // ||b - A x|| / ||b|| in FP64
  int converged = 0;
  int stalled = 0;
  int fallback = 0;  // 1: operator raised to FP32, 2: finished in plain FP64
  int precond_dropped = 0;  // the preconditioner stalled and was switched off
  int precond_steps = 0;    // Krylov steps taken with the preconditioner
  int rejected = 0;         // inner corrections with a NaN or infinite entry, not added to x
};

// Failures forced from the host, for tests of the fallbacks: every kernel
// still runs to completion, only values are replaced.  Process-wide; reset
// with fault_plan() = FaultPlan().
struct FaultPlan {
  int precond = 0;        // output of the learned preconditioner: 1 NaN, 2 zero, 3 random
  int corrupt_outer = 0;  // mixed solver: the inner corrections of the first n outer steps set to NaN
  int batch_rhs = -1;     // batched solver: this right-hand side's correction set to NaN in the
                          // first corrupt_outer outer steps
  int maxiter = 0;        // > 0: cap on the Krylov steps of a mixed solve
};
FaultPlan &fault_plan();
// statistics of the most recent solve_mixed call
MixedStats &last_mixed_stats();

MixedStats solve_mixed(Context &c, const double kappa[2], const Spinor &b, Spinor &x, const MixedParams &p,
                       Deflation *defl);

// MixedParams from the C-ABI arguments (GMRES = (m, k); basis 0 FP64 / 1 FP32;
// op_prec 0/1/2; inner_tol <= 0 picks 1e-2 (FP16), 1e-6 (FP32), 1e-10 (FP64))
MixedParams mixed_params(const int GMRES[2], double tol, int basis, int op_prec, int op_recon, double inner_tol);

// the library's trained network as a flexible right preconditioner (FP32
// spinors, even checkerboard) on the context's links (qqcd_learned.cpp)
LinOp nn_precond(Context &c);

// unit source at 1-based global site src[4], Dirac d, colour col (both checkerboards zeroed first by the caller)
void point_source_dev(Context &c, Spinor &b_e, Spinor &b_o, const int src[4], int d, int col);

// even-odd source preparation / reconstruction / full-system residual of
// fermprop (qqcd-test/cfgsprops/quark/quark.f90:600-617, 797-819), FP64
void prep_source(Context &c, const double kappa[2], const Spinor &b_e, const Spinor &b_o, Spinor &src);
void reconstruct_odd(Context &c, const double kappa[2], const Spinor &b_o, const Spinor &x_e, Spinor &x_o);
double full_residual(Context &c, const double kappa[2], const Spinor &b_e, const Spinor &b_o, const Spinor &x_e,
                     const Spinor &x_o);

}  // namespace qqcd
