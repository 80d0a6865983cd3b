// qqcd_dense.h -- small dense complex linear algebra on the host (FP64).
//
// The (m+1) x m Hessenberg problems of the Krylov solvers (m <= ~128) are
// far too small for the GPU.  These routines replace the reference's own
// LAPACK adaptations in qqcd-test/cfgsprops/quark/pseudolapack.f90
// (linearsolver :45 ~ ZGESV, hessenberg :700 ~ ZGEHRD, qgenerator :774 ~
// ZUNGHR, evalues/QRdouble :868/:2196 ~ ZHSEQR, orgschur/schurswap
// :252/:389 ~ ZTRSEN) with textbook algorithms (Golub & Van Loan, Matrix
// Computations, 4th ed., 5.1, 7.4-7.6): Householder reduction to Hessenberg
// form, implicit single-shift complex QR with Wilkinson shifts, Givens
// swaps of adjacent Schur diagonal entries, Householder least squares,
// LU with partial pivoting.  All matrices are column-major.
#pragma once
#include <complex>
#include <vector>

namespace qqcd {
namespace dense {

typedef std::complex<double> zd;

// Complex Schur form A = Q T Q^H of the n x n matrix a (lda); returns T in t
// (n x n) and Q in q (n x n).  Returns false if QR did not converge.
bool schur(int n, const zd *a, int lda, zd *t, zd *q);

// Reorder the Schur form so that the k eigenvalues of smallest modulus lead
// (stable selection: ties keep their order).  Updates t and q in place.
void reorder_smallest(int n, zd *t, zd *q, int k);

// Least squares min ||c - H y||, H rows x cols (ldh), rows >= cols.  Returns
// the residual norm; y gets cols entries.
double lsq(int rows, int cols, const zd *h, int ldh, const zd *c, zd *y);

// Solve A x = b (n x n, one right-hand side) by LU with partial pivoting.
bool lu_solve(int n, const zd *a, int lda, const zd *b, zd *x);

// Harmonic Ritz step of GMRES-DR (Morgan 2002): from the (m+1) x m matrix
// hbar (ldh) form F = H_m + |h_{m+1,m}|^2 H_m^{-H} e_m e_m^T, take its Schur
// form with the k harmonic Ritz values of smallest modulus first and return
// the k leading Schur vectors in p (m x k, ld m) and all m values in theta.
bool harmonic_ritz(int m, const zd *hbar, int ldh, int k, zd *p, zd *theta);

}  // namespace dense
}  // namespace qqcd
