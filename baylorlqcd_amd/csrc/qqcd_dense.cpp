// qqcd_dense.cpp -- see qqcd_dense.h.
#include "qqcd_dense.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace qqcd {
namespace dense {

static inline double abs1(zd z) { return std::fabs(z.real()) + std::fabs(z.imag()); }

// Givens G = [[c, s], [-conj(s), c]] with c real so that G [x; y] = [r; 0].
static void givens(zd x, zd y, double &c, zd &s) {
  const double ax = std::abs(x), ay = std::abs(y);
  if (ay == 0.0) {
    c = 1.0;
    s = 0.0;
    return;
  }
  if (ax == 0.0) {
    c = 0.0;
    s = std::conj(y) / ay;
    return;
  }
  const double rho = std::hypot(ax, ay);
  c = ax / rho;
  s = (x / ax) * std::conj(y) / rho;
}

// rows i, i+1 of t, columns c0..c1-1:  [ti; tj] <- G [ti; tj]
static void rot_rows(int n, zd *t, int i, int c0, int c1, double c, zd s) {
  for (int j = c0; j < c1; j++) {
    const zd a = t[i + (size_t)n * j], b = t[i + 1 + (size_t)n * j];
    t[i + (size_t)n * j] = c * a + s * b;
    t[i + 1 + (size_t)n * j] = -std::conj(s) * a + c * b;
  }
}

// columns i, i+1 of a (ld n), rows r0..r1-1:  [ci, cj] <- [ci, cj] G^H
static void rot_cols(int n, zd *a, int i, int r0, int r1, double c, zd s) {
  for (int r = r0; r < r1; r++) {
    const zd x = a[r + (size_t)n * i], y = a[r + (size_t)n * (i + 1)];
    a[r + (size_t)n * i] = c * x + std::conj(s) * y;
    a[r + (size_t)n * (i + 1)] = -s * x + c * y;
  }
}

bool schur(int n, const zd *a, int lda, zd *t, zd *q) {
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      t[i + (size_t)n * j] = a[i + (size_t)lda * j];
      q[i + (size_t)n * j] = i == j ? 1.0 : 0.0;
    }
  if (n <= 1) return true;
  // Householder reduction to upper Hessenberg form, accumulating Q
  std::vector<zd> v(n);
  for (int k = 0; k < n - 2; k++) {
    double xn = 0;
    for (int i = k + 1; i < n; i++) xn += std::norm(t[i + (size_t)n * k]);
    xn = std::sqrt(xn);
    if (xn == 0.0) continue;
    const zd x0 = t[k + 1 + (size_t)n * k];
    const zd ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : zd(1.0);
    for (int i = 0; i < n; i++) v[i] = 0;
    for (int i = k + 1; i < n; i++) v[i] = t[i + (size_t)n * k];
    v[k + 1] += ph * xn;
    double vn = 0;
    for (int i = k + 1; i < n; i++) vn += std::norm(v[i]);
    if (vn == 0.0) continue;
    const double beta = 2.0 / vn;
    // t <- (I - beta v v^H) t
    for (int j = 0; j < n; j++) {
      zd s = 0;
      for (int i = k + 1; i < n; i++) s += std::conj(v[i]) * t[i + (size_t)n * j];
      s *= beta;
      for (int i = k + 1; i < n; i++) t[i + (size_t)n * j] -= v[i] * s;
    }
    // t <- t (I - beta v v^H), q <- q (I - beta v v^H)
    for (int i = 0; i < n; i++) {
      zd s = 0, sq = 0;
      for (int j = k + 1; j < n; j++) {
        s += t[i + (size_t)n * j] * v[j];
        sq += q[i + (size_t)n * j] * v[j];
      }
      s *= beta;
      sq *= beta;
      for (int j = k + 1; j < n; j++) {
        t[i + (size_t)n * j] -= s * std::conj(v[j]);
        q[i + (size_t)n * j] -= sq * std::conj(v[j]);
      }
    }
    for (int i = k + 2; i < n; i++) t[i + (size_t)n * k] = 0;
  }
  // implicit single-shift QR on the Hessenberg matrix
  const double eps = std::numeric_limits<double>::epsilon();
  int hi = n - 1, iter = 0, total = 0;
  while (hi > 0) {
    int l = hi;
    for (; l > 0; l--) {
      const double sub = abs1(t[l + (size_t)n * (l - 1)]);
      const double dia = abs1(t[l - 1 + (size_t)n * (l - 1)]) + abs1(t[l + (size_t)n * l]);
      if (sub <= eps * dia || sub < std::numeric_limits<double>::min()) {
        t[l + (size_t)n * (l - 1)] = 0;
        break;
      }
    }
    if (l == hi) {
      hi--;
      iter = 0;
      continue;
    }
    if (++total > 100 * n) return false;
    iter++;
    // Wilkinson shift from the trailing 2x2 of the active block
    zd mu;
    const zd aa = t[hi - 1 + (size_t)n * (hi - 1)], bb = t[hi - 1 + (size_t)n * hi];
    const zd cc = t[hi + (size_t)n * (hi - 1)], dd = t[hi + (size_t)n * hi];
    if (iter % 11 == 10) {
      mu = dd + std::abs(cc) * 0.75;  // exceptional shift
    } else {
      const zd tr = 0.5 * (aa - dd);
      const zd disc = std::sqrt(tr * tr + bb * cc);
      const zd m1 = dd + tr + disc, m2 = dd + tr - disc;
      mu = std::abs(m1 - dd) < std::abs(m2 - dd) ? m1 : m2;
    }
    zd x = t[l + (size_t)n * l] - mu, y = t[l + 1 + (size_t)n * l];
    for (int k = l; k < hi; k++) {
      double c;
      zd s;
      givens(x, y, c, s);
      rot_rows(n, t, k, k > l ? k - 1 : l, n, c, s);
      rot_cols(n, t, k, 0, std::min(k + 3, hi + 1), c, s);
      rot_cols(n, q, k, 0, n, c, s);
      if (k > l) t[k + 1 + (size_t)n * (k - 1)] = 0;
      if (k < hi - 1) {
        x = t[k + 1 + (size_t)n * k];
        y = t[k + 2 + (size_t)n * k];
      }
    }
  }
  for (int j = 0; j < n; j++)
    for (int i = j + 1; i < n; i++) t[i + (size_t)n * j] = 0;
  return true;
}

// swap the adjacent diagonal entries k, k+1 of the Schur form
static void swap_adjacent(int n, zd *t, zd *q, int k) {
  const zd a = t[k + (size_t)n * k], d = t[k + 1 + (size_t)n * (k + 1)], b = t[k + (size_t)n * (k + 1)];
  if (a == d) return;
  double c;
  zd s;
  // eigenvector of the 2x2 block for d is (b, d - a): rotate it onto e_1
  givens(b, d - a, c, s);
  rot_rows(n, t, k, k, n, c, s);
  rot_cols(n, t, k, 0, k + 2, c, s);
  rot_cols(n, q, k, 0, n, c, s);
  t[k + 1 + (size_t)n * k] = 0;
  t[k + (size_t)n * k] = d;
  t[k + 1 + (size_t)n * (k + 1)] = a;
}

void reorder_smallest(int n, zd *t, zd *q, int k) {
  for (int p = 0; p < k && p < n; p++) {
    int best = p;
    for (int j = p + 1; j < n; j++)
      if (std::abs(t[j + (size_t)n * j]) < std::abs(t[best + (size_t)n * best])) best = j;
    for (int j = best - 1; j >= p; j--) swap_adjacent(n, t, q, j);
  }
}

double lsq(int rows, int cols, const zd *h, int ldh, const zd *c, zd *y) {
  std::vector<zd> a((size_t)rows * cols), b(c, c + rows), v(rows);
  for (int j = 0; j < cols; j++)
    for (int i = 0; i < rows; i++) a[i + (size_t)rows * j] = h[i + (size_t)ldh * j];
  for (int k = 0; k < cols; k++) {
    double xn = 0;
    for (int i = k; i < rows; i++) xn += std::norm(a[i + (size_t)rows * k]);
    xn = std::sqrt(xn);
    if (xn == 0.0) continue;
    const zd x0 = a[k + (size_t)rows * k];
    const zd ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : zd(1.0);
    for (int i = k; i < rows; i++) v[i] = a[i + (size_t)rows * k];
    v[k] += ph * xn;
    double vn = 0;
    for (int i = k; i < rows; i++) vn += std::norm(v[i]);
    const double beta = 2.0 / vn;
    for (int j = k; j < cols; j++) {
      zd s = 0;
      for (int i = k; i < rows; i++) s += std::conj(v[i]) * a[i + (size_t)rows * j];
      s *= beta;
      for (int i = k; i < rows; i++) a[i + (size_t)rows * j] -= v[i] * s;
    }
    zd s = 0;
    for (int i = k; i < rows; i++) s += std::conj(v[i]) * b[i];
    s *= beta;
    for (int i = k; i < rows; i++) b[i] -= v[i] * s;
  }
  for (int k = cols - 1; k >= 0; k--) {
    zd s = b[k];
    for (int j = k + 1; j < cols; j++) s -= a[k + (size_t)rows * j] * y[j];
    const zd d = a[k + (size_t)rows * k];
    y[k] = std::abs(d) > 0 ? s / d : zd(0.0);
  }
  double r = 0;
  for (int i = cols; i < rows; i++) r += std::norm(b[i]);
  return std::sqrt(r);
}

bool lu_solve(int n, const zd *a, int lda, const zd *b, zd *x) {
  std::vector<zd> m((size_t)n * n), r(b, b + n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) m[i + (size_t)n * j] = a[i + (size_t)lda * j];
  for (int k = 0; k < n; k++) {
    int p = k;
    for (int i = k + 1; i < n; i++)
      if (std::abs(m[i + (size_t)n * k]) > std::abs(m[p + (size_t)n * k])) p = i;
    if (std::abs(m[p + (size_t)n * k]) == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; j++) std::swap(m[k + (size_t)n * j], m[p + (size_t)n * j]);
      std::swap(r[k], r[p]);
    }
    for (int i = k + 1; i < n; i++) {
      const zd f = m[i + (size_t)n * k] / m[k + (size_t)n * k];
      m[i + (size_t)n * k] = f;
      for (int j = k + 1; j < n; j++) m[i + (size_t)n * j] -= f * m[k + (size_t)n * j];
      r[i] -= f * r[k];
    }
  }
  for (int k = n - 1; k >= 0; k--) {
    zd s = r[k];
    for (int j = k + 1; j < n; j++) s -= m[k + (size_t)n * j] * x[j];
    x[k] = s / m[k + (size_t)n * k];
  }
  return true;
}

bool harmonic_ritz(int m, const zd *hbar, int ldh, int k, zd *p, zd *theta) {
  std::vector<zd> F((size_t)m * m), HmH((size_t)m * m), em(m, 0.0), f(m), T((size_t)m * m), Q((size_t)m * m);
  for (int j = 0; j < m; j++)
    for (int i = 0; i < m; i++) {
      F[i + (size_t)m * j] = hbar[i + (size_t)ldh * j];
      HmH[i + (size_t)m * j] = std::conj(hbar[j + (size_t)ldh * i]);
    }
  em[m - 1] = 1.0;
  if (!lu_solve(m, HmH.data(), m, em.data(), f.data())) return false;
  const double h2 = std::norm(hbar[m + (size_t)ldh * (m - 1)]);
  for (int i = 0; i < m; i++) F[i + (size_t)m * (m - 1)] += h2 * f[i];
  if (!schur(m, F.data(), m, T.data(), Q.data())) return false;
  reorder_smallest(m, T.data(), Q.data(), k);
  for (int i = 0; i < m; i++) theta[i] = T[i + (size_t)m * i];
  for (int j = 0; j < k; j++)
    for (int i = 0; i < m; i++) p[i + (size_t)m * j] = Q[i + (size_t)m * j];
  return true;
}

}  // namespace dense
}  // namespace qqcd
