"""CPU restatement of the reference's inverters and of fermprop (numpy).

TEST INFRASTRUCTURE ONLY.  The operator is the C oracle's Schur operator
(``oracle.schur``); the small dense problems use numpy / scipy.

* ``bicgstab``  -- qqcd-test/cfgsprops/quark/inverters.f90:302-487
* ``gmresdr``   -- inverters.f90:2871-3572 (Morgan's GMRES-DR: harmonic Ritz
                   vectors of smallest modulus via a sorted complex Schur form,
                   as orgschur does with myselect, :3356-3378)
* ``gmresproj`` -- inverters.f90:7183-8304
* ``fermprop``  -- qqcd-test/cfgsprops/quark/quark.f90:570-617, 786-819

Vectors are flattened even-site arrays (complex128).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

from . import full_m, gamma_table, hop, parity_mask, schur


class SchurOperator:
    """M_hat on the even sites (canonical whole-lattice arrays inside)."""

    def __init__(self, n, U, bc, kappa, c1=(1, 1, 1, 1), c2=(1, 1, 1, 1)):
        self.n, self.U, self.bc, self.kappa = n, U, bc, kappa
        self.c1, self.c2 = c1, c2
        self.even = ~parity_mask(n)
        self.V = int(np.prod(n))
        self.count = 0

    def __call__(self, v):
        full = np.zeros((self.V, 4, 3), dtype=np.complex128)
        full[self.even] = v.reshape(-1, 4, 3)
        out = schur(self.n, self.U, self.bc, self.c1, self.c2, 0, self.kappa, full)
        self.count += 1
        return out[self.even].ravel()


def bicgstab(A, b, x0, tol, maxit=100000):
    """BiCGStab; stops when the recursive residual drops below tol*||r0||."""
    x = x0.copy()
    r = b - A(x)
    rhat = r.copy()
    r0 = np.linalg.norm(r)
    rho = alpha = omega = 1.0 + 0j
    v = np.zeros_like(b)
    p = np.zeros_like(b)
    it = 0
    while it < maxit:
        oldrho = rho
        rho = np.vdot(rhat, r)
        beta = (rho / oldrho) * (alpha / omega)
        p = r + beta * (p - omega * v)
        v = A(p)
        alpha = rho / np.vdot(rhat, v)
        r = r - alpha * v
        t = A(r)
        omega = np.vdot(t, r) / np.vdot(t, t)
        x = x + omega * r + alpha * p
        r = r - omega * t
        it += 1
        if np.linalg.norm(r) < tol * r0:
            break
    return x, it


def _arnoldi(A, V, H, j):
    w = A(V[:, j])
    for _ in range(2):  # classical Gram-Schmidt, twice
        h = V[:, :j + 1].conj().T @ w
        w = w - V[:, :j + 1] @ h
        H[:j + 1, j] += h
    H[j + 1, j] = np.linalg.norm(w)
    V[:, j + 1] = w / H[j + 1, j]


def gmresdr(A, b, m, k, tol, x0=None, maxit=100000):
    """GMRES-DR(m, k); returns x, iterations and the last deflation space
    (V_{k+1}, Hbar_k) of the final restart."""
    N = b.size
    x = np.zeros_like(b) if x0 is None else x0.copy()
    r = b - A(x)
    r0 = np.linalg.norm(r)
    V = np.zeros((N, m + 1), dtype=np.complex128)
    H = np.zeros((m + 1, m), dtype=np.complex128)
    c = np.zeros(m + 1, dtype=np.complex128)
    V[:, 0] = r / r0
    c[0] = r0
    j0, it, defl = 0, 0, None
    while True:
        for j in range(j0, m):
            _arnoldi(A, V, H, j)
            it += 1
        y = np.linalg.lstsq(H, c, rcond=None)[0]
        x = x + V[:, :m] @ y
        r = b - A(x)
        if np.linalg.norm(r) < tol * r0 or it >= maxit:
            return x, it, defl
        srv = c - H @ y
        hm = H[:m, :m]
        f = np.linalg.solve(hm.conj().T, np.eye(m)[:, m - 1])
        F = hm.copy()
        F[:, m - 1] += abs(H[m, m - 1]) ** 2 * f
        theta = np.linalg.eigvals(F)
        cut = np.sort(np.abs(theta))[k - 1]
        T, Z, sdim = sla.schur(F, output="complex", sort=lambda z: abs(z) <= cut * (1 + 1e-12))
        P = np.zeros((m + 1, k + 1), dtype=np.complex128)
        P[:m, :k] = Z[:, :k]
        s = srv.copy()
        for _ in range(2):
            s = s - P[:, :k] @ (P[:, :k].conj().T @ s)
        P[:, k] = s / np.linalg.norm(s)
        Hn = P.conj().T @ H @ P[:m, :k]
        cn = P.conj().T @ srv
        Vn = V @ P
        vk = Vn[:, k] - Vn[:, :k] @ (Vn[:, :k].conj().T @ Vn[:, k])
        Vn[:, k] = vk / np.linalg.norm(vk)
        V[:] = 0
        V[:, :k + 1] = Vn
        H[:] = 0
        H[:k + 1, :k] = Hn
        c[:] = 0
        c[:k + 1] = cn
        defl = (Vn.copy(), Hn.copy())
        j0 = k


def gmresproj(A, b, defl, m, tol, maxit=100000):
    """GMRES-Proj: projection over (V_{k+1}, Hbar_k), then GMRES(m), repeated."""
    Vd, Hd = defl
    k = Hd.shape[1]
    x = np.zeros_like(b)
    r = b.copy()
    r0 = np.linalg.norm(r)
    it = 0
    while it < maxit:
        d = np.linalg.lstsq(Hd, Vd.conj().T @ r, rcond=None)[0]
        x = x + Vd[:, :k] @ d
        r = b - A(x)
        if np.linalg.norm(r) < tol * r0:
            break
        V = np.zeros((b.size, m + 1), dtype=np.complex128)
        H = np.zeros((m + 1, m), dtype=np.complex128)
        beta = np.linalg.norm(r)
        V[:, 0] = r / beta
        for j in range(m):
            _arnoldi(A, V, H, j)
            it += 1
        e = np.zeros(m + 1, dtype=np.complex128)
        e[0] = beta
        y = np.linalg.lstsq(H, e, rcond=None)[0]
        x = x + V[:, :m] @ y
        r = b - A(x)
        if np.linalg.norm(r) < tol * r0:
            break
    return x, it


def source_prep(n, U, bc, kappa, b):
    """vecsrc_e = b_e/k^2 + H_eo (1 - a i g5) b_o / (k (1 + a^2)), a = 2 k mu
    (quark.f90:600-617).  b is a whole-lattice canonical array."""
    k, a = kappa[0], 2 * kappa[0] * kappa[1]
    odd = parity_mask(n)
    ig5 = 1j * gamma_table()[4]
    t = np.einsum("de,sec->sdc", np.eye(4) - a * ig5, b)
    t[~odd] = 0
    hb = hop(n, U, bc, [1] * 4, [1] * 4, 0, 0, t)
    src = b / k ** 2 + hb / (k * (1 + a * a))
    return src[~odd].ravel()


def reconstruct(n, U, bc, kappa, b, xe):
    """x_o = (1 - a i g5)(b_o + k H_oe x_e)/(1 + a^2) (quark.f90:786-819)."""
    k, a = kappa[0], 2 * kappa[0] * kappa[1]
    odd = parity_mask(n)
    V = int(np.prod(n))
    x = np.zeros((V, 4, 3), dtype=np.complex128)
    x[~odd] = xe.reshape(-1, 4, 3)
    hx = hop(n, U, bc, [1] * 4, [1] * 4, 0, 1, x)
    ig5 = 1j * gamma_table()[4]
    xo = np.einsum("de,sec->sdc", np.eye(4) - a * ig5, b + k * hx) / (1 + a * a)
    x[odd] = xo[odd]
    return x


def fermprop(n, U, bc, kappa, b, tol=1e-12, solver="gmresdr", m=40, k=10):
    """Full-lattice solve M x = b through the even-odd system (isotropic
    coefficients).  Returns x and the true full-system relative residual."""
    src = source_prep(n, U, bc, kappa, b)
    A = SchurOperator(n, U, bc, kappa)
    if solver == "gmresdr":
        xe, _, _ = gmresdr(A, src, m, k, tol)
    else:
        xe, _ = bicgstab(A, src, np.zeros_like(src), tol)
    x = reconstruct(n, U, bc, kappa, b, xe)
    r = b - full_m(n, U, bc, [1] * 4, [1] * 4, kappa, x)
    return x, np.linalg.norm(r) / np.linalg.norm(b)
