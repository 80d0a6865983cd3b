"""Closed-form and invariant checks that pin the oracle independently of the
reference build: each would fail for a dropped term, a wrong sign or index,
or a transposed link.

* free field (U = 1): plane waves are eigenvectors of H with symbol
  sum_mu [2 c1 cos p_mu - 2 i c2 gamma_mu sin p_mu]; antiperiodic directions
  shift p by pi/n;
* gamma5-hermiticity: H(idag=1) = g5 H g5 = H^dagger;
* gauge covariance: H[U^W](W psi) = W H[U] psi (the D/debug.f90:31-82 test);
* dense 4^4 brute force: the even-odd Schur solve with source preparation
  and reconstruction equals the dense solve of the full M = G - kappa H.
"""
import numpy as np
import pytest

import oracle as O
from baylorlqcd_amd import gauge


def coords(n):
    V = int(np.prod(n))
    s = np.arange(V)
    return np.stack([s % n[0], (s // n[0]) % n[1], (s // (n[0] * n[1])) % n[2], s // (n[0] * n[1] * n[2])], 1)


def spinor(n, seed):
    rng = np.random.default_rng(seed)
    V = int(np.prod(n))
    return rng.standard_normal((V, 4, 3)) + 1j * rng.standard_normal((V, 4, 3))


@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (1, 1, 1, -1), (-1, 1, -1, 1)])
def test_free_field_plane_waves(bc):
    n = (4, 4, 8, 8)
    V = int(np.prod(n))
    U = np.broadcast_to(np.eye(3, dtype=complex), (V, 4, 3, 3)).copy()
    G = O.gamma_table()
    X = coords(n)
    rng = np.random.default_rng(1)
    c1, c2 = (1.0, 0.8, 1.0, 0.6), (1.0, 0.9, 1.0, 0.6)
    for trial in range(3):
        k = rng.integers(0, 4, size=4)
        p = np.array([(2 * np.pi * k[m] + (np.pi if bc[m] == -1 else 0.0)) / n[m] for m in range(4)])
        u = rng.standard_normal((4, 3)) + 1j * rng.standard_normal((4, 3))
        phase = np.exp(1j * (X @ p))
        psi = phase[:, None, None] * u[None]
        sym = sum(2 * c1[m] * np.cos(p[m]) * np.eye(4) - 2j * c2[m] * np.sin(p[m]) * G[m] for m in range(4))
        want = phase[:, None, None] * (sym @ u)[None]
        for par in (0, 1):
            got = O.hop(n, U, bc, c1, c2, 0, par, psi)
            mask = O.parity_mask(n) == bool(par)
            assert np.max(np.abs(got[mask] - want[mask])) < 1e-14 * np.max(np.abs(want[mask]))


def test_gamma5_hermiticity():
    n, bc = (4, 4, 4, 8), (1, 1, 1, -1)
    U = gauge.haar(n, 3)
    G5 = O.gamma_table()[4]
    x, y = spinor(n, 1), spinor(n, 2)
    odd = O.parity_mask(n)
    x[~odd] = 0
    Hx = O.hop(n, U, bc, [1] * 4, [1] * 4, 0, 0, x)
    Hdy = O.hop(n, U, bc, [1] * 4, [1] * 4, 1, 1, np.where(odd[:, None, None], 0, y))
    # <y_e, H x_o> = <H^dag y_e, x_o>
    lhs = np.vdot(y[~odd], Hx[~odd])
    rhs = np.vdot(Hdy[odd], x[odd])
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # H^dag = g5 H g5
    g5x = np.einsum("de,sec->sdc", G5, x)
    want = np.einsum("de,sec->sdc", G5, O.hop(n, U, bc, [1] * 4, [1] * 4, 0, 0, g5x))
    got = O.hop(n, U, bc, [1] * 4, [1] * 4, 1, 0, x)
    assert np.max(np.abs(got[~odd] - want[~odd])) < 1e-13


def test_gauge_covariance():
    n, bc = (4, 4, 4, 8), (1, 1, 1, -1)
    V = int(np.prod(n))
    U = gauge.haar(n, 5)
    W = gauge.haar(n, 6)[:, 0]  # one SU(3) matrix per site
    X = coords(n)
    fwd = np.zeros((4, V), dtype=int)
    for m in range(4):
        Y = X.copy()
        Y[:, m] = (Y[:, m] + 1) % n[m]
        fwd[m] = Y[:, 0] + n[0] * (Y[:, 1] + n[1] * (Y[:, 2] + n[2] * Y[:, 3]))
    UW = np.empty_like(U)
    for m in range(4):
        UW[:, m] = W @ U[:, m] @ W[fwd[m]].conj().transpose(0, 2, 1)
    psi = spinor(n, 7)
    Wpsi = np.einsum("sab,sdb->sda", W, psi)
    for par in (0, 1):
        for kap in ((0.13, 0.0), (0.13, 0.04)):
            a = O.hop(n, UW, bc, [1] * 4, [1] * 4, 0, par, Wpsi)
            b = np.einsum("sab,sdb->sda", W, O.hop(n, U, bc, [1] * 4, [1] * 4, 0, par, psi))
            mask = O.parity_mask(n) == bool(par)
            assert np.max(np.abs(a[mask] - b[mask])) < 1e-13


def test_dense_even_odd_solve_4x4x4x4():
    """Brute force on 4^4 (rank 3072): assemble M = G - kappa H column by
    column from the oracle's hopping term and check that the even-odd path
    (source preparation, Schur solve, reconstruction) gives the dense
    solution, for Wilson and twisted mass."""
    from oracle import solvers
    n, bc = (4, 4, 4, 4), (1, 1, 1, -1)
    V = int(np.prod(n))
    U = gauge.haar(n, 9)
    b = spinor(n, 10)
    for kap in ((0.12, 0.0), (0.12, 0.03)):
        N = V * 12
        M = np.zeros((N, N), dtype=complex)
        e = np.zeros(N, dtype=complex)
        for j in range(N):
            e[:] = 0
            e[j] = 1
            M[:, j] = O.full_m(n, U, bc, [1] * 4, [1] * 4, kap, e.reshape(V, 4, 3)).ravel()
        x_dense = np.linalg.solve(M, b.ravel()).reshape(V, 4, 3)
        x, res = solvers.fermprop(n, U, bc, kap, b, tol=1e-13, solver="gmresdr", m=20, k=5)
        assert res < 1e-11
        assert np.max(np.abs(x - x_dense)) < 1e-9 * np.max(np.abs(x_dense))
        # the structure of M: unit diagonal for Wilson, 1 + a i g5 for twisted
        assert np.allclose(np.diag(M), 1.0)
