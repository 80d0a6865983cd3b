"""Host-side checks that need no GPU: the C ABI exports, the solvers' dense
linear algebra against numpy/scipy, the gauge generators, the geometry."""
import os
import re

import numpy as np
import pytest
import scipy.linalg as sla

import oracle as O
from baylorlqcd_amd import _lib, gauge, lattice

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    with open(os.path.join(ROOT, "include", "qqcd_hip.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(qqcd_\w+)\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) > 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # and the Python signature table covers them all
    assert not [s for s in syms if s not in _lib.SIGNATURES]


def test_no_gpu_is_an_error_not_a_fallback():
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be present")
    with pytest.raises(_lib.QQCDError):
        lattice.init((4, 4, 4, 4))


def cm(M):
    """column-major buffer of M (a C array of M^T)"""
    return np.ascontiguousarray(np.asarray(M, dtype=complex).T)


def dense_schur(A):
    n = A.shape[0]
    a, t, q = cm(A), np.zeros((n, n), complex), np.zeros((n, n), complex)
    _lib.check(_lib.lib().qqcd_dense_schur(n, _lib.dp(a.view(np.float64)), n, _lib.dp(t.view(np.float64)),
                                           _lib.dp(q.view(np.float64))))
    return t, q


@pytest.mark.parametrize("n", [1, 2, 7, 30, 61])
def test_dense_schur_and_reorder(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    t, q = dense_schur(A)
    T, Q = t.T, q.T
    assert np.linalg.norm(Q @ T @ Q.conj().T - A) < 1e-13 * np.linalg.norm(A)
    assert np.linalg.norm(Q.conj().T @ Q - np.eye(n)) < 1e-13 * n
    assert np.all(np.tril(T, -1) == 0)
    k = min(6, n)
    _lib.check(_lib.lib().qqcd_dense_reorder(n, _lib.dp(t.view(np.float64)), _lib.dp(q.view(np.float64)), k))
    T, Q = t.T, q.T
    assert np.linalg.norm(Q @ T @ Q.conj().T - A) < 1e-13 * np.linalg.norm(A)
    ev = np.sort(np.abs(np.linalg.eigvals(A)))
    assert np.allclose(np.sort(np.abs(np.diag(T))[:k]), ev[:k], rtol=1e-10)


def test_dense_schur_degenerate_and_hermitian():
    # repeated eigenvalues and a Hermitian (normal) matrix
    rng = np.random.default_rng(3)
    Q0 = sla.qr(rng.standard_normal((12, 12)) + 1j * rng.standard_normal((12, 12)))[0]
    A = Q0 @ np.diag([1, 1, 2, 2, 2, 3, 4, 5, 6, 7, 8, 9]).astype(complex) @ Q0.conj().T
    t, q = dense_schur(A)
    T, Q = t.T, q.T
    assert np.linalg.norm(Q @ T @ Q.conj().T - A) < 1e-12 * np.linalg.norm(A)
    assert np.allclose(np.sort(np.diag(T).real), [1, 1, 2, 2, 2, 3, 4, 5, 6, 7, 8, 9], atol=1e-10)


@pytest.mark.parametrize("rows,cols", [(5, 4), (41, 40), (21, 20), (11, 3)])
def test_dense_lsq(rows, cols):
    rng = np.random.default_rng(rows)
    H = rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))
    c = rng.standard_normal(rows) + 1j * rng.standard_normal(rows)
    y = np.zeros(cols, complex)
    res = np.zeros(1)
    _lib.check(_lib.lib().qqcd_dense_lsq(rows, cols, _lib.dp(cm(H).view(np.float64)), rows,
                                         _lib.dp(np.ascontiguousarray(c).view(np.float64)),
                                         _lib.dp(y.view(np.float64)), _lib.dp(res)))
    want = np.linalg.lstsq(H, c, rcond=None)[0]
    assert np.allclose(y, want, rtol=1e-10, atol=1e-12)
    assert abs(res[0] - np.linalg.norm(c - H @ want)) < 1e-10


def test_harmonic_ritz_against_scipy():
    """The k harmonic Ritz values of smallest modulus: eigenvalues of
    H_m + |h|^2 H_m^{-H} e_m e_m^T (Morgan 2002), and the returned Schur
    vectors span their eigenvectors."""
    m, k = 30, 8
    rng = np.random.default_rng(11)
    Hb = np.zeros((m + 1, m), complex)
    for j in range(m):
        Hb[: j + 2, j] = rng.standard_normal(j + 2) + 1j * rng.standard_normal(j + 2)
        Hb[j + 1, j] = abs(Hb[j + 1, j])
    Hb[:m, :m] += 3 * np.eye(m)
    p = np.zeros((k, m), complex)
    th = np.zeros(m, complex)
    _lib.check(_lib.lib().qqcd_dense_harmonic_ritz(m, _lib.dp(cm(Hb).view(np.float64)), m + 1, k,
                                                   _lib.dp(p.view(np.float64)), _lib.dp(th.view(np.float64))))
    P = p.T
    Hm = Hb[:m, :m]
    f = np.linalg.solve(Hm.conj().T, np.eye(m)[:, -1])
    F = Hm.copy()
    F[:, -1] += abs(Hb[m, m - 1]) ** 2 * f
    w, vecs = np.linalg.eig(F)
    order = np.argsort(np.abs(w))
    assert np.allclose(np.sort(np.abs(th[:k])), np.abs(w[order[:k]]), rtol=1e-9)
    assert np.linalg.norm(P.conj().T @ P - np.eye(k)) < 1e-12
    for i in order[:k]:
        v = vecs[:, i]
        assert np.linalg.norm(v - P @ (P.conj().T @ v)) < 1e-8


def test_gauge_generators():
    n = (4, 4, 4, 8)
    for U in (gauge.haar(n, 1), gauge.near_identity(n, 0.1, 2)):
        I = np.eye(3)
        assert np.max(np.abs(U @ U.conj().swapaxes(-1, -2) - I)) < 1e-13
        assert np.max(np.abs(np.linalg.det(U) - 1)) < 1e-13
        assert np.max(np.abs(gauge.reconstruct12(gauge.compress12(U)) - U)) < 1e-13
    # cold field: plaquette 1; Haar field: plaquette ~ 0
    V = int(np.prod(n))
    assert abs(gauge.plaquette(n, np.broadcast_to(np.eye(3), (V, 4, 3, 3))) - 1) < 1e-14
    assert abs(gauge.plaquette(n, gauge.haar(n, 5))) < 0.05


def test_reference_sizes():
    for n in ((4, 4, 4, 4), (8, 8, 8, 16), (24, 24, 24, 48), (32, 32, 32, 64), (48, 48, 48, 96)):
        assert lattice.ref_sizes(n) == O.ref_sizes(n)
    assert lattice.ref_sizes((32, 32, 32, 64)) == (65536, 69632)
    assert lattice.ref_sizes((48, 48, 48, 96)) == (331776, 345600)


def test_geometry_check_rejects_multi_process():
    vb, vbi = lattice.gblchker()
    lattice.check_geometry(vb, vbi, 0, [False] * 4)
    with pytest.raises(_lib.QQCDError):
        lattice.check_geometry(myid=1)
    with pytest.raises(_lib.QQCDError):
        lattice.check_geometry(ldiv=[True, False, False, False])
    with pytest.raises(_lib.QQCDError):
        lattice.check_geometry(vecbl=vb[::-1])


def test_bench_refuses_more_than_one_gpu():
    # the bench measures one GPU; a request for more exits with a message (status 2)
    import subprocess
    import sys
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 2
    assert "one GPU" in out.stderr and out.stdout.strip() == ""
