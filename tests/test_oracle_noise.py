"""CPU checks of the oracle's noise sources and trace estimators.

The noise generator has no reference counterpart to compare with (the
reference draws from its lagged-Fibonacci stream rnds); what the reference
fixes is the value set and the mapping of a uniform number to it
(qqcd-test/cfgsprops/quark/disconloops.f90:4518-4569), which these tests pin
together with the statistical properties a Z4 source must have.  The 16
local Dirac matrices are pinned by the Clifford algebra; the trace
estimator by the exact trace it must reproduce when the sources run over a
basis.
"""
import numpy as np

import oracle as O

N = (4, 4, 4, 8)


def test_z4_values_and_statistics():
    a = O.z4(N, 1234, 0)
    vals = {1: 0, -1: 0, 1j: 0, -1j: 0}
    for v in a.ravel():
        vals[complex(v)] += 1
    tot = a.size
    for v, c in vals.items():
        # each value with probability 1/4: within 5 sigma
        assert abs(c / tot - 0.25) < 5 * np.sqrt(0.25 * 0.75 / tot), (v, c)
    assert np.all(np.abs(a) == 1.0)
    # E[eta eta^dag] = 1: off-diagonal products average to zero
    flat = a.reshape(-1)
    assert abs(np.mean(flat[:-1] * np.conj(flat[1:]))) < 5 / np.sqrt(tot)


def test_z4_deterministic_and_independent():
    a = O.z4(N, 99, 3)
    assert np.array_equal(a, O.z4(N, 99, 3))
    b = O.z4(N, 99, 4)
    c = O.z4(N, 100, 3)
    for other in (b, c):
        agree = np.mean(a == other)
        assert abs(agree - 0.25) < 0.05
    s = O.z4(N, 99, 3, scale=1 / np.sqrt(2))
    assert np.allclose(s, a / np.sqrt(2), rtol=0, atol=1e-15)


def test_z4_threshold_mapping():
    # the value follows from u alone: recompute u in numpy from the same counter
    V = int(np.prod(N))
    seed, rhs = 5, 2
    G = np.uint64(0x9E3779B97F4A7C15)

    def mix(z):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))

    with np.errstate(over="ignore"):
        c = (np.uint64(rhs) * np.uint64(V) + np.arange(V, dtype=np.uint64)[:, None]) * np.uint64(12) + \
            np.arange(12, dtype=np.uint64)[None, :]
        u = (mix(np.uint64(seed) + G * (np.uint64(1) + c)) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    want = np.where(u > 0.75, 1, np.where((u > 0.5) & (u < 0.75), -1,
                                           np.where((u > 0.25) & (u < 0.5), 1j, -1j)))
    assert np.array_equal(O.z4(N, seed, rhs).reshape(V, 12), want)
    # Z2 (z2source): the same uniform numbers, 1 above 0.5, else -1, real
    assert np.array_equal(O.z2(N, seed, rhs).reshape(V, 12), np.where(u > 0.5, 1.0, -1.0) + 0j)


def test_gamma16_algebra():
    G = O.gamma16()
    g = O.gamma_table()
    # trace orthogonality: Tr(G_a^dag G_b) = 4 delta_ab (a basis of 4x4 matrices)
    M = np.einsum("aji,bjk->abik", np.conj(G), G).trace(axis1=2, axis2=3)
    assert np.allclose(M, 4 * np.eye(16), atol=1e-14)
    # Clifford algebra of gamma_1..4 and gamma_5 = g1 g2 g3 g4
    for m in range(4):
        for n in range(4):
            ac = g[m] @ g[n] + g[n] @ g[m]
            assert np.allclose(ac, 2 * np.eye(4) * (m == n), atol=1e-14)
    assert np.allclose(g[0] @ g[1] @ g[2] @ g[3], g[4], atol=1e-14)
    # monomial rows (what the device kernel relies on)
    for a in range(16):
        assert np.all(np.count_nonzero(np.abs(G[a]) > 1e-12, axis=1) == 1)
    # sigma_12 = i g1 g2 (anticommuting)
    assert np.allclose(G[10], 1j * g[0] @ g[1], atol=1e-14)


def test_trace16_exact_trace_over_a_basis():
    # with eta over all unit vectors and x = B eta for a site-local 12x12 B(x),
    # the estimator summed over the basis is sum_{x in t} Tr(Gamma_n B(x))
    n = (2, 2, 2, 4)
    V = int(np.prod(n))
    rng = np.random.default_rng(3)
    B = rng.standard_normal((V, 12, 12)) + 1j * rng.standard_normal((V, 12, 12))
    tot = np.zeros((n[3], 16), dtype=np.complex128)
    for s in range(V):
        for j in range(12):
            eta = np.zeros((V, 4, 3), dtype=np.complex128)
            eta[s].reshape(12)[j] = 1.0
            x = np.zeros_like(eta)
            x[s] = (B[s] @ eta[s].reshape(12)).reshape(4, 3)
            tot += O.trace16(n, eta, x)
    G = O.gamma16()
    want = np.zeros_like(tot)
    t = np.arange(V) // (n[0] * n[1] * n[2])
    for s in range(V):
        for a in range(16):
            Gfull = np.kron(G[a], np.eye(3))  # (spin, colour) row order = 3*spin + colour
            want[t[s], a] += np.trace(Gfull @ B[s])
    assert np.allclose(tot, want, rtol=1e-12, atol=1e-12)
