"""CPU checks of the oracle's loop operators (loopops, momfacs, spacesum).

Pinned two ways:

* against the reference Fortran itself: golden_loops_8x8x8x16.npz holds the
  output of the reference's loopops over every (ieo, ibleo) slice (per-site
  operators on timeslices 0 and nt-1), of spacesum with momfacs for two
  momentum origins, and of nsaverage, for two boundary-condition sets
  (tests/golden/make_golden.py loops);
* by what the mathematics fixes: every operator is gauge invariant (a
  random gauge rotation of U, b and x leaves each site's value unchanged --
  a link on the wrong side or a missing dagger breaks it); the local scalar
  of x = b is |b|^2 (+12 with ksub = 1); the momentum sums at momentum 1 add
  up to the total over the lattice.
"""
import os

import numpy as np

import oracle as O
from baylorlqcd_amd import gauge

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "golden_loops_8x8x8x16.npz")
N = (8, 8, 8, 16)
CASES = {"a": ((1, 1, 1, -1), 0, (1, 1, 1)), "b": ((-1, 1, 0, 1), 1, (3, 2, 5))}


def spinor(n, seed):
    rng = np.random.default_rng(seed)
    V = int(np.prod(n))
    return rng.standard_normal((V, 4, 3)) + 1j * rng.standard_normal((V, 4, 3))


def sites(n):
    ls = n[0] * n[1] * n[2]
    return np.concatenate([np.arange(ls), np.arange((n[3] - 1) * ls, n[3] * ls)])


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def inputs():
    return gauge.haar(N, 3), spinor(N, 301), spinor(N, 302)


def test_loopops_matches_reference_golden():
    g = np.load(GOLD)
    U, b, x = inputs()
    for name, (bc, ksub, io) in CASES.items():
        J = O.loopops(N, U, bc, b, x, ksub)
        assert rel(J[sites(N)], g[f"loopops_{name}"]) < 1e-13, name
        assert rel(O.spacesum(N, J, io), g[f"spacesum_{name}"]) < 1e-13, name


def test_nsaverage_matches_reference_golden():
    g = np.load(GOLD)
    U, b, x = inputs()
    bc = CASES["a"][0]
    J = O.spacesum(N, O.loopops(N, U, bc, b, x, 0), nmom=1)[:, 0, :]
    J[:, :4] *= 0.15
    assert rel(J, g["nsaverage_a_k150"]) < 1e-13


def test_loopops_gauge_invariant():
    n = (4, 4, 4, 8)
    bc = (-1, 1, 0, -1)
    U = gauge.haar(n, 5)
    b, x = spinor(n, 1), spinor(n, 2)
    # random SU(3) rotation per site: U_mu(x) -> G(x) U_mu(x) G(x+mu)^dag, psi -> G psi
    G = gauge.haar(n, 6)[:, 0]
    Ug = np.empty_like(U)
    for mu in range(4):
        nb, _ = O.forward(n, (1, 1, 1, 1), mu)
        Ug[:, mu] = G @ U[:, mu] @ np.conj(np.transpose(G[nb], (0, 2, 1)))
    bg = np.einsum("sij,saj->sai", G, b)
    xg = np.einsum("sij,saj->sai", G, x)
    J = O.loopops(n, U, bc, b, x, 0)
    Jg = O.loopops(n, Ug, bc, bg, xg, 0)
    assert rel(Jg, J) < 1e-13
    # and a wrong placement of the link is not invariant (the check has teeth)
    Jbad = O.loopops(n, np.conj(np.transpose(Ug, (0, 1, 3, 2))), bc, bg, xg, 0)
    assert rel(Jbad[:, :4], J[:, :4]) > 1e-3


def test_local_scalar_and_ksub():
    n = (4, 4, 4, 8)
    U = gauge.haar(n, 7)
    b = spinor(n, 3)
    J0 = O.loopops(n, U, (1, 1, 1, -1), b, b, 0)
    J1 = O.loopops(n, U, (1, 1, 1, -1), b, b, 1)
    assert np.allclose(J0[:, 4], np.sum(np.abs(b) ** 2, axis=(1, 2)), rtol=1e-14, atol=0)
    assert np.allclose(J1[:, 4] - J0[:, 4], 12.0, rtol=0, atol=1e-12)
    assert np.array_equal(J1[:, [0, 1, 2, 3, 5, 6, 7, 8]], J0[:, [0, 1, 2, 3, 5, 6, 7, 8]])


def test_spacesum_momentum_one_is_the_timeslice_sum():
    n = (4, 4, 4, 8)
    rng = np.random.default_rng(9)
    J = rng.standard_normal((O.volume(n), 9)) + 1j * rng.standard_normal((O.volume(n), 9))
    out = O.spacesum(n, J, (2, 1, 3))
    t = O.coords(n)[:, 3]
    for tt in range(n[3]):
        assert np.allclose(out[tt, 0], J[t == tt].sum(0), rtol=1e-14, atol=1e-13)
    # the local vector currents carry no momentum factors
    assert np.all(out[:, 1:, 5:] == 0)
    # zero-mode factors: sin terms vanish at the origin site and cos terms are 1
    F = O.momfacs(n, (1, 1, 1))
    assert np.allclose(F[0, :, :3], 0.0, atol=1e-6) and np.allclose(F[0, :, 3:], 1.0, atol=1e-12)
