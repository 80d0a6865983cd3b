"""GPU parity of the operator layer against the C oracle (same seeded inputs).

Bar: FP64 results agree with the oracle to 1e-13 relative (max-norm over the
array, ||a-b||_max / ||b||_max); lower storage precisions to the tolerance
written at each test.  The oracle itself is pinned to the reference Fortran
by the golden fixtures (tests/test_oracle_golden.py).
"""
import numpy as np
import pytest

import oracle as O
from baylorlqcd_amd import diracops, gauge, lattice

pytestmark = pytest.mark.gpu

TOL64 = 1e-13


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def spinor(n, seed):
    rng = np.random.default_rng(seed)
    V = int(np.prod(n))
    return rng.standard_normal((V, 4, 3)) + 1j * rng.standard_normal((V, 4, 3))


def rows(n, a):
    nvh, ntot = O.ref_sizes(n)
    return np.asarray(a).reshape(6, ntot, 64, order="F")[:, :nvh]


@pytest.fixture(scope="module", params=[((8, 8, 8, 16), (1, 1, 1, -1)), ((4, 4, 4, 8), (1, 1, 1, 1)),
                                        ((8, 4, 4, 12), (-1, 1, 0, -1))])
def lat(request):
    n, bc = request.param
    print(f"[operators] lattice {n} bc {bc}", flush=True)
    lattice.init(n, bc)
    U = gauge.haar(n, 11)
    u = O.links_canon_to_ref(n, U)
    co = O.coact_array()
    diracops.gauge_load(u, co)
    yield n, bc, U, u, co
    lattice.finalize()


def test_hsingle_fp64(lat):
    n, bc, U, u, co = lat
    psi = spinor(n, 3)
    be, bo = O.spinor_canon_to_ref(n, psi)
    for gblclr, gin in ((1, bo), (2, be)):
        for idag in (0, 1):
            want = O.hsingle_ref(n, u, co, bc, gblclr, idag, gin)
            got = lattice.new_spinor(n).ravel(order="F")
            diracops.Hsingle(got, u, gin, idag, co, bc, gblclr)
            assert rel(rows(n, got), rows(n, want)) < TOL64, (gblclr, idag)
            # Hsingle writes every element: nothing outside rows 1..nvhalf
            assert np.all(np.asarray(got).reshape(6, -1, 64, order="F")[:, O.ref_sizes(n)[0]:] == 0)


@pytest.mark.parametrize("kappa", [(0.15, 0.0), (0.156, 0.005), (0.156, -0.005), (0.1563, 0.0015)])
@pytest.mark.parametrize("idag", [0, 1])
def test_hdbletm_fp64(lat, kappa, idag):
    n, bc, U, u, co = lat
    be, _ = O.spinor_canon_to_ref(n, spinor(n, 5))
    want = O.hdbletm_ref(n, u, co, bc, idag, kappa, be)
    got = np.zeros_like(want)
    diracops.Hdbletm(got, u, None, be, idag, co, kappa, -1, bc)
    assert rel(rows(n, got), rows(n, want)) < TOL64
    if kappa[1] <= 0:
        got2 = np.zeros_like(want)
        diracops.Hdouble(got2, u, be, idag, co, kappa[0], bc)
        assert rel(rows(n, got2), rows(n, want)) < TOL64


@pytest.mark.parametrize("prec,recon,tol", [(0, 18, 1e-13), (0, 12, 1e-12), (1, 18, 2e-6), (1, 12, 2e-6),
                                            (2, 18, 5e-3), (2, 12, 5e-3)])
def test_hop_precisions(lat, prec, recon, tol):
    """Storage-precision ladder: the relative L2 error against the FP64 oracle
    stays within the stated tolerance (FP32 ~ 1e-7 per element, FP16 with a
    per-site norm ~ 1e-3)."""
    n, bc, U, u, co = lat
    psi = spinor(n, 7)
    for par in (0, 1):
        want = O.hop(n, U, bc, [1] * 4, [1] * 4, 0, par, psi)
        got = diracops.hop_canonical(psi, par, 0, prec, recon)
        mask = O.parity_mask(n) == bool(par)
        err = np.linalg.norm(got[mask] - want[mask]) / np.linalg.norm(want[mask])
        assert err < tol, (par, err)


def test_general_coefficients():
    """Anisotropic / tadpole coefficients (c1 != c2) take the split path."""
    n, bc = (4, 4, 8, 8), (1, -1, 1, -1)
    lattice.init(n, bc)
    try:
        U = gauge.haar(n, 2)
        c1, c2 = (0.9, 1.1, 1.0, 0.7), (1.2, 0.8, 1.0, 0.7)
        co = O.coact_array(c1, c2)
        diracops.gauge_load_canonical(U, co)
        psi = spinor(n, 9)
        for idag in (0, 1):
            for par in (0, 1):
                want = O.hop(n, U, bc, c1, c2, idag, par, psi)
                got = diracops.hop_canonical(psi, par, idag)
                mask = O.parity_mask(n) == bool(par)
                assert rel(got[mask], want[mask]) < TOL64
    finally:
        lattice.finalize()


def test_gamma_ops(lat):
    n, bc, U, u, co = lat
    G = O.gamma_table()
    psi = spinor(n, 13)
    be, bo = O.spinor_canon_to_ref(n, psi)
    a = be.copy()
    diracops.gam5x(a)
    want, _ = O.spinor_canon_to_ref(n, np.einsum("de,sec->sdc", G[4], psi))
    assert rel(rows(n, a), rows(n, want)) < TOL64
    for mu in range(1, 6):
        eo, oo = np.zeros_like(be), np.zeros_like(bo)
        diracops.gammamult(be, bo, eo, oo, mu, 2, 5)
        we, wo = O.spinor_canon_to_ref(n, np.einsum("de,sec->sdc", G[mu - 1], psi))
        nvh, ntot = O.ref_sizes(n)
        for got, w in ((eo, we), (oo, wo)):
            g5 = got.reshape(6, ntot, 4, 2, 8, order="F")
            w5 = w.reshape(6, ntot, 4, 2, 8, order="F")
            assert np.array_equal(g5[:, :nvh, :, 1, 4], w5[:, :nvh, :, 1, 4])
            g5c = g5.copy()
            g5c[:, :, :, 1, 4] = 0
            assert not np.any(g5c)
