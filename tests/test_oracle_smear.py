"""The oracle's Gaussian smearing (oracle.smear) against the reference's own
smear (golden_smear_8x8x8x16.npz, tests/golden/make_golden.py smear: three
(bc, asmear, nsmear) cases incl. antiperiodic and fixed spatial edges and
nsmear = 0); the free field (plane waves are eigenvectors with eigenvalue
(1 - 6a + 2a sum_i cos p_i)^nsmear) and gauge covariance beside it.  The
reference's twMshift and gamma5mult outputs in the same fixture are checked
against closed forms built from the oracle's Hsingle / Hdbletm:
twMshift g = g + (-i g5)/(2 kappa mu) (g + kappa H g_other),
gamma5mult g = (-i g5)/(2 kappa mu) Hdbletm g."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from baylorlqcd_amd import gauge

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
from make_golden import SHIFT_KAPPA, SMEAR_CASES, slice_summary, smear_inputs  # noqa: E402

N = (8, 8, 8, 16)


@pytest.fixture(scope="module")
def g():
    return np.load(os.path.join(ROOT, "tests", "golden", "golden_smear_8x8x8x16.npz"))


def close(got, gold, tol):
    sl, nrm, sm = got
    return (np.max(np.abs(sl - gold[0])) < tol * np.max(np.abs(gold[0])) + 1e-300
            and np.max(np.abs(nrm - gold[1])) < tol * np.max(gold[1])
            and np.max(np.abs(sm - gold[2])) < tol * max(1.0, np.max(np.abs(gold[2]))) * 10)


@pytest.mark.parametrize("case", SMEAR_CASES, ids=[c[0] for c in SMEAR_CASES])
def test_smear_reference(g, case):
    name, bc, a, ns = case
    U, psi = smear_inputs(N)
    got = slice_summary(N, O.smear(N, U, bc, psi, a, ns))
    gold = (g[f"smear_{name}_slices"], g[f"smear_{name}_norms"], g[f"smear_{name}_sums"])
    assert close(got, gold, 1e-13)


def test_smear_free_field():
    n = (8, 8, 4, 4)
    V = int(np.prod(n))
    U = np.broadcast_to(np.eye(3, dtype=np.complex128), (V, 4, 3, 3)).copy()
    c = O.coords(n)
    p = 2 * np.pi * np.array([1, 3, 1, 0]) / np.array(n)
    wave = np.exp(1j * (c @ p))
    psi = wave[:, None, None] * np.ones((1, 4, 3))
    a, ns = 0.17, 3
    lam = (1 - 6 * a + 2 * a * np.sum(np.cos(p[:3]))) ** ns
    assert np.allclose(O.smear(n, U, (1, 1, 1, 1), psi, a, ns), lam * psi, atol=1e-13)
    # antiperiodic in x: momenta shifted by pi/nx
    pa = p + np.array([np.pi / n[0], 0, 0, 0])
    wave = np.exp(1j * (c @ pa))
    psi = wave[:, None, None] * np.ones((1, 4, 3))
    lam = (1 - 6 * a + 2 * a * np.sum(np.cos(pa[:3]))) ** ns
    assert np.allclose(O.smear(n, U, (-1, 1, 1, 1), psi, a, ns), lam * psi, atol=1e-13)


def test_smear_gauge_covariant():
    n = (8, 4, 4, 4)
    U, psi = gauge.haar(n, 5), smear_inputs(n)[1]
    Om = gauge.haar(n, 6)[:, 0]
    fw = [O.forward(n, (1, 1, 1, 1), mu)[0] for mu in range(4)]
    Ur = np.stack([Om @ U[:, mu] @ np.conj(np.swapaxes(Om[fw[mu]], -1, -2)) for mu in range(4)], 1)
    rot = lambda f: np.einsum("vab,vdb->vda", Om, f)  # noqa: E731
    assert np.allclose(O.smear(n, Ur, (1, 1, 1, 1), rot(psi), 0.2, 2), rot(O.smear(n, U, (1, 1, 1, 1), psi, 0.2, 2)),
                       atol=1e-12)


@pytest.mark.parametrize("idag", [0, 1])
def test_twmshift_gamma5mult_closed_forms(g, idag):
    U, psi = smear_inputs(N)
    u = O.links_canon_to_ref(N, U)
    ge, go = O.spinor_canon_to_ref(N, psi)
    co = O.coact_array()
    k, mu = SHIFT_KAPPA
    bc = (1, 1, 1, -1)
    Gam = -1j * O.gamma_table()[4]
    he = O.spinor_ref_to_canon(N, O.hsingle_ref(N, u, co, bc, 1, idag, go), None)
    ho = O.spinor_ref_to_canon(N, None, O.hsingle_ref(N, u, co, bc, 2, idag, ge))
    odd = O.parity_mask(N)
    H = np.where(odd[:, None, None], ho, he)
    twm = psi + np.einsum("de,sec->sdc", Gam, psi + k * H) / (2 * k * mu)
    assert close(slice_summary(N, twm), (g[f"twm{idag}_slices"], g[f"twm{idag}_norms"], g[f"twm{idag}_sums"]), 1e-12)
    hd = O.spinor_ref_to_canon(N, O.hdbletm_ref(N, u, co, bc, idag, (k, mu), ge), None)
    g5m = np.einsum("de,sec->sdc", Gam, hd) / (2 * k * mu)
    g5m[odd] = 0
    assert close(slice_summary(N, g5m), (g[f"g5m{idag}_slices"], g[f"g5m{idag}_norms"], g[f"g5m{idag}_sums"]),
                 1e-12)
