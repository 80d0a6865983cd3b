"""Gaussian smearing (qqcd_smear, qqcd_smear.hip) and the shifted twisted
operators twMshift / gamma5mult on the GPU, against the reference's own
outputs (golden_smear_8x8x8x16.npz) and, on a larger lattice, against the
oracle's smearing (pinned on the same fixture).  Bar: 1e-13 relative (FP64,
sums in another order)."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from baylorlqcd_amd import diracops, gauge, lattice

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
from make_golden import SHIFT_KAPPA, SMEAR_CASES, slice_summary, smear_inputs  # noqa: E402

N = (8, 8, 8, 16)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def g():
    lattice.init(N, (1, 1, 1, -1))
    yield np.load(os.path.join(ROOT, "tests", "golden", "golden_smear_8x8x8x16.npz"))
    lattice.finalize()


@pytest.mark.parametrize("case", SMEAR_CASES, ids=[c[0] for c in SMEAR_CASES])
def test_smear_reference(g, case):
    name, bc, a, ns = case
    print(f"[smear] case {name}", flush=True)
    U, psi = smear_inputs(N)
    u = O.links_canon_to_ref(N, U)
    ge, go = O.spinor_canon_to_ref(N, psi)
    he, ho = np.zeros_like(ge), np.zeros_like(go)
    diracops.smear(he, ho, u, ge, go, a, ns, bc)
    sl, nrm, sm = slice_summary(N, O.spinor_ref_to_canon(N, he, ho))
    assert rel(sl, g[f"smear_{name}_slices"]) < 1e-13
    assert rel(nrm, g[f"smear_{name}_norms"]) < 1e-13
    assert np.max(np.abs(sm - g[f"smear_{name}_sums"])) < 1e-11


def test_smear_against_oracle(g):
    n, bc, a, ns = (16, 16, 16, 32), (1, -1, 1, 1), 0.2, 5
    print(f"[smear] {n} against the oracle", flush=True)
    lattice.init(n, bc)
    try:
        U = gauge.haar(n, 8)
        rng = np.random.default_rng(9)
        V = int(np.prod(n))
        psi = rng.standard_normal((V, 4, 3)) + 1j * rng.standard_normal((V, 4, 3))
        ge, go = O.spinor_canon_to_ref(n, psi)
        he, ho = np.zeros_like(ge), np.zeros_like(go)
        diracops.smear(he, ho, O.links_canon_to_ref(n, U), ge, go, a, ns, bc)
        assert rel(O.spinor_ref_to_canon(n, he, ho), O.smear(n, U, bc, psi, a, ns)) < 1e-13
    finally:
        lattice.init(N, (1, 1, 1, -1))


@pytest.mark.parametrize("idag", [0, 1])
def test_twmshift_gamma5mult_reference(g, idag):
    print(f"[smear] twMshift / gamma5mult idag={idag}", flush=True)
    U, psi = smear_inputs(N)
    u = O.links_canon_to_ref(N, U)
    ge, go = O.spinor_canon_to_ref(N, psi)
    co = O.coact_array()
    gw = np.concatenate([ge, go])
    hw = np.zeros_like(gw)
    gd = np.zeros((18, 1, 1, 1, 1))
    diracops.twMshift(hw, u, gd, gw, idag, co, SHIFT_KAPPA, -1, (1, 1, 1, -1))
    half = ge.size
    sl, nrm, sm = slice_summary(N, O.spinor_ref_to_canon(N, hw[:half], hw[half:]))
    assert rel(sl, g[f"twm{idag}_slices"]) < 1e-13 and rel(nrm, g[f"twm{idag}_norms"]) < 1e-13
    he = np.zeros_like(ge)
    diracops.gamma5mult(he, u, gd, ge, idag, co, SHIFT_KAPPA, -1, (1, 1, 1, -1))
    sl, nrm, sm = slice_summary(N, O.spinor_ref_to_canon(N, he, None))
    assert rel(sl, g[f"g5m{idag}_slices"]) < 1e-13 and rel(nrm, g[f"g5m{idag}_norms"]) < 1e-13


def test_shift_operators_reject_clover(g):
    U, psi = smear_inputs(N)
    u = O.links_canon_to_ref(N, U)
    ge, _ = O.spinor_canon_to_ref(N, psi)
    with pytest.raises(Exception):
        diracops.gamma5mult(np.zeros_like(ge), u, None, ge, 0, O.coact_array(), SHIFT_KAPPA, -2, (1, 1, 1, -1))
