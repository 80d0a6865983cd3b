"""The oracle's pseudo-heatbath sweep (oracle/heatbath.py) against the
reference's own sweeps (golden_sweep_8x8x8x16.npz, written by
tests/golden/make_golden.py sweep through qqcd-test/cfgsprops/cfgsprops.f90's
sweep loop, staple and phbup): the stream state after each sweep must be
identical (every Kennedy-Pendleton accept decision, hence every draw, the
same), the links equal to rounding of cos / sin / log (1e-13).  An
anisotropic coact with coact(1,mu,nu) != coact(1,nu,mu) pins the index
order; links stay unitary."""
import os

import numpy as np
import pytest

import oracle as O
from oracle import heatbath as HB
from oracle import lagfib as LF
from baylorlqcd_amd import gauge

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N = (8, 8, 8, 16)
CASES = [("iso", 3, 2), ("aniso", 4, 1), ("rect", 8, 1), ("mixed", 9, 1)]


@pytest.fixture(scope="module")
def gold():
    g = np.load(os.path.join(ROOT, "tests", "golden", "golden_sweep_8x8x8x16.npz"))
    c = np.load(os.path.join(ROOT, "tests", "golden", "golden_8x8x8x16.npz"))
    return g, O.links_canon_to_ref(N, gauge.reconstruct12(c["links12"]))


def summary(u):
    nvh = 256
    U = u.reshape(16, 2, 4, 320, 18)[:, :, :, :nvh]
    return U.sum(axis=3), U[0]


@pytest.mark.parametrize("name,seed,nsw", CASES)
def test_sweeps_match_reference(gold, name, seed, nsw):
    g, u = gold
    st = LF.init(0, seed)
    co = g[f"{name}_coact"]
    ga = tuple(int(v) for v in g[f"{name}_gaction"])
    for k in range(1, nsw + 1):
        u, st = HB.sweep(N, u, co, st, 1, gaction=ga)
        assert np.array_equal(st[0], g[f"{name}_{k}_iw"]) and tuple(st[1:]) == tuple(g[f"{name}_{k}_jrkr"])
        sums, blk = summary(u)
        assert np.max(np.abs(sums - g[f"{name}_{k}_sums"])) < 1e-11
        assert np.max(np.abs(blk - g[f"{name}_{k}_block1"])) < 1e-12
    U = O.links_ref_to_canon(N, u)
    eye = np.broadcast_to(np.eye(3), U.shape)
    assert np.allclose(U @ np.conj(np.swapaxes(U, -1, -2)), eye, atol=1e-12)


def test_coact_order_matters(gold):
    # the transposed anisotropic coact gives a different field: the check above pins mu, nu
    g, u = gold
    co = g["aniso_coact"].reshape(4, 4, 4, order="F")
    cot = np.asfortranarray(np.swapaxes(co, 1, 2)).ravel(order="F")
    ut, _ = HB.sweep(N, u, cot, LF.init(0, 4), 1)
    sums, _ = summary(ut)
    assert np.max(np.abs(sums - g["aniso_1_sums"])) > 1e-3


def test_staples_bit_identical(gold):
    import hashlib
    g, _ = gold
    n = N
    U = O.links_canon_to_ref(n, gauge.haar(n, 4)).reshape(16, 2, 4, 320, 18)
    co = g["rect_coact"].reshape(4, 4, 4, order="F")
    parts = []
    for ieo in range(2):
        for ibl in range(16):
            x = HB._sites(n, ieo, ibl)
            for mu in range(4):
                c = HB._staple(n, U, x, mu, co, np.array(n), [2, 2, 2, 2])
                parts.append(np.ascontiguousarray(c).ravel())  # (nvhalf, 18) C order = c(18, nvhalf) Fortran
    assert hashlib.sha256(np.concatenate(parts).tobytes()).hexdigest() == str(g["staples_rect_sha256"])
