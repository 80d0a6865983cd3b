"""The reference's gmresproject (called with its whole argument list) and
fermpropR (ir = -1) outputs, pinned by the oracle's operators on the CPU
(golden_boundary_8x8x8x16.npz, written by tests/golden/make_golden.py
boundary from the reference compiled from source).

* fermpropR with ir = -1 negates coact(3,1,:) before fermprop
  (qqcd-test/cfgsprops/quark/quark.f90:921-926): its solution must solve the
  full system with c1 = -1 and must not solve the c1 = +1 one.
* gmresproject (qqcd-test/cfgsprops/quark/inverters.f90:7183-8304) zeroes
  every shift of xshift and solves into the first; gdr(mvp, 1) holds the
  least-squares residual after each Arnoldi step, 0 at the true-residual
  operator applications.
"""
import os

import numpy as np
import pytest

import oracle as O
from baylorlqcd_amd import gauge

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
N, BC = (8, 8, 8, 16), (1, 1, 1, -1)


@pytest.fixture(scope="module")
def gb():
    return np.load(os.path.join(GOLD, "golden_boundary_8x8x8x16.npz"))


@pytest.fixture(scope="module")
def U():
    g = np.load(os.path.join(GOLD, "golden_8x8x8x16.npz"))
    return gauge.reconstruct12(g["links12"])


def test_fermpropR_solves_the_r_minus_one_system(gb, U):
    kappa = (float(gb["case"][0]), 0.0)
    V = int(np.prod(N))
    b = np.zeros((V, 4, 3), dtype=complex)
    b[0, 0, 0] = 1.0
    odd = O.parity_mask(N)
    x = np.zeros_like(b)
    x[~odd] = gb["fermpropR_xe"]
    x[odd] = gb["fermpropR_xo"]
    res_r = np.linalg.norm(b - O.full_m(N, U, BC, [-1.0] * 4, [1.0] * 4, kappa, x)) / np.linalg.norm(b)
    res_w = np.linalg.norm(b - O.full_m(N, U, BC, [1.0] * 4, [1.0] * 4, kappa, x)) / np.linalg.norm(b)
    assert res_r < 1e-9
    assert res_w > 1e-2


def test_gmresproject_golden(gb, U):
    kappa = (float(gb["case"][0]), 0.0)
    odd = O.parity_mask(N)
    V = int(np.prod(N))
    phi = np.zeros((V, 4, 3), dtype=complex)
    x = np.zeros_like(phi)
    phi[~odd] = gb["proj_phi_e"]
    x[~odd] = gb["proj_x"]
    r = phi - O.schur(N, U, BC, [1] * 4, [1] * 4, 0, kappa, x)
    nphi = np.linalg.norm(phi[~odd])
    assert np.linalg.norm(r[~odd]) / nphi < 1e-9
    assert float(gb["proj_x_shift2_absmax"]) == 0.0
    gdr = gb["proj_gdr"]
    mvp, it = int(gb["proj_mvp"]), int(gb["proj_itercount"])
    assert gdr.shape == (mvp,)
    # one nonzero entry per Arnoldi step; zeros at the true-residual products
    assert np.count_nonzero(gdr) == it and gdr[0] == 0 and gdr[-1] == 0
    assert 0 < gdr[gdr > 0].min() < 1e-10 * nphi
