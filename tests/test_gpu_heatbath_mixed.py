"""GPU heatbath and the mixed-precision solver.

Heatbath: links stay in SU(3) to rounding; the device plaquette equals the
numpy plaquette of the downloaded field to rounding; the equilibrium
plaquette at beta = 6.0 on 8^3x16 agrees with the reference heatbath's
(0.5935 after 200 sweeps, tests/golden) within 0.004 (the run-to-run spread
of an 8^3x16 plaquette is ~0.001); at small beta it follows the strong
coupling series <P> = beta/18 + beta^2/216 + O(beta^3)... to 2%.

Mixed precision (FP32 Krylov basis; FP32 or FP16 storage operator; FP64
reliable updates): the config-1 solution agrees with the reference's
GMRES-DR solution to 1e-8 and the true full residual is below 1e-10.
"""
import os

import numpy as np
import pytest

import oracle as O
from baylorlqcd_amd import _lib, diracops, gauge, inverters, lattice

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_heatbath_beta6():
    n = (8, 8, 8, 16)
    g = np.load(os.path.join(GOLD, "golden_8x8x8x16.npz"))
    lattice.init(n, (1, 1, 1, -1))
    try:
        co = O.coact_array()
        diracops.gauge_heatbath(6.0, 150, 1234, 1, co)
        plaqs = []
        for _ in range(5):
            diracops.gauge_heatbath(6.0, 10, 1234, -1, co)
            plaqs.append(diracops.gauge_plaquette())
        U = diracops.gauge_get_canonical(n)
        assert np.max(np.abs(U @ U.conj().swapaxes(-1, -2) - np.eye(3))) < 1e-13
        assert np.max(np.abs(np.linalg.det(U) - 1)) < 1e-13
        assert abs(gauge.plaquette(n, U) - plaqs[-1]) < 1e-12
        mean = float(np.mean(plaqs))
        print(f"heatbath beta=6.0 8^3x16: <P> = {mean:.4f} (reference heatbath {float(g['plaquette_ref']):.4f})")
        assert abs(mean - float(g["plaquette_ref"])) < 0.004
    finally:
        lattice.finalize()


def test_heatbath_strong_coupling():
    n = (8, 8, 8, 8)
    lattice.init(n, (1, 1, 1, 1))
    try:
        co = O.coact_array()
        beta = 0.6
        diracops.gauge_heatbath(beta, 40, 99, 1, co)
        p = []
        for _ in range(10):
            diracops.gauge_heatbath(beta, 4, 99, -1, co)
            p.append(diracops.gauge_plaquette())
        want = beta / 18 + beta ** 2 / 216
        print(f"heatbath beta={beta}: <P> = {np.mean(p):.5f}, strong coupling {want:.5f}")
        assert abs(np.mean(p) - want) < 0.02 * want + 0.002
    finally:
        lattice.finalize()


@pytest.mark.parametrize("basis,op_prec,recon", [(1, 1, 12), (1, 1, 18), (1, 2, 12), (2, 2, 12), (2, 1, 18)])
def test_mixed_precision_config1(basis, op_prec, recon):
    # basis 2: the Krylov basis and deflation vectors stored as FP16 fields
    from oracle import solvers
    n, bc = (8, 8, 8, 16), (1, 1, 1, -1)
    g = np.load(os.path.join(GOLD, "golden_8x8x8x16.npz"))
    lattice.init(n, bc)
    try:
        U = gauge.reconstruct12(g["links12"])
        u = O.links_canon_to_ref(n, U)
        co = O.coact_array()
        diracops.gauge_load(u, co)
        kap = (0.150, 0.0)
        V = int(np.prod(n))
        b = np.zeros((V, 4, 3), complex)
        b[0, 0, 0] = 1.0
        odd = O.parity_mask(n)
        src = np.zeros_like(b)
        src[~odd] = solvers.source_prep(n, U, bc, kap, b).reshape(-1, 4, 3)
        phi, _ = O.spinor_canon_to_ref(n, src)
        x = np.zeros_like(phi)
        it = np.zeros(1, np.int32)
        _lib.check(_lib.lib().qqcd_solve_mixed(
            _lib.dp(phi), _lib.dp(x), _lib.ints((40, 10)).ctypes.data_as(_lib._ip), 1e-11, basis, op_prec, recon, 0.0,
            0, _lib.dp(u), _lib.dp(co), _lib.dp(np.array(kap)), _lib.ints(bc).ctypes.data_as(_lib._ip),
            it.ctypes.data_as(_lib._ip)))
        st = inverters.last_solve_stats()
        xe = O.spinor_ref_to_canon(n, x, None)[~odd]
        xf = solvers.reconstruct(n, U, bc, kap, b, xe.ravel())
        r = b - O.full_m(n, U, bc, [1] * 4, [1] * 4, kap, xf)
        res = np.linalg.norm(r) / np.linalg.norm(b)
        print(f"mixed basis={basis} op=({op_prec},{recon}): {int(it[0])} inner iterations, {st['cycles']} outer, "
              f"residual {res:.2e}")
        assert res < 1e-10
        assert float(np.max(np.abs(xe - g["gmresdr_k150_xe"])) / np.max(np.abs(g["gmresdr_k150_xe"]))) < 1e-8
    finally:
        lattice.finalize()


def test_bench_solve_entry_point():
    """Device-resident timing entry point: every method reaches 1e-10 on a
    16^3x32 heatbath field (beta 6, twisted kappa 0.150, mu 0.01)."""
    n = (16, 16, 16, 32)
    lattice.init(n, (1, 1, 1, -1))
    try:
        co = O.coact_array()
        diracops.gauge_heatbath(6.0, 100, 5, 1, co)
        for method, prec in ((0, 0), (1, 1), (1, 2), (2, 2)):
            sec, res = np.zeros(1), np.zeros(1)
            its, outer = np.zeros(1, np.int32), np.zeros(1, np.int32)
            _lib.check(_lib.lib().qqcd_bench_solve(
                method, _lib.ints((1, 1, 1, 1)).ctypes.data_as(_lib._ip), 1, 1, _lib.dp(np.array([0.150, 0.01])),
                _lib.ints((40, 16)).ctypes.data_as(_lib._ip), 1e-10, prec, 12, 0.0, 0, _lib.dp(sec),
                its.ctypes.data_as(_lib._ip), outer.ctypes.data_as(_lib._ip), _lib.dp(res)))
            print(f"bench_solve method {method} op_prec {prec}: {sec[0]:.3f}s, {its[0]} its, {outer[0]} outer, "
                  f"residual {res[0]:.2e}")
            assert res[0] < 1e-10
    finally:
        lattice.finalize()
