"""GPU parity at the workload sizes, against the C oracle on the same inputs.

* 32^3x64 (2.1M sites): the hopping term in all six storage formats (FP64,
  FP32, FP16 with per-site norms; 18- and 12-real links) onto both parities,
  and the twisted-mass Schur operator (Hdbletm, kappa 0.156, mu 0.005), on
  the reunitarised 1 + 0.1 Gaussian field of the device generator.  Bars:
  FP64 1e-13 relative (max norm; 1e-12 with 12-real links, whose third row
  is rebuilt), FP32 2e-6 and FP16 5e-3 relative L2.
* 24^3x48: the multi-RHS hop in its default launch form (two slots per
  lane, one workgroup per tile) for 12 right-hand sides in blocks of 8 (the
  second block half filled) against the oracle per right-hand side, FP32
  with 12-real links (2e-6 relative L2) and FP64 (1e-13 relative, max norm);
  the batched noise loop (8 Z4 sources, the first by
  GMRES-DR(60,20), the rest in one multi-RHS GMRES-Proj) on a beta = 6.0
  heatbath field at kappa 0.1565: every source reaches a true residual below
  1e-10 by the oracle's own full-lattice operator, and the 16 local trace
  estimators agree with the oracle's evaluation on the solutions (1e-8, the
  order of the solve tolerance).  (The reference's loop operators are pinned
  against its own outputs at 8^3x16, tests/test_gpu_loops.py.)
"""
import numpy as np
import pytest

import oracle as O
from baylorlqcd_amd import diracops, disconloops, lattice

pytestmark = pytest.mark.gpu

BC = (1, 1, 1, -1)


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def rel2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def gaussian_spinor(n, seed):
    V = int(np.prod(n))
    return np.random.default_rng(seed).standard_normal(2 * 12 * V).view(np.complex128).reshape(V, 4, 3)


@pytest.fixture(scope="module")
def lat32():
    n = (32, 32, 32, 64)
    print(f"[config sizes] lattice {n}", flush=True)
    lattice.init(n, BC)
    co = O.coact_array()
    diracops.gauge_random(1, 0.1, 2024, co)
    U = diracops.gauge_get_canonical(n)
    psi = gaussian_spinor(n, 12345)
    yield n, U, co, psi
    lattice.finalize()


@pytest.mark.parametrize("par", [1, 0])
def test_dslash_32x32x32x64_six_formats(lat32, par):
    n, U, co, psi = lat32
    want = O.hop(n, U, BC, [1] * 4, [1] * 4, 0, par, psi)
    mask = O.parity_mask(n) == bool(par)
    w = want[mask]
    for prec, recon, tol in ((0, 18, 1e-13), (0, 12, 1e-12), (1, 18, 2e-6), (1, 12, 2e-6), (2, 18, 5e-3),
                             (2, 12, 5e-3)):
        print(f"[config sizes] hop par {par} prec {prec} recon {recon}", flush=True)
        got = diracops.hop_canonical(psi, par, 0, prec, recon)[mask]
        err = relmax(got, w) if prec == 0 else rel2(got, w)
        print(f"[config sizes]   error {err:.2e}", flush=True)
        assert err < tol, (prec, recon, err)
        del got


def test_hdbletm_32x32x32x64_twisted(lat32):
    n, U, co, psi = lat32
    kappa = (0.156, 0.005)
    print("[config sizes] twisted Schur operator", flush=True)
    want = O.schur(n, U, BC, [1] * 4, [1] * 4, 0, kappa, psi)
    got = diracops.schur_canonical(psi, kappa, 0, 0, 18)
    even = ~O.parity_mask(n)
    assert relmax(got[even], want[even]) < 1e-13


def test_hop_multi_24x24x24x48():
    n = (24, 24, 24, 48)
    nrhs = 12
    print(f"[config sizes] multi-RHS hop {n}, {nrhs} right-hand sides", flush=True)
    lattice.init(n, BC)
    try:
        co = O.coact_array()
        diracops.gauge_random(1, 0.1, 2024, co)
        U = diracops.gauge_get_canonical(n)
        V = int(np.prod(n))
        rng = np.random.default_rng(4321)
        psi = (rng.standard_normal((nrhs, V, 4, 3)) + 1j * rng.standard_normal((nrhs, V, 4, 3)))
        odd = O.parity_mask(n)
        for prec, recon in ((1, 12), (0, 18)):
            print(f"[config sizes] multi-RHS hop prec {prec} recon {recon}", flush=True)
            out = np.zeros_like(psi)
            disconloops.hop_multi_canonical(psi, out, 1, prec=prec, recon=recon, nb=8, order=0)
            for r in range(nrhs):
                want = O.hop(n, U, BC, [1] * 4, [1] * 4, 0, 1, psi[r])[odd]
                got = out[r][odd]
                err = rel2(got, want) if prec == 1 else relmax(got, want)
                assert err < (2e-6 if prec == 1 else 1e-13), (prec, r, err)
            del out
    finally:
        lattice.finalize()


def test_noise_loop_24x24x24x48_eight_sources():
    n = (24, 24, 24, 48)
    kappa, nrhs, seed = (0.1565, 0.0), 8, 77
    print(f"[config sizes] noise loop {n}, {nrhs} sources", flush=True)
    lattice.init(n, BC)
    try:
        co = O.coact_array()
        diracops.gauge_heatbath(6.0, 200, 2024, 1, co)
        U = diracops.gauge_get_canonical(n)
        u = diracops.gauge_get_ref(n)
        disconloops.batch_set()
        run = disconloops.noise_loop(nrhs, seed, kappa, (60, 20), tol=1e-10, coact=co, bc=BC)
        print("[config sizes] noise loop", run["times"], run["iters"], flush=True)
        assert np.all(run["res_full"] < 1e-10)
        srcs_e, srcs_o = [], []
        for r in range(nrhs):
            be = lattice.new_spinor(n).ravel(order="F")
            bo = lattice.new_spinor(n).ravel(order="F")
            disconloops.z4source(be, bo, seed, r, 1.0)
            srcs_e.append(be)
            srcs_o.append(bo)
        xe, xo, it, rr = disconloops.fermprop_multi(srcs_e, srcs_o, kappa, co, BC, 1e-10, (60, 20), u)
        want = np.zeros((n[3], 16), dtype=np.complex128)
        for r in range(nrhs):
            eta = O.spinor_ref_to_canon(n, srcs_e[r], srcs_o[r])
            x = O.spinor_ref_to_canon(n, xe[r], xo[r])
            res = eta - O.full_m(n, U, BC, [1] * 4, [1] * 4, kappa, x)
            assert np.linalg.norm(res) / np.linalg.norm(eta) < 1e-10, r
            want += O.trace16(n, eta, x) / nrhs
        assert relmax(run["traces"], want) < 1e-8
    finally:
        lattice.finalize()
