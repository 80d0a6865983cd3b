"""CPU checks of the preconditioner-network oracle (oracle/nn.py).

What pins it: the adjoint identity <g, L psi> = <L^dag g, psi> (fixes the
input gradient, including the reversed paths and W^dag); central finite
differences of the loss in Re W and Im W (fixes the weight gradient and its
factor 2); gauge covariance of a layer under a random local SU(3) rotation
(fixes which link and which end of it each path uses); and the free-field
case, where the +mu transport of a constant field is the constant field.
"""
import numpy as np

import oracle.nn as NN
from baylorlqcd_amd import gauge

N = (4, 4, 4, 4)
BC = (1, 1, 1, -1)


def field(C, seed):
    rng = np.random.default_rng(seed)
    V = int(np.prod(N))
    return rng.standard_normal((C, V, 4, 3)) + 1j * rng.standard_normal((C, V, 4, 3))


def test_transport_free_field():
    V = int(np.prod(N))
    U = np.tile(np.eye(3, dtype=np.complex128), (V, 4, 1, 1))
    psi = np.ones((V, 4, 3), dtype=np.complex128)
    for p in range(1, 9):
        mu = (p - 1) // 2
        T = NN.transport(N, U, (1, 1, 1, 1), p, psi)
        assert np.allclose(T, 1.0)
        T = NN.transport(N, U, BC, p, psi)
        # antiperiodic t: the sites whose hop crosses the edge pick up -1
        if mu == 3:
            assert np.isclose(T.real.min(), -1.0) and np.isclose(T.real.max(), 1.0)


def test_layer_adjoint_identity():
    U = gauge.haar(N, 3)
    W = NN.init_weights((3, 2), 5)[0]
    psi, g = field(3, 1), field(2, 2)
    lhs = np.vdot(g, NN.layer(N, U, BC, W, psi))
    rhs = np.vdot(NN.layer_adjoint(N, U, BC, W, g), psi)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_gauge_covariance():
    U = gauge.haar(N, 4)
    W = NN.init_weights((2, 2), 6)[0]
    psi = field(2, 7)
    V = int(np.prod(N))
    Om = gauge.haar(N, 8)[:, 0]  # one SU(3) matrix per site
    # U'_mu(s) = Om(s) U_mu(s) Om(s+mu)^dag, psi'(s) = Om(s) psi(s)
    Up = np.empty_like(U)
    for mu in range(4):
        t, _ = NN._shift(N, mu, +1)
        Up[:, mu] = np.einsum("sij,sjk,slk->sil", Om, U[:, mu], np.conj(Om[t]))
    psip = np.einsum("sij,bsdj->bsdi", Om, psi)
    out = NN.layer(N, U, BC, W, psi)
    outp = NN.layer(N, Up, BC, W, psip)
    assert np.allclose(outp, np.einsum("sij,asdj->asdi", Om, out), atol=1e-11)
    assert V == Om.shape[0]


def test_weight_gradient_finite_differences():
    U = gauge.haar(N, 9)
    Ws = NN.init_weights((1, 2, 1), 10)
    r, e = field(1, 11)[0], field(1, 12)[0]
    L, grads, gr = NN.loss_and_grads(N, U, BC, Ws, r, e)
    rng = np.random.default_rng(13)
    h = 1e-6
    for l in range(len(Ws)):
        for _ in range(4):
            idx = tuple(rng.integers(0, d) for d in Ws[l].shape)
            for part in (1.0, 1j):
                Wp = [w.copy() for w in Ws]
                Wm = [w.copy() for w in Ws]
                Wp[l][idx] += h * part
                Wm[l][idx] -= h * part
                Lp = NN.loss_and_grads(N, U, BC, Wp, r, e)[0]
                Lm = NN.loss_and_grads(N, U, BC, Wm, r, e)[0]
                fd = (Lp - Lm) / (2 * h)
                an = 2 * (grads[l][idx].real if part == 1.0 else grads[l][idx].imag)
                assert abs(fd - an) < 1e-6 * max(1.0, abs(an)), (l, idx, part, fd, an)
    # input gradient: dL/dRe r
    for _ in range(4):
        idx = tuple(rng.integers(0, d) for d in r.shape)
        rp, rm = r.copy(), r.copy()
        rp[idx] += h
        rm[idx] -= h
        fd = (NN.loss_and_grads(N, U, BC, Ws, rp, e)[0] - NN.loss_and_grads(N, U, BC, Ws, rm, e)[0]) / (2 * h)
        assert abs(fd - 2 * gr[idx].real) < 1e-6 * max(1.0, abs(fd))
