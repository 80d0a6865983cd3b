"""GPU checks of the learned preconditioner.

* forward layer, adjoint layer and weight gradient of the device kernels
  (BF16 matrix cores, FP32 accumulation) against the float64 CPU oracle
  (oracle/nn.py, itself pinned by finite differences and gauge covariance
  in tests/test_oracle_nn.py); bar: relative Frobenius error below 1.5e-2,
  the BF16 rounding of the operands;
* training lowers the loss on harvested (residual, error) pairs;
* the trained network (four 16^3x32 training configurations) as the
  flexible right preconditioner reaches the 1e-10 true residual on a
  held-out configuration and needs fewer than 0.8x the Krylov steps of the
  same mixed solver without it.
"""
import numpy as np
import pytest

import oracle as O
import oracle.nn as NN
from baylorlqcd_amd import diracops, gauge, lattice, learned

pytestmark = pytest.mark.gpu

TOL = 1.5e-2


def frob(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def field(n, C, seed):
    rng = np.random.default_rng(seed)
    V = int(np.prod(n))
    return rng.standard_normal((C, V, 4, 3)) + 1j * rng.standard_normal((C, V, 4, 3))


@pytest.fixture(scope="module", params=[((4, 4, 4, 4), (1, 1, 1, -1)), ((8, 4, 4, 8), (-1, 1, 0, -1))])
def lat(request):
    n, bc = request.param
    print(f"[nn] lattice {n} bc {bc}", flush=True)
    lattice.init(n, bc)
    from oracle import coact_array, links_canon_to_ref
    U = gauge.haar(n, 5)
    diracops.gauge_load(links_canon_to_ref(n, U), coact_array())
    Ws = NN.init_weights((1, 8, 8, 1), 3)
    learned.set_weights(Ws)
    yield n, bc, U, Ws
    lattice.finalize()


@pytest.mark.parametrize("layer", [0, 1, 2])
def test_layer_forward(lat, layer):
    n, bc, U, Ws = lat
    W = Ws[layer]
    psi = field(n, W.shape[1], 10 + layer)
    print(f"[nn] forward layer {layer} {W.shape[:2]}", flush=True)
    got = learned.layer_canonical(layer, psi)
    want = NN.layer(n, U, bc, W, psi)
    assert frob(got, want) < TOL


@pytest.mark.parametrize("layer", [0, 1, 2])
def test_layer_adjoint(lat, layer):
    n, bc, U, Ws = lat
    W = Ws[layer]
    g = field(n, W.shape[0], 20 + layer)
    print(f"[nn] adjoint layer {layer}", flush=True)
    got = learned.layer_canonical(layer, g, adjoint=1)
    want = NN.layer_adjoint(n, U, bc, W, g)
    assert frob(got, want) < TOL


@pytest.mark.parametrize("layer", [0, 1, 2])
def test_weight_gradient(lat, layer):
    n, bc, U, Ws = lat
    W = Ws[layer]
    psi = field(n, W.shape[1], 30 + layer)
    g = field(n, W.shape[0], 40 + layer)
    print(f"[nn] weight gradient layer {layer}", flush=True)
    got = learned.wgrad_canonical(layer, psi, g)
    want = NN.layer_wgrad(n, U, bc, psi, g)
    assert frob(got, want) < TOL


def test_network_gradient_matches_oracle(lat):
    # the chain the trainer runs: forward, output gradient, adjoint layers and
    # weight gradients, layer by layer against the oracle's loss_and_grads
    n, bc, U, Ws = lat
    r, e = field(n, 1, 50)[0], field(n, 1, 51)[0]
    L, grads, _ = NN.loss_and_grads(n, U, bc, Ws, r, e)
    acts = [r[None]]
    for l in range(len(Ws)):
        acts.append(learned.layer_canonical(l, acts[-1]))
    g = (acts[-1][0] - e)[None]
    for l in range(len(Ws) - 1, -1, -1):
        dW = learned.wgrad_canonical(l, acts[l], g)
        assert frob(dW, grads[l]) < 3 * TOL, l
        if l > 0:
            g = learned.layer_canonical(l, g, adjoint=1)


def bf16(a):
    """Round to the nearest BF16 value (ties to even), real and imaginary parts."""
    def r(x):
        u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
        return u.astype(np.uint32).view(np.float32).astype(np.float64)
    return r(a.real) + 1j * r(a.imag)


def test_apply_matches_layer_chain(lat):
    # the solver's application skips the transports out of the (zero) odd
    # sites in the first layer and forms only the even sites in the last one;
    # on BF16-exact input it must give the dense layer chain's even sites exactly
    n, bc, U, Ws = lat
    odd = O.parity_mask(n)
    psi = bf16(field(n, 1, 60)[0])
    psi[odd] = 0
    print("[nn] sparse application vs dense layer chain", flush=True)
    got = learned.apply_canonical(psi)
    acts = psi[None]
    for l in range(len(Ws)):
        acts = learned.layer_canonical(l, acts)
    assert frob(got[~odd], acts[0][~odd]) < 1e-6
    # and the network itself against the oracle
    want = NN.forward(n, U, bc, Ws, psi)[-1][0]
    assert frob(got[~odd], want[~odd]) < 3 * TOL


@pytest.fixture(scope="module")
def trained():
    n, bc = (16, 16, 16, 32), (1, 1, 1, -1)
    print("[nn] training on 16^3x32", flush=True)
    lattice.init(n, bc)
    out = learned.train(nconf=4, beta=6.0, nsweep=200, kappa=(0.150, 0.157), nsrc=1, m=40, k=16,
                        harvest_tol=1e-8, max_pairs=128, channels=(1, 8, 8, 8, 1), steps=300, batch=4, lr=1e-3,
                        seed=7)
    print("[nn] train", {k: v for k, v in out.items() if k != "loss"}, "loss[0], loss[-1]", out["loss"][0],
          out["loss"][-1], flush=True)
    yield n, bc, out
    lattice.finalize()


def test_training_lowers_loss(trained):
    n, bc, out = trained
    assert out["pairs"] > 8
    assert out["loss_after"] < 0.8 * out["loss_before"]


def test_learned_solve_gains(trained):
    n, bc, out = trained
    # held-out: a new heatbath configuration, point source, near-critical kappa
    co = np.zeros(64)
    for mu in range(4):
        co[2 + 4 * (0 + 4 * mu)] = 1.0
        co[2 + 4 * (1 + 4 * mu)] = 1.0
    diracops.gauge_heatbath(6.0, 200, 4242, 1, co)
    kappa = (0.1565, 0.0)
    plain = learned.bench(1, kappa, GMRES=(40, 16))
    nn = learned.bench(2, kappa, GMRES=(40, 16))
    print("[nn] mixed without network", plain, flush=True)
    print("[nn] mixed with network", nn, flush=True)
    assert plain["true_residual"] < 1e-10 and nn["true_residual"] < 1e-10
    assert nn["nn_steps"] > 0
    # measured on this field: about half the Krylov steps
    assert nn["iterations"] < 0.8 * plain["iterations"]
