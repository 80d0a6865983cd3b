"""CPU oracle of the gauge-equivariant preconditioner network.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): float64 numpy, used by
tests/ to check the device kernels; never called by the product.

The network (no reference counterpart: the reference has no learned
preconditioner) is a stack of linear parallel-transport convolution layers
over the paths P = {0, +x, -x, +y, -y, +z, -z, +t, -t}:

    out_a(s) = sum_{b, p} W[a, b, p] . (T_p psi_b)(s)

with W[a, b, p] a complex 4x4 matrix acting on spin (identity on colour, so
that the layer commutes with gauge transformations), and the transports

    (T_0 psi)(s)   = psi(s)
    (T_+mu psi)(s) = U_mu(s) psi(s + mu)          (* bc[mu] across the edge)
    (T_-mu psi)(s) = U_mu(s - mu)^dag psi(s - mu) (* bc[mu] across the edge)

Path index p = 0 for the identity, 1 + 2 mu for +mu and 2 + 2 mu for -mu.
Fields are canonical: psi[C, V, 4, 3] complex with s = x + nx (y + ny (z +
nz t)), links U[V, 4, 3, 3].

Gradients follow the Wirtinger convention of a real loss L: g = dL/dconj(.).
For L = ||N(r) - e||^2 the output gradient is N(r) - e; the input gradient of
a layer is the adjoint layer (reversed paths, W^dag), and the weight gradient
is dW[a, b, p] = sum_{s, colour} g_a(s) (T_p psi_b)(s)^dag; the gradient with
respect to Re W and Im W is (2 Re dW, 2 Im dW).
"""
import numpy as np

NPATH = 9


def _shift(n, mu, sign):
    """Index map s -> s + sign*mu and the edge mask (True where the hop
    crosses the lattice boundary)."""
    V = int(np.prod(n))
    s = np.arange(V)
    x = [s % n[0], (s // n[0]) % n[1], (s // (n[0] * n[1])) % n[2], s // (n[0] * n[1] * n[2])]
    y = list(x)
    y[mu] = (x[mu] + sign) % n[mu]
    edge = (x[mu] == n[mu] - 1) if sign > 0 else (x[mu] == 0)
    t = y[0] + n[0] * (y[1] + n[1] * (y[2] + n[2] * y[3]))
    return t, edge


def transport(n, U, bc, p, psi):
    """T_p psi for one channel psi[V, 4, 3]."""
    if p == 0:
        return psi.copy()
    mu, back = (p - 1) // 2, (p - 1) % 2
    if not back:
        t, edge = _shift(n, mu, +1)
        out = np.einsum("sij,sdj->sdi", U[:, mu], psi[t])
    else:
        t, edge = _shift(n, mu, -1)
        out = np.einsum("sji,sdj->sdi", np.conj(U[t, mu]), psi[t])
    f = np.where(edge, float(bc[mu]), 1.0)
    return out * f[:, None, None]


def reverse(p):
    return 0 if p == 0 else (p + 1 if p % 2 == 1 else p - 1)


def layer(n, U, bc, W, psi):
    """W[Cout, Cin, 9, 4, 4], psi[Cin, V, 4, 3] -> [Cout, V, 4, 3]."""
    cout, cin = W.shape[:2]
    V = psi.shape[1]
    out = np.zeros((cout, V, 4, 3), dtype=np.complex128)
    for b in range(cin):
        for p in range(NPATH):
            T = transport(n, U, bc, p, psi[b])
            out += np.einsum("aij,sjc->asic", W[:, b, p], T)
    return out


def layer_adjoint(n, U, bc, W, g):
    """The adjoint of layer(): g[Cout, V, 4, 3] -> [Cin, V, 4, 3]."""
    cout, cin = W.shape[:2]
    V = g.shape[1]
    out = np.zeros((cin, V, 4, 3), dtype=np.complex128)
    for a in range(cout):
        for p in range(NPATH):
            T = transport(n, U, bc, reverse(p), g[a])
            out += np.einsum("bji,sjc->bsic", np.conj(W[a, :, p]), T)
    return out


def layer_wgrad(n, U, bc, psi, g):
    """dL/dconj(W)[a, b, p] = sum_s,c g_a(s) (T_p psi_b)(s)^dag."""
    cin, cout = psi.shape[0], g.shape[0]
    dW = np.zeros((cout, cin, NPATH, 4, 4), dtype=np.complex128)
    for b in range(cin):
        for p in range(NPATH):
            T = transport(n, U, bc, p, psi[b])
            dW[:, b, p] = np.einsum("asic,sjc->aij", g, np.conj(T))
    return dW


def forward(n, U, bc, Ws, r):
    """Network output for one input channel r[V, 4, 3]; returns all activations."""
    acts = [r[None]]
    for W in Ws:
        acts.append(layer(n, U, bc, W, acts[-1]))
    return acts


def loss_and_grads(n, U, bc, Ws, r, e):
    """L = ||N(r) - e||^2 with N's output channel 0; returns (L, [dW per layer], dL/dconj(r))."""
    acts = forward(n, U, bc, Ws, r)
    diff = acts[-1][0] - e
    L = float(np.sum(np.abs(diff) ** 2))
    g = diff[None]
    grads = [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        grads[l] = layer_wgrad(n, U, bc, acts[l], g)
        g = layer_adjoint(n, U, bc, Ws[l], g)
    return L, grads, g[0]


def init_weights(channels, seed, scale=None):
    """Random complex weights for channel counts e.g. (1, 8, 8, 8, 1)."""
    rng = np.random.default_rng(seed)
    Ws = []
    for l in range(len(channels) - 1):
        cin, cout = channels[l], channels[l + 1]
        s = scale if scale is not None else 1.0 / np.sqrt(cin * NPATH * 4)
        Ws.append(s * (rng.standard_normal((cout, cin, NPATH, 4, 4)) +
                       1j * rng.standard_normal((cout, cin, NPATH, 4, 4))) / np.sqrt(2))
    return Ws
