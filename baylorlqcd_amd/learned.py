"""The learned preconditioner: a gauge-equivariant network of
parallel-transport convolution layers (no counterpart in the reference).

Weights are complex arrays W[l] of shape (cout, cin, 9, 4, 4): the spin
matrix of output channel a, input channel b and path p (0, +x, -x, +y, -y,
+z, -z, +t, -t).  See include/qqcd_hip.h for the layer definition.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

CHANNELS = (1, 8, 8, 8, 1)


def _ch(channels):
    return _lib.ints(channels)


def init(channels=CHANNELS, scale=0.0245, noise=0.02, seed=1):
    ch = _ch(channels)
    _lib.check(_lib.lib().qqcd_nn_init(len(channels) - 1, ch.ctypes.data_as(_lib._ip), float(scale), float(noise),
                                       int(seed)))


def set_weights(Ws):
    channels = [Ws[0].shape[1]] + [W.shape[0] for W in Ws]
    flat = np.concatenate([np.ascontiguousarray(W, dtype=np.complex128).ravel() for W in Ws])
    ch = _ch(channels)
    _lib.check(_lib.lib().qqcd_nn_set_weights(len(Ws), ch.ctypes.data_as(_lib._ip), _lib.dp(flat.view(np.float64))))


def get_weights():
    nl = ctypes.c_int()
    ch = np.zeros(17, np.int32)
    _lib.check(_lib.lib().qqcd_nn_get_weights(ctypes.byref(nl), ch.ctypes.data_as(_lib._ip), None))
    n = nl.value
    sizes = [int(ch[l + 1] * ch[l] * 9 * 16) for l in range(n)]
    flat = np.zeros(2 * sum(sizes))
    _lib.check(_lib.lib().qqcd_nn_get_weights(ctypes.byref(nl), ch.ctypes.data_as(_lib._ip), _lib.dp(flat)))
    z = flat.view(np.complex128)
    out, off = [], 0
    for l in range(n):
        out.append(z[off:off + sizes[l]].reshape(ch[l + 1], ch[l], 9, 4, 4).copy())
        off += sizes[l]
    return out


def layer_canonical(layer, psi, adjoint=0):
    """psi: (C, V, 4, 3) complex -> layer (or adjoint) output (C', V, 4, 3)."""
    W = get_weights()[layer]
    co = W.shape[1] if adjoint else W.shape[0]
    out = np.zeros((co,) + psi.shape[1:], dtype=np.complex128)
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    _lib.check(_lib.lib().qqcd_nn_layer_canonical(int(layer), int(adjoint), _lib.dp(psi.view(np.float64)),
                                                  _lib.dp(out.view(np.float64))))
    return out


def wgrad_canonical(layer, psi, g):
    W = get_weights()[layer]
    dW = np.zeros(W.shape, dtype=np.complex128)
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    g = np.ascontiguousarray(g, dtype=np.complex128)
    _lib.check(_lib.lib().qqcd_nn_wgrad_canonical(int(layer), _lib.dp(psi.view(np.float64)),
                                                  _lib.dp(g.view(np.float64)), _lib.dp(dW.view(np.float64))))
    return dW


def apply_canonical(r):
    r = np.ascontiguousarray(r, dtype=np.complex128)
    z = np.zeros_like(r)
    _lib.check(_lib.lib().qqcd_nn_apply_canonical(_lib.dp(r.view(np.float64)), _lib.dp(z.view(np.float64))))
    return z


def train(nconf=4, beta=6.0, nsweep=200, kappa=(0.150, 0.157), nsrc=1, m=40, k=16, harvest_tol=1e-8,
          max_pairs=256, channels=CHANNELS, keep=0, steps=200, batch=4, lr=1e-3, seed=1):
    loss = np.zeros(max(steps, 1))
    stats = np.zeros(7)
    ch = _ch(channels)
    _lib.check(_lib.lib().qqcd_nn_train(
        int(nconf), float(beta), int(nsweep), float(kappa[0]), float(kappa[1]), int(nsrc), int(m), int(k),
        float(harvest_tol), int(max_pairs), len(channels) - 1, ch.ctypes.data_as(_lib._ip), int(keep), int(steps),
        int(batch), float(lr), int(seed), _lib.dp(loss), _lib.dp(stats)))
    return {"loss": loss[:steps], "pairs": int(stats[0]), "harvest_s": stats[1], "train_s": stats[2],
            "loss_before": stats[3], "loss_after": stats[4],
            "validation_start": _validation(stats[5]), "validation_kept": _validation(stats[6])}


def _validation(v):
    """Decode a validation solve count: Krylov steps, plus 1,000,000 when it did not converge."""
    v = int(v)
    return {"krylov_steps": v % 1000000, "converged": v < 1000000}


def bench(method, kappa, GMRES=(40, 16), tol=1e-10, op_prec=1, op_recon=12, src=(1, 1, 1, 1), spin=1, colour=1):
    sec, res = np.zeros(1), np.zeros(1)
    it, outer, nns, drop = (np.zeros(1, np.int32) for _ in range(4))
    _lib.check(_lib.lib().qqcd_bench_learned(
        int(method), _lib.ints(src).ctypes.data_as(_lib._ip), int(spin), int(colour),
        _lib.dp(np.asarray(kappa, dtype=np.float64)), _lib.ints(GMRES).ctypes.data_as(_lib._ip), float(tol),
        int(op_prec), int(op_recon), _lib.dp(sec), it.ctypes.data_as(_lib._ip), outer.ctypes.data_as(_lib._ip),
        _lib.dp(res), nns.ctypes.data_as(_lib._ip), drop.ctypes.data_as(_lib._ip)))
    return {"seconds": float(sec[0]), "iterations": int(it[0]), "outer": int(outer[0]),
            "true_residual": float(res[0]), "nn_steps": int(nns[0]), "dropped": int(drop[0])}


def bench_nn(nwarm=2, niter=10):
    ms = ctypes.c_double()
    _lib.check(_lib.lib().qqcd_bench_nn(int(nwarm), int(niter), ctypes.byref(ms)))
    return ms.value
