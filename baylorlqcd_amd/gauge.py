"""Seeded SU(3) gauge fields for the solve path (host side, numpy).

Fields are complex128 arrays of shape (V, 4, 3, 3) in canonical site order
s = x + nx*(y + ny*(z + nz*t)); U[s, mu] is the link from s to s + mu.
The 12-real form keeps rows 0 and 1; row 2 is conj(row0 x row1), the
convention of the reference's GFwrite/fixsu3
(qqcd-test/cfgsprops/glue/gaugetools.f90:1280-1290, 1338-1432).

Generators:
  * haar(n, seed)            -- Haar-random SU(3): complex Gaussian 3x3, QR,
                                phase-fixed so R has a positive diagonal,
                                then divided by det^(1/3);
  * near_identity(n, eps, s) -- reunitarised 1 + eps * Gaussian;
  * heatbath fields come from the GPU heatbath (``qqcd_heatbath``) or the
    reference's own sweep for fixtures.
"""
from __future__ import annotations

import numpy as np


def volume(n):
    return int(np.prod(n))


def _su3_from_gauss(z):
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=-2, axis2=-1)
    ph = d / np.abs(d)
    q = q * ph[..., None, :]
    det = np.linalg.det(q)
    return q / (det ** (1.0 / 3.0))[..., None, None]


def haar(n, seed):
    rng = np.random.default_rng(seed)
    V = volume(n)
    z = rng.standard_normal((V, 4, 3, 3)) + 1j * rng.standard_normal((V, 4, 3, 3))
    return _su3_from_gauss(z / np.sqrt(2.0))


def reunitarise(m):
    """Project 3x3 matrices back to SU(3) by Gram-Schmidt on rows 0, 1 and the
    cross-product third row (the fixsu3 construction)."""
    r0 = m[..., 0, :]
    r0 = r0 / np.linalg.norm(r0, axis=-1, keepdims=True)
    r1 = m[..., 1, :]
    r1 = r1 - np.sum(np.conj(r0) * r1, axis=-1, keepdims=True) * r0
    r1 = r1 / np.linalg.norm(r1, axis=-1, keepdims=True)
    r2 = np.conj(np.cross(r0, r1))
    return np.stack([r0, r1, r2], axis=-2)


def near_identity(n, eps, seed):
    rng = np.random.default_rng(seed)
    V = volume(n)
    z = rng.standard_normal((V, 4, 3, 3)) + 1j * rng.standard_normal((V, 4, 3, 3))
    return reunitarise(np.eye(3) + eps * z)


def compress12(U):
    return np.ascontiguousarray(U[..., :2, :])


def reconstruct12(U12):
    r0 = U12[..., 0, :]
    r1 = U12[..., 1, :]
    return np.stack([r0, r1, np.conj(np.cross(r0, r1))], axis=-2)


def plaquette(n, U):
    """Average (1/3) Re Tr of the 6V elementary plaquettes."""
    n = tuple(int(v) for v in n)
    Ul = U.reshape(n[3], n[2], n[1], n[0], 4, 3, 3)
    axis = {0: 3, 1: 2, 2: 1, 3: 0}
    tot = 0.0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            a = Ul[..., mu, :, :]
            b = np.roll(Ul[..., nu, :, :], -1, axis=axis[mu])
            c = np.roll(Ul[..., mu, :, :], -1, axis=axis[nu])
            d = Ul[..., nu, :, :]
            p = a @ b @ np.conj(np.swapaxes(c, -1, -2)) @ np.conj(np.swapaxes(d, -1, -2))
            tot += np.trace(p, axis1=-2, axis2=-1).real.sum()
    return tot / (6 * volume(n) * 3)
