"""Test infrastructure only (never imported by the product): a numpy
restatement of the reference's pseudo-heatbath sweep with its random stream.

  sweep            qqcd-test/cfgsprops/cfgsprops.f90:1973-2003
  staple           qqcd-test/cfgsprops/glue/gaugetools.f90:39-345 (plaquette terms, gaction = 1)
  mm, mmd, mdm     qqcd-test/cfgsprops/glue/gaugetools.f90:741-817, 908-984, 1074-1151
  phbup            qqcd-test/cfgsprops/glue/heatbath.f90:33-76
  phbsub1-3        qqcd-test/cfgsprops/glue/heatbath.f90:80-235
  su2up            qqcd-test/cfgsprops/glue/heatbath.f90:239-323
  sprd1            qqcd-test/cfgsprops/glue/heatbath.f90:327-467

Links in the reference layout u(18, ntotal, 4, 2, 16) (flat Fortran order);
the state is the lagfib state (iw, jr, kr) of oracle/lagfib.py.  Each link set
(ieo, ibl, mu) is updated with numpy over its nvhalf sites in the reference's
site order, the sums and products written in the reference's order (numpy
does not fuse multiply-adds), so staples and projections are bit-identical to
the reference's for identical links; numpy's cos / sin / log may differ from
the reference's library in the last place.  Pinned against the reference's
own sweeps in tests/test_oracle_heatbath.py.
"""
import numpy as np

from . import lagfib as LF

TWOPI = float(np.float32(6.283185307179586))


def _dims(n):
    nvh = int(np.prod(n)) // 32
    return nvh, nvh + max(2 * nvh // m for m in n)


def _sites(n, ieo, ibl):
    """coordinates (nvh, 4) of reference sites i = 0..nvh-1 of block ibl, parity ieo (0-based)."""
    nvh, _ = _dims(n)
    ns = [m // 2 for m in n]
    L = 2 * np.arange(nvh)
    s = np.empty((nvh, 4), dtype=np.int64)
    s[:, 0] = L % ns[0]
    q = L // ns[0]
    s[:, 1] = q % ns[1]
    q //= ns[1]
    s[:, 2] = q % ns[2]
    s[:, 3] = q // ns[2]
    s[:, 0] += (s.sum(1) & 1) != ieo
    b = np.array([(ibl >> m) & 1 for m in range(4)])
    return 2 * s + b


def _slot(n, x):
    """(ibl, ieo, i) of each coordinate row."""
    s, b = x >> 1, x & 1
    ns = [m // 2 for m in n]
    L = s[:, 0] + ns[0] * (s[:, 1] + ns[1] * (s[:, 2] + ns[2] * s[:, 3]))
    return (b * np.array([1, 2, 4, 8])).sum(1), s.sum(1) & 1, L // 2


def _mm(y, v):
    w = np.empty_like(y)
    for r in range(3):
        a = [y[:, 6 * r + k] for k in range(6)]
        for c in range(3):
            w[:, 6 * r + 2 * c] = (a[0] * v[:, 2 * c] - a[1] * v[:, 2 * c + 1] + a[2] * v[:, 6 + 2 * c]
                                   - a[3] * v[:, 7 + 2 * c] + a[4] * v[:, 12 + 2 * c] - a[5] * v[:, 13 + 2 * c])
            w[:, 6 * r + 2 * c + 1] = (a[0] * v[:, 2 * c + 1] + a[1] * v[:, 2 * c] + a[2] * v[:, 7 + 2 * c]
                                       + a[3] * v[:, 6 + 2 * c] + a[4] * v[:, 13 + 2 * c] + a[5] * v[:, 12 + 2 * c])
    return w


def _mmd(y, v):
    w = np.empty_like(y)
    for r in range(3):
        a = [y[:, 6 * r + k] for k in range(6)]
        for c in range(3):
            b = [v[:, 6 * c + k] for k in range(6)]
            w[:, 6 * r + 2 * c] = a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3] + a[4] * b[4] + a[5] * b[5]
            w[:, 6 * r + 2 * c + 1] = (-(a[0] * b[1]) + a[1] * b[0] - a[2] * b[3] + a[3] * b[2] - a[4] * b[5]
                                       + a[5] * b[4])
    return w


def _mdm(y, v):
    w = np.empty_like(y)
    for r in range(3):
        for c in range(3):
            w[:, 6 * r + 2 * c] = (y[:, 2 * r] * v[:, 2 * c] + y[:, 2 * r + 1] * v[:, 2 * c + 1]
                                   + y[:, 6 + 2 * r] * v[:, 6 + 2 * c] + y[:, 7 + 2 * r] * v[:, 7 + 2 * c]
                                   + y[:, 12 + 2 * r] * v[:, 12 + 2 * c] + y[:, 13 + 2 * r] * v[:, 13 + 2 * c])
            w[:, 6 * r + 2 * c + 1] = (y[:, 2 * r] * v[:, 2 * c + 1] - y[:, 2 * r + 1] * v[:, 2 * c]
                                       + y[:, 6 + 2 * r] * v[:, 7 + 2 * c] - y[:, 7 + 2 * r] * v[:, 6 + 2 * c]
                                       + y[:, 12 + 2 * r] * v[:, 13 + 2 * c] - y[:, 13 + 2 * r] * v[:, 12 + 2 * c])
    return w


class _Stream:
    """rnds(m) calls on a lagfib state."""

    def __init__(self, state):
        self.state = state

    def rnds(self, m):
        rn, self.state = LF.rnds(self.state, m)
        return rn


def sweep(n, u, coact, state, nsweep=1, nhit=1, gaction=(1, 1, 1, 1)):
    """nsweep sweeps; returns (u, state).  u flat Fortran (18, ntotal, 4, 2, 16)."""
    if nhit < 1:
        raise ValueError("nhit >= 1")
    ga = [int(v) for v in gaction]
    nvh, ntot = _dims(n)
    U = np.array(u, dtype=np.float64, copy=True).reshape(16, 2, 4, ntot, 18)
    co = np.asarray(coact, dtype=np.float64).reshape(4, 4, 4, order="F")  # co[i, mu, nu]
    rs = _Stream(state)
    nn = np.array(n)
    for _ in range(nsweep):
        for ieo in range(2):
            for ibl in range(16):
                x = _sites(n, ieo, ibl)
                for mu in range(4):
                    c = _staple(n, U, x, mu, co, nn, ga)
                    for isu2 in (1, 2, 3):
                        _phbup_sub(U, ieo, ibl, mu, isu2, c, rs, nvh)
    return U.reshape(-1), rs.state


def _link(n, U, x, mu):
    b, e, i = _slot(n, x)
    return U[b, e, mu, i, :]


def _staple(n, U, x, mu, co, nn, ga):
    """plaquette staples (1, 2) and, for gaction = 2, the rectangles: 3-6 long in mu (coact(2,mu,nu)),
    7-8 long in nu (coact(2,nu,mu)), each a path from x to x+mu in the reference's product order."""
    c = np.zeros((x.shape[0], 18))
    L = lambda y, d: _link(n, U, y % nn, d)  # noqa: E731
    for nu in range(4):
        if nu == mu:
            continue
        dn = np.zeros(4, dtype=np.int64)
        dn[nu] = 1
        dm = np.zeros(4, dtype=np.int64)
        dm[mu] = 1
        w = _mmd(_mm(L(x, nu), L(x + dn, mu)), L(x + dm, nu))
        c = c + co[0, mu, nu] * w
        w = _mm(_mdm(L(x - dn, nu), L(x - dn, mu)), L(x - dn + dm, nu))
        c = c + co[0, mu, nu] * w
        if ga[mu] != 1:
            c2 = co[1, mu, nu]
            # 3: x, x+nu, x+nu+mu, x+nu+2mu, x+2mu, x+mu
            w = _mmd(_mm(L(x, nu), L(x + dn, mu)), _mmd(_mm(L(x + dm, mu), L(x + 2 * dm, nu)), L(x + dm + dn, mu)))
            c = c + c2 * w
            # 4: x, x-nu, x-nu+mu, x-nu+2mu, x+2mu, x+mu
            w = _mm(_mdm(L(x - dn, nu), L(x - dn, mu)),
                    _mmd(_mm(L(x - dn + dm, mu), L(x - dn + 2 * dm, nu)), L(x + dm, mu)))
            c = c + c2 * w
            # 5: x, x-mu, x-mu+nu, x+nu, x+nu+mu, x+mu
            w = _mmd(_mm(_mm(_mdm(L(x - dm, mu), L(x - dm, nu)), L(x - dm + dn, mu)), L(x + dn, mu)), L(x + dm, nu))
            c = c + c2 * w
            # 6: x, x-mu, x-mu-nu, x-nu, x-nu+mu, x+mu
            z = x - dn - dm
            w = _mm(_mdm(_mm(_mdm(L(z, mu), L(z, nu)), L(z + dn, mu)), L(x - dn, mu)), L(x - dn + dm, nu))
            c = c + c2 * w
        if ga[nu] != 1:
            c2 = co[1, nu, mu]
            # 7: x, x+nu, x+2nu, x+2nu+mu, x+nu+mu, x+mu
            w = _mmd(_mm(L(x, nu), _mmd(_mm(L(x + dn, nu), L(x + 2 * dn, mu)), L(x + dn + dm, nu))), L(x + dm, nu))
            c = c + c2 * w
            # 8: x, x-nu, x-2nu, x-2nu+mu, x-nu+mu, x+mu
            w = _mm(_mdm(L(x - dn, nu), _mm(_mdm(L(x - 2 * dn, nu), L(x - 2 * dn, mu)), L(x - 2 * dn + dm, nu))),
                    L(x - dn + dm, nu))
            c = c + c2 * w
    return c


def _phbup_sub(U, ieo, ibl, mu, isu2, c, rs, nvh):
    u = U[ibl, ieo, mu, :nvh, :]
    P = 6 if isu2 == 2 else 0
    Q = 6 if isu2 == 1 else 12
    s3 = -1.0 if isu2 == 2 else 1.0
    # phbsub
    z1 = u[:, P] * c[:, P]
    for k in range(1, 6):
        z1 = z1 + u[:, P + k] * c[:, P + k]
    for k in range(6):
        z1 = z1 + u[:, Q + k] * c[:, Q + k]
    z2 = -(u[:, P] * c[:, Q + 1]) + u[:, P + 1] * c[:, Q]
    for m in (1, 2):
        z2 = z2 - u[:, P + 2 * m] * c[:, Q + 2 * m + 1]
        z2 = z2 + u[:, P + 2 * m + 1] * c[:, Q + 2 * m]
    for m in range(3):
        z2 = z2 - u[:, Q + 2 * m] * c[:, P + 2 * m + 1]
        z2 = z2 + u[:, Q + 2 * m + 1] * c[:, P + 2 * m]
    z3 = u[:, P] * c[:, Q]
    for k in range(1, 6):
        z3 = z3 + u[:, P + k] * c[:, Q + k]
    for k in range(6):
        z3 = z3 - u[:, Q + k] * c[:, P + k]
    z3 = s3 * z3
    z4 = -(u[:, P] * c[:, P + 1]) + u[:, P + 1] * c[:, P]
    for m in (1, 2):
        z4 = z4 - u[:, P + 2 * m] * c[:, P + 2 * m + 1]
        z4 = z4 + u[:, P + 2 * m + 1] * c[:, P + 2 * m]
    for m in range(3):
        z4 = z4 + u[:, Q + 2 * m] * c[:, Q + 2 * m + 1]
        z4 = z4 - u[:, Q + 2 * m + 1] * c[:, Q + 2 * m]
    # su2up: Kennedy-Pendleton passes over the sites not yet accepted, in site order
    dk = 1.0 / np.sqrt(z1 * z1 + z2 * z2 + z3 * z3 + z4 * z4)
    done = np.zeros(nvh, dtype=bool)
    a0 = np.zeros(nvh)
    while not done.all():
        idx = np.flatnonzero(~done)
        rn = rs.rnds(4 * idx.size).reshape(-1, 4)
        tt = np.cos(TWOPI * rn[:, 2])
        d2 = -(np.log(rn[:, 0]) + np.log(rn[:, 1]) * tt * tt) * dk[idx] * 3.0
        ok = rn[:, 3] * rn[:, 3] <= (1.0 - 0.5 * d2)
        done[idx[ok]] = True
        a0[idx[ok]] = 1.0 - d2[ok]
    rn = rs.rnds(2 * nvh).reshape(-1, 2)
    rad = 1.0 - a0 * a0
    a3 = np.sqrt(rad) * (2.0 * rn[:, 0] - 1.0)
    rad2 = np.sqrt(np.abs(rad - a3 * a3))
    theta = TWOPI * rn[:, 1]
    a1 = rad2 * np.cos(theta)
    a2 = rad2 * np.sin(theta)
    n1 = dk * (a0 * z1 + a1 * z2 + a2 * z3 + a3 * z4)
    n2 = dk * (-(a0 * z2) + a1 * z1 + a2 * z4 - a3 * z3)
    n3 = dk * (-(a0 * z3) - a1 * z4 + a2 * z1 + a3 * z2)
    n4 = dk * (-(a0 * z4) + a1 * z3 - a2 * z2 + a3 * z1)
    t3 = s3 * n3
    # sprd1
    for col in range(3):
        pr, pi = u[:, P + 2 * col].copy(), u[:, P + 2 * col + 1].copy()
        qr, qi = u[:, Q + 2 * col].copy(), u[:, Q + 2 * col + 1].copy()
        u[:, P + 2 * col] = n1 * pr - n4 * pi + t3 * qr - n2 * qi
        u[:, P + 2 * col + 1] = n1 * pi + n4 * pr + t3 * qi + n2 * qr
        u[:, Q + 2 * col] = -(t3 * pr) - n2 * pi + n1 * qr + n4 * qi
        u[:, Q + 2 * col + 1] = -(t3 * pi) + n2 * pr + n1 * qi - n4 * qr
