"""Python mirror of the reference's ``quark`` module procedures on the path:
pointsource (qqcd-test/cfgsprops/quark/quark.f90:46-108) and fermprop for
Wilson / twisted mass (:458-842)."""
from __future__ import annotations

import numpy as np

from . import _lib


def pointsource(be, bo, src, icsrc, idsrc, myid=0, iblv=None, vecblinv=None):
    _lib.check(_lib.lib().qqcd_pointsource(_lib.dp(be), _lib.dp(bo), _lib.ints(src).ctypes.data_as(_lib._ip),
                                           int(icsrc), int(idsrc)))


def fermprop(rwdir, be, bo, nkappa, kappa, cSW, inverter, coact, bc, resmax, itermin, omegaM3R, u, vectype,
             kGMRES, xe, xo, *geometry, isignal=1):
    """fermprop(rwdir,be,bo,nkappa,kappa,cSW,inverter,coact,bc,resmax,itermin,
    omegaM3R,u,vectype,kGMRES,xe,xo,...); nkappa must be -1.  As in the
    reference, vectype is n of GMRES(n) / GMRES-DR(n,k) for inverter > 3 and
    the initial-vector flag for BiCGStab.  Returns itercount."""
    if nkappa != -1:
        raise _lib.QQCDError("only nkappa = -1 (Wilson / twisted mass) is on this path")
    it = np.zeros(1, np.int32)
    ngm = _lib.ints((vectype, kGMRES))
    _lib.check(_lib.lib().qqcd_fermprop(
        _lib.dp(be), _lib.dp(bo), _lib.dp(np.ascontiguousarray(kappa, dtype=np.float64)), int(inverter),
        _lib.dp(np.ascontiguousarray(coact, dtype=np.float64).ravel(order="F")),
        _lib.ints(bc).ctypes.data_as(_lib._ip), float(resmax), int(itermin), ngm.ctypes.data_as(_lib._ip),
        int(vectype), _lib.dp(u), _lib.dp(xe), _lib.dp(xo), int(isignal), it.ctypes.data_as(_lib._ip)))
    return int(it[0])


def propagator(nrhs, src, kappa, GMRES=(40, 16), tol=1e-10, op_prec=2, op_recon=12, method=0, solutions=False):
    """The point-source propagator on the resident links: nrhs (<= 12)
    spin-colour sources at the 1-based site src, the first by the mixed solver
    (kept deflation space), the rest batched (method 0), one by one over the
    deflation space (1) or one by one with the learned preconditioner (2).
    Returns a dict with iters, res_full, times (first, rest, total), stats and,
    with solutions=True, lists xe, xo of reference-layout solutions."""
    from . import lattice
    it = np.zeros(nrhs, np.int32)
    res = np.zeros(nrhs)
    times = np.zeros(3)
    stats = np.zeros(8, np.int32)
    xe = xo = None
    pe = po = None
    if solutions:
        xe = [lattice.new_spinor().ravel(order="F") for _ in range(nrhs)]
        xo = [lattice.new_spinor().ravel(order="F") for _ in range(nrhs)]
        pe, _ke = _lib.dpp(xe)
        po, _ko = _lib.dpp(xo)
    _lib.check(_lib.lib().qqcd_propagator(
        int(nrhs), _lib.ints(src).ctypes.data_as(_lib._ip), _lib.dp(np.ascontiguousarray(kappa, dtype=np.float64)),
        _lib.ints(GMRES).ctypes.data_as(_lib._ip), float(tol), int(op_prec), int(op_recon), int(method), pe, po,
        it.ctypes.data_as(_lib._ip), _lib.dp(res), _lib.dp(times), stats.ctypes.data_as(_lib._ip)))
    keys = ("outer", "cycles", "batched_matvecs", "graph_replays", "fallback", "nb", "precond_dropped",
            "first_outer")
    out = {"iters": it, "res_full": res, "times": times, "stats": dict(zip(keys, (int(s) for s in stats)))}
    if solutions:
        out["xe"], out["xo"] = xe, xo
    return out


def fpwrite(n, xe, xo, filename, kappa, tmqcd=0.0, cSW=0.0, icsrc=1, idsrc=1, icfg=1, myid=0):
    """FPwrite (qqcd-test/cfgsprops/quark/quark.f90:3379-3429) of reference-layout
    xe / xo (flat float64); host-only."""
    _lib.check(_lib.lib().qqcd_fpwrite(_lib.ints(n).ctypes.data_as(_lib._ip), _lib.dp(xe), _lib.dp(xo),
                                       str(filename).encode(), float(kappa), float(tmqcd), float(cSW), int(icsrc),
                                       int(idsrc), int(icfg), int(myid)))


def fpread(n, xe, xo, filename):
    """FPread (quark.f90:3433-3481) into xe / xo; returns the trailer as a dict."""
    d = np.zeros(3)
    i = np.zeros(4, np.int32)
    ip = [i[k:k + 1].ctypes.data_as(_lib._ip) for k in range(4)]
    _lib.check(_lib.lib().qqcd_fpread(_lib.ints(n).ctypes.data_as(_lib._ip), _lib.dp(xe), _lib.dp(xo),
                                      str(filename).encode(), _lib.dp(d[0:1]), _lib.dp(d[1:2]), _lib.dp(d[2:3]),
                                      *ip))
    return {"kappa": float(d[0]), "tmqcd": float(d[1]), "cSW": float(d[2]), "icsrc": int(i[0]), "idsrc": int(i[1]),
            "icfg": int(i[2]), "myid": int(i[3])}
