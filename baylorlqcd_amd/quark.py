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
