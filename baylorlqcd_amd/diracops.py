"""Python mirror of the reference's ``diracops`` module on the HIP library.

Same procedure names, argument order and meaning as
qqcd-test/cfgsprops/quark/diracops.f90 (smear :712, gamma5mult :1750, twMshift :1863,
Hsingle :554, Hdouble :264,
Hdbletm :435, gammamult :105) and basics.f90 (gam5x :408); arrays are numpy
float64 arrays in the reference order (Fortran (6, ntotal, 4, 2, 8) spinors,
(18, ntotal, 4, 2, 16) links), results are written into the output argument
as the Fortran ``intent(out)`` dummies are.  The trailing geometry arguments
are optional and checked, not used.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .lattice import check_geometry


def _f64(a, name):
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous)):
        raise TypeError(f"{name} must be a contiguous float64 array")
    return a


def Hsingle(gout, u, gin, idag, coact, bc, gblclr, vecbl=None, vecblinv=None, myid=None, nn=None, ldiv=None,
            nms=None, lvbc=None, ib=None, lbd=None, iblv=None, MRT=None):
    check_geometry(vecbl, vecblinv, myid, ldiv, lbd, iblv)
    L = _lib.lib()
    _lib.check(L.qqcd_hsingle(_lib.dp(_f64(gout, "gout")), _lib.dp(_f64(u, "u")), _lib.dp(_f64(gin, "gin")),
                              int(idag), _lib.dp(np.ascontiguousarray(coact, dtype=np.float64).ravel(order="F")),
                              _lib.ints(bc).ctypes.data_as(_lib._ip), int(gblclr)))


def Hdouble(he, u, ge, idag, coact, kappa, bc, vecbl=None, vecblinv=None, myid=None, nn=None, ldiv=None, nms=None,
            lvbc=None, ib=None, lbd=None, iblv=None, MRT=None):
    check_geometry(vecbl, vecblinv, myid, ldiv, lbd, iblv)
    L = _lib.lib()
    _lib.check(L.qqcd_hdouble(_lib.dp(_f64(he, "he")), _lib.dp(_f64(u, "u")), _lib.dp(_f64(ge, "ge")), int(idag),
                              _lib.dp(np.ascontiguousarray(coact, dtype=np.float64).ravel(order="F")),
                              float(kappa), _lib.ints(bc).ctypes.data_as(_lib._ip)))


def Hdbletm(he, u, GeeGooinv, ge, idag, coact, kappa, iflag, bc, vecbl=None, vecblinv=None, myid=None, nn=None,
            ldiv=None, nms=None, lvbc=None, ib=None, lbd=None, iblv=None, MRT=None):
    check_geometry(vecbl, vecblinv, myid, ldiv, lbd, iblv)
    L = _lib.lib()
    k = np.ascontiguousarray(kappa, dtype=np.float64)
    _lib.check(L.qqcd_hdbletm(_lib.dp(_f64(he, "he")), _lib.dp(_f64(u, "u")), _lib.dp(_f64(ge, "ge")), int(idag),
                              _lib.dp(np.ascontiguousarray(coact, dtype=np.float64).ravel(order="F")), _lib.dp(k),
                              int(iflag), _lib.ints(bc).ctypes.data_as(_lib._ip)))


def smear(he, ho, u, ge, go, asmear, nsmear, bc, vecbl=None, vecblinv=None, myid=None, nn=None, ldiv=None,
          nms=None, lvbc=None, ib=None, lbd=None, iblv=None, MRT=None):
    """Gaussian smearing (diracops.f90:712): he, ho = (1 + a sum_{mu<=3} H_mu)^nsmear (ge, go)."""
    check_geometry(vecbl, vecblinv, myid, ldiv, lbd, iblv)
    L = _lib.lib()
    _lib.check(L.qqcd_smear(_lib.dp(_f64(he, "he")), _lib.dp(_f64(ho, "ho")), _lib.dp(_f64(u, "u")),
                            _lib.dp(_f64(ge, "ge")), _lib.dp(_f64(go, "go")), float(asmear), int(nsmear),
                            _lib.ints(bc).ctypes.data_as(_lib._ip)))


def twMshift(hwhole, u, GeeGooinv, gwhole, idag, coact, kappa, iflag, bc, vecbl=None, vecblinv=None, myid=None,
             nn=None, ldiv=None, nms=None, lvbc=None, ib=None, lbd=None, iblv=None, MRT=None):
    """twMshift (diracops.f90:1863): the full-lattice shifted twisted operator; hwhole, gwhole hold
    (6, ntotal, 4, 2, 8, 2) (both checkerboards)."""
    check_geometry(vecbl, vecblinv, myid, ldiv, lbd, iblv)
    L = _lib.lib()
    _lib.check(L.qqcd_twmshift(_lib.dp(_f64(hwhole, "hwhole")), _lib.dp(_f64(u, "u")), _lib.dp(_f64(gwhole, "gwhole")),
                               int(idag), _lib.dp(np.ascontiguousarray(coact, dtype=np.float64).ravel(order="F")),
                               _lib.dp(np.ascontiguousarray(kappa, dtype=np.float64)), int(iflag),
                               _lib.ints(bc).ctypes.data_as(_lib._ip)))


def gamma5mult(he, u, GeeGooinv, ge, idag, coact, kappa, iflag, bc, vecbl=None, vecblinv=None, myid=None, nn=None,
               ldiv=None, nms=None, lvbc=None, ib=None, lbd=None, iblv=None, MRT=None):
    """gamma5mult (diracops.f90:1750): (-i g5 / (2 kappa mu)) times the twisted Schur operator."""
    check_geometry(vecbl, vecblinv, myid, ldiv, lbd, iblv)
    L = _lib.lib()
    _lib.check(L.qqcd_gamma5mult(_lib.dp(_f64(he, "he")), _lib.dp(_f64(u, "u")), _lib.dp(_f64(ge, "ge")), int(idag),
                                 _lib.dp(np.ascontiguousarray(coact, dtype=np.float64).ravel(order="F")),
                                 _lib.dp(np.ascontiguousarray(kappa, dtype=np.float64)), int(iflag),
                                 _lib.ints(bc).ctypes.data_as(_lib._ip)))


def gammamult(evenIN, oddIN, evenOUT, oddOUT, mu, ieo, ibleo):
    L = _lib.lib()
    _lib.check(L.qqcd_gammamult(_lib.dp(_f64(evenIN, "evenIN")), _lib.dp(_f64(oddIN, "oddIN")),
                                _lib.dp(_f64(evenOUT, "evenOUT")), _lib.dp(_f64(oddOUT, "oddOUT")), int(mu),
                                int(ieo), int(ibleo)))


def gam5x(a):
    _lib.check(_lib.lib().qqcd_gam5x(_lib.dp(_f64(a, "a"))))


def hop_canonical(psi, par_out, idag=0, prec=0, recon=18, out=None):
    """Hopping term on canonical (V, 4, 3) complex spinors (native interface)."""
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.zeros_like(psi) if out is None else np.ascontiguousarray(out, dtype=np.complex128)
    _lib.check(_lib.lib().qqcd_hop_canonical(_lib.dp(psi.view(np.float64)), _lib.dp(out.view(np.float64)),
                                             int(par_out), int(idag), int(prec), int(recon)))
    return out


def schur_canonical(psi, kappa, idag=0, prec=0, recon=18):
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.zeros_like(psi)
    k = np.ascontiguousarray(kappa, dtype=np.float64)
    _lib.check(_lib.lib().qqcd_schur_canonical(_lib.dp(psi.view(np.float64)), _lib.dp(out.view(np.float64)),
                                               _lib.dp(k), int(idag), int(prec), int(recon)))
    return out


def gauge_load(u, coact):
    _lib.check(_lib.lib().qqcd_gauge_load(_lib.dp(_f64(u, "u")),
                                          _lib.dp(np.ascontiguousarray(coact, dtype=np.float64).ravel(order="F"))))


def gauge_load_canonical(U, coact):
    U = np.ascontiguousarray(U, dtype=np.complex128)
    _lib.check(_lib.lib().qqcd_gauge_load_canonical(
        _lib.dp(U.view(np.float64)), _lib.dp(np.ascontiguousarray(coact, dtype=np.float64).ravel(order="F"))))


def gauge_random(kind, eps, seed, coact):
    """Seeded SU(3) field generated on the device (0 Haar, 1 near identity)."""
    _lib.check(_lib.lib().qqcd_gauge_random(int(kind), float(eps), int(seed),
                                            _lib.dp(np.ascontiguousarray(coact, dtype=np.float64).ravel(order="F"))))


def gauge_get_canonical(n):
    V = int(np.prod(n))
    U = np.zeros((V, 4, 3, 3), dtype=np.complex128)
    _lib.check(_lib.lib().qqcd_gauge_get_canonical(_lib.dp(U.view(np.float64))))
    return U


def gauge_get_ref(n=None):
    """The resident links in the reference layout (flat Fortran order of u(18, ntotal, 4, 2, 16))."""
    from . import lattice
    u = np.zeros(int(np.prod(lattice.link_shape(n))))
    _lib.check(_lib.lib().qqcd_gauge_get_ref(_lib.dp(u)))
    return u


def gauge_heatbath(beta, nsweep, seed, start, coact):
    """Device pseudo-heatbath (start: 0 cold, 1 hot, -1 continue)."""
    _lib.check(_lib.lib().qqcd_gauge_heatbath(float(beta), int(nsweep), int(seed), int(start),
                                              _lib.dp(np.ascontiguousarray(coact, dtype=np.float64).ravel(order="F"))))


def gauge_plaquette():
    p = np.zeros(1)
    _lib.check(_lib.lib().qqcd_gauge_plaquette(_lib.dp(p)))
    return float(p[0])
