"""Python mirror of the reference's ``inverters`` module on the HIP library.

Procedure names and argument order of qqcd-test/cfgsprops/quark/
inverters.f90 (bicgstab :302, gmresdr :2871, psc :4310, gmresproject :7183,
vecdot :13485).  ``itercount`` is returned (it is an intent(out) scalar in
the Fortran); ``x`` is updated in place.  GeeGooinv and the geometry
arguments are accepted and ignored (Wilson/twisted, one process).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def _arr(a):
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous)):
        raise TypeError("arrays must be contiguous float64 in the reference order")
    return a


def _co(coact):
    return _lib.dp(np.ascontiguousarray(coact, dtype=np.float64).ravel(order="F"))


def _k(kappa):
    return np.ascontiguousarray(kappa, dtype=np.float64)


def vecdot(a, b):
    out = np.zeros(2)
    _lib.check(_lib.lib().qqcd_vecdot(_lib.dp(_arr(a)), _lib.dp(_arr(b)), _lib.dp(out)))
    return out


def gmresdr(rwdir, phi, x, GMRES, resmax, itermin, u, GeeGooinv, iflag, kappa, coact, bc, *geometry):
    it = np.zeros(1, np.int32)
    k = _k(kappa)
    _lib.check(_lib.lib().qqcd_gmresdr(_lib.dp(_arr(phi)), _lib.dp(_arr(x)), _lib.ints(GMRES).ctypes.data_as(_lib._ip),
                                       float(resmax), int(itermin), it.ctypes.data_as(_lib._ip), _lib.dp(_arr(u)),
                                       _co(coact), _lib.dp(k), int(iflag), _lib.ints(bc).ctypes.data_as(_lib._ip)))
    return int(it[0])


def gmresproject(rwdir, b, x, GMRES, resmax, itermin, u, GeeGooinv, iflag, kappa, coact, bc, *geometry):
    it = np.zeros(1, np.int32)
    k = _k(kappa)
    _lib.check(_lib.lib().qqcd_gmresproject(_lib.dp(_arr(b)), _lib.dp(_arr(x)),
                                            _lib.ints(GMRES).ctypes.data_as(_lib._ip), float(resmax), int(itermin),
                                            it.ctypes.data_as(_lib._ip), _lib.dp(_arr(u)), _co(coact), _lib.dp(k),
                                            int(iflag), _lib.ints(bc).ctypes.data_as(_lib._ip)))
    return int(it[0])


def psc(rwdir, b, x, GMRES, resmax, itermin, u, GeeGooinv, iflag, kappa, coact, bc, *geometry, isignal=1,
        resmax_proj=0.0):
    it = np.zeros(1, np.int32)
    k = _k(kappa)
    _lib.check(_lib.lib().qqcd_psc(_lib.dp(_arr(b)), _lib.dp(_arr(x)), _lib.ints(GMRES).ctypes.data_as(_lib._ip),
                                   float(resmax), int(itermin), it.ctypes.data_as(_lib._ip), _lib.dp(_arr(u)),
                                   _co(coact), _lib.dp(k), int(iflag), _lib.ints(bc).ctypes.data_as(_lib._ip),
                                   int(isignal), float(resmax_proj)))
    return int(it[0])


def bicgstab(rwdir, phi, x, resmax, itermin, u, GeeGooinv, iflag, kappa, coact, bc, *geometry):
    it = np.zeros(1, np.int32)
    k = _k(kappa)
    _lib.check(_lib.lib().qqcd_bicgstab(_lib.dp(_arr(phi)), _lib.dp(_arr(x)), float(resmax), int(itermin),
                                        it.ctypes.data_as(_lib._ip), _lib.dp(_arr(u)), _co(coact), _lib.dp(k),
                                        int(iflag), _lib.ints(bc).ctypes.data_as(_lib._ip)))
    return int(it[0])


def last_solve_stats():
    """dict(rel_res, iters, matvecs, cycles, converged) of the last solve."""
    r = np.zeros(1)
    v = np.zeros(4, np.int32)
    _lib.check(_lib.lib().qqcd_last_solve_stats(_lib.dp(r), v[0:].ctypes.data_as(_lib._ip),
                                                v[1:].ctypes.data_as(_lib._ip), v[2:].ctypes.data_as(_lib._ip),
                                                v[3:].ctypes.data_as(_lib._ip)))
    return dict(rel_res=float(r[0]), iters=int(v[0]), matvecs=int(v[1]), cycles=int(v[2]), converged=bool(v[3]))


def deflation_clear():
    _lib.check(_lib.lib().qqcd_deflation_clear())


def solve_log():
    """The last solve's CFGSPROPS.LOG lines as the reference writes them (format (a12,i9,es17.10)):
    [(label, count, value)] -- "gmresdr-norm" per true residual, "bicgstab" per iteration."""
    L = _lib.lib()
    n = ctypes.c_int()
    _lib.check(L.qqcd_solve_log(0, ctypes.byref(n), None, None, None))
    cap = n.value
    lab = ctypes.create_string_buffer(12 * max(cap, 1))
    cnt = np.zeros(max(cap, 1), dtype=np.int32)
    val = np.zeros(max(cap, 1))
    _lib.check(L.qqcd_solve_log(cap, ctypes.byref(n), lab, cnt.ctypes.data_as(_lib._ip), _lib.dp(val)))
    raw = lab.raw
    return [(raw[12 * i:12 * i + 12].decode().strip(), int(cnt[i]), float(val[i])) for i in range(cap)]
