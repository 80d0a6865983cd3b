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


def gmresproject(rwdir, b, xshift, GMRES, resmax, itermin, u, GeeGooinv, iflag, kappa, coact, bc, *geometry,
                 isignal=2, mvp=None, gdr=None):
    """gmresproject with the reference's arguments (inverters.f90:7183-7271):
    xshift holds one or more reference-order spinors (its size a multiple of
    one), all zeroed, the solution in the first; mvp (int32 array of one)
    receives the operator applications; gdr (float64, Fortran order, rows =
    its first extent) receives the residual history in column 1.  Returns
    itercount."""
    it = np.zeros(1, np.int32)
    k = _k(kappa)
    from . import lattice
    slen = int(np.prod(lattice.spinor_shape()))
    xs = _arr(xshift)
    if xs.size % slen:
        raise ValueError("xshift must hold whole spinors")
    mv = np.zeros(1, np.int32) if mvp is None else mvp
    if gdr is None:
        gd, ld = None, 0
    else:
        if not (gdr.dtype == np.float64 and gdr.flags.f_contiguous):
            raise TypeError("gdr must be a Fortran-ordered float64 array")
        gd, ld = _lib.dp(gdr), int(gdr.shape[0])
    _lib.check(_lib.lib().qqcd_gmresproject(
        _lib.dp(_arr(b)) if b is not None else None, _lib.dp(xs), xs.size // slen,
        _lib.ints(GMRES).ctypes.data_as(_lib._ip), float(resmax), int(itermin), it.ctypes.data_as(_lib._ip),
        _lib.dp(_arr(u)), _co(coact), _lib.dp(k), int(iflag), _lib.ints(bc).ctypes.data_as(_lib._ip), int(isignal),
        mv.ctypes.data_as(_lib._ip), gd, ld))
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


def solve_mixed(phi, x, GMRES, tol, basis_prec, op_prec, op_recon, inner_tol, keep_deflation, u, coact, kappa, bc):
    """Mixed-precision solve of the even-odd system M_hat x = phi (reference-order
    even checkerboards; x is the starting guess and receives the solution).
    Returns the Krylov steps; mixed_stats() has the rest."""
    it = np.zeros(1, np.int32)
    _lib.check(_lib.lib().qqcd_solve_mixed(
        _lib.dp(_arr(phi)), _lib.dp(_arr(x)), _lib.ints(GMRES).ctypes.data_as(_lib._ip), float(tol), int(basis_prec),
        int(op_prec), int(op_recon), float(inner_tol), int(keep_deflation), _lib.dp(_arr(u)), _co(coact),
        _lib.dp(_k(kappa)), _lib.ints(bc).ctypes.data_as(_lib._ip), it.ctypes.data_as(_lib._ip)))
    return int(it[0])


def set_fault(precond=0, corrupt_outer=0, batch_rhs=-1, maxiter=0):
    """Force failures for tests of the fallbacks (qqcd_set_fault); all
    defaults switch them off."""
    _lib.check(_lib.lib().qqcd_set_fault(int(precond), int(corrupt_outer), int(batch_rhs), int(maxiter)))


def mixed_stats():
    """dict(outer, fallback, precond_dropped, rejected, converged, rel_res) of the last mixed solve."""
    v = np.zeros(5, np.int32)
    r = np.zeros(1)
    _lib.check(_lib.lib().qqcd_mixed_stats(*(v[k:].ctypes.data_as(_lib._ip) for k in range(5)), _lib.dp(r)))
    return dict(outer=int(v[0]), fallback=int(v[1]), precond_dropped=int(v[2]), rejected=int(v[3]),
                converged=bool(v[4]), rel_res=float(r[0]))


def bench_solve(method, kappa, op_prec=2, recon=12, GMRES=(40, 16), tol=1e-10, src=(1, 1, 1, 1), spin=1, colour=1,
                inner_tol=0.0, keep_deflation=0):
    """Device-resident solve of M x = point source (qqcd_bench_solve): method 0
    plain FP64 GMRES-DR, 1 mixed precision (FP32 basis), 2 mixed precision
    (FP16 basis).  Returns seconds, Krylov steps, outer steps and the true
    full-system residual."""
    sec, res = np.zeros(1), np.zeros(1)
    its, outer = np.zeros(1, np.int32), np.zeros(1, np.int32)
    _lib.check(_lib.lib().qqcd_bench_solve(
        int(method), _lib.ints(src).ctypes.data_as(_lib._ip), int(spin), int(colour),
        _lib.dp(np.asarray(kappa, dtype=np.float64)), _lib.ints(GMRES).ctypes.data_as(_lib._ip), float(tol),
        int(op_prec), int(recon), float(inner_tol), int(keep_deflation), _lib.dp(sec),
        its.ctypes.data_as(_lib._ip), outer.ctypes.data_as(_lib._ip), _lib.dp(res)))
    return {"seconds": float(sec[0]), "iterations": int(its[0]), "outer": int(outer[0]),
            "true_residual": float(res[0])}


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
