"""Python mirror of the noise-source path of the reference's ``disconloops``
module (qqcd-test/cfgsprops/quark/disconloops.f90): Z4/Z2 sources (z4source
:4359, z2source :4583), the loop over noise solves (testFUNC :1980-2569,
discon :3445-3608), the reference's loop operators (loopops :7769-7910,
spacesum :9012-9124, momfacs :8925-9010, nsaverage :3079-3143) and 16
local trace estimators -- on the GPU, with all right-hand sides after the
first solved together by the batched GMRES-Proj.

Local Dirac matrices of the trace estimators, in order (the reference's
gamma basis, qqcd-test/cfgsprops/quark/diracops.f90:61-71):
``GAMMA_NAMES``.
"""
from __future__ import annotations

import numpy as np

from . import _lib

GAMMA_NAMES = ("1", "g1", "g2", "g3", "g4", "g5", "g1g5", "g2g5", "g3g5", "g4g5",
               "s12", "s13", "s14", "s23", "s24", "s34")


def z4source(be, bo, seed, rhs=0, scale=1.0):
    """Z4 noise {1,-1,i,-i}*scale on both checkerboards (either may be None)."""
    _lib.check(_lib.lib().qqcd_z4source(_lib.dp(be), _lib.dp(bo), int(seed), int(rhs), float(scale)))


def z2source(be, bo, seed, rhs=0, scale=1.0):
    """Z2 noise {1,-1}*scale in the real part on both checkerboards."""
    _lib.check(_lib.lib().qqcd_z2source(_lib.dp(be), _lib.dp(bo), int(seed), int(rhs), float(scale)))


def trace16(eta_e, eta_o, x_e, x_o):
    """(nt, 16) complex: sum over each timeslice of eta^dag Gamma_n x."""
    L = _lib.lib()
    nt = _nt()
    out = np.zeros(nt * 32)
    _lib.check(L.qqcd_trace16(_lib.dp(eta_e), _lib.dp(eta_o), _lib.dp(x_e), _lib.dp(x_o), _lib.dp(out)))
    return out.reshape(nt, 16, 2) @ np.array([1.0, 1j])


LOOP_OPS = ("J1", "J2", "J3", "J4", "scalar", "V1", "V2", "V3", "V4")


def loopops(ksub, xe, xo, be, bo, u=None, bc=None):
    """The reference's loop operators (loopops, disconloops.f90:7769-7910) of
    sink x and source b (reference-layout arrays) on every (ieo, ibleo) slice:
    returns Je, Jo as flat Fortran (2, nvhalf, 9, 2, 8) arrays."""
    from . import lattice
    nvh = lattice.ref_sizes()[0]
    Je = np.zeros(2 * nvh * 9 * 16)
    Jo = np.zeros(2 * nvh * 9 * 16)
    bcp = None if bc is None else _lib.ints(bc).ctypes.data_as(_lib._ip)
    _lib.check(_lib.lib().qqcd_loopops(int(ksub), _lib.dp(xe), _lib.dp(xo), _lib.dp(u), _lib.dp(be), _lib.dp(bo),
                                       bcp, _lib.dp(Je), _lib.dp(Jo)))
    return Je, Jo


def loopsum(ksub, xe, xo, be, bo, io=(1, 1, 1), cur_scale=1.0, u=None, bc=None):
    """Timeslice and momentum sums of the loop operators (spacesum with the
    factors of momfacs for origin io; io None: momentum 1 only), currents
    times cur_scale: (nt, 5, 9) complex."""
    nt = _nt()
    out = np.zeros(2 * nt * 5 * 9)
    bcp = None if bc is None else _lib.ints(bc).ctypes.data_as(_lib._ip)
    iop = None if io is None else _lib.ints(io).ctypes.data_as(_lib._ip)
    _lib.check(_lib.lib().qqcd_loopsum(int(ksub), _lib.dp(xe), _lib.dp(xo), _lib.dp(u), _lib.dp(be), _lib.dp(bo),
                                       bcp, iop, float(cur_scale), _lib.dp(out)))
    J = out.reshape(9, 5, nt, 2) @ np.array([1.0, 1j])
    return J.transpose(2, 1, 0)


def _nt():
    from . import lattice
    return lattice.dims()[3]


def fermprop_multi(be, bo, kappa, coact, bc, resmax, GMRES, u):
    """Solve M x_r = b_r for lists of reference-layout sources; returns
    (xe list, xo list, itercount array, whole-lattice relative residuals)."""
    n = len(be)
    xe = [np.zeros_like(b) for b in be]
    xo = [np.zeros_like(b) for b in bo]
    pbe, kb = _lib.dpp(be)
    pbo, kbo = _lib.dpp(bo)
    pxe, kxe = _lib.dpp(xe)
    pxo, kxo = _lib.dpp(xo)
    it = np.zeros(n, np.int32)
    rr = np.zeros(n)
    co = np.ascontiguousarray(coact, dtype=np.float64).ravel(order="F")
    _lib.check(_lib.lib().qqcd_fermprop_multi(
        n, pbe, pbo, _lib.dp(np.ascontiguousarray(kappa, dtype=np.float64)), _lib.dp(co),
        _lib.ints(bc).ctypes.data_as(_lib._ip), float(resmax), _lib.ints(GMRES).ctypes.data_as(_lib._ip),
        _lib.dp(u), pxe, pxo, it.ctypes.data_as(_lib._ip), _lib.dp(rr)))
    return xe, xo, it, rr


def noise_loop(nrhs, seed, kappa, GMRES, tol=1e-10, scale=1.0, coact=None, bc=(1, 1, 1, -1), u=None, nb=0,
               method=0, loops=True, ksub=0, io=None, lagfib=None):
    """The device noise loop.  Returns a dict with traces ((nt, 16) complex),
    loops ((nt, 5, 9) complex: the reference's operators averaged over the
    sources, currents times kappa; None when loops is False), res_full, iters,
    times (first, rest, total) and stats.  method 0: batched; 1: sequential
    plain FP64 (GMRES-DR then GMRES-Proj per source); 2: sequential
    single-RHS mixed precision.  lagfib (a LagFib state): the sources are the
    ones its z4source would return on nrhs successive calls (seed unused), and
    the state is advanced past them, as the reference's per-source loop
    (disconloops.f90:3032-3053) leaves it."""
    L = _lib.lib()
    nt = _nt()
    tr = np.zeros(nt * 32)
    lo = np.zeros(nt * 90) if loops else None
    iop = None if io is None else _lib.ints(io).ctypes.data_as(_lib._ip)
    res = np.zeros(nrhs)
    it = np.zeros(nrhs, np.int32)
    times = np.zeros(3)
    stats = np.zeros(8, np.int32)
    co = None if coact is None else np.ascontiguousarray(coact, dtype=np.float64).ravel(order="F")
    rest = (_lib.dp(np.ascontiguousarray(kappa, dtype=np.float64)), _lib.dp(co),
            _lib.ints(bc).ctypes.data_as(_lib._ip), _lib.dp(u), _lib.ints(GMRES).ctypes.data_as(_lib._ip),
            float(tol), int(nb), int(method), _lib.dp(tr), int(ksub), iop, _lib.dp(lo), _lib.dp(res),
            it.ctypes.data_as(_lib._ip), _lib.dp(times), stats.ctypes.data_as(_lib._ip))
    if lagfib is None:
        _lib.check(L.qqcd_noise_loop(int(nrhs), int(seed), float(scale), *rest))
    else:
        _lib.check(L.qqcd_noise_loop_lagfib(int(nrhs), *lagfib._args(), float(scale), *rest))
    keys = ("outer", "cycles", "batched_matvecs", "graph_replays", "fallback", "nb", "order", "first_outer")
    return {"traces": tr.reshape(nt, 16, 2) @ np.array([1.0, 1j]),
            "loops": None if lo is None else lo.reshape(nt, 5, 9, 2) @ np.array([1.0, 1j]),
            "res_full": res, "iters": it, "times": times, "stats": dict(zip(keys, (int(s) for s in stats)))}


def batch_set(reorth=0, use_graph=1, inner_tol=1e-5, inner_cycles=4, max_outer=30, op_prec=1, op_recon=12,
              basis_prec=2):
    """Knobs of the batched solver; basis_prec 2 (FP16, the default) or 1
    (FP32) for the inner GMRES's Krylov basis."""
    _lib.check(_lib.lib().qqcd_batch_set(int(reorth), int(use_graph), float(inner_tol), int(inner_cycles),
                                         int(max_outer), int(op_prec), int(op_recon), int(basis_prec)))


def batch_stats():
    import ctypes
    o, c, g, f = (ctypes.c_int() for _ in range(4))
    sm, bm = ctypes.c_long(), ctypes.c_long()
    _lib.check(_lib.lib().qqcd_batch_stats(ctypes.byref(o), ctypes.byref(c), ctypes.byref(sm), ctypes.byref(bm),
                                           ctypes.byref(g), ctypes.byref(f)))
    return {"outer": o.value, "cycles": c.value, "slot_matvecs": sm.value, "batched_matvecs": bm.value,
            "graph_replays": g.value, "fallback": f.value}


def batch_rejected():
    """Per-slot corrections of the last batched solve that were not added to x (NaN or infinite)."""
    import ctypes
    r = ctypes.c_int()
    _lib.check(_lib.lib().qqcd_batch_rejected(ctypes.byref(r)))
    return r.value


def hop_multi_canonical(psi, out, par_out, idag=0, prec=0, recon=18, nact=None, nb=0, order=1):
    """Multi-RHS hop on canonical arrays psi/out of shape (nrhs, V, 4, 3) complex."""
    nrhs = psi.shape[0]
    nact = nrhs if nact is None else nact
    _lib.check(_lib.lib().qqcd_hop_multi_canonical(
        nrhs, int(nact), int(nb), int(prec), int(recon), int(idag), int(par_out), int(order),
        _lib.dp(psi.view(np.float64)), _lib.dp(out.view(np.float64))))


def schur_multi_canonical(psi, out, kappa, idag=0, prec=0, recon=18, nact=None, nb=0, order=1):
    nrhs = psi.shape[0]
    nact = nrhs if nact is None else nact
    _lib.check(_lib.lib().qqcd_schur_multi_canonical(
        nrhs, int(nact), int(nb), int(prec), int(recon), _lib.dp(np.ascontiguousarray(kappa, dtype=np.float64)),
        int(idag), int(order), _lib.dp(psi.view(np.float64)), _lib.dp(out.view(np.float64))))


def autotune_multi(prec, recon, nrhs):
    import ctypes
    nb, order, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    _lib.check(_lib.lib().qqcd_autotune_multi(int(prec), int(recon), int(nrhs), ctypes.byref(nb),
                                              ctypes.byref(order), ctypes.byref(ms)))
    return nb.value, order.value, ms.value


def bench_dslash_multi(prec, recon, nrhs, nact=None, nb=0, order=1, nwarm=3, niter=10):
    import ctypes
    ms = ctypes.c_double()
    nact = nrhs if nact is None else nact
    _lib.check(_lib.lib().qqcd_bench_dslash_multi(int(prec), int(recon), int(nrhs), int(nact), int(nb), int(order),
                                                  int(nwarm), int(niter), ctypes.byref(ms)))
    return ms.value
