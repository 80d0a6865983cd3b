"""Lattice set-up and the reference's single-process geometry tables.

``init(n, bc)`` starts the HIP library for an n = (nx, ny, nz, nt) lattice.
``initlat(bc)`` returns the tables the reference's ``initlat`` builds
(qqcd-test/cfgsprops/glue/lattice.f90:33-59) for one process, computed from
their closed forms; the operator mirrors use them only to check that a
caller's geometry arguments describe one process with the canonical block
order, as the C library supports nothing else.
"""
from __future__ import annotations

import numpy as np

from . import _lib

_STATE = {"n": None, "bc": None}


def init(n, bc=(1, 1, 1, -1), device=0):
    n = tuple(int(v) for v in n)
    _lib.check(_lib.lib().qqcd_init(_lib.ints(n).ctypes.data_as(_lib._ip),
                                    _lib.ints(bc).ctypes.data_as(_lib._ip), int(device)))
    _STATE["n"] = n
    _STATE["bc"] = tuple(int(v) for v in bc)


def finalize():
    _lib.check(_lib.lib().qqcd_finalize())
    _STATE["n"] = None


def dims():
    if _STATE["n"] is None:
        raise _lib.QQCDError("call lattice.init first")
    return _STATE["n"]


def ref_sizes(n=None):
    """(nvhalf, ntotal) of the reference arrays (basics.f90:20-30)."""
    n = dims() if n is None else n
    V = int(np.prod(n))
    nvh = V // 32
    nb = max(2 * nvh // int(m) for m in n)
    return nvh, nvh + nb


def spinor_shape(n=None):
    return (6, ref_sizes(n)[1], 4, 2, 8)


def link_shape(n=None):
    return (18, ref_sizes(n)[1], 4, 2, 16)


def new_spinor(n=None):
    return np.zeros(spinor_shape(n), order="F")


def new_links(n=None):
    return np.zeros(link_shape(n), order="F")


def setblk():
    """lbd(16,4), iblv(16,4) of setblk (lattice.f90:298-331), 1-based values."""
    lbd = np.zeros((16, 4), dtype=bool)
    iblv = np.zeros((16, 4), dtype=np.int32)
    for ibl in range(16):
        b = [(ibl >> m) & 1 for m in range(4)]
        for mu in range(4):
            lbd[ibl, mu] = b[mu] == 0
            iblv[ibl, mu] = (ibl ^ (1 << mu)) + 1
    return lbd, iblv


def gblchker():
    """vecbl(2,8), vecblinv(2,16) of gblchker (lattice.f90:335-354)."""
    even = [ibl + 1 for ibl in range(16) if bin(ibl).count("1") % 2 == 0]
    odd = [ibl + 1 for ibl in range(16) if bin(ibl).count("1") % 2 == 1]
    vecbl = np.array([even, odd], dtype=np.int32)
    vecblinv = np.zeros((2, 16), dtype=np.int32)
    for ibl in range(16):
        vecblinv[0, ibl] = 1 + bin(ibl).count("1") % 2
        vecblinv[1, ibl] = ibl // 2 + 1
    return vecbl, vecblinv


def check_geometry(vecbl=None, vecblinv=None, myid=None, ldiv=None, lbd=None, iblv=None):
    """Raise unless the given reference geometry arguments describe one
    process with the canonical block tables (the only supported case)."""
    if myid not in (None, 0):
        raise _lib.QQCDError("only a single process (myid = 0) is supported")
    if ldiv is not None and np.any(np.asarray(ldiv)):
        raise _lib.QQCDError("ldiv must be all .false. (one process per direction)")
    vb, vbi = gblchker()
    lb, ib = setblk()
    for got, want, name in ((vecbl, vb, "vecbl"), (vecblinv, vbi, "vecblinv"), (lbd, lb, "lbd"),
                            (iblv, ib, "iblv")):
        if got is not None and not np.array_equal(np.asarray(got).astype(want.dtype), want):
            raise _lib.QQCDError(f"{name} is not the canonical single-process table")
