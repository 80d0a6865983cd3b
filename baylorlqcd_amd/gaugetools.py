"""Python mirror of the reference's gauge-field files: GFwrite / GFread
(qqcd-test/cfgsprops/glue/gaugetools.f90:1259-1334) on reference-layout link
arrays u(18, ntotal, 4, 2, 16).  Host-only: no lattice context is needed."""
from __future__ import annotations

import numpy as np

from . import _lib


def _n(n):
    return _lib.ints(n).ctypes.data_as(_lib._ip)


def gfwrite(n, u, filename, gaction=(1, 1, 1, 1), beta=6.0, alat=(1.0, 1.0, 1.0, 1.0), tad=(1.0, 1.0, 1.0, 1.0),
            icfgsave=1, myid=0):
    """Write u (flat float64, Fortran order) to a new file; fixsu3 is applied
    to u in place first, as the reference does."""
    if u.dtype != np.float64 or not u.flags.c_contiguous:
        raise _lib.QQCDError("u must be a contiguous float64 array (updated in place)")
    _lib.check(_lib.lib().qqcd_gfwrite(_n(n), _lib.dp(u), str(filename).encode(), _lib.ints(gaction).ctypes.data_as(
        _lib._ip), float(beta), _lib.dp(np.ascontiguousarray(alat, dtype=np.float64)),
        _lib.dp(np.ascontiguousarray(tad, dtype=np.float64)), int(icfgsave), int(myid)))


def gfread(n, u, filename):
    """Read into u (flat float64, rows 1-2 from the file, the links completed
    by fixsu3); returns the trailer as a dict."""
    ga = np.zeros(4, np.int32)
    beta, alat, tad = np.zeros(1), np.zeros(4), np.zeros(4)
    ic, mi = np.zeros(1, np.int32), np.zeros(1, np.int32)
    _lib.check(_lib.lib().qqcd_gfread(_n(n), _lib.dp(u), str(filename).encode(), ga.ctypes.data_as(_lib._ip),
                                      _lib.dp(beta), _lib.dp(alat), _lib.dp(tad), ic.ctypes.data_as(_lib._ip),
                                      mi.ctypes.data_as(_lib._ip)))
    return {"gaction": ga, "beta": float(beta[0]), "alat": alat, "tad": tad, "icfgsave": int(ic[0]),
            "myid": int(mi[0])}
