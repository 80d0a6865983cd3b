"""Python mirror of the reference's random-number module ``lagfib``
(qqcd-test/cfgsprops/glue/lagfib.f90): ``initlagfib`` (:80) and ``rnds``
(:219), plus the reference's Z4 / Z2 sources drawn from that stream
(``z4source`` / ``z2source``, qqcd-test/cfgsprops/quark/disconloops.f90:4359,
:4583) and its pseudo-heatbath ``sweep`` (qqcd-test/cfgsprops/cfgsprops.f90:
1973-2003).  A ``LagFib`` holds what the module holds -- iwlagfib(0:520),
jrlagfib, krlagfib -- and every call advances it exactly as the reference's
calls would, so draws, sources and states are bit-identical to the
reference's.  Draws are generated on the GPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class LagFib:
    def __init__(self, myid: int = 0, iseed: int = 1):
        """initlagfib(myid, iseed) (host only)."""
        self.iw = np.zeros(521, dtype=np.int32)
        self._jr, self._kr = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.lib().qqcd_lagfib_init(int(myid), int(iseed), self._iwp(), ctypes.byref(self._jr),
                                               ctypes.byref(self._kr)))

    @classmethod
    def from_state(cls, iw, jr, kr):
        s = cls.__new__(cls)
        s.iw = np.array(iw, dtype=np.int32, copy=True)
        if s.iw.shape != (521,):
            raise ValueError("iw must hold 521 words")
        s._jr, s._kr = ctypes.c_int(int(jr)), ctypes.c_int(int(kr))
        return s

    @property
    def state(self):
        """(iwlagfib copy, jrlagfib, krlagfib)."""
        return self.iw.copy(), self._jr.value, self._kr.value

    def _iwp(self):
        return self.iw.ctypes.data_as(ctypes.POINTER(ctypes.c_int))

    def _args(self):
        return self._iwp(), ctypes.byref(self._jr), ctypes.byref(self._kr)

    def rnds(self, m: int) -> np.ndarray:
        """rnds(m, rn): the next m numbers in (0, 1)."""
        rn = np.zeros(max(int(m), 1))
        _lib.check(_lib.lib().qqcd_rnds(int(m), _lib.dp(rn), *self._args()))
        return rn[:int(m)]

    def z4source(self, be, bo):
        """z4source(be, bo, nvhalf, myid): Z4 noise into the reference-layout checkerboards."""
        _lib.check(_lib.lib().qqcd_z4source_lagfib(_lib.dp(be), _lib.dp(bo), *self._args()))

    def z2source(self, be, bo):
        """z2source(be, bo, nvhalf): Z2 noise into the reference-layout checkerboards."""
        _lib.check(_lib.lib().qqcd_z2source_lagfib(_lib.dp(be), _lib.dp(bo), *self._args()))

    def sweep(self, u, coact, nsweep=1, gaction=(1, 1, 1, 1), nhit=1):
        """sweep(gaction, coact, nhit, u, ...) nsweep times: the reference-layout links u
        (18, ntotal, 4, 2, 16) updated in place from this stream."""
        ga, _ga = _lib.ip(gaction)
        co = np.ascontiguousarray(coact, dtype=np.float64)
        _lib.check(_lib.lib().qqcd_sweep(ga, _lib.dp(co), int(nhit), int(nsweep), _lib.dp(u), *self._args()))
