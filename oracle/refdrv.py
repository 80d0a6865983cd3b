"""ctypes access to the reference Fortran compiled by oracle/ref/Makefile.

TEST INFRASTRUCTURE ONLY (golden-vector generation and the reference CPU
baseline).  A build exists only where /root/reference was present when
oracle/ref/Makefile ran; ``available()`` says whether one is on disk.
"""
from __future__ import annotations

import concurrent.futures
import ctypes
import os
import subprocess
import tempfile
import threading

import numpy as np

from . import HERE

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
REF_SRC = "/root/reference/qqcd-test"


# The reference keeps large work arrays as local variables (for example
# htmp(6,ntotal,4,2,8) in Hdouble, qqcd-test/cfgsprops/quark/diracops.f90:
# 326) and is run with `ulimit -s unlimited`.  Every call into it therefore
# runs on one worker thread with a 3 GiB stack.
_EXEC = None


def _run(fn, *args):
    global _EXEC
    if _EXEC is None:
        old = threading.stack_size()
        threading.stack_size(3 << 30)
        try:
            _EXEC = concurrent.futures.ThreadPoolExecutor(max_workers=1)
            _EXEC.submit(lambda: None).result()
        finally:
            threading.stack_size(old)
    return _EXEC.submit(fn, *args).result()


def _p(a):
    return a.ctypes.data_as(_dp)


def _ip_(a):
    return a.ctypes.data_as(_ip)


def tag(n, solvers=True):
    return "x".join(str(int(v)) for v in n) + ("" if solvers else "-ds")


def lib_path(n, solvers=True):
    return os.path.join(HERE, "_ref", tag(n, solvers), "libqqcdref.so")


def available(n, solvers=True):
    return os.path.exists(lib_path(n, solvers))


def build(n, solvers=True, jobs=4):
    """Compile the reference for lattice n (needs /root/reference and flang)."""
    if not os.path.isdir(REF_SRC):
        raise FileNotFoundError("reference sources not present: " + REF_SRC)
    subprocess.check_call(
        ["make", "-s", "-f", os.path.join(HERE, "ref", "Makefile"), f"-j{jobs}",
         f"NX={n[0]}", f"NY={n[1]}", f"NZ={n[2]}", f"NT={n[3]}", f"SOLVERS={1 if solvers else 0}"],
        stdout=subprocess.DEVNULL)
    return lib_path(n, solvers)


class RefLattice:
    """One compiled reference build (fixed lattice) with its geometry tables."""

    def __init__(self, n, bc=(1, 1, 1, -1), solvers=True, workdir=None):
        path = lib_path(n, solvers)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.L = ctypes.CDLL(path)
        nn = np.zeros(4, np.int32)
        nvh = ctypes.c_int()
        ntot = ctypes.c_int()
        _run(self.L.ref_dims, _ip_(nn), ctypes.byref(nvh), ctypes.byref(ntot))
        self.n = tuple(int(v) for v in nn)
        if self.n != tuple(int(v) for v in n):
            raise ValueError(f"build holds {self.n}, asked for {tuple(n)}")
        self.nvhalf, self.ntotal = nvh.value, ntot.value
        self.solvers = solvers
        self.workdir = workdir or tempfile.mkdtemp(prefix="qqcdref_")
        # the reference solvers append to <rwdir>CFGSPROPS.LOG, opened status="old"
        open(os.path.join(self.workdir, "CFGSPROPS.LOG"), "a").close()
        self.set_bc(bc)

    def set_bc(self, bc):
        self.bc = np.ascontiguousarray(bc, dtype=np.int32)
        d = (self.workdir.rstrip("/") + "/").encode()
        _run(self.L.ref_init, _ip_(self.bc), ctypes.c_char_p(d), ctypes.c_int(len(d)))

    @property
    def slen(self):
        return 6 * self.ntotal * 64

    @property
    def ulen(self):
        return 18 * self.ntotal * 128

    def tables(self):
        nvh = self.nvhalf
        lvbc = np.zeros((2, 4, nvh), np.int32)
        vecbl = np.zeros((8, 2), np.int32)
        vecblinv = np.zeros((16, 2), np.int32)
        iblv = np.zeros((4, 16), np.int32)
        lbd = np.zeros((4, 16), np.int32)
        nms = np.zeros(4, np.int32)
        _run(self.L.ref_tables, _ip_(lvbc), _ip_(vecbl), _ip_(vecblinv), _ip_(iblv), _ip_(lbd), _ip_(nms))
        return dict(lvbc=lvbc, vecbl=vecbl, vecblinv=vecblinv, iblv=iblv, lbd=lbd, nms=nms)

    def hsingle(self, u, gin, coact, gblclr, idag=0):
        gout = np.zeros(self.slen)
        g = np.array(gin, dtype=np.float64, copy=True)
        _run(self.L.ref_hsingle, _p(gout), _p(np.ascontiguousarray(u)), _p(g), int(idag),
                           _p(np.ascontiguousarray(coact)), int(gblclr))
        return gout

    def hdouble(self, u, ge, coact, kappa, idag=0):
        he = np.zeros(self.slen)
        g = np.array(ge, dtype=np.float64, copy=True)
        _run(self.L.ref_hdouble, _p(he), _p(np.ascontiguousarray(u)), _p(g), int(idag),
                           _p(np.ascontiguousarray(coact)), ctypes.c_double(kappa))
        return he

    def hdbletm(self, u, ge, coact, kappa, idag=0, iflag=-1):
        he = np.zeros(self.slen)
        g = np.array(ge, dtype=np.float64, copy=True)
        k = np.ascontiguousarray(kappa, dtype=np.float64)
        _run(self.L.ref_hdbletm, _p(he), _p(np.ascontiguousarray(u)), _p(g), int(idag),
                           _p(np.ascontiguousarray(coact)), _p(k), int(iflag))
        return he

    def smear(self, u, ge, go, asmear, nsmear):
        he, ho = np.zeros(self.slen), np.zeros(self.slen)
        _run(self.L.ref_smear, _p(he), _p(ho), _p(np.ascontiguousarray(u)), _p(np.ascontiguousarray(ge)),
             _p(np.ascontiguousarray(go)), ctypes.c_double(asmear), int(nsmear))
        return he, ho

    def twmshift(self, u, ge, go, coact, kappa, idag=0):
        g = np.concatenate([np.asarray(ge, dtype=np.float64), np.asarray(go, dtype=np.float64)])
        h = np.zeros(2 * self.slen)
        _run(self.L.ref_twmshift, _p(h), _p(np.ascontiguousarray(u)), _p(g), int(idag),
             _p(np.ascontiguousarray(coact)), _p(np.ascontiguousarray(kappa, dtype=np.float64)))
        return h[:self.slen], h[self.slen:]

    def gamma5mult(self, u, ge, coact, kappa, idag=0):
        he = np.zeros(self.slen)
        g = np.array(ge, dtype=np.float64, copy=True)
        _run(self.L.ref_gamma5mult, _p(he), _p(np.ascontiguousarray(u)), _p(g), int(idag),
             _p(np.ascontiguousarray(coact)), _p(np.ascontiguousarray(kappa, dtype=np.float64)))
        return he

    def gammamult(self, ein, oin, mu, ieo, ibleo, eout=None, oout=None):
        eout = np.zeros(self.slen) if eout is None else np.array(eout, copy=True)
        oout = np.zeros(self.slen) if oout is None else np.array(oout, copy=True)
        _run(self.L.ref_gammamult, _p(np.ascontiguousarray(ein)), _p(np.ascontiguousarray(oin)),
                             _p(eout), _p(oout), int(mu), int(ieo), int(ibleo))
        return eout, oout

    def gam5x(self, a):
        a = np.array(a, dtype=np.float64, copy=True)
        _run(self.L.ref_gam5x, _p(a))
        return a

    # ---- solver build only -------------------------------------------
    def vecdot(self, a, b):
        out = np.zeros(2)
        _run(self.L.ref_vecdot, _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(b)), _p(out))
        return complex(out[0], out[1])

    def solve(self, which, phi, u, coact, kappa, nk=(40, 10), resmax=1e-10, itermin=0,
              x0=None, isignal=1):
        """which: 0 bicgstab, 5 gmresdr, 6 psc (with isignal 1 or 2)."""
        x = np.zeros(self.slen) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        ph = np.array(phi, dtype=np.float64, copy=True)
        it = ctypes.c_int()
        nkk = np.ascontiguousarray(nk, dtype=np.int32)
        _run(self.L.ref_solve, int(which), _p(ph), _p(x), _ip_(nkk), ctypes.c_double(resmax),
                         int(itermin), ctypes.byref(it), _p(np.ascontiguousarray(u)),
                         _p(np.ascontiguousarray(coact)),
                         _p(np.ascontiguousarray(kappa, dtype=np.float64)), int(isignal))
        return x, it.value

    def gmresproject(self, phi, u, coact, kappa, nk, resmax, itermin=0, isignal=2, nshift=1, ngdr=2000, ncol=1):
        """The reference's gmresproject called directly (after psc with isignal=1 built
        its deflation space): returns xshift (slen, nshift), itercount, mvp, gdr (ngdr, ncol)."""
        xs = np.full(self.slen * nshift, 7.0)  # the reference zeroes every shift
        ph = np.array(phi, dtype=np.float64, copy=True)
        it = ctypes.c_int()
        mvp = ctypes.c_int(-5)
        gdr = np.zeros((ngdr, ncol), order="F")
        nkk = np.ascontiguousarray(nk, dtype=np.int32)
        _run(self.L.ref_gmresproject, _p(ph), _p(xs), int(nshift), _ip_(nkk), ctypes.c_double(resmax),
             int(itermin), ctypes.byref(it), _p(np.ascontiguousarray(u)), _p(np.ascontiguousarray(coact)),
             _p(np.ascontiguousarray(kappa, dtype=np.float64)), int(isignal), ctypes.byref(mvp), _p(gdr),
             int(ngdr), int(ncol))
        return xs.reshape(nshift, self.slen), it.value, mvp.value, gdr

    def fermpropr(self, be, bo, u, coact, kappa, inverter, ir, nk=(40, 10), resmax=1e-10, itermin=0, isignal=1):
        xe = np.zeros(self.slen)
        xo = np.zeros(self.slen)
        b_e = np.array(be, dtype=np.float64, copy=True)
        b_o = np.array(bo, dtype=np.float64, copy=True)
        uu = np.array(u, dtype=np.float64, copy=True)
        nkk = np.ascontiguousarray(nk, dtype=np.int32)
        _run(self.L.ref_fermpropr, _p(b_e), _p(b_o), _p(np.ascontiguousarray(kappa, dtype=np.float64)),
             int(inverter), _p(np.ascontiguousarray(coact)), ctypes.c_double(resmax), int(itermin), _ip_(nkk),
             _p(uu), _p(xe), _p(xo), int(isignal), int(ir))
        return xe, xo

    def fermprop(self, be, bo, u, coact, kappa, inverter, nk=(40, 10), resmax=1e-10, itermin=0,
                 isignal=1, xe0=None):
        xe = np.zeros(self.slen) if xe0 is None else np.array(xe0, dtype=np.float64, copy=True)
        xo = np.zeros(self.slen)
        b_e = np.array(be, dtype=np.float64, copy=True)
        b_o = np.array(bo, dtype=np.float64, copy=True)
        uu = np.array(u, dtype=np.float64, copy=True)
        nkk = np.ascontiguousarray(nk, dtype=np.int32)
        _run(self.L.ref_fermprop, _p(b_e), _p(b_o), _p(np.ascontiguousarray(kappa, dtype=np.float64)),
                            int(inverter), _p(np.ascontiguousarray(coact)), ctypes.c_double(resmax),
                            int(itermin), _ip_(nkk), _p(uu), _p(xe), _p(xo), int(isignal))
        return xe, xo

    def pointsource(self, src, icsrc, idsrc):
        be = np.zeros(self.slen)
        bo = np.zeros(self.slen)
        s = np.ascontiguousarray(src, dtype=np.int32)
        _run(self.L.ref_pointsource, _p(be), _p(bo), _ip_(s), int(icsrc), int(idsrc))
        return be, bo

    def heatbath(self, u, coact, nsweep, nhit=1, iseed=1):
        uu = np.array(u, dtype=np.float64, copy=True)
        _run(self.L.ref_heatbath, _p(uu), _p(np.ascontiguousarray(coact)), int(nsweep), int(nhit), int(iseed))
        return uu

    def lagfib_init(self, myid, iseed):
        iw = np.zeros(521, dtype=np.int32)
        jr, kr = ctypes.c_int(), ctypes.c_int()
        _run(self.L.ref_lagfib_init, int(myid), int(iseed), _ip_(iw), ctypes.byref(jr), ctypes.byref(kr))
        return iw, jr.value, kr.value

    def lagfib_get(self):
        iw = np.zeros(521, dtype=np.int32)
        jr, kr = ctypes.c_int(), ctypes.c_int()
        _run(self.L.ref_lagfib_get, _ip_(iw), ctypes.byref(jr), ctypes.byref(kr))
        return iw, jr.value, kr.value

    def lagfib_set(self, iw, jr, kr):
        _run(self.L.ref_lagfib_set, _ip_(np.ascontiguousarray(iw, dtype=np.int32)), int(jr), int(kr))

    def rnds(self, m):
        rn = np.zeros(int(m))
        _run(self.L.ref_rnds, int(m), _p(rn))
        return rn

    def z4source(self):
        be, bo = np.zeros(self.slen), np.zeros(self.slen)
        _run(self.L.ref_z4source, _p(be), _p(bo))
        return be, bo

    def z2source(self):
        be, bo = np.zeros(self.slen), np.zeros(self.slen)
        _run(self.L.ref_z2source, _p(be), _p(bo))
        return be, bo

    def staple(self, u, coact, mu, ieo, ibl, gaction=(1, 1, 1, 1)):
        """c(18, nvhalf) of link set (mu, ieo, ibl) (1-based), flat Fortran order."""
        c = np.zeros(18 * self.nvhalf)
        uu = np.array(u, dtype=np.float64, copy=True)
        _run(self.L.ref_staple, int(mu), int(ieo), int(ibl), _p(np.ascontiguousarray(coact)),
             _ip_(np.ascontiguousarray(gaction, dtype=np.int32)), _p(c), _p(uu))
        return c

    def sweeps(self, u, coact, nsweep, gaction=(1, 1, 1, 1), nhit=1):
        uu = np.array(u, dtype=np.float64, copy=True)
        _run(self.L.ref_sweeps, _p(uu), _p(np.ascontiguousarray(coact)),
             _ip_(np.ascontiguousarray(gaction, dtype=np.int32)), int(nsweep), int(nhit))
        return uu

    def aveplaq(self, u):
        uu = np.array(u, dtype=np.float64, copy=True)
        p = ctypes.c_double()
        _run(self.L.ref_aveplaq, _p(uu), ctypes.byref(p))
        return p.value

    def loopops(self, ksub, xe, xo, u, be, bo):
        """Je, Jo as (2, 8, 9, nvhalf, 2) C-order views of the Fortran
        (2, nvhalf, 9, 2, 8) arrays."""
        nvh = self.nvhalf
        Je = np.zeros(2 * nvh * 9 * 16)
        Jo = np.zeros(2 * nvh * 9 * 16)
        args = [np.array(a, dtype=np.float64, copy=True) for a in (xe, xo, u, be, bo)]
        _run(self.L.ref_loopops, int(ksub), *[_p(a) for a in args], _p(Je), _p(Jo))
        return Je, Jo

    def spacesum(self, Je, Jo, io=(1, 1, 1)):
        """Jtime flat Fortran (2, nt, 5, 9)."""
        out = np.zeros(2 * self.n[3] * 5 * 9)
        iov = np.ascontiguousarray(io, dtype=np.int32)
        _run(self.L.ref_spacesum, _p(np.ascontiguousarray(Je)), _p(np.ascontiguousarray(Jo)), _ip_(iov), _p(out))
        return out

    def nsaverage(self, u, be, bo, xe, xo, kappa, coact):
        out = np.zeros(2 * self.n[3] * 5 * 9)
        args = [np.array(a, dtype=np.float64, copy=True) for a in (u, be, bo, xe, xo)]
        _run(self.L.ref_nsaverage, _p(out), *[_p(a) for a in args], ctypes.c_double(kappa),
             _p(np.ascontiguousarray(coact)))
        return out

    def gfwrite(self, u, path, gaction=(1, 1, 1, 1), beta=6.0, alat=(1, 1, 1, 1), tad=(1, 1, 1, 1), icfg=1):
        """GFwrite (fixsu3 is applied to a copy of u, as the reference does in place)."""
        uu = np.array(u, dtype=np.float64, copy=True)
        f = path.encode()
        _run(self.L.ref_gfwrite, _p(uu), ctypes.c_char_p(f), ctypes.c_int(len(f)),
             _ip_(np.ascontiguousarray(gaction, dtype=np.int32)), ctypes.c_double(beta),
             _p(np.ascontiguousarray(alat, dtype=np.float64)), _p(np.ascontiguousarray(tad, dtype=np.float64)),
             ctypes.c_int(icfg))
        return uu

    def gfread(self, path):
        u = np.zeros(self.ulen)
        ga = np.zeros(4, np.int32)
        beta = ctypes.c_double()
        alat, tad = np.zeros(4), np.zeros(4)
        icfg, mi = ctypes.c_int(), ctypes.c_int()
        f = path.encode()
        _run(self.L.ref_gfread, _p(u), ctypes.c_char_p(f), ctypes.c_int(len(f)), _ip_(ga), ctypes.byref(beta),
             _p(alat), _p(tad), ctypes.byref(icfg), ctypes.byref(mi))
        return u, dict(gaction=ga, beta=beta.value, alat=alat, tad=tad, icfg=icfg.value, myid=mi.value)

    def fpwrite(self, xe, xo, path, kappa=0.15, tmqcd=0.0, csw=0.0, icsrc=1, idsrc=1, icfg=1):
        a = np.array(xe, dtype=np.float64, copy=True)
        b = np.array(xo, dtype=np.float64, copy=True)
        f = path.encode()
        _run(self.L.ref_fpwrite, _p(a), _p(b), ctypes.c_char_p(f), ctypes.c_int(len(f)), ctypes.c_double(kappa),
             ctypes.c_double(tmqcd), ctypes.c_double(csw), ctypes.c_int(icsrc), ctypes.c_int(idsrc),
             ctypes.c_int(icfg))

    def fpread(self, path):
        xe, xo = np.zeros(self.slen), np.zeros(self.slen)
        k, t, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        ints = [ctypes.c_int() for _ in range(4)]
        f = path.encode()
        _run(self.L.ref_fpread, _p(xe), _p(xo), ctypes.c_char_p(f), ctypes.c_int(len(f)), ctypes.byref(k),
             ctypes.byref(t), ctypes.byref(c), *[ctypes.byref(v) for v in ints])
        return xe, xo, dict(kappa=k.value, tmqcd=t.value, cSW=c.value, icsrc=ints[0].value, idsrc=ints[1].value,
                            icfg=ints[2].value, myid=ints[3].value)

    def fixaction(self, beta):
        co = np.zeros(64)
        _run(self.L.ref_fixaction, ctypes.c_double(beta), _p(co))
        return co
