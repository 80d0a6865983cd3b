"""CPU oracle for the Wilson / twisted-mass solve path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker: the product library never calls it.

Two legs:

* ``liboracle.so`` (``oracle/qqcd_oracle.c``): a plain-C restatement of the
  reference's operators (Hsingle, Hdouble, Hdbletm, gammamult, vecdot, the
  full-system operator) and of its blocked/checkerboarded storage order.
* ``oracle/_ref/<lattice>/libqqcdref.so`` (``oracle/ref/Makefile``): the
  reference Fortran itself, compiled from its sources with flang, reached
  through the project-written ``oracle/ref/ref_driver.f90``.  It exists only
  where the reference sources were present at build time.

Canonical (oracle) arrays: spinors are complex128 of shape (V, 4, 3)
(site, Dirac, colour), links complex128 of shape (V, 4, 3, 3) (site, mu,
row, col), with site = x + nx*(y + ny*(z + nz*t)).  Reference arrays are
flat float64 vectors in Fortran order: spinors (6, ntotal, 4, 2, 8), links
(18, ntotal, 4, 2, 16).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _iptr(a):
    return a.ctypes.data_as(_ip)


def lib():
    """Load (building on first use if needed) the C oracle."""
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        src = os.path.join(HERE, "qqcd_oracle.c")
        if not os.path.exists(path) or os.path.getmtime(path) < os.path.getmtime(src):
            subprocess.check_call(["make", "-s", "-f", os.path.join(HERE, "Makefile")])
        L = ctypes.CDLL(path)
        L.orc_ref_nvhalf.restype = ctypes.c_long
        L.orc_ref_ntotal.restype = ctypes.c_long
        _LIB = L
    return _LIB


def dims(n):
    return np.ascontiguousarray(n, dtype=np.int32)


def volume(n):
    return int(np.prod(n))


def ref_sizes(n):
    """(nvhalf, ntotal) of the reference's single-process layout (basics.f90:20-30)."""
    L = lib()
    d = dims(n)
    return int(L.orc_ref_nvhalf(_iptr(d))), int(L.orc_ref_ntotal(_iptr(d)))


def spinor_ref_len(n):
    return 6 * ref_sizes(n)[1] * 64


def link_ref_len(n):
    return 18 * ref_sizes(n)[1] * 128


def ref_site(n, x):
    """1-based (i, ieo, ibl, ibleo, gblclr) of global site x in the reference layout."""
    L = lib()
    d = dims(n)
    xx = np.ascontiguousarray(x, dtype=np.int32)
    out = [ctypes.c_int() for _ in range(5)]
    L.orc_ref_site(_iptr(d), _iptr(xx), *[ctypes.byref(o) for o in out])
    return tuple(o.value for o in out)


def spinor_canon_to_ref(n, psi):
    L = lib()
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    be = np.zeros(spinor_ref_len(n))
    bo = np.zeros(spinor_ref_len(n))
    L.orc_spinor_canon_to_ref(_iptr(dims(n)), _ptr(psi.view(np.float64)), _ptr(be), _ptr(bo))
    return be, bo


def spinor_ref_to_canon(n, be=None, bo=None):
    L = lib()
    psi = np.zeros((volume(n), 4, 3), dtype=np.complex128)
    null = ctypes.cast(None, _dp)
    be_p = _ptr(np.ascontiguousarray(be, dtype=np.float64)) if be is not None else null
    bo_p = _ptr(np.ascontiguousarray(bo, dtype=np.float64)) if bo is not None else null
    L.orc_spinor_ref_to_canon(_iptr(dims(n)), be_p, bo_p, _ptr(psi.view(np.float64)))
    return psi


def links_ref_to_canon(n, u):
    L = lib()
    U = np.zeros((volume(n), 4, 3, 3), dtype=np.complex128)
    L.orc_links_ref_to_canon(_iptr(dims(n)), _ptr(np.ascontiguousarray(u, dtype=np.float64)),
                             _ptr(U.view(np.float64)))
    return U


def links_canon_to_ref(n, U):
    L = lib()
    U = np.ascontiguousarray(U, dtype=np.complex128)
    u = np.zeros(link_ref_len(n))
    L.orc_links_canon_to_ref(_iptr(dims(n)), _ptr(U.view(np.float64)), _ptr(u))
    return u


def gamma_table():
    """gamma_1..gamma_5 as a (5, 4, 4) complex array (diracops.f90:61-71)."""
    out = np.zeros(5 * 4 * 4 * 2)
    lib().orc_gamma_table(_ptr(out))
    return out.reshape(5, 4, 4, 2) @ np.array([1.0, 1.0j])


def coact_array(c1=(1.0, 1.0, 1.0, 1.0), c2=(1.0, 1.0, 1.0, 1.0)):
    """Flat Fortran-order coact(4,4,4) with coact(3,1,mu)=c1, coact(3,2,mu)=c2."""
    co = np.zeros((4, 4, 4), order="F")
    for mu in range(4):
        co[2, 0, mu] = c1[mu]
        co[2, 1, mu] = c2[mu]
    return np.asfortranarray(co).ravel(order="F")


def _cc(c):
    return np.ascontiguousarray(c, dtype=np.float64)


def hop(n, U, bc, c1, c2, idag, par_out, psi, out=None):
    """Hopping term onto the sites of global parity par_out (0 even, 1 odd).

    Sites of the other parity in ``out`` are left as they were (zeros for a
    fresh output)."""
    L = lib()
    U = np.ascontiguousarray(U, dtype=np.complex128)
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    if out is None:
        out = np.zeros_like(psi)
    L.orc_hop(_iptr(dims(n)), _ptr(U.view(np.float64)), _iptr(dims(bc)), _ptr(_cc(c1)),
              _ptr(_cc(c2)), int(idag), int(par_out), _ptr(psi.view(np.float64)),
              _ptr(out.view(np.float64)))
    return out


def schur(n, U, bc, c1, c2, idag, kappa, psi):
    """Even-odd Schur operator (Hdouble / Hdbletm) on the even sites of psi."""
    L = lib()
    U = np.ascontiguousarray(U, dtype=np.complex128)
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.zeros_like(psi)
    t1 = np.zeros_like(psi)
    t2 = np.zeros_like(psi)
    L.orc_schur(_iptr(dims(n)), _ptr(U.view(np.float64)), _iptr(dims(bc)), _ptr(_cc(c1)),
                _ptr(_cc(c2)), int(idag), _ptr(_cc(kappa)), _ptr(psi.view(np.float64)),
                _ptr(out.view(np.float64)), _ptr(t1.view(np.float64)), _ptr(t2.view(np.float64)))
    return out


def full_m(n, U, bc, c1, c2, kappa, x):
    """Full-lattice M x = (1 + 2 kappa mu i g5) x - kappa H x."""
    L = lib()
    U = np.ascontiguousarray(U, dtype=np.complex128)
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(x)
    tmp = np.zeros_like(x)
    L.orc_full_m(_iptr(dims(n)), _ptr(U.view(np.float64)), _iptr(dims(bc)), _ptr(_cc(c1)),
                 _ptr(_cc(c2)), _ptr(_cc(kappa)), _ptr(x.view(np.float64)),
                 _ptr(out.view(np.float64)), _ptr(tmp.view(np.float64)))
    return out


def gammamult(n, mu, psi):
    L = lib()
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.zeros_like(psi)
    L.orc_gammamult(_iptr(dims(n)), int(mu), _ptr(psi.view(np.float64)), _ptr(out.view(np.float64)))
    return out


def hsingle_ref(n, u, coact, bc, gblclr, idag, gin):
    L = lib()
    gout = np.zeros(spinor_ref_len(n))
    L.orc_hsingle_ref(_iptr(dims(n)), _ptr(np.ascontiguousarray(u, dtype=np.float64)),
                      _ptr(_cc(coact)), _iptr(dims(bc)), int(gblclr), int(idag),
                      _ptr(np.ascontiguousarray(gin, dtype=np.float64)), _ptr(gout))
    return gout


def hdbletm_ref(n, u, coact, bc, idag, kappa, ge):
    L = lib()
    he = np.zeros(spinor_ref_len(n))
    L.orc_hdbletm_ref(_iptr(dims(n)), _ptr(np.ascontiguousarray(u, dtype=np.float64)),
                      _ptr(_cc(coact)), _iptr(dims(bc)), int(idag), _ptr(_cc(kappa)),
                      _ptr(np.ascontiguousarray(ge, dtype=np.float64)), _ptr(he))
    return he


def parity_mask(n):
    """Boolean (V,) mask of odd global parity."""
    V = volume(n)
    s = np.arange(V)
    x = s % n[0]
    y = (s // n[0]) % n[1]
    z = (s // (n[0] * n[1])) % n[2]
    t = s // (n[0] * n[1] * n[2])
    return ((x + y + z + t) & 1).astype(bool)


def z4(n, seed, rhs=0, scale=1.0):
    """Z4 noise of right-hand side ``rhs`` on the whole lattice, canonical
    (V, 4, 3) complex (the library's counter-based z4source)."""
    L = lib()
    V = volume(n)
    out = np.zeros((V, 4, 3), dtype=np.complex128)
    L.orc_z4(_iptr(dims(n)), ctypes.c_uint64(int(seed)), ctypes.c_long(int(rhs)), ctypes.c_double(scale),
             _ptr(out.view(np.float64)))
    return out


def z2(n, seed, rhs=0, scale=1.0):
    """Z2 noise (the library's counter-based z2source), canonical (V, 4, 3)."""
    L = lib()
    V = volume(n)
    out = np.zeros((V, 4, 3), dtype=np.complex128)
    L.orc_noise(_iptr(dims(n)), ctypes.c_uint64(int(seed)), ctypes.c_long(int(rhs)), ctypes.c_double(scale),
                ctypes.c_int(1), _ptr(out.view(np.float64)))
    return out


GAMMA16_NAMES = ("1", "g1", "g2", "g3", "g4", "g5", "g1g5", "g2g5", "g3g5", "g4g5",
                 "s12", "s13", "s14", "s23", "s24", "s34")


def gamma16():
    """The 16 local Dirac matrices of the trace estimators as dense (16, 4, 4)
    complex matrices, built by matrix products of the reference's gamma_mu:
    1; gamma_1..4; gamma_5; gamma_mu gamma_5; sigma_munu = (i/2)[g_mu, g_nu]."""
    g = gamma_table()
    out = [np.eye(4, dtype=np.complex128)] + [g[m] for m in range(5)]
    out += [g[m] @ g[4] for m in range(4)]
    for a, b in ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)):
        out.append(0.5j * (g[a] @ g[b] - g[b] @ g[a]))
    return np.array(out)


def trace16(n, eta, x):
    """(nt, 16) complex: sum over each timeslice of eta(x)^dag Gamma_n x(x) for
    canonical (V, 4, 3) fields (dense 4x4 products per site)."""
    G = gamma16()
    V = volume(n)
    t = np.arange(V) // (n[0] * n[1] * n[2])
    # per site and Gamma: sum_{a,b,c} conj(eta_ac) G_ab x_bc
    per = np.einsum("sac,nab,sbc->sn", np.conj(eta), G, x, optimize=True)
    out = np.zeros((n[3], 16), dtype=np.complex128)
    np.add.at(out, t, per)
    return out


# ---------------------------------------------------------------------------
# Loop operators of the disconnected-loop path


def coords(n):
    """(V, 4) integer coordinates (x, y, z, t) of the canonical sites."""
    s = np.arange(volume(n))
    return np.stack([s % n[0], (s // n[0]) % n[1], (s // (n[0] * n[1])) % n[2], s // (n[0] * n[1] * n[2])], 1)


def forward(n, bc, mu):
    """Canonical index of x + mu for every site x, and the factor of that hop:
    bc[mu] where it crosses the lattice edge, else 1."""
    c = coords(n)
    c[:, mu] += 1
    wrap = c[:, mu] == n[mu]
    c[wrap, mu] = 0
    idx = c[:, 0] + n[0] * (c[:, 1] + n[1] * (c[:, 2] + n[2] * c[:, 3]))
    return idx, np.where(wrap, float(bc[mu]), 1.0)


LOOP_OPS = ("J1", "J2", "J3", "J4", "scalar", "V1", "V2", "V3", "V4")


def loopops(n, U, bc, b, x, ksub=0):
    """(V, 9) complex: the reference's loop operators per site for source b and
    sink x (canonical fields), qqcd-test/cfgsprops/quark/disconloops.f90:
    7769-7910 with mulfor2 (:8374-8465), mulbac2 (:8574-8675), vv (:8788-8825)
    and vgv (:8829-8921):
      J_mu(x)   = conj( f x(x+mu)^dag (1 - g_mu) U_mu(x)^dag b(x)
                      - f x(x)^dag (1 + g_mu) U_mu(x) b(x+mu) ),
      scalar(x) = conj( x(x)^dag b(x) ) + 12 ksub,
      V_mu(x)   = conj( x(x)^dag g_mu b(x) ),
    f = bc[mu] on hops across the lattice edge."""
    g = gamma_table()
    one = np.eye(4)
    J = np.zeros((volume(n), 9), dtype=np.complex128)
    for mu in range(4):
        nb, f = forward(n, bc, mu)
        Um = U[:, mu]
        udb = np.einsum("sji,saj->sai", np.conj(Um), b, optimize=True)  # U^dag b(x)
        t1 = f * np.einsum("sac,ab,sbc->s", np.conj(x[nb]), one - g[mu], udb, optimize=True)
        ub = np.einsum("sij,saj->sai", Um, b[nb], optimize=True) * f[:, None, None]  # U b(x+mu)
        t2 = -np.einsum("sac,ab,sbc->s", np.conj(x), one + g[mu], ub, optimize=True)
        J[:, mu] = np.conj(t1 + t2)
        J[:, 5 + mu] = np.conj(np.einsum("sac,ab,sbc->s", np.conj(x), g[mu], b, optimize=True))
    J[:, 4] = np.conj(np.einsum("sac,sac->s", np.conj(x), b)) + 12.0 * ksub
    return J


# twoPi in momfacs is a default-real literal: single precision, then widened
TWO_PI = float(np.float32(6.283185307179586))


def momfacs(n, io=(1, 1, 1)):
    """(V, 4, 5) momentum factors of momfacs (disconloops.f90:8925-9010) for
    the origin io (1-based global coordinates); d_k = twoPi (x_k + 1 - io_k) / n_k."""
    c = coords(n)
    d = [TWO_PI * (c[:, k] + 1 - io[k]) / n[k] for k in range(3)]
    s = [np.sin(v) for v in d]
    co = [np.cos(v) for v in d]
    s2 = [np.sin(2.0 * v) for v in d]
    c2 = [np.cos(2.0 * v) for v in d]
    F = np.zeros((volume(n), 4, 5))
    F[:, 0, :4] = np.stack([s[0], s[1], s[2], (co[0] + co[1] + co[2]) / 3.0], 1)
    F[:, 1, :4] = np.stack([s[0] * (co[1] + co[2]) / 2.0, s[1] * (co[0] + co[2]) / 2.0,
                            s[2] * (co[0] + co[1]) / 2.0,
                            (co[0] * co[1] + co[0] * co[2] + co[1] * co[2]) / 3.0], 1)
    F[:, 2, :4] = np.stack([s[0] * co[1] * co[2], s[1] * co[0] * co[2], s[2] * co[0] * co[1],
                            co[0] * co[1] * co[2]], 1)
    F[:, 3, :4] = np.stack([s2[0], s2[1], s2[2], (c2[0] + c2[1] + c2[2]) / 3.0], 1)
    F[:, :, 4] = F[:, :, 3]
    return F


def spacesum(n, J, io=(1, 1, 1), nmom=5):
    """(nt, nmom, 9) complex timeslice sums of per-site operators J (V, 9),
    spacesum (disconloops.f90:9012-9124): momentum 1 sums every operator;
    momentum m = 2..nmom weights the currents by the factors of momfacs:
      J_1 -> F3 Im J_1 + i F2 Im J_1,  J_2 -> F1 Im J_2 + i F3 Im J_2,
      J_3 -> F2 Im J_3 + i F1 Im J_3,  J_4 -> F4 J_4,  scalar -> F5 scalar,
    the local vector currents only at momentum 1."""
    t = coords(n)[:, 3]
    out = np.zeros((n[3], nmom, 9), dtype=np.complex128)
    np.add.at(out[:, 0, :], t, J)
    if nmom > 1:
        F = momfacs(n, io)
        for m in range(1, nmom):
            f = F[:, m - 1, :]
            per = np.zeros((volume(n), 5), dtype=np.complex128)
            per[:, 0] = f[:, 2] * J[:, 0].imag + 1j * f[:, 1] * J[:, 0].imag
            per[:, 1] = f[:, 0] * J[:, 1].imag + 1j * f[:, 2] * J[:, 1].imag
            per[:, 2] = f[:, 1] * J[:, 2].imag + 1j * f[:, 0] * J[:, 2].imag
            per[:, 3] = f[:, 3] * J[:, 3]
            per[:, 4] = f[:, 4] * J[:, 4]
            np.add.at(out[:, m, :5], t, per)
    return out


def loops_ref_to_canon(n, Je, Jo):
    """Per-site operators (V, 9) complex from the reference's flat Fortran
    Je/Jo(2, nvhalf, 9, 2, 8) over every (ieo, ibleo) slice (even sites from
    Je, odd from Jo): each is placed in a reference-layout spinor at
    component 3 spin + colour = op and read back through the layout map."""
    nvh, ntot = ref_sizes(n)
    out = np.zeros((volume(n), 9), dtype=np.complex128)
    par = parity_mask(n)
    for J, odd in ((Je, False), (Jo, True)):
        J5 = np.asarray(J).reshape((2, nvh, 9, 2, 8), order="F")
        A = np.zeros((6, ntot, 4, 2, 8), order="F")
        for op in range(9):
            sp, col = divmod(op, 3)
            A[2 * col, :nvh, sp] = J5[0, :, op]
            A[2 * col + 1, :nvh, sp] = J5[1, :, op]
        flat = A.ravel(order="F")
        psi = spinor_ref_to_canon(n, None, flat) if odd else spinor_ref_to_canon(n, flat, None)
        sel = par if odd else ~par
        out[sel] = psi[sel].reshape(-1, 12)[:, :9]
    return out


def smear(n, U, bc, psi, asmear, nsmear):
    """Gaussian smearing (smear, qqcd-test/cfgsprops/quark/diracops.f90:712-856):
    nsmear times h <- (1 - 6a) h + a sum_{mu=1..3} [U_mu(x) h(x+mu) + U_mu(x-mu)^dag h(x-mu)],
    both terms times bc(mu) across the edge, spin-diagonal; psi (V, 4, 3)."""
    h = np.array(psi, dtype=np.complex128, copy=True)
    fw = [forward(n, bc, mu) for mu in range(3)]
    bw = []
    for mu in range(3):
        idx, f = fw[mu]
        inv = np.empty_like(idx)
        inv[idx] = np.arange(idx.size)
        bw.append((inv, f[inv]))
    for _ in range(int(nsmear)):
        acc = np.zeros_like(h)
        for mu in range(3):
            i, f = fw[mu]
            acc = acc + f[:, None, None] * np.einsum("vab,vdb->vda", U[:, mu], h[i])
            j, g = bw[mu]
            acc = acc + g[:, None, None] * np.einsum("vba,vdb->vda", np.conj(U[j, mu]), h[j])
        h = (1.0 - 6.0 * asmear) * h + asmear * acc
    return h
