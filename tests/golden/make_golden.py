"""Write the golden fixtures of the solve path with the reference Fortran.

Runs only where /root/reference is present (this script builds it with
oracle/ref/Makefile, then calls it through oracle/refdrv.py).  Output:
tests/golden/golden_4x4x4x4.npz and tests/golden/golden_8x8x8x16.npz;
with the argument ``loops`` only tests/golden/golden_loops_8x8x8x16.npz
(the loop operators loopops / spacesum / nsaverage of the noise loop); with
``io`` only tests/golden/golden_io_8x8x8x16.json (SHA-256 digests of the
reference's GFwrite / FPwrite files and of what GFread / FPread return); with
``lagfib`` only tests/golden/golden_lagfib_8x8x8x16.npz (initlagfib states,
rnds draws including the zero skip, z4source / z2source digests); with
``sweep`` only tests/golden/golden_sweep_8x8x8x16.npz (the reference's
pseudo-heatbath sweeps from the config-1 field: stream states, per-set link
sums and the links of block 1 after each sweep); with ``smear`` only
tests/golden/golden_smear_8x8x8x16.npz (the reference's smear, twMshift and
gamma5mult: timeslices 0 and nt-1 of each output, per-timeslice norms and
sums); with ``noiseloop`` only tests/golden/golden_noiseloop_8x8x8x16.npz
(the reference's own per-source loop z4source -> fermprop -> nsaverage from
initlagfib(0, 5) on the config-1 field: the mean of nsaverage's momentum-1
operators over the sources and the lagfib state after the loop).

Inputs are regenerated from seeds by the tests (spinors: numpy
default_rng(seed) Gaussians; links: baylorlqcd_amd.gauge.haar), except the
8^3x16 heatbath field, which is stored (12 reals per link; row 2 is
conj(row0 x row1) and the reference consumed that reconstructed field).

Canonical orders: spinor psi[s, spin, colour], s = x + nx*(y + ny*(z + nz*t));
"even" arrays hold the global-parity-0 sites in increasing s.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from oracle import refdrv  # noqa: E402
from baylorlqcd_amd import gauge  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def spinor(n, seed):
    rng = np.random.default_rng(seed)
    V = int(np.prod(n))
    return rng.standard_normal((V, 4, 3)) + 1j * rng.standard_normal((V, 4, 3))


def canon_even(n, be):
    psi = O.spinor_ref_to_canon(n, be, None)
    return psi[~O.parity_mask(n)]


def canon_odd(n, bo):
    psi = O.spinor_ref_to_canon(n, None, bo)
    return psi[O.parity_mask(n)]


def ops_block(R, n, U, tag, out):
    u = O.links_canon_to_ref(n, U)
    co = O.coact_array()
    psi = spinor(n, 101)
    be, bo = O.spinor_canon_to_ref(n, psi)
    for gbl, gin in ((1, bo), (2, be)):
        for idag in (0, 1):
            g = R.hsingle(u, gin, co, gbl, idag)
            par = gbl - 1
            pm = O.parity_mask(n) == bool(par)
            out[f"{tag}hsingle_g{gbl}_d{idag}"] = O.spinor_ref_to_canon(
                n, g if gbl == 1 else None, g if gbl == 2 else None)[pm]
    for name, kap in (("wilson", (0.150, 0.0)), ("tm", (0.156, 0.005)), ("tmneg", (0.156, -0.005))):
        for idag in (0, 1):
            he = R.hdbletm(u, be, co, kap, idag)
            out[f"{tag}hdbletm_{name}_d{idag}"] = canon_even(n, he)
    out[f"{tag}hdouble_d0"] = canon_even(n, R.hdouble(u, be, co, 0.150, 0))
    # anisotropic coefficients (c1 != c2)
    co2 = O.coact_array((0.9, 1.1, 1.0, 0.7), (1.2, 0.8, 1.0, 0.7))
    g = R.hsingle(u, be, co2, 2, 0)
    out[f"{tag}hsingle_aniso_g2_d0"] = canon_odd(n, g)


def main():
    # 4^4 builds without the inverters: their off-path debug prints index
    # sites up to 128 with constant subscripts, which flang rejects when
    # ntotal is smaller (qqcd-test/cfgsprops/quark/inverters.f90:2325-2333).
    for n, solvers in (((4, 4, 4, 4), False), ((8, 8, 8, 16), True)):
        refdrv.build(n, solvers=solvers)

    # ---------------- 4^4: operators in both boundary conditions, solves --------
    n = (4, 4, 4, 4)
    out = {}
    for bc in ((1, 1, 1, -1), (1, 1, 1, 1)):
        R = refdrv.RefLattice(n, bc=bc, solvers=False)
        tag = "ap_" if bc[3] == -1 else "p_"
        U = gauge.haar(n, 7)
        ops_block(R, n, U, tag, out)
        if bc[3] == -1:
            t = R.tables()
            for k, v in t.items():
                out["tables_" + k] = v
            psi = spinor(n, 202)
            be, bo = O.spinor_canon_to_ref(n, psi)
            out["gam5x"] = canon_even(n, R.gam5x(be))
            for mu in range(1, 6):
                eo, oo = R.gammamult(be, bo, mu, 2, 5)
                out[f"gammamult_mu{mu}_e"] = canon_even(n, eo)
                out[f"gammamult_mu{mu}_o"] = canon_odd(n, oo)
    np.savez_compressed(os.path.join(HERE, "golden_4x4x4x4.npz"), **out)
    print("wrote golden_4x4x4x4.npz", len(out), "arrays")

    # ---------------- 8^3x16: the config-1 field, operators and solves ----------
    n = (8, 8, 8, 16)
    bc = (1, 1, 1, -1)
    R = refdrv.RefLattice(n, bc=bc)
    out = {}
    co6 = R.fixaction(6.0)
    U0 = gauge.haar(n, 20261202)
    u = R.heatbath(O.links_canon_to_ref(n, U0), co6, 200, nhit=1, iseed=11)
    U = O.links_ref_to_canon(n, u)
    U12 = gauge.compress12(U)
    U = gauge.reconstruct12(U12)
    out["links12"] = U12
    out["plaquette_ref"] = np.array(R.aveplaq(O.links_canon_to_ref(n, U)))
    u = O.links_canon_to_ref(n, U)
    co = O.coact_array()
    psi = spinor(n, 101)
    be, bo = O.spinor_canon_to_ref(n, psi)
    out["hsingle_g2_d0"] = canon_odd(n, R.hsingle(u, be, co, 2, 0))
    out["hdbletm_tm_d0"] = canon_even(n, R.hdbletm(u, be, co, (0.156, 0.005), 0))
    psi2 = spinor(n, 202)
    be2, bo2 = O.spinor_canon_to_ref(n, psi2)
    out["vecdot"] = np.array(R.vecdot(be2, bo2))
    ps_be, ps_bo = R.pointsource((3, 2, 4, 1), 2, 3)
    out["pointsource_be"] = ps_be
    out["pointsource_bo"] = ps_bo
    # config 1: point source at the origin, spin 1 colour 1, Wilson kappa = 0.150
    ps_e, ps_o = R.pointsource((1, 1, 1, 1), 1, 1)
    cases = {
        "gmresdr_k150": dict(kap=(0.150, 0.0), inverter=5, nk=(40, 10), src=(ps_e, ps_o)),
        "bicgstab_k150": dict(kap=(0.150, 0.0), inverter=0, nk=(1, 0), src=(ps_e, ps_o)),
        "gmresdr_tm": dict(kap=(0.156, 0.005), inverter=5, nk=(40, 10), src=(ps_e, ps_o)),
    }
    for name, cs in cases.items():
        b_e, b_o = cs["src"]
        R.set_bc(bc)
        log0 = open(os.path.join(R.workdir, "CFGSPROPS.LOG")).read()
        xe, xo = R.fermprop(b_e, b_o, u, co, cs["kap"], cs["inverter"], nk=cs["nk"], resmax=1e-11, isignal=1)
        log = open(os.path.join(R.workdir, "CFGSPROPS.LOG")).read()[len(log0):]
        out[f"{name}_xe"] = canon_even(n, xe)
        if name == "gmresdr_k150":
            out[f"{name}_xo"] = canon_odd(n, xo)
        print(name, "log tail:", log.strip().splitlines()[-1:] if log.strip() else "")
    # psc: first RHS by GMRES-DR (isignal=1), a second RHS by GMRES-Proj
    b2_e, b2_o = R.pointsource((1, 1, 1, 1), 2, 3)
    xe1, _ = R.fermprop(ps_e, ps_o, u, co, (0.150, 0.0), 6, nk=(40, 10), resmax=1e-11, isignal=1)
    xe2, xo2 = R.fermprop(b2_e, b2_o, u, co, (0.150, 0.0), 6, nk=(40, 10), resmax=1e-11, isignal=2)
    out["psc_rhs1_xe"] = canon_even(n, xe1)[:64]
    out["psc_rhs2_xe"] = canon_even(n, xe2)
    np.savez_compressed(os.path.join(HERE, "golden_8x8x8x16.npz"), **out)
    print("wrote golden_8x8x8x16.npz", len(out), "arrays; plaquette", float(out["plaquette_ref"]))


LOOP_CASES = {
    # name: (bc, ksub, momentum origin io)
    "a": ((1, 1, 1, -1), 0, (1, 1, 1)),
    "b": ((-1, 1, 0, 1), 1, (3, 2, 5)),
}


def loop_sites(n):
    """Canonical sites whose per-site operators are stored: timeslices 0 and
    nt-1 (the time hop of the last one crosses the lattice edge)."""
    ls = n[0] * n[1] * n[2]
    return np.concatenate([np.arange(ls), np.arange((n[3] - 1) * ls, n[3] * ls)])


def loops():
    n = (8, 8, 8, 16)
    refdrv.build(n, solvers=True)
    U = gauge.haar(n, 3)
    u = O.links_canon_to_ref(n, U)
    b, x = spinor(n, 301), spinor(n, 302)
    be, bo = O.spinor_canon_to_ref(n, b)
    xe, xo = O.spinor_canon_to_ref(n, x)
    c = O.coords(n)
    sites = loop_sites(n)
    out = {}
    for name, (bc, ksub, io) in LOOP_CASES.items():
        R = refdrv.RefLattice(n, bc=bc)
        Je, Jo = R.loopops(ksub, xe, xo, u, be, bo)
        nvh = R.nvhalf
        Je4, Jo4 = Je.reshape(8, 2, 9, nvh, 2), Jo.reshape(8, 2, 9, nvh, 2)
        J = np.zeros((len(sites), 9), dtype=np.complex128)
        for k, s in enumerate(sites):
            i, ieo, _, ibleo, gc = O.ref_site(n, c[s])
            A = Je4 if gc == 1 else Jo4
            J[k] = A[ibleo - 1, ieo - 1, :, i - 1, 0] + 1j * A[ibleo - 1, ieo - 1, :, i - 1, 1]
        out[f"loopops_{name}"] = J
        Jt = R.spacesum(Je, Jo, io).reshape(9, 5, n[3], 2)
        out[f"spacesum_{name}"] = (Jt[..., 0] + 1j * Jt[..., 1]).transpose(2, 1, 0)
    R = refdrv.RefLattice(n, bc=LOOP_CASES["a"][0])
    Jn = R.nsaverage(u, be, bo, xe, xo, 0.15, O.coact_array()).reshape(9, 5, n[3], 2)
    out["nsaverage_a_k150"] = (Jn[..., 0] + 1j * Jn[..., 1]).transpose(2, 1, 0)[:, 0, :]
    np.savez_compressed(os.path.join(HERE, "golden_loops_8x8x8x16.npz"), **out)
    print("wrote golden_loops_8x8x8x16.npz", len(out), "arrays")


IO_CASE = dict(links_seed=17, spinor_seed=5, gaction=(1, 1, 1, 2), beta=6.0, alat=(1.0, 1.0, 1.0, 0.5),
               tad=(0.9, 0.9, 0.9, 1.0), icfg=7, kappa=0.156, tmqcd=0.005, cSW=0.0, icsrc=2, idsrc=3)


def io_inputs(n):
    u = O.links_canon_to_ref(n, gauge.haar(n, IO_CASE["links_seed"]))
    rng = np.random.default_rng(IO_CASE["spinor_seed"])
    V = int(np.prod(n))
    psi = rng.standard_normal((V, 4, 3)) + 1j * rng.standard_normal((V, 4, 3))
    xe, xo = O.spinor_canon_to_ref(n, psi)
    return u, xe, xo


def io():
    import hashlib
    import json
    import tempfile
    n = (8, 8, 8, 16)
    refdrv.build(n, solvers=True)
    R = refdrv.RefLattice(n, bc=(1, 1, 1, -1))
    u, xe, xo = io_inputs(n)
    d = tempfile.mkdtemp()
    c = IO_CASE
    gf = os.path.join(d, "cfg.gf")
    uw = R.gfwrite(u, gf, c["gaction"], c["beta"], c["alat"], c["tad"], c["icfg"])
    ur, _ = R.gfread(gf)
    fp = os.path.join(d, "prop.fp")
    R.fpwrite(xe, xo, fp, c["kappa"], c["tmqcd"], c["cSW"], c["icsrc"], c["idsrc"], c["icfg"])
    ye, yo, _ = R.fpread(fp)
    sha = lambda b: hashlib.sha256(b).hexdigest()  # noqa: E731
    out = {"lattice": list(n), "case": {k: list(v) if isinstance(v, tuple) else v for k, v in c.items()},
           "gf_file_sha256": sha(open(gf, "rb").read()), "gf_file_bytes": os.path.getsize(gf),
           "gfwrite_u_sha256": sha(uw.tobytes()), "gfread_u_sha256": sha(ur.tobytes()),
           "fp_file_sha256": sha(open(fp, "rb").read()), "fp_file_bytes": os.path.getsize(fp),
           "fpread_xe_sha256": sha(ye.tobytes()), "fpread_xo_sha256": sha(yo.tobytes())}
    with open(os.path.join(HERE, "golden_io_8x8x8x16.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote golden_io_8x8x8x16.json")


LAG_INIT = [(0, 1), (0, 2), (3, 1), (0, 1234567), (5, 99)]


def lagfib():
    import hashlib
    n = (8, 8, 8, 16)
    refdrv.build(n, solvers=True)
    R = refdrv.RefLattice(n, bc=(1, 1, 1, -1))
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()  # noqa: E731
    out = {}
    states = [R.lagfib_init(my, sd) for my, sd in LAG_INIT]
    out["init_cases"] = np.array(LAG_INIT, dtype=np.int32)
    out["init_iw"] = np.stack([s[0] for s in states])
    out["init_jrkr"] = np.array([s[1:] for s in states], dtype=np.int32)
    # from initlagfib(0, 1): 1000 draws, then 200003 more (digest), then 100 from a state whose
    # next draw is zero
    R.lagfib_init(0, 1)
    out["rn_1000"] = R.rnds(1000)
    iw, jr, kr = R.lagfib_get()
    out["after_1000_iw"], out["after_1000_jrkr"] = iw, np.array([jr, kr], np.int32)
    out["rn_200003_sha256"] = np.array(sha(R.rnds(200003)))
    iw, jr, kr = R.lagfib_get()
    out["after_200003_iw"], out["after_200003_jrkr"] = iw, np.array([jr, kr], np.int32)
    iw = iw.copy()
    iw[kr] = iw[jr]
    R.lagfib_set(iw, jr, kr)
    out["zero_in_iw"], out["zero_in_jrkr"] = iw, np.array([jr, kr], np.int32)
    out["rn_zero_100"] = R.rnds(100)
    iw, jr, kr = R.lagfib_get()
    out["zero_out_iw"], out["zero_out_jrkr"] = iw, np.array([jr, kr], np.int32)
    # z4source then z2source from initlagfib(0, 5)
    R.lagfib_init(0, 5)
    be, bo = R.z4source()
    out["z4_be_sha256"], out["z4_bo_sha256"] = np.array(sha(be)), np.array(sha(bo))
    be, bo = R.z2source()
    out["z2_be_sha256"], out["z2_bo_sha256"] = np.array(sha(be)), np.array(sha(bo))
    iw, jr, kr = R.lagfib_get()
    out["after_z_iw"], out["after_z_jrkr"] = iw, np.array([jr, kr], np.int32)
    np.savez_compressed(os.path.join(HERE, "golden_lagfib_8x8x8x16.npz"), **out)
    print("wrote golden_lagfib_8x8x8x16.npz")


def sweep_coact_aniso():
    """coact with coact(1, mu, nu) different from coact(1, nu, mu) (pins the index order)."""
    co = np.zeros((4, 4, 4), order="F")
    for mu in range(4):
        for nu in range(4):
            if mu != nu:
                co[0, mu, nu] = 2.0 + 0.1 * mu + 0.03 * nu
    return np.asfortranarray(co).ravel(order="F")


def sweep_summary(u, n):
    nvh = int(np.prod(n)) // 32
    ntot = nvh + max(2 * nvh // m for m in n)
    U = u.reshape(16, 2, 4, ntot, 18)[:, :, :, :nvh]
    return U.sum(axis=3), U[0].copy()


def sweep_coact_rect():
    """an improved-action coact: plaquette and rectangle coefficients, both asymmetric in (mu, nu)."""
    co = np.zeros((4, 4, 4), order="F")
    for mu in range(4):
        for nu in range(4):
            if mu != nu:
                co[0, mu, nu] = 2.0 + 0.1 * mu + 0.03 * nu
                co[1, mu, nu] = -0.1 * (1 + 0.1 * mu + 0.03 * nu)
    return np.asfortranarray(co).ravel(order="F")


def sweep_cases():
    """(name, coact builder, lagfib seed, sweeps, gaction)."""
    return [("iso", None, 3, 2, (1, 1, 1, 1)), ("aniso", "aniso", 4, 1, (1, 1, 1, 1)),
            ("rect", "rect", 8, 1, (2, 2, 2, 2)), ("mixed", "rect", 9, 1, (1, 1, 1, 2))]


def sweep():
    n = (8, 8, 8, 16)
    refdrv.build(n, solvers=True)
    R = refdrv.RefLattice(n, bc=(1, 1, 1, 1))
    g = np.load(os.path.join(HERE, "golden_8x8x8x16.npz"))
    u0 = O.links_canon_to_ref(n, gauge.reconstruct12(g["links12"]))
    out = {}
    for name, cob, seed, nsw, ga in sweep_cases():
        co = {None: lambda: R.fixaction(6.0), "aniso": sweep_coact_aniso, "rect": sweep_coact_rect}[cob]()
        out[f"{name}_coact"] = np.asarray(co)
        out[f"{name}_gaction"] = np.array(ga, np.int32)
        iw, jr, kr = R.lagfib_init(0, seed)
        u = u0
        for k in range(1, nsw + 1):
            u = R.sweeps(u, co, 1, gaction=ga)
            iw, jr, kr = R.lagfib_get()
            sums, blk = sweep_summary(u, n)
            out[f"{name}_{k}_iw"], out[f"{name}_{k}_jrkr"] = iw, np.array([jr, kr], np.int32)
            out[f"{name}_{k}_sums"], out[f"{name}_{k}_block1"] = sums, blk
            out[f"{name}_{k}_plaq"] = np.array(R.aveplaq(u))
    # the staples of every link set of a Haar field with rectangles: bit-exact digest
    import hashlib
    uh = O.links_canon_to_ref(n, gauge.haar(n, 4))
    st = np.concatenate([R.staple(uh, sweep_coact_rect(), mu, ieo, ibl, (2, 2, 2, 2))
                         for ieo in (1, 2) for ibl in range(1, 17) for mu in range(1, 5)])
    out["staples_rect_sha256"] = np.array(hashlib.sha256(st.tobytes()).hexdigest())
    np.savez_compressed(os.path.join(HERE, "golden_sweep_8x8x8x16.npz"), **out)
    print("wrote golden_sweep_8x8x8x16.npz")


SMEAR_CASES = [("a", (1, 1, 1, -1), 0.25, 3), ("b", (-1, 1, 0, 1), 0.1, 2), ("c", (1, 1, 1, 1), 0.3, 0)]
SHIFT_KAPPA = (0.15, 0.02)


def slice_summary(n, psi):
    """canonical field -> (timeslices 0 and nt-1, per-timeslice |.|^2, per-timeslice sums)."""
    ls = n[0] * n[1] * n[2]
    t = psi.reshape(n[3], ls, 4, 3)
    return (np.concatenate([t[0], t[-1]]), np.sum(np.abs(t) ** 2, axis=(1, 2, 3)), t.sum(axis=(1, 2, 3)))


def smear_inputs(n):
    U = gauge.haar(n, 3)
    rng = np.random.default_rng(1)
    V = int(np.prod(n))
    psi = rng.standard_normal((V, 4, 3)) + 1j * rng.standard_normal((V, 4, 3))
    return U, psi


def smear():
    n = (8, 8, 8, 16)
    refdrv.build(n, solvers=True)
    U, psi = smear_inputs(n)
    u = O.links_canon_to_ref(n, U)
    ge, go = O.spinor_canon_to_ref(n, psi)
    out = {}
    for name, bc, a, ns in SMEAR_CASES:
        R = refdrv.RefLattice(n, bc=bc)
        he, ho = R.smear(u, ge, go, a, ns)
        sl, nrm, sm = slice_summary(n, O.spinor_ref_to_canon(n, he, ho))
        out[f"smear_{name}_slices"], out[f"smear_{name}_norms"], out[f"smear_{name}_sums"] = sl, nrm, sm
    R = refdrv.RefLattice(n, bc=(1, 1, 1, -1))
    co = O.coact_array()
    for idag in (0, 1):
        h1, h2 = R.twmshift(u, ge, go, co, SHIFT_KAPPA, idag)
        sl, nrm, sm = slice_summary(n, O.spinor_ref_to_canon(n, h1, h2))
        out[f"twm{idag}_slices"], out[f"twm{idag}_norms"], out[f"twm{idag}_sums"] = sl, nrm, sm
        g5 = R.gamma5mult(u, ge, co, SHIFT_KAPPA, idag)
        sl, nrm, sm = slice_summary(n, O.spinor_ref_to_canon(n, g5, None))
        out[f"g5m{idag}_slices"], out[f"g5m{idag}_norms"], out[f"g5m{idag}_sums"] = sl, nrm, sm
    np.savez_compressed(os.path.join(HERE, "golden_smear_8x8x8x16.npz"), **out)
    print("wrote golden_smear_8x8x8x16.npz")


NOISE_CASE = dict(iseed=5, nrhs=3, kappa=0.150, nk=(40, 10), resmax=1e-12)


def noiseloop():
    """The reference's own noise loop on the config-1 field: initlagfib(0, iseed), then per
    source z4source, fermprop (GMRES-DR) and nsaverage, averaged over the sources (what a
    testFUNC-style loop accumulates), and the lagfib state after the loop."""
    n = (8, 8, 8, 16)
    refdrv.build(n, solvers=True)
    g = np.load(os.path.join(HERE, "golden_8x8x8x16.npz"))
    u = O.links_canon_to_ref(n, gauge.reconstruct12(g["links12"]))
    co = O.coact_array()
    cs = NOISE_CASE
    R = refdrv.RefLattice(n, bc=(1, 1, 1, -1))
    R.lagfib_init(0, cs["iseed"])
    J = np.zeros((n[3], 5, 9), dtype=np.complex128)
    for r in range(cs["nrhs"]):
        be, bo = R.z4source()
        xe, xo = R.fermprop(be, bo, u, co, (cs["kappa"], 0.0), 5, nk=cs["nk"], resmax=cs["resmax"])
        Jn = R.nsaverage(u, be, bo, xe, xo, cs["kappa"], co).reshape(9, 5, n[3], 2)
        J += (Jn[..., 0] + 1j * Jn[..., 1]).transpose(2, 1, 0)
    J /= cs["nrhs"]
    iw, jr, kr = R.lagfib_get()
    np.savez_compressed(os.path.join(HERE, "golden_noiseloop_8x8x8x16.npz"), nsaverage_mean=J[:, 0, :],
                        lagfib_after_iw=iw, lagfib_after_jrkr=np.array([jr, kr], dtype=np.int32),
                        case=np.array([cs["iseed"], cs["nrhs"]], dtype=np.int32), kappa=np.float64(cs["kappa"]))
    print("wrote golden_noiseloop_8x8x8x16.npz", np.abs(J[:, 0, :]).max())


BOUNDARY_CASE = dict(kappa=0.150, nk=(40, 10), resmax_dr=1e-11, resmax_proj=1e-10)


def boundary():
    """The reference's gmresproject called directly with its whole argument list
    (inverters.f90:7183-7271) after psc (isignal=1) built the deflation space, and
    fermpropR with ir = -1 (quark.f90:846-931), on the config-1 field."""
    n = (8, 8, 8, 16)
    refdrv.build(n, solvers=True)
    g = np.load(os.path.join(HERE, "golden_8x8x8x16.npz"))
    u = O.links_canon_to_ref(n, gauge.reconstruct12(g["links12"]))
    co = O.coact_array()
    cs = BOUNDARY_CASE
    kap = (cs["kappa"], 0.0)
    R = refdrv.RefLattice(n, bc=(1, 1, 1, -1))
    out = {}
    # even-odd right-hand sides: the even halves of two point sources
    p1e, _ = R.pointsource((1, 1, 1, 1), 1, 1)
    p2e, _ = R.pointsource((3, 2, 4, 5), 2, 3)
    x1, it1 = R.solve(6, p1e, u, co, kap, nk=cs["nk"], resmax=cs["resmax_dr"], isignal=1)
    m, k = cs["nk"]
    xs, it, mvp, gdr = R.gmresproject(p2e, u, co, kap, (m - k, k), cs["resmax_proj"], isignal=2, nshift=2)
    out["proj_phi_e"] = canon_even(n, p2e)
    out["proj_x"] = canon_even(n, xs[0])
    out["proj_x_shift2_absmax"] = np.float64(np.abs(xs[1]).max())
    out["proj_itercount"] = np.int32(it)
    out["proj_mvp"] = np.int32(mvp)
    out["proj_gdr"] = gdr[:max(mvp, 1), 0].copy()
    out["case"] = np.array([cs["kappa"], m, k, cs["resmax_dr"], cs["resmax_proj"]])
    # fermpropR, Wilson r = -1, GMRES-DR(40,10): the config-1 point source
    ps_e, ps_o = R.pointsource((1, 1, 1, 1), 1, 1)
    xe, xo = R.fermpropr(ps_e, ps_o, u, co, kap, 5, -1, nk=cs["nk"], resmax=cs["resmax_dr"], isignal=1)
    out["fermpropR_xe"] = canon_even(n, xe)
    out["fermpropR_xo"] = canon_odd(n, xo)
    np.savez_compressed(os.path.join(HERE, "golden_boundary_8x8x8x16.npz"), **out)
    print("wrote golden_boundary_8x8x8x16.npz: gmresproject itercount", it, "mvp", mvp, "gdr[-1]",
          out["proj_gdr"][-1], "rhs1 iters", it1)


if __name__ == "__main__":
    if sys.argv[1:] == ["boundary"]:
        boundary()
    elif sys.argv[1:] == ["loops"]:
        loops()
    elif sys.argv[1:] == ["noiseloop"]:
        noiseloop()
    elif sys.argv[1:] == ["io"]:
        io()
    elif sys.argv[1:] == ["lagfib"]:
        lagfib()
    elif sys.argv[1:] == ["sweep"]:
        sweep()
    elif sys.argv[1:] == ["smear"]:
        smear()
    else:
        main()
