"""Benchmark: the Wilson hopping operator (Dslash) on a 32^3x64 lattice.

One step = one application of the hopping term H_oe (V/2 = 1,048,576 output
sites) on the resident gauge field, FP64 storage and arithmetic with
18-real links -- the reference's precision.  Inputs are resident in HBM
before timing starts; the time is the HIP-event average per launch on the
library's stream, the launches cycling over four input/output pairs so that
no launch finds its input in the last-level cache.  The other storage
precisions (FP32, FP16 with per-site norms; 18- and 12-real links) are
measured on the same input and reported in ``sweep`` with their relative
error against FP64.

Unique bytes per output site (the value's model): the input spinor and the
output once and the eight links of the site's hops once,
2*24*s_psi + 8*L*s_U (L = 18 or 12 link reals; FP16 adds 8 bytes of norms);
the neighbour-counted rate (8*24*s_psi + 8*L*s_U + 24*s_psi) is reported
beside it.

``config1`` times config 1 (8^3x16) beside the reference's own solvers.
``solve`` times the solve path on a seeded beta = 6.0 heatbath field (device
heatbath), point source, to a true full-system residual ||b - M x|| / ||b||
< 1e-10: at Wilson kappa = 0.1565 the plain FP64 GMRES-DR(40,16), the
mixed-precision solver (FP32 or FP16 12-real operator, FP32 or FP16 Krylov
basis, FP64 reliable updates) and the mixed solver with the learned
preconditioner (trained first on 16^3x32 heatbath configurations, flexible
GMRES-DR; four held-out fields); at twisted mass kappa = 0.1565, mu = 0.005
plain and mixed.  ``noise_batch`` runs the disconnected-loop workload on
24^3x48: 64 Z4 sources, the first by GMRES-DR(60,20), the rest together by
the batched GMRES-Proj, against the sequential plain FP64 and mixed solves.
``propagator`` solves the twelve spin-colour point sources on 48^3x96
(twisted mass), the first by the mixed solver and the other eleven batched,
against plain FP64 GMRES-DR on the first.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-solve] [--no-noise]
                    [--no-rows] [--no-propagator] [--propagator-learned N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from baylorlqcd_amd import _lib, diracops, disconloops, gauge, inverters, lattice, learned, quark  # noqa: E402
from baylorlqcd_amd.lagfib import LagFib  # noqa: E402

N = (32, 32, 32, 64)
BC = (1, 1, 1, -1)
HBM_PEAK_GBS = 8000.0


def site_bytes(prec, recon):
    """Neighbour-counted bytes per output site: every one of the eight hops
    reads its neighbour spinor and link from memory (caches ignored)."""
    s_psi = {0: 8, 1: 4, 2: 2}[prec]
    s_u = s_psi
    extra = 36 if prec == 2 else 0
    return 8 * 24 * s_psi + 8 * recon * s_u + 24 * s_psi + extra


def unique_bytes(prec, recon):
    """Bytes a launch must move at least, per output site: the input spinor
    (other parity) and the output once, and the eight links of the site's
    hops once (four stored at the site, four at its backward neighbours);
    FP16 spinors add one 4-byte norm each for input and output."""
    s = {0: 8, 1: 4, 2: 2}[prec]
    return 2 * 24 * s + 8 * recon * s + (8 if prec == 2 else 0)


def gpu_dslash(prec, recon, steps, warmup, which=0):
    ms = np.zeros(1)
    err = np.zeros(1)
    _lib.check(_lib.lib().qqcd_bench_dslash(int(prec), int(recon), int(which), int(warmup), int(steps),
                                            _lib.dp(ms), _lib.dp(err)))
    return float(ms[0]), float(err[0])


def gpu_dslash_rot(prec, recon, steps, warmup, nbuf=4):
    """ms per hop launch, cycling over nbuf input/output pairs (caches cold)."""
    ms = np.zeros(1)
    _lib.check(_lib.lib().qqcd_bench_dslash_rot(int(prec), int(recon), int(nbuf), int(warmup), int(steps),
                                                _lib.dp(ms)))
    return float(ms[0])


def cpu_baseline(U, vh):
    """The reference Fortran Hsingle (compiled from source into oracle/_ref)
    timed on one host core; falls back to the C oracle port when no
    reference build travelled with the tree."""
    sys.path.insert(0, ROOT)
    import oracle as O
    from oracle import refdrv
    psi = np.random.default_rng(5).standard_normal((int(np.prod(N)), 4, 3)) * (1 + 0j)
    be, bo = O.spinor_canon_to_ref(N, psi)
    co = O.coact_array()
    b = unique_bytes(0, 18) * vh
    if refdrv.available(N, solvers=False):
        R = refdrv.RefLattice(N, bc=BC, solvers=False)
        u = O.links_canon_to_ref(N, U)
        R.hsingle(u, be, co, 2)  # warm-up
        reps, t0 = 0, time.perf_counter()
        while True:
            R.hsingle(u, be, co, 2)
            reps += 1
            if time.perf_counter() - t0 > 10.0 or reps >= 8:
                break
        dt = (time.perf_counter() - t0) / reps
        return {"value": b / dt / 1e9, "unit": "GB/s", "cores": 1, "kind": "reference",
                "sample": f"{reps} Hsingle(gblclr=2) applications of the reference Fortran on 32^3x64 "
                          f"(FP64, {dt*1e3:.0f} ms each, flang -O3, single process)"}
    nth = int(os.environ.get("OMP_NUM_THREADS", "1"))
    O.hop(N, U, BC, [1] * 4, [1] * 4, 0, 1, psi)
    reps, t0 = 0, time.perf_counter()
    while True:
        O.hop(N, U, BC, [1] * 4, [1] * 4, 0, 1, psi)
        reps += 1
        if time.perf_counter() - t0 > 10.0 or reps >= 8:
            break
    dt = (time.perf_counter() - t0) / reps
    return {"value": b / dt / 1e9, "unit": "GB/s", "cores": nth, "kind": "port",
            "sample": f"{reps} applications of the C oracle hopping term on 32^3x64 ({dt*1e3:.0f} ms each)"}


def cpu_sweep_baseline():
    """The reference's own sweep (oracle/_ref, 8^3x16 build, one core): link updates per second."""
    sys.path.insert(0, ROOT)
    import oracle as O
    from oracle import refdrv
    n = (8, 8, 8, 16)
    if not refdrv.available(n, solvers=True):
        return None
    R = refdrv.RefLattice(n, bc=(1, 1, 1, 1))
    u = O.links_canon_to_ref(n, gauge.haar(n, 7))
    co = np.zeros(64)
    for mu in range(4):
        for nu in range(4):
            if mu != nu:
                co[4 * mu + 16 * nu] = 6.0
    R.lagfib_init(0, 1)
    u = R.sweeps(u, co, 1)
    reps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 3.0:
        u = R.sweeps(u, co, 1)
        reps += 1
    dt = (time.perf_counter() - t0) / reps
    return {"value": round(4 * int(np.prod(n)) / dt / 1e6, 2), "unit": "Mlink updates/s", "cores": 1,
            "kind": "reference", "sample": f"{reps} sweeps of the reference Fortran on 8^3x16 ({dt*1e3:.1f} ms each)"}


def cpu_rows_baselines():
    """The reference's own smear and rnds (oracle/_ref, 8^3x16 build, one core), about two seconds each:
    smearing in GB/s algorithmic (2400 B per site and iteration, as the GPU row counts it), draws per second."""
    sys.path.insert(0, ROOT)
    import oracle as O
    from oracle import refdrv
    n = (8, 8, 8, 16)
    if not refdrv.available(n, solvers=True):
        return None, None
    R = refdrv.RefLattice(n, bc=(1, 1, 1, -1))
    V = int(np.prod(n))
    u = O.links_canon_to_ref(n, gauge.haar(n, 7))
    rng = np.random.default_rng(5)
    ge, go = rng.standard_normal(R.slen), rng.standard_normal(R.slen)
    nsm = 20
    R.smear(u, ge, go, 0.25, nsm)
    reps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 2.0:
        R.smear(u, ge, go, 0.25, nsm)
        reps += 1
    dt = (time.perf_counter() - t0) / (reps * nsm)
    smear = {"value": round(2400 * V / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "reference",
             "sample": f"{reps * nsm} smearing iterations of the reference Fortran on 8^3x16 ({dt*1e3:.2f} ms each)"}
    R.lagfib_init(0, 1)
    m = 1 << 22
    R.rnds(m)
    reps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 2.0:
        R.rnds(m)
        reps += 1
    dt = (time.perf_counter() - t0) / reps
    lag = {"value": round(m / dt / 1e9, 4), "unit": "Gdraws/s", "cores": 1, "kind": "reference",
           "sample": f"{reps} calls of the reference's rnds for 2^22 numbers ({dt*1e3:.1f} ms each)"}
    return smear, lag


def rows_section(no_cpu=False):
    """The other rows of the path on 32^3x64: Gaussian smearing, the reference's lagged-Fibonacci
    stream and the reference's heatbath sweep drawing from it."""
    V = int(np.prod(N))
    L = _lib.lib()
    # HBM bytes per launch from the hardware counters (profiles/traffic_smear_lagfib_32x32x32x64.md)
    traffic = {}
    tpath = os.path.join(ROOT, "profiles", "traffic_smear_lagfib_32x32x32x64.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)
    ms = np.zeros(1)
    _lib.check(L.qqcd_bench_smear(3, 20, _lib.dp(ms)))
    # unique bytes per output site and checkerboard launch: the site's own and the other
    # checkerboard's input spinor, the six spatial links of its hops and the output once (FP64)
    g = 1440 * V / (ms[0] * 1e-3) / 1e9
    smear = {"us_per_iteration": round(ms[0] * 1e3, 1), "bytes_per_site": 1440, "bytes_model": "unique",
             "GB/s": round(g, 1), "frac": round(g / HBM_PEAK_GBS, 3),
             "neighbour_counted_GB/s": round(2400 * V / (ms[0] * 1e-3) / 1e9, 1),
             "traffic_per_launch": traffic.get("k_smear", {}).get("bytes_per_launch")}
    print(f"[bench] smear {smear}", file=sys.stderr, flush=True)
    m = 2048 * 32768
    _lib.check(L.qqcd_bench_lagfib(m, 10, 0, _lib.dp(ms)))
    lag = {"draws": m, "us": round(ms[0] * 1e3, 1), "Gdraws/s": round(m / (ms[0] * 1e-3) / 1e9, 2),
           "GB/s_written": round(4 * m / (ms[0] * 1e-3) / 1e9, 1), "frac": round(4 * m / (ms[0] * 1e-3) / 1e9 / HBM_PEAK_GBS, 3),
           "traffic_per_launch": traffic.get("k_lag_gen", {}).get("bytes_per_launch")}
    print(f"[bench] lagfib {lag}", file=sys.stderr, flush=True)
    u = diracops.gauge_get_ref()
    co = np.zeros(64)
    for mu in range(4):
        for nu in range(4):
            if mu != nu:
                co[4 * mu + 16 * nu] = 6.0
    st = LagFib(0, 1)
    t = []
    for nsw in (1, 3):
        t0 = time.perf_counter()
        st.sweep(u, co, nsw)
        t.append(time.perf_counter() - t0)
    per = (t[1] - t[0]) / 2
    sweep = {"s_per_sweep": round(per, 3), "Mlink_updates/s": round(4 * V / per / 1e6, 1),
             "call_with_transfers_s": round(t[0], 3), "stream": "the reference's lagfib, initlagfib(0, 1)"}
    if not no_cpu:
        try:
            sweep["cpu_baseline"] = cpu_sweep_baseline()
        except Exception as e:  # reported, never required
            sweep["cpu_baseline"] = {"value": None, "sample": f"unavailable: {e}"}
        try:
            smear["cpu_baseline"], lag["cpu_baseline"] = cpu_rows_baselines()
        except Exception as e:  # reported, never required
            smear["cpu_baseline"] = lag["cpu_baseline"] = {"value": None, "sample": f"unavailable: {e}"}
        print(f"[bench] smear / lagfib cpu baselines {smear['cpu_baseline']} {lag['cpu_baseline']}",
              file=sys.stderr, flush=True)
    print(f"[bench] sweep {sweep}", file=sys.stderr, flush=True)
    return {"smear": smear, "lagfib": lag, "sweep_reference_order": sweep}


def bench_solve(method, kappa, op_prec=2, recon=12, mk=(40, 16), tol=1e-10):
    sec, res = np.zeros(1), np.zeros(1)
    its, outer = np.zeros(1, np.int32), np.zeros(1, np.int32)
    _lib.check(_lib.lib().qqcd_bench_solve(
        int(method), _lib.ints((1, 1, 1, 1)).ctypes.data_as(_lib._ip), 1, 1, _lib.dp(np.asarray(kappa, float)),
        _lib.ints(mk).ctypes.data_as(_lib._ip), float(tol), int(op_prec), int(recon), 0.0, 0, _lib.dp(sec),
        its.ctypes.data_as(_lib._ip), outer.ctypes.data_as(_lib._ip), _lib.dp(res)))
    return {"seconds": round(float(sec[0]), 4), "iterations": int(its[0]), "outer": int(outer[0]),
            "ms_per_iteration": round(1e3 * float(sec[0]) / max(1, int(its[0])), 4),
            "true_residual": float(f"{res[0]:.3e}")}


def heatbath_field(seed=2024, sweeps=200):
    co = np.zeros(64)
    for mu in range(4):
        co[2 + 4 * (0 + 4 * mu)] = 1.0
        co[2 + 4 * (1 + 4 * mu)] = 1.0
    t0 = time.perf_counter()
    diracops.gauge_heatbath(6.0, sweeps, seed, 1, co)
    plaq = diracops.gauge_plaquette()
    print(f"[bench] heatbath {sweeps} sweeps {time.perf_counter() - t0:.1f}s, plaquette {plaq:.4f}",
          file=sys.stderr, flush=True)
    return co, plaq


_CPU_SOLVE = []


def cpu_solve_baseline():
    """The reference's own gmresdr (oracle/_ref, 8^3x16 build, one core) on the config-1 field (the golden
    fixture), Wilson kappa 0.150, point source, GMRES-DR(40,16) to 1e-10, repeated for about three seconds:
    even-site Schur iterations x sites per second, the rate the GPU rows are compared with.  Run in a child
    process: the reference's gmresdr writes its own progress lines to standard output."""
    if _CPU_SOLVE:
        return _CPU_SOLVE[0]
    import subprocess
    code = "import json, sys, bench; sys.stderr.write('CPU_SOLVE ' + json.dumps(bench._cpu_solve_child()) + '\\n')"
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=300)
    lines = [ln for ln in out.stderr.splitlines() if ln.startswith("CPU_SOLVE ")]
    if out.returncode != 0 or not lines:
        raise RuntimeError(out.stderr.strip().splitlines()[-1] if out.stderr.strip() else "child failed")
    _CPU_SOLVE.append(json.loads(lines[-1][len("CPU_SOLVE "):]))
    return _CPU_SOLVE[0]


def _cpu_solve_child():
    sys.path.insert(0, ROOT)
    import oracle as O
    from oracle import refdrv
    n = (8, 8, 8, 16)
    if not refdrv.available(n, solvers=True):
        return None
    R = refdrv.RefLattice(n, bc=(1, 1, 1, -1))
    g = np.load(os.path.join(ROOT, "tests", "golden", "golden_8x8x8x16.npz"))
    u = O.links_canon_to_ref(n, gauge.reconstruct12(g["links12"]))
    co = O.coact_array()
    be, _ = R.pointsource((1, 1, 1, 1), 1, 1)
    its, reps, t0 = 0, 0, time.perf_counter()
    while time.perf_counter() - t0 < 3.0 or reps == 0:
        _, it = R.solve(5, be, u, co, (0.150, 0.0), nk=(40, 16), resmax=1e-10)
        its += it
        reps += 1
    dt = time.perf_counter() - t0
    vh = int(np.prod(n)) // 2
    cfg1 = {"cores": 1, "kind": "reference"}
    for name, which, nk in (("gmresdr_40_10", 5, (40, 10)), ("bicgstab", 0, (40, 10))):
        t1 = time.perf_counter()
        _, it1 = R.solve(which, be, u, co, (0.150, 0.0), nk=nk, resmax=1e-10)
        cfg1[name] = {"seconds": round(time.perf_counter() - t1, 4), "iterations": it1}
    return {"value": round(vh * its / dt / 1e6, 3), "unit": "M site-iterations/s", "cores": 1, "kind": "reference",
            "sample": f"{reps} GMRES-DR(40,16) solves of the reference Fortran on 8^3x16, {its // reps} iterations "
                      f"each ({dt / reps:.2f} s each)", "config1": cfg1}


def config1_section(no_cpu=False):
    """Config 1 beside the reference: 8^3x16, the golden fixture's beta = 6.0
    heatbath field (200 of the reference's own sweeps), Wilson kappa 0.150, the
    even half of the point source at the origin (spin 1, colour 1) as the
    even-odd right-hand side, GMRES-DR(40,10) and BiCGStab to the reference's
    relative tolerance 1e-10; the GPU drop-ins timed by the host clock around
    the call (transfers included, median of 5) against the reference's own
    gmresdr and bicgstab on one host core (oracle/_ref, child process)."""
    lattice.finalize()
    n = (8, 8, 8, 16)
    lattice.init(n, BC)
    g = np.load(os.path.join(ROOT, "tests", "golden", "golden_8x8x8x16.npz"))
    co = np.zeros(64)
    for mu in range(4):
        co[2 + 4 * (0 + 4 * mu)] = 1.0
        co[2 + 4 * (1 + 4 * mu)] = 1.0
    diracops.gauge_load_canonical(gauge.reconstruct12(g["links12"]), co)
    u = diracops.gauge_get_ref(n)
    be = lattice.new_spinor(n).ravel(order="F")
    bo = lattice.new_spinor(n).ravel(order="F")
    quark.pointsource(be, bo, (1, 1, 1, 1), 1, 1)
    kap = (0.150, 0.0)
    out = {"lattice": list(n), "kappa": kap[0], "source": "point, origin, spin 1 colour 1 (even half)",
           "tolerance": "1e-10 relative to the initial even-odd residual (the reference's test)"}
    for name, call in (("gmresdr_40_10", lambda x: inverters.gmresdr(None, be, x, (40, 10), 1e-10, 0, u, None, -1,
                                                                        kap, co, BC)),
                       ("bicgstab", lambda x: inverters.bicgstab(None, be, x, 1e-10, 0, u, None, -1, kap, co, BC))):
        ts, its = [], 0
        for _ in range(5):
            x = np.zeros_like(be)
            _lib.check(_lib.lib().qqcd_deflation_clear())
            t0 = time.perf_counter()
            its = call(x)
            ts.append(time.perf_counter() - t0)
        st = inverters.last_solve_stats()
        out[name] = {"gpu_seconds": round(float(np.median(ts)), 4), "iterations": its,
                     "relative_residual": float(f"{st['rel_res']:.3e}")}
    if not no_cpu:
        try:
            cpu = cpu_solve_baseline()
            out["cpu_baseline"] = cpu.get("config1") if cpu else None
        except Exception as e:  # reported, never required
            out["cpu_baseline"] = {"sample": f"unavailable: {e}"}
        ref = out.get("cpu_baseline") or {}
        for name in ("gmresdr_40_10", "bicgstab"):
            if name in ref and ref[name].get("seconds"):
                out[name]["reference_cpu_seconds"] = ref[name]["seconds"]
                out[name]["speedup_vs_reference"] = round(ref[name]["seconds"] / out[name]["gpu_seconds"], 1)
    print(f"[bench] config 1: {out}", file=sys.stderr, flush=True)
    lattice.finalize()
    return out


def solve_section(train_steps=1500, no_cpu=False):
    """Time to a 1e-10 true residual on a 32^3x64 beta=6.0 heatbath field:
    plain FP64 GMRES-DR, the mixed solver, and the mixed solver with the
    learned preconditioner trained on 16^3x32 configurations (Wilson kappa in
    [0.150, 0.157], pairs from plain GMRES-DR iterates), at Wilson kappa
    0.1565; plus the twisted-mass point kappa 0.1565, mu 0.005."""
    train = train_network(train_steps)
    lattice.init(N, BC)
    co, plaq = heatbath_field()
    nn_ms = learned.bench_nn(2, 10)
    print(f"[bench] network application {nn_ms:.3f} ms", file=sys.stderr, flush=True)
    wil = (0.1565, 0.0)
    plain = learned.bench(0, wil, GMRES=(40, 16))
    print(f"[bench] Wilson plain FP64 GMRES-DR(40,16): {plain}", file=sys.stderr, flush=True)
    mixed = learned.bench(1, wil, GMRES=(40, 16), op_prec=1, op_recon=12)
    print(f"[bench] Wilson mixed FP32: {mixed}", file=sys.stderr, flush=True)
    mixed16 = learned.bench(1, wil, GMRES=(40, 16), op_prec=2, op_recon=12)
    print(f"[bench] Wilson mixed FP16 operator: {mixed16}", file=sys.stderr, flush=True)
    mixed16h = bench_solve(2, wil, 2, 12)
    print(f"[bench] Wilson mixed FP16 operator, FP16 basis: {mixed16h}", file=sys.stderr, flush=True)
    lrn = learned.bench(2, wil, GMRES=(40, 16), op_prec=1, op_recon=12)
    print(f"[bench] Wilson mixed FP32 + learned: {lrn}", file=sys.stderr, flush=True)
    tm = (0.1565, 0.005)
    tplain = bench_solve(0, tm)
    print(f"[bench] twisted plain FP64: {tplain}", file=sys.stderr, flush=True)
    tmixed = bench_solve(1, tm, 1, 12)
    print(f"[bench] twisted mixed FP32: {tmixed}", file=sys.stderr, flush=True)
    # three more held-out 32^3x64 fields (four with the first): the mixed FP32 solve with and without
    # the network on each
    held = [{"seed": 2024, "plaquette": round(plaq, 5), "mixed_fp32": mixed, "mixed_fp32_learned": lrn}]
    for seed in (2025, 2026, 2027):
        _, pq = heatbath_field(seed)
        hm = learned.bench(1, wil, GMRES=(40, 16), op_prec=1, op_recon=12)
        hl = learned.bench(2, wil, GMRES=(40, 16), op_prec=1, op_recon=12)
        held.append({"seed": seed, "plaquette": round(pq, 5), "mixed_fp32": hm, "mixed_fp32_learned": hl})
        print(f"[bench] held-out field {seed}: mixed {hm} learned {hl}", file=sys.stderr, flush=True)
    for h in held:
        h["iteration_ratio"] = round(h["mixed_fp32_learned"]["iterations"] / max(1, h["mixed_fp32"]["iterations"]), 3)
        h["time_ratio"] = round(h["mixed_fp32_learned"]["seconds"] / h["mixed_fp32"]["seconds"], 3)
    best = min((mixed, mixed16, mixed16h, lrn), key=lambda r: r["seconds"])
    vh = int(np.prod(N)) // 2
    rates = {"plain_fp64": round(vh * plain["iterations"] / plain["seconds"] / 1e6, 1),
             "mixed_fp16_operator_fp16_basis": round(vh * mixed16h["iterations"] / mixed16h["seconds"] / 1e6, 1)}
    cpu = None
    if not no_cpu:
        try:
            cpu = cpu_solve_baseline()
        except Exception as e:  # reported, never required
            cpu = {"value": None, "sample": f"unavailable: {e}"}
        print(f"[bench] solve cpu baseline {cpu}", file=sys.stderr, flush=True)
    return {"lattice": list(N), "beta": 6.0, "heatbath_sweeps": 200, "plaquette": round(plaq, 5),
            "M_site_iterations_per_s": rates, "cpu_baseline": cpu,
            "source": "point, origin, spin 1 colour 1",
            "target": "||b - M x|| / ||b|| < 1e-10 (FP64, full system)",
            "wilson_kappa_0.1565": {"plain_fp64_gmresdr_40_16": plain, "mixed_fp32": mixed,
                                    "mixed_fp16_operator": mixed16, "mixed_fp16_operator_fp16_basis": mixed16h,
                                    "mixed_fp32_learned": lrn,
                                    "speedup_best_vs_plain": round(plain["seconds"] / best["seconds"], 2),
                                    "speedup_learned_vs_plain": round(plain["seconds"] / lrn["seconds"], 2),
                                    "iteration_ratio_learned_vs_mixed": round(lrn["iterations"] /
                                                                               max(1, mixed["iterations"]), 3)},
            "twisted_kappa_0.1565_mu_0.005": {"plain_fp64_gmresdr_40_16": tplain, "mixed_fp32": tmixed,
                                              "speedup_mixed_vs_plain": round(tplain["seconds"] /
                                                                              tmixed["seconds"], 2)},
            "learned_preconditioner": {"training": train, "network_ms_per_application": round(nn_ms, 3),
                                       "held_out_fields_kappa_0.1565": held,
                                       "layers": "4 parallel-transport convolutions, channels 1-8-8-8-1, "
                                                 "BF16 MFMA, FP32 accumulation"}}


PROFILE_SOLVES = {"plain": (0, 0), "fp32": (1, 1), "fp16": (1, 2), "fp16basis": (2, 2)}


def profile_solve_section(which):
    """One solve of the solve section alone (no training), for kernel traces:
    the 32^3x64 heatbath field, Wilson kappa 0.1565, point source, 1e-10."""
    lattice.finalize()
    lattice.init(N, BC)
    heatbath_field()
    method, op_prec = PROFILE_SOLVES[which]
    r = bench_solve(method, (0.1565, 0.0), op_prec, 12 if op_prec else 18)
    print(f"[bench] profile solve {which}: {r}", file=sys.stderr, flush=True)
    return {"which": which, **r}


def train_network(train_steps):
    """The learned preconditioner trained on ten 16^3x32 beta=6.0 heatbath
    configurations (Wilson kappa uniform in [0.150, 0.157]) on 20,000
    (residual, error) pairs from every Arnoldi step of plain GMRES-DR(40,16)
    solves; returns the training record."""
    lattice.finalize()
    lattice.init((16, 16, 16, 32), BC)
    t0 = time.perf_counter()
    tr = learned.train(nconf=10, beta=6.0, nsweep=200, kappa=(0.150, 0.157), nsrc=24, m=40, k=16, harvest_tol=1e-8,
                       max_pairs=20000, channels=(1, 8, 8, 8, 1), steps=train_steps, batch=8, lr=1e-3, seed=11)
    train = {"lattice": [16, 16, 16, 32], "configs": 10, "pairs": tr["pairs"], "adam_steps": train_steps,
             "batch": 8, "lr": 1e-3, "loss_before": round(float(tr["loss_before"]), 4),
             "loss_after": round(float(tr["loss_after"]), 4),
             "validation_solve": {"start": tr["validation_start"], "kept": tr["validation_kept"]},
             "harvest_s": round(float(tr["harvest_s"]), 2),
             "train_s": round(float(tr["train_s"]), 2), "wall_s": round(time.perf_counter() - t0, 2)}
    print(f"[bench] training {train}", file=sys.stderr, flush=True)
    lattice.finalize()
    return train


def propagator_section(train_steps=1500, n5=(48, 48, 48, 96), nlearned=0, trained=None):
    """The full propagator at scale: a 48^3x96 beta=6.0 heatbath field,
    twisted mass kappa 0.1563, mu 0.0015, the 12 spin-colour point sources at
    the origin to a 1e-10 true residual each.  The first by the mixed solver
    (FP16 storage, 12-real links, FP32 basis, GMRES-DR(40,16) inner solves,
    FP64 reliable updates), the other 11 batched through GMRES-Proj(24) over
    its deflation space; against plain FP64 GMRES-DR(40,16) on the first
    source; and the learned preconditioner on the first `nlearned` sources."""
    # the network the solve section trained stays loaded in the library
    train = trained if trained is not None or nlearned == 0 else train_network(train_steps)
    lattice.finalize()
    lattice.init(n5, BC)
    co, plaq = heatbath_field()
    tm = (0.1563, 0.0015)
    plain = learned.bench(0, tm, GMRES=(40, 16))
    print(f"[bench] propagator: plain FP64 GMRES-DR(40,16), source 1: {plain}", file=sys.stderr, flush=True)
    b = quark.propagator(12, (1, 1, 1, 1), tm, GMRES=(40, 16), tol=1e-10, op_prec=2, op_recon=12, method=0)
    print(f"[bench] propagator batched: iters {b['iters']} res {b['res_full'].max():.2e} times {b['times']} "
          f"stats {b['stats']}", file=sys.stderr, flush=True)
    lr = None
    if nlearned > 0:
        lr = quark.propagator(nlearned, (1, 1, 1, 1), tm, GMRES=(40, 16), tol=1e-10, op_prec=2, op_recon=12,
                              method=2)
        print(f"[bench] propagator learned: iters {lr['iters']} res {lr['res_full'].max():.2e} times {lr['times']} "
              f"stats {lr['stats']}", file=sys.stderr, flush=True)
    return {"lattice": list(n5), "beta": 6.0, "plaquette": round(plaq, 5), "kappa": tm[0], "mu": tm[1],
            "sources": "12 spin-colour point sources at the origin", "target": "1e-10 true residual each",
            "plain_fp64_gmresdr_40_16_source1": plain,
            "mixed_fp16_12_batched": {"iterations": [int(v) for v in b["iters"]],
                                      "max_true_residual": float(b["res_full"].max()),
                                      "seconds_first": round(float(b["times"][0]), 3),
                                      "seconds_other_11": round(float(b["times"][1]), 3),
                                      "seconds_total": round(float(b["times"][2]), 3), "stats": b["stats"]},
            "learned_first_sources": None if lr is None else {
                "iterations": [int(v) for v in lr["iters"]], "max_true_residual": float(lr["res_full"].max()),
                "seconds": [round(float(v), 3) for v in lr["times"]], "stats": lr["stats"]},
            "plain_seconds_x12_estimate": round(12 * plain["seconds"], 1),
            "speedup_vs_plain_x12": round(12 * plain["seconds"] / float(b["times"][2]), 2),
            "training": train}


def noise_section(nrhs=64, seq=3, graph=1, fp32_basis=False):
    """The disconnected-loop workload: 24^3x48 beta=6.0 heatbath field,
    Wilson kappa 0.1565, nrhs Z4 sources, GMRES-DR(60,20) on the first and
    the batched GMRES-Proj(40) over its deflation space on the rest (Krylov
    basis in FP16, then again in FP32), every source to a 1e-10 true
    residual, 16 local trace estimators; against the same sources one after
    the other in plain FP64 (psc sequence), timed on `seq` sources and
    scaled to nrhs - 1."""
    lattice.finalize()
    n3 = (24, 24, 24, 48)
    lattice.init(n3, BC)
    co, plaq = heatbath_field()
    kappa = (0.1565, 0.0)
    disconloops.batch_set(use_graph=graph)
    r = disconloops.noise_loop(nrhs, 99, kappa, (60, 20), tol=1e-10, coact=co, bc=BC)
    print(f"[bench] batched noise loop: times {r['times']} stats {r['stats']}", file=sys.stderr, flush=True)
    h = None
    if fp32_basis:
        disconloops.batch_set(use_graph=graph, basis_prec=1)
        h = disconloops.noise_loop(nrhs, 99, kappa, (60, 20), tol=1e-10, coact=co, bc=BC)
        disconloops.batch_set(use_graph=graph)
        print(f"[bench] batched noise loop, FP32 basis: times {h['times']} stats {h['stats']}", file=sys.stderr,
              flush=True)
    s = disconloops.noise_loop(seq, 99, kappa, (60, 20), tol=1e-10, coact=co, bc=BC, method=1)
    per = s["times"][1] / max(1, seq - 1)
    print(f"[bench] sequential plain FP64: times {s['times']}", file=sys.stderr, flush=True)
    m = disconloops.noise_loop(seq, 99, kappa, (60, 20), tol=1e-10, coact=co, bc=BC, method=2)
    perm = m["times"][1] / max(1, seq - 1)
    print(f"[bench] sequential mixed: times {m['times']}", file=sys.stderr, flush=True)
    nb, order, tms = disconloops.autotune_multi(1, 12, nrhs - 1)
    return {"lattice": list(n3), "kappa": kappa[0], "sources": nrhs, "noise": "Z4 {1,-1,i,-i}",
            "plaquette": round(plaq, 5),
            "batched": {"first_source_s": round(r["times"][0], 3), "batched_s": round(r["times"][1], 3),
                        "total_s": round(r["times"][2], 3), "max_rel_residual": float(np.max(r["res_full"])),
                        "iterations_per_source": int(np.median(r["iters"][1:])) if nrhs > 1 else None,
                        "stats": r["stats"]},
            "batched_fp32_basis": None if h is None else {
                "batched_s": round(h["times"][1], 3), "total_s": round(h["times"][2], 3),
                "max_rel_residual": float(np.max(h["res_full"])),
                "iterations_per_source": int(np.median(h["iters"][1:])) if nrhs > 1 else None,
                "traces_max_rel_diff_vs_fp16_basis":
                    float(np.max(np.abs(h["traces"] - r["traces"])) / np.max(np.abs(r["traces"]))),
                "speedup_vs_sequential_mixed": round(perm * (nrhs - 1) / h["times"][1], 2),
                "stats": h["stats"]},
            "autotune": {"rhs_block": nb, "tile_order": order, "ms_per_hop": round(tms, 4)},
            "sequential_plain_fp64_s_per_source": round(per, 3),
            "sequential_mixed_s_per_source": round(perm, 3),
            "speedup_batched_vs_sequential_plain": round(per * (nrhs - 1) / r["times"][1], 2),
            "speedup_batched_vs_sequential_mixed": round(perm * (nrhs - 1) / r["times"][1], 2),
            "traces_timeslice0_gamma5": [float(r["traces"][0, 5].real), float(r["traces"][0, 5].imag)]}


def multi_probe(spec):
    """One 1 GiB copy (the counters' calibration) and four launches of the
    63-slot multi-RHS hop at 24^3x48 (FP32 spinors, 12-real links, RHS block
    8, block outermost) in launch form KERNEL with bz BZ."""
    kernel, bz = (int(v) for v in spec.split(":"))
    lattice.init((24, 24, 24, 48), BC)
    co = np.zeros(64)
    for mu in range(4):
        co[2 + 4 * (0 + 4 * mu)] = 1.0
        co[2 + 4 * (1 + 4 * mu)] = 1.0
    diracops.gauge_random(1, 0.1, 2024, co)
    L = _lib.lib()
    ms = np.zeros(1)
    _lib.check(L.qqcd_bench_copy(float(1 << 30), 3, _lib.dp(ms)))
    copy_ms = ms[0]
    _lib.check(L.qqcd_set_multi_tune(kernel, bz, 0))
    _lib.check(L.qqcd_bench_dslash_multi(1, 12, 63, 63, 8, 0, 1, 3, _lib.dp(ms)))
    _lib.check(L.qqcd_set_multi_tune(-1, -1, -1))
    print(json.dumps({"probe": "multi-RHS hop, 24x24x24x48, 63 slots, FP32/12, nb 8", "kernel": kernel, "bz": bz,
                      "copy_ms": round(copy_ms, 4), "hop_ms": round(ms[0], 4)}), flush=True)
    lattice.finalize()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-rows", action="store_true", help="skip smearing, the lagfib stream and the reference sweep")
    ap.add_argument("--no-noise", action="store_true", help="skip the 24^3x48 batched noise-source section")
    ap.add_argument("--train-steps", type=int, default=1500)
    ap.add_argument("--noise-rhs", type=int, default=64)
    ap.add_argument("--no-graph", action="store_true",
                    help="no HIP graph replays in the batched and the single-RHS solvers (kernel traces)")
    ap.add_argument("--copy-probe", action="store_true", help="also time a 1 GiB plain copy kernel")
    ap.add_argument("--multi-probe", metavar="KERNEL:BZ",
                    help="only: a 1 GiB copy, then the 63-slot multi-RHS hop at 24^3x48 (FP32/12, RHS block 8) "
                         "in the given launch form (hardware-counter runs)")
    ap.add_argument("--profile-solve", choices=sorted(PROFILE_SOLVES),
                    help="also run one 32^3x64 solve alone (for kernel traces)")
    ap.add_argument("--no-propagator", action="store_true",
                    help="skip the 48^3x96 twisted-mass 12-source propagator (about 90 s)")
    ap.add_argument("--propagator-learned", type=int, default=0,
                    help="also solve the first N propagator sources with the learned preconditioner (about a minute each)")
    ap.add_argument("--noise-fp32-basis", action="store_true",
                    help="also run the batched noise loop with an FP32 Krylov basis")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus above 1 is not supported "
              "(the multi-GPU decomposition is a design note only)", file=sys.stderr)
        sys.exit(2)
    if args.multi_probe:
        multi_probe(args.multi_probe)
        return
    if args.no_graph:
        _lib.check(_lib.lib().qqcd_set_step_graphs(0))

    t_setup = time.perf_counter()
    lattice.init(N, BC)
    co = np.zeros(64)
    for mu in range(4):
        co[2 + 4 * (0 + 4 * mu)] = 1.0
        co[2 + 4 * (1 + 4 * mu)] = 1.0
    diracops.gauge_random(1, 0.1, 2024, co)
    vh = int(np.prod(N)) // 2
    print(f"[bench] setup {time.perf_counter() - t_setup:.1f}s", file=sys.stderr, flush=True)

    # the headline: launches cycled over 4 input/output pairs (1.6 GB of spinors), so no launch finds its
    # fields in the 256 MB last-level cache; the same-buffer time is reported beside it
    ms = gpu_dslash_rot(0, 18, args.steps, args.warmup)
    ms_hot, err = gpu_dslash(0, 18, args.steps, args.warmup)
    gbs = unique_bytes(0, 18) * vh / (ms * 1e-3) / 1e9
    print(f"[bench] FP64/18 dslash {ms*1e3:.1f} us (same buffers {ms_hot*1e3:.1f} us)  {gbs:.0f} GB/s unique",
          file=sys.stderr, flush=True)

    sweep = []
    if not args.no_sweep:
        for prec, recon in ((0, 12), (1, 18), (1, 12), (2, 18), (2, 12)):
            m = gpu_dslash_rot(prec, recon, max(20, args.steps // 2), args.warmup)
            mh, e = gpu_dslash(prec, recon, max(20, args.steps // 2), args.warmup)
            g = unique_bytes(prec, recon) * vh / (m * 1e-3) / 1e9
            sweep.append({"storage": {0: "f64", 1: "f32", 2: "f16"}[prec], "links": recon,
                          "us": round(m * 1e3, 2), "us_same_buffers": round(mh * 1e3, 2),
                          "unique_bytes_per_site": unique_bytes(prec, recon),
                          "GB/s": round(g, 1), "frac": round(g / HBM_PEAK_GBS, 3),
                          "neighbour_counted_GB/s": round(site_bytes(prec, recon) * vh / (m * 1e-3) / 1e9, 1),
                          "relerr_vs_f64": float(f"{e:.3g}")})
            print(f"[bench] {sweep[-1]}", file=sys.stderr, flush=True)
        m2, e2 = gpu_dslash(0, 18, max(20, args.steps // 2), args.warmup, which=1)
        sweep.append({"op": "Hdbletm (2 hops + twisted epilogue), per hop", "storage": "f64", "links": 18,
                      "us": round(m2 * 1e3, 2)})

    copy = None
    if args.copy_probe:
        cms = np.zeros(1)
        _lib.check(_lib.lib().qqcd_bench_copy(float(1 << 30), 20, _lib.dp(cms)))
        copy = {"bytes_per_launch": 2 * (1 << 30), "us": round(cms[0] * 1e3, 1),
                "GB/s": round(2 * (1 << 30) / (cms[0] * 1e-3) / 1e9, 1)}
        print(f"[bench] copy {copy}", file=sys.stderr, flush=True)

    prof, prof_src = None, None
    ppath = os.path.join(ROOT, "profiles", "traffic_dslash_f64_18_32x32x32x64.json")
    if os.path.exists(ppath):
        with open(ppath) as f:
            pj = json.load(f)
        prof = pj.get("bytes_per_launch")
        prof_src = (f"rocprofv3 FETCH_SIZE/WRITE_SIZE of {pj.get('kernel')}, measured on commit "
                    f"{pj.get('measured_on_commit')} (profiles/traffic_dslash_f64_18_32x32x32x64.md)")

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(diracops.gauge_get_canonical(N), vh)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"value": None, "unit": "GB/s", "cores": 0, "kind": "port", "sample": f"unavailable: {e}"}

    rows = None
    if not args.no_rows:
        rows = rows_section(args.no_cpu)

    config1 = None
    if not args.no_solve:
        config1 = config1_section(args.no_cpu)
    solve = None
    if not args.no_solve:
        solve = solve_section(args.train_steps, args.no_cpu)
    noise = None
    if not args.no_noise:
        noise = noise_section(args.noise_rhs, graph=0 if args.no_graph else 1, fp32_basis=args.noise_fp32_basis)
    psolve = None
    if args.profile_solve:
        psolve = profile_solve_section(args.profile_solve)
    propagator = None
    if not args.no_propagator:
        propagator = propagator_section(args.train_steps, nlearned=args.propagator_learned,
                                        trained=solve["learned_preconditioner"]["training"] if solve else None)

    line = {
        "metric": "Dslash HBM GB/s (unique bytes / kernel time), 32^3x64 Wilson hopping term",
        "value": round(gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: reunitarised 1+0.1*Gaussian SU(3) links (device generator, seed 2024), Gaussian spinors (seed 12345)",
        "config": {"workload": "hopping term H_oe, 32x32x32x64, bc (1,1,1,-1), FP64 spinors, 18-real links",
                   "lattice": list(N), "sites_per_launch": vh, "bytes_per_site": unique_bytes(0, 18),
                   "buffers": "4 input/output pairs in turn"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 3), "traffic": prof, "traffic_source": prof_src,
                     "bytes_model": "unique: input and output spinor and the eight links once per output site",
                     "neighbour_counted_GB/s": round(site_bytes(0, 18) * vh / (ms * 1e-3) / 1e9, 1),
                     "us_same_buffers": round(ms_hot * 1e3, 2)},
        "cpu_baseline": cpu,
        "sweep": sweep,
        "config1": config1,
        "solve": solve,
        "noise_batch": noise,
        "rows": rows,
    }
    if copy:
        line["copy_probe"] = copy
    if propagator:
        line["propagator"] = propagator
    if psolve:
        line["profile_solve"] = psolve
    print(json.dumps(line))
    lattice.finalize()


if __name__ == "__main__":
    main()
