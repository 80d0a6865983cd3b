"""Benchmark: the Wilson hopping operator (Dslash) on a 32^3x64 lattice.

One step = one application of the hopping term H_oe (V/2 = 1,048,576 output
sites) on the resident gauge field, FP64 storage and arithmetic with
18-real links -- the reference's precision.  Inputs are resident in HBM
before timing starts; the time is the HIP-event average per launch on the
library's stream.  The other storage precisions (FP32, FP16 with per-site
norms; 18- and 12-real links) are measured on the same input and reported in
``sweep`` with their relative error against FP64.

Algorithmic bytes per output site: 8*24*s_psi + 8*L*s_U + 24*s_psi
(L = 18 or 12 link reals); FP16 spinors add 9*4 bytes of per-site norms.

A second section times the solve path on the same lattice: a seeded
beta = 6.0 heatbath field (device heatbath), a point source, twisted-mass
kappa = 0.1565, mu = 0.005, solved to a true full-system residual
||b - M x|| / ||b|| < 1e-10 by (a) the plain FP64 GMRES-DR(40,16) and (b) the
mixed-precision solver (FP32 Krylov basis, FP16 / 12-real operator, FP64
reliable updates); both times, iteration counts and residuals are reported
in ``solve``.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-solve]
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

from baylorlqcd_amd import _lib, diracops, gauge, lattice  # noqa: E402

N = (32, 32, 32, 64)
BC = (1, 1, 1, -1)
HBM_PEAK_GBS = 8000.0


def site_bytes(prec, recon):
    s_psi = {0: 8, 1: 4, 2: 2}[prec]
    s_u = s_psi
    extra = 36 if prec == 2 else 0
    return 8 * 24 * s_psi + 8 * recon * s_u + 24 * s_psi + extra


def gpu_dslash(prec, recon, steps, warmup, which=0):
    ms = np.zeros(1)
    err = np.zeros(1)
    _lib.check(_lib.lib().qqcd_bench_dslash(int(prec), int(recon), int(which), int(warmup), int(steps),
                                            _lib.dp(ms), _lib.dp(err)))
    return float(ms[0]), float(err[0])


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
    b = site_bytes(0, 18) * vh
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


def bench_solve(method, kappa, op_prec=2, recon=12, mk=(40, 16), tol=1e-10):
    sec, res = np.zeros(1), np.zeros(1)
    its, outer = np.zeros(1, np.int32), np.zeros(1, np.int32)
    _lib.check(_lib.lib().qqcd_bench_solve(
        int(method), _lib.ints((1, 1, 1, 1)).ctypes.data_as(_lib._ip), 1, 1, _lib.dp(np.asarray(kappa, float)),
        _lib.ints(mk).ctypes.data_as(_lib._ip), float(tol), int(op_prec), int(recon), 0.0, 0, _lib.dp(sec),
        its.ctypes.data_as(_lib._ip), outer.ctypes.data_as(_lib._ip), _lib.dp(res)))
    return {"seconds": round(float(sec[0]), 4), "iterations": int(its[0]), "outer": int(outer[0]),
            "true_residual": float(f"{res[0]:.3e}")}


def solve_section():
    """Time to a 1e-10 true residual on a 32^3x64 beta=6.0 heatbath field."""
    co = np.zeros(64)
    for mu in range(4):
        co[2 + 4 * (0 + 4 * mu)] = 1.0
        co[2 + 4 * (1 + 4 * mu)] = 1.0
    t0 = time.perf_counter()
    diracops.gauge_heatbath(6.0, 200, 2024, 1, co)
    plaq = diracops.gauge_plaquette()
    t_hb = time.perf_counter() - t0
    print(f"[bench] heatbath 200 sweeps {t_hb:.1f}s, plaquette {plaq:.4f}", file=sys.stderr, flush=True)
    kappa = (0.1565, 0.005)
    plain = bench_solve(0, kappa)
    print(f"[bench] plain FP64 GMRES-DR(40,16): {plain}", file=sys.stderr, flush=True)
    mixed = bench_solve(1, kappa, 2, 12)
    print(f"[bench] mixed FP32 basis / FP16-12 operator: {mixed}", file=sys.stderr, flush=True)
    mixed32 = bench_solve(1, kappa, 1, 12)
    print(f"[bench] mixed FP32 basis / FP32-12 operator: {mixed32}", file=sys.stderr, flush=True)
    return {"lattice": list(N), "beta": 6.0, "heatbath_sweeps": 200, "plaquette": round(plaq, 5),
            "kappa": kappa[0], "mu": kappa[1], "source": "point, origin, spin 1 colour 1",
            "target": "||b - M x|| / ||b|| < 1e-10 (FP64, full system)",
            "plain_fp64_gmresdr_40_16": plain, "mixed_fp16op_fp32basis": mixed, "mixed_fp32op_fp32basis": mixed32,
            "speedup_mixed_vs_plain": round(plain["seconds"] / mixed["seconds"], 2),
            "learned_preconditioner": "not built yet"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--copy-probe", action="store_true", help="also time a 1 GiB plain copy kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus above 1 is not supported "
              "(the multi-GPU decomposition is a design note only)", file=sys.stderr)
        sys.exit(2)

    t_setup = time.perf_counter()
    lattice.init(N, BC)
    co = np.zeros(64)
    for mu in range(4):
        co[2 + 4 * (0 + 4 * mu)] = 1.0
        co[2 + 4 * (1 + 4 * mu)] = 1.0
    diracops.gauge_random(1, 0.1, 2024, co)
    vh = int(np.prod(N)) // 2
    print(f"[bench] setup {time.perf_counter() - t_setup:.1f}s", file=sys.stderr, flush=True)

    ms, err = gpu_dslash(0, 18, args.steps, args.warmup)
    gbs = site_bytes(0, 18) * vh / (ms * 1e-3) / 1e9
    print(f"[bench] FP64/18 dslash {ms*1e3:.1f} us  {gbs:.0f} GB/s", file=sys.stderr, flush=True)

    sweep = []
    if not args.no_sweep:
        for prec, recon in ((0, 12), (1, 18), (1, 12), (2, 18), (2, 12)):
            m, e = gpu_dslash(prec, recon, max(20, args.steps // 2), args.warmup)
            g = site_bytes(prec, recon) * vh / (m * 1e-3) / 1e9
            sweep.append({"storage": {0: "f64", 1: "f32", 2: "f16"}[prec], "links": recon,
                          "us": round(m * 1e3, 2), "GB/s": round(g, 1), "frac": round(g / HBM_PEAK_GBS, 3),
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

    solve = None
    if not args.no_solve:
        solve = solve_section()

    prof = None
    ppath = os.path.join(ROOT, "profiles", "traffic_dslash_f64_18_32x32x32x64.json")
    if os.path.exists(ppath):
        with open(ppath) as f:
            prof = json.load(f).get("bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(diracops.gauge_get_canonical(N), vh)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"value": None, "unit": "GB/s", "cores": 0, "kind": "port", "sample": f"unavailable: {e}"}

    line = {
        "metric": "Dslash HBM GB/s (algorithmic bytes / kernel time), 32^3x64 Wilson hopping term",
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
                   "lattice": list(N), "sites_per_launch": vh, "bytes_per_site": site_bytes(0, 18)},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 3), "traffic": prof},
        "cpu_baseline": cpu,
        "sweep": sweep,
        "solve": solve,
    }
    if copy:
        line["copy_probe"] = copy
    print(json.dumps(line))
    lattice.finalize()


if __name__ == "__main__":
    main()
