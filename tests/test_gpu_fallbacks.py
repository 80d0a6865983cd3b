"""The solvers' fallbacks, with the failures forced from the host
(qqcd_set_fault): every kernel still runs to completion, only values are
replaced, so no GPU fault or hang is involved.

The reference has no fallback: it writes INVERTERS.ERROR when its small dense
solve meets a zero row (qqcd-test/cfgsprops/quark/inverters.f90:13128-13133)
and otherwise runs to its iteration limit.  The mixed-precision solver here
must instead (i) never add a correction with a NaN or infinite entry to x,
(ii) drop the learned preconditioner when an outer step stalls, (iii) climb
FP16 -> FP32 -> plain FP64 GMRES-DR when the low-precision operator stops
reducing the FP64 true residual, and (iv) still reach ||b - M x|| / ||b|| <
1e-10 on the full system.  The batched solver must keep a right-hand side's
x when its correction is not finite and finish it by the single-RHS solver.

Lattice: the config-1 field (8^3x16, beta 6.0 heatbath of the reference,
golden fixture), Wilson kappa 0.150, point source at the origin.
"""
import numpy as np
import pytest

import oracle as O
from baylorlqcd_amd import diracops, disconloops, gauge, inverters, lattice, learned

pytestmark = pytest.mark.gpu
GOLD = "tests/golden/golden_8x8x8x16.npz"
KAPPA = (0.150, 0.0)


@pytest.fixture(scope="module")
def cfg1():
    import os
    n, bc = (8, 8, 8, 16), (1, 1, 1, -1)
    print("[fallbacks] config-1 lattice", flush=True)
    lattice.init(n, bc)
    g = np.load(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), GOLD))
    U = gauge.reconstruct12(g["links12"])
    u = O.links_canon_to_ref(n, U)
    co = O.coact_array()
    diracops.gauge_load(u, co)
    learned.init(seed=3)
    yield n, bc, U, u, co
    inverters.set_fault()
    lattice.finalize()


@pytest.fixture(autouse=True)
def no_fault_after():
    yield
    inverters.set_fault()


@pytest.mark.parametrize("mode,name", [(1, "NaN"), (2, "zero"), (3, "random")])
def test_preconditioner_failure_is_dropped(cfg1, mode, name):
    print(f"[fallbacks] preconditioner output {name}", flush=True)
    inverters.set_fault(precond=mode)
    r = learned.bench(2, KAPPA, GMRES=(40, 16), op_prec=1)
    st = inverters.mixed_stats()
    print("[fallbacks]", r, st, flush=True)
    assert r["true_residual"] < 1e-10 and st["converged"]
    if mode in (1, 2):
        # a NaN network output makes the inner correction non-finite; a zero
        # one leaves the flexible basis empty: either way the first outer
        # step stalls and the network is switched off
        assert r["dropped"] == 1 and st["precond_dropped"] == 1
    if mode == 1:
        assert st["rejected"] >= 1


def test_corrupted_fp16_step_climbs_to_fp32(cfg1):
    print("[fallbacks] FP16 operator, first correction NaN", flush=True)
    inverters.set_fault(corrupt_outer=1)
    r = inverters.bench_solve(1, KAPPA, op_prec=2, recon=12, GMRES=(40, 10))
    st = inverters.mixed_stats()
    print("[fallbacks]", r, st, flush=True)
    assert st["rejected"] == 1 and st["fallback"] == 1
    assert r["true_residual"] < 1e-10 and st["converged"]


def test_two_corrupted_steps_finish_in_fp64(cfg1):
    print("[fallbacks] FP16 operator, first two corrections NaN", flush=True)
    inverters.set_fault(corrupt_outer=2)
    r = inverters.bench_solve(1, KAPPA, op_prec=2, recon=12, GMRES=(40, 10))
    st = inverters.mixed_stats()
    print("[fallbacks]", r, st, flush=True)
    assert st["rejected"] == 2 and st["fallback"] == 2
    assert r["true_residual"] < 1e-10 and st["converged"]


def test_capped_budget_reports_not_converged(cfg1):
    print("[fallbacks] Krylov budget capped at 12 steps", flush=True)
    inverters.set_fault(maxiter=12)
    r = inverters.bench_solve(1, KAPPA, op_prec=1, recon=12, GMRES=(40, 10))
    st = inverters.mixed_stats()
    print("[fallbacks]", r, st, flush=True)
    assert not st["converged"]
    assert r["iterations"] <= 12
    assert np.isfinite(r["true_residual"]) and 1e-10 < r["true_residual"] < 1.0
    # and without the cap the same solve converges
    inverters.set_fault()
    r = inverters.bench_solve(1, KAPPA, op_prec=1, recon=12, GMRES=(40, 10))
    assert r["true_residual"] < 1e-10 and inverters.mixed_stats()["converged"]


def test_nan_starting_guess_restarts_from_zero(cfg1):
    n, bc, U, u, co = cfg1
    print("[fallbacks] mixed solve from a NaN starting guess", flush=True)
    V = int(np.prod(n))
    ps = np.zeros((V, 4, 3), dtype=np.complex128)
    ps[0, 0, 0] = 1.0
    pe, _ = O.spinor_canon_to_ref(n, ps)
    x = np.full_like(pe, np.nan)
    inverters.solve_mixed(pe, x, (40, 10), 1e-10, 1, 1, 12, 0.0, 0, u, co, KAPPA, bc)
    st = inverters.mixed_stats()
    print("[fallbacks]", st, flush=True)
    odd = O.parity_mask(n)
    xc = O.spinor_ref_to_canon(n, x, None)
    assert np.all(np.isfinite(xc[~odd])) and st["converged"] and st["rel_res"] < 1e-10
    # the even-odd system's residual, recomputed by the oracle
    bcan = O.spinor_ref_to_canon(n, pe, None)
    r = bcan - O.schur(n, U, bc, [1] * 4, [1] * 4, 0, KAPPA, xc)
    assert np.linalg.norm(r[~odd]) / np.linalg.norm(bcan[~odd]) < 1e-10


@pytest.mark.parametrize("steps", [1, 1000])
def test_batched_corrupted_slot(cfg1, steps):
    # steps 1: the right-hand side's first correction is NaN and is skipped;
    # steps 1000: every batched correction of it is NaN, so it reaches the
    # single-RHS solver (which the fault does not touch)
    n, bc, U, u, co = cfg1
    print(f"[fallbacks] batched solve, right-hand side 2 corrupted in {steps} outer steps", flush=True)
    srcs = []
    for r in range(5):
        be = lattice.new_spinor(n).ravel(order="F")
        bo = lattice.new_spinor(n).ravel(order="F")
        disconloops.z4source(be, bo, 99, r, 1.0)
        srcs.append((be, bo))
    inverters.set_fault(corrupt_outer=steps, batch_rhs=2)
    xe, xo, it, rr = disconloops.fermprop_multi([s[0] for s in srcs], [s[1] for s in srcs], KAPPA, co, bc, 1e-10,
                                                (40, 10), u)
    bst = disconloops.batch_stats()
    rej = disconloops.batch_rejected()
    print("[fallbacks] batch", bst, "rejected", rej, "res", rr, flush=True)
    assert rej >= 1
    if steps > 1:
        assert bst["fallback"] >= 1
    one = [1.0] * 4
    for r, (be, bo) in enumerate(srcs):
        b = O.spinor_ref_to_canon(n, be, bo)
        x = O.spinor_ref_to_canon(n, xe[r], xo[r])
        res = np.linalg.norm(b - O.full_m(n, U, bc, one, one, KAPPA, x)) / np.linalg.norm(b)
        assert res < 1e-10, (r, res)
