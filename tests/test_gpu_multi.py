"""GPU parity of the many-right-hand-side path.

* the multi-RHS hop and Schur operator (RHS-innermost fields, every RHS
  block and tile order, partial activity) against the C oracle per RHS;
* Z4 sources bit-exact against the oracle's restatement;
* the 16 local trace estimators against the oracle;
* the reference's loop operators (qqcd_loopops per site, qqcd_loopsum per
  timeslice and momentum) against the oracle, bar 1e-12 relative;
* batched solves (qqcd_fermprop_multi, qqcd_noise_loop): every source
  reaches a true FP64 residual below 1e-10, the point source among them
  matches the reference Fortran's solution (golden fixture) within 1e-8,
  and graph replays reproduce the eager launches.

Bars: FP64 operators 1e-13 relative (max norm), FP32 storage 2e-6.
"""
import os

import numpy as np
import pytest

import oracle as O
from baylorlqcd_amd import _lib, diracops, disconloops, gauge, lattice, learned, quark

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "golden_8x8x8x16.npz")


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def spinors(n, nrhs, seed):
    rng = np.random.default_rng(seed)
    V = int(np.prod(n))
    return rng.standard_normal((nrhs, V, 4, 3)) + 1j * rng.standard_normal((nrhs, V, 4, 3))


@pytest.fixture(scope="module", params=[((8, 8, 8, 16), (1, 1, 1, -1)), ((4, 4, 4, 8), (-1, 1, 0, -1))])
def lat(request):
    n, bc = request.param
    print(f"[multi] lattice {n} bc {bc}", flush=True)
    lattice.init(n, bc)
    U = gauge.haar(n, 21)
    u = O.links_canon_to_ref(n, U)
    co = O.coact_array()
    diracops.gauge_load(u, co)
    yield n, bc, U, u, co
    lattice.finalize()


@pytest.mark.parametrize("nrhs,nb,order", [(1, 0, 1), (5, 8, 0), (16, 16, 1), (37, 8, 1), (37, 64, 0)])
def test_hop_multi_fp64(lat, nrhs, nb, order):
    n, bc, U, u, co = lat
    psi = spinors(n, nrhs, 7 + nrhs)
    one = [1.0] * 4
    for par_out in (0, 1):
        for idag in (0, 1):
            out = np.zeros_like(psi)
            print(f"[multi] hop nrhs={nrhs} nb={nb} order={order} par={par_out} idag={idag}", flush=True)
            disconloops.hop_multi_canonical(psi, out, par_out, idag=idag, prec=0, recon=18, nb=nb, order=order)
            for r in range(nrhs):
                want = O.hop(n, U, bc, one, one, idag, par_out, psi[r])
                assert rel(out[r], want) < 1e-13, (r, par_out, idag)


def test_hop_multi_fp32_12(lat):
    n, bc, U, u, co = lat
    psi = spinors(n, 12, 3)
    out = np.zeros_like(psi)
    one = [1.0] * 4
    disconloops.hop_multi_canonical(psi, out, 0, prec=1, recon=12, nb=4, order=1)
    for r in range(12):
        want = O.hop(n, U, bc, one, one, 0, 0, psi[r])
        assert rel(out[r], want) < 2e-6, r


def test_hop_multi_partial_activity(lat):
    # retired slots (>= nact) are neither read into nor written
    n, bc, U, u, co = lat
    nrhs, nact = 21, 13
    psi = spinors(n, nrhs, 5)
    sentinel = np.full_like(psi, 3.25 - 1.5j)
    out = sentinel.copy()
    one = [1.0] * 4
    disconloops.hop_multi_canonical(psi, out, 1, prec=0, recon=18, nact=nact, nb=8, order=0)
    odd = O.parity_mask(n)
    for r in range(nrhs):
        if r < nact:
            want = O.hop(n, U, bc, one, one, 0, 1, psi[r])
            assert rel(out[r][odd], want[odd]) < 1e-13, r
        else:
            assert np.array_equal(out[r][odd], sentinel[r][odd]) or np.all(out[r][odd] == 0), r


@pytest.mark.parametrize("kernel", [0, 1, 2, 3, 4])
def test_hop_multi_kernels(lat, kernel):
    # every launch form of the multi-RHS hop (hops in sequence; pipelined on a
    # persistent grid; pipelined with one workgroup per tile; two slots per
    # lane with two or one hops in flight), with retired slots (nact odd, so
    # one slot pair is half retired), in both storage formats and both tile
    # orders
    n, bc, U, u, co = lat
    nrhs, nact = 37, 29
    psi = spinors(n, nrhs, 11)
    one = [1.0] * 4
    odd = O.parity_mask(n)
    lib = _lib.lib()
    _lib.check(lib.qqcd_set_multi_tune(kernel, -1, -1))
    # form 2 also with its residency cap (dynamic LDS: two workgroups per CU)
    if kernel == 2:
        _lib.check(lib.qqcd_set_multi_lds(57344))
    try:
        for prec, recon, bar, nb, order in ((0, 18, 1e-13, 8, 0), (1, 12, 2e-6, 8, 0), (1, 18, 2e-6, 4, 1),
                                            (1, 12, 2e-6, 2, 0)):
            print(f"[multi] hop kernel={kernel} prec={prec} recon={recon} nb={nb} order={order}", flush=True)
            sentinel = np.full_like(psi, 3.25 - 1.5j)
            out = sentinel.copy()
            disconloops.hop_multi_canonical(psi, out, 1, prec=prec, recon=recon, nact=nact, nb=nb, order=order)
            for r in range(nrhs):
                if r < nact:
                    want = O.hop(n, U, bc, one, one, 0, 1, psi[r])
                    assert rel(out[r][odd], want[odd]) < bar, (kernel, prec, r)
                else:
                    assert np.array_equal(out[r][odd], sentinel[r][odd]) or np.all(out[r][odd] == 0), (kernel, r)
    finally:
        _lib.check(lib.qqcd_set_multi_tune(-1, -1, -1))
        _lib.check(lib.qqcd_set_multi_lds(-1))


def test_tuning_knobs_reject_bad_values(lat):
    lib = _lib.lib()
    assert lib.qqcd_set_multi_tune(5, -1, -1) != 0
    assert lib.qqcd_set_multi_lds(1 << 30) != 0
    assert lib.qqcd_set_step_graphs(3) != 0
    assert lib.qqcd_set_pair_norms(2) != 0
    assert lib.qqcd_set_batch_mid(0) != 0
    # the defaults are left in place
    _lib.check(lib.qqcd_set_multi_tune(-1, -1, -1))
    _lib.check(lib.qqcd_set_multi_lds(-1))
    _lib.check(lib.qqcd_set_step_graphs(2))
    _lib.check(lib.qqcd_set_pair_norms(1))
    _lib.check(lib.qqcd_set_batch_mid(2))


@pytest.mark.parametrize("recon", [12, 18])
def test_schur_multi_fp16_intermediate(lat, recon):
    """The FP32 multi-RHS Schur operator with its odd-checkerboard
    intermediate in FP16 (one norm per site and slot: the first hop's FP16
    output form, the second's FP16 input form) or in FP32, against the
    oracle's FP64 Schur operator: FP16 rounding (2^-11 relative per site)
    against FP32 rounding, both far apart, the FP16 one well inside 1e-3."""
    n, bc, U, u, co = lat
    nrhs = 7
    psi = spinors(n, nrhs, 23)
    odd = O.parity_mask(n)
    psi[:, odd] = 0
    one = [1.0] * 4
    kappa = (0.156, 0.005)
    lib = _lib.lib()
    errs = {}
    try:
        for mid in (1, 2):
            _lib.check(lib.qqcd_set_batch_mid(mid))
            out = np.zeros_like(psi)
            disconloops.schur_multi_canonical(psi, out, kappa, prec=1, recon=recon, nb=4, order=0)
            errs[mid] = max(rel(out[r][~odd], O.schur(n, U, bc, one, one, 0, kappa, psi[r])[~odd])
                            for r in range(nrhs))
    finally:
        _lib.check(lib.qqcd_set_batch_mid(2))
    print(f"[multi] Schur operator, FP32 spinors, recon {recon}: max rel error {errs}", flush=True)
    assert errs[1] < 1e-6, errs
    assert 1e-5 < errs[2] < 1e-3, errs


@pytest.mark.parametrize("kappa", [(0.150, 0.0), (0.156, 0.005)])
def test_schur_multi_fp64(lat, kappa):
    n, bc, U, u, co = lat
    nrhs = 9
    psi = spinors(n, nrhs, 11)
    odd = O.parity_mask(n)
    psi[:, odd] = 0
    out = np.zeros_like(psi)
    one = [1.0] * 4
    disconloops.schur_multi_canonical(psi, out, kappa, prec=0, recon=18, nb=4, order=1)
    for r in range(nrhs):
        want = O.schur(n, U, bc, one, one, 0, kappa, psi[r])
        assert rel(out[r][~odd], want[~odd]) < 1e-13, r


def test_z4source_bitexact(lat):
    n, bc, U, u, co = lat
    for rhs, scale in ((0, 1.0), (5, 1 / np.sqrt(2))):
        be = lattice.new_spinor(n).ravel(order="F")
        bo = lattice.new_spinor(n).ravel(order="F")
        disconloops.z4source(be, bo, 4242, rhs, scale)
        got = O.spinor_ref_to_canon(n, be, bo)
        want = O.z4(n, 4242, rhs, scale)
        assert np.array_equal(got, want), rhs


def test_z2source_bitexact(lat):
    n, bc, U, u, co = lat
    be = lattice.new_spinor(n).ravel(order="F")
    bo = lattice.new_spinor(n).ravel(order="F")
    disconloops.z2source(be, bo, 77, 3, 1.0)
    assert np.array_equal(O.spinor_ref_to_canon(n, be, bo), O.z2(n, 77, 3))


def test_trace16(lat):
    n, bc, U, u, co = lat
    eta, x = spinors(n, 2, 17)
    ee, eo = O.spinor_canon_to_ref(n, eta)
    xe, xo = O.spinor_canon_to_ref(n, x)
    got = disconloops.trace16(ee, eo, xe, xo)
    want = O.trace16(n, eta, x)
    assert rel(got, want) < 1e-13


@pytest.mark.parametrize("ksub", [0, 1])
def test_loopops_and_loopsum(lat, ksub):
    n, bc, U, u, co = lat
    b, x = spinors(n, 2, 23 + ksub)
    be, bo = O.spinor_canon_to_ref(n, b)
    xe, xo = O.spinor_canon_to_ref(n, x)
    print(f"[multi] loopops ksub={ksub}", flush=True)
    Je, Jo = disconloops.loopops(ksub, xe, xo, be, bo)
    want = O.loopops(n, U, bc, b, x, ksub)
    assert rel(O.loops_ref_to_canon(n, Je, Jo), want) < 1e-12
    for io, cur in (((1, 1, 1), 1.0), ((2, 1, 3), 0.5), (None, 0.15)):
        print(f"[multi] loopsum io={io}", flush=True)
        got = disconloops.loopsum(ksub, xe, xo, be, bo, io=io, cur_scale=cur)
        w = O.spacesum(n, want, io if io else (1, 1, 1), nmom=5 if io else 1)
        w[:, :, :4] *= cur
        if io is None:
            assert np.all(got[:, 1:] == 0)
            got = got[:, :1]
        assert rel(got, w) < 1e-12


@pytest.fixture(scope="module")
def cfg1():
    n, bc = (8, 8, 8, 16), (1, 1, 1, -1)
    print("[multi] config-1 lattice", flush=True)
    lattice.init(n, bc)
    g = np.load(GOLD)
    U = gauge.reconstruct12(g["links12"])
    u = O.links_canon_to_ref(n, U)
    co = O.coact_array()
    diracops.gauge_load(u, co)
    yield n, bc, U, u, co, g
    lattice.finalize()


def full_rel_res(n, bc, U, kappa, b, x):
    one = [1.0] * 4
    r = b - O.full_m(n, U, bc, one, one, kappa, x)
    return np.linalg.norm(r) / np.linalg.norm(b)


@pytest.mark.parametrize("reorth,graph,basis", [(1, 1, 1), (0, 0, 1), (2, 1, 1), (1, 1, 2), (0, 0, 2)])
def test_fermprop_multi(cfg1, reorth, graph, basis):
    # basis 2: the inner GMRES keeps its Krylov basis in FP16; the FP64
    # reliable updates still bring every source to the same true residual
    n, bc, U, u, co, g = cfg1
    kappa = (0.150, 0.0)
    disconloops.batch_set(reorth=reorth, use_graph=graph, basis_prec=basis)
    # sources: the config-1 point source (origin, spin 1, colour 1; its reference
    # solution is a golden fixture) first and again as the fourth, Z4 noise otherwise
    srcs = []
    V = int(np.prod(n))
    ps = np.zeros((V, 4, 3), dtype=np.complex128)
    ps[0, 0, 0] = 1.0
    pe, po = O.spinor_canon_to_ref(n, ps)
    for r in range(7):
        if r in (0, 3):
            srcs.append((pe.copy(), po.copy()))
        else:
            be = lattice.new_spinor(n).ravel(order="F")
            bo = lattice.new_spinor(n).ravel(order="F")
            disconloops.z4source(be, bo, 77, r, 1.0)
            srcs.append((be, bo))
    print(f"[multi] fermprop_multi reorth={reorth} graph={graph} basis={basis}", flush=True)
    xe, xo, it, rr = disconloops.fermprop_multi([s[0] for s in srcs], [s[1] for s in srcs], kappa, co, bc, 1e-10,
                                                (40, 10), u)
    st = disconloops.batch_stats()
    print("[multi] stats", st, "iters", it, "res", rr, flush=True)
    odd = O.parity_mask(n)
    for r, (be, bo) in enumerate(srcs):
        b = O.spinor_ref_to_canon(n, be, bo)
        x = O.spinor_ref_to_canon(n, xe[r], xo[r])
        res = full_rel_res(n, bc, U, kappa, b, x)
        assert res < 1e-10, (r, res)
        assert abs(rr[r] - res) < 1e-11, (r, rr[r], res)
        if r in (0, 3):
            assert rel(x[~odd], g["gmresdr_k150_xe"]) < 1e-8, r
            assert rel(x[odd], g["gmresdr_k150_xo"]) < 1e-8, r
    if graph:
        assert st["graph_replays"] > 0
    disconloops.batch_set()


def test_fermprop_multi_normalisation_in_operator(cfg1):
    """With the FP16 basis and the two-slot operator each Arnoldi step's
    normalisation is done by the next operator application (the second hop
    scales its result by 1 / ||w|| and writes the FP16 basis vector), and the
    steps run in pairs (one Gram-Schmidt pass over the basis for the dots of
    A v_j and A^2 v_j, one for both updates, column j + 1 from the Arnoldi
    relation) or one at a time; with the one-slot launch form (2) a separate
    pass normalises.  Paired steps take the norms of the updated vectors from
    the projections (the update pass then writes the next basis vector) or
    from the updated vectors (a third pass).  The operator's intermediate
    field is FP16 with per-(site, slot) norms where the two-slot kernel runs,
    or FP32.  All reach the same true residuals and solutions to the solve's
    tolerance, with graph replays and eagerly."""
    n, bc, U, u, co, g = cfg1
    kappa = (0.150, 0.0)
    srcs = []
    for r in range(7):
        be = lattice.new_spinor(n).ravel(order="F")
        bo = lattice.new_spinor(n).ravel(order="F")
        disconloops.z4source(be, bo, 91, r, 1.0)
        srcs.append((be, bo))
    lib = _lib.lib()
    sols = {}
    try:
        for form, steps, norms, mid in ((3, 2, 1, 2), (3, 2, 0, 2), (3, 2, 1, 1), (3, 1, 1, 2), (2, 1, 1, 2)):
            for graph in (1, 0):
                _lib.check(lib.qqcd_set_multi_tune(form, -1, -1))
                _lib.check(lib.qqcd_set_arnoldi_steps(steps))
                _lib.check(lib.qqcd_set_pair_norms(norms))
                _lib.check(lib.qqcd_set_batch_mid(mid))
                disconloops.batch_set(use_graph=graph, basis_prec=2)
                print(f"[multi] normalisation in the operator: form {form} steps {steps} norms {norms} mid {mid} "
                      f"graph {graph}", flush=True)
                xe, xo, it, rr = disconloops.fermprop_multi([s[0] for s in srcs], [s[1] for s in srcs], kappa, co,
                                                            bc, 1e-10, (40, 10), u)
                xs = []
                for r, (be, bo) in enumerate(srcs):
                    b = O.spinor_ref_to_canon(n, be, bo)
                    x = O.spinor_ref_to_canon(n, xe[r], xo[r])
                    assert full_rel_res(n, bc, U, kappa, b, x) < 1e-10, (form, steps, norms, mid, graph, r)
                    xs.append(x)
                sols[(form, steps, norms, mid, graph)] = xs
    finally:
        _lib.check(lib.qqcd_set_multi_tune(-1, -1, -1))
        _lib.check(lib.qqcd_set_arnoldi_steps(2))
        _lib.check(lib.qqcd_set_pair_norms(1))
        _lib.check(lib.qqcd_set_batch_mid(2))
        disconloops.batch_set()
    for key, xs in sols.items():
        for r in range(len(srcs)):
            assert rel(xs[r], sols[(2, 1, 1, 2, 0)][r]) < 1e-8, (key, r)
    # graph replays give the eager launches' result bit for bit
    for key in ((3, 2, 1, 2), (3, 2, 0, 2), (3, 2, 1, 1), (3, 1, 1, 2)):
        for r in range(len(srcs)):
            assert np.array_equal(sols[key + (1,)][r], sols[key + (0,)][r]), (key, r)


@pytest.mark.parametrize("nsrc,gm", [(2, (40, 10)), (6, (41, 10))])
def test_fermprop_multi_unpaired_shapes(cfg1, nsrc, gm):
    # shapes where the batched solver cannot pair its Arnoldi steps: one
    # batched source (RHS block 1: no two-slot operator) and an odd cycle
    # length (m - k = 31); every source still reaches the true residual
    n, bc, U, u, co, g = cfg1
    kappa = (0.150, 0.0)
    srcs = []
    for r in range(nsrc):
        be = lattice.new_spinor(n).ravel(order="F")
        bo = lattice.new_spinor(n).ravel(order="F")
        disconloops.z4source(be, bo, 55, r, 1.0)
        srcs.append((be, bo))
    disconloops.batch_set(basis_prec=2)
    print(f"[multi] unpaired shapes: {nsrc} sources, GMRES {gm}", flush=True)
    xe, xo, it, rr = disconloops.fermprop_multi([s[0] for s in srcs], [s[1] for s in srcs], kappa, co, bc, 1e-10,
                                                gm, u)
    for r, (be, bo) in enumerate(srcs):
        b = O.spinor_ref_to_canon(n, be, bo)
        x = O.spinor_ref_to_canon(n, xe[r], xo[r])
        assert full_rel_res(n, bc, U, kappa, b, x) < 1e-10, r


def test_noise_loop_traces_and_graph(cfg1):
    n, bc, U, u, co, g = cfg1
    kappa = (0.150, 0.0)
    nrhs, seed = 6, 31
    runs = {}
    for graph in (1, 0):
        disconloops.batch_set(use_graph=graph)
        print(f"[multi] noise_loop graph={graph}", flush=True)
        runs[graph] = disconloops.noise_loop(nrhs, seed, kappa, (40, 10), tol=1e-10, coact=co, bc=bc)
    disconloops.batch_set()
    a, b = runs[1], runs[0]
    print("[multi] noise loop", a["times"], a["stats"], a["iters"], flush=True)
    assert np.all(a["res_full"] < 1e-10) and np.all(b["res_full"] < 1e-10)
    # graph replays run the same kernels with the same arguments as the eager launches
    assert np.array_equal(a["iters"], b["iters"])
    assert rel(a["traces"], b["traces"]) < 1e-12
    assert rel(a["loops"], b["loops"]) < 1e-12
    assert a["stats"]["graph_replays"] > 0 and b["stats"]["graph_replays"] == 0
    # the traces are those of the solutions of the same sources
    srcs_e, srcs_o = [], []
    for r in range(nrhs):
        be = lattice.new_spinor(n).ravel(order="F")
        bo = lattice.new_spinor(n).ravel(order="F")
        disconloops.z4source(be, bo, seed, r, 1.0)
        srcs_e.append(be)
        srcs_o.append(bo)
    xe, xo, it, rr = disconloops.fermprop_multi(srcs_e, srcs_o, kappa, co, bc, 1e-10, (40, 10), u)
    want = np.zeros((n[3], 16), dtype=np.complex128)
    wloop = np.zeros((n[3], 9), dtype=np.complex128)
    for r in range(nrhs):
        eta = O.spinor_ref_to_canon(n, srcs_e[r], srcs_o[r])
        x = O.spinor_ref_to_canon(n, xe[r], xo[r])
        want += O.trace16(n, eta, x) / nrhs
        wloop += O.spacesum(n, O.loopops(n, U, bc, eta, x, 0), nmom=1)[:, 0, :] / nrhs
    wloop[:, :4] *= kappa[0]
    # both sets of solutions meet 1e-10 relative residual; traces agree to that order
    assert rel(a["traces"], want) < 1e-8
    # the reference's operators as nsaverage forms them (currents times kappa)
    assert rel(a["loops"][:, 0, :], wloop) < 1e-8
    assert np.all(a["loops"][:, 1:] == 0)


def test_noise_loop_fp16_basis(cfg1):
    # the batched loop with an FP16 Krylov basis against the FP32 one: the same
    # sources solved to the same true residual, so traces agree to that order
    n, bc, U, u, co, g = cfg1
    kappa = (0.150, 0.0)
    runs = {}
    for basis in (1, 2):
        disconloops.batch_set(basis_prec=basis)
        print(f"[multi] noise_loop basis={basis}", flush=True)
        runs[basis] = disconloops.noise_loop(6, 41, kappa, (40, 10), tol=1e-10, coact=co, bc=bc)
    disconloops.batch_set()
    a, b = runs[2], runs[1]
    print("[multi] FP16 basis", a["times"], a["stats"], a["iters"], "FP32", b["times"], b["iters"], flush=True)
    assert np.all(a["res_full"] < 1e-10)
    assert rel(a["traces"], b["traces"]) < 1e-8
    assert rel(a["loops"], b["loops"]) < 1e-8
    with pytest.raises(RuntimeError):
        disconloops.batch_set(basis_prec=0)
    disconloops.batch_set()


@pytest.mark.parametrize("method,scale", [(0, 1.0), (1, 2.0)])
def test_noise_loop_lagfib_sources(cfg1, method, scale):
    # the loop on the reference's stream: its sources are the ones successive
    # z4source calls return from the same state, and it leaves the state where
    # those calls leave it
    from baylorlqcd_amd.lagfib import LagFib
    n, bc, U, u, co, g = cfg1
    kappa = (0.150, 0.0)
    nrhs = 4
    L = LagFib(0, 7)
    print(f"[multi] noise_loop on the lagfib stream method={method} scale={scale}", flush=True)
    a = disconloops.noise_loop(nrhs, 0, kappa, (40, 10), tol=1e-10, scale=scale, coact=co, bc=bc, method=method,
                               lagfib=L)
    assert np.all(a["res_full"] < 1e-10)
    R = LagFib(0, 7)
    srcs_e, srcs_o = [], []
    for r in range(nrhs):
        be = lattice.new_spinor(n).ravel(order="F")
        bo = lattice.new_spinor(n).ravel(order="F")
        R.z4source(be, bo)
        srcs_e.append(be)
        srcs_o.append(bo)
    iw, jr, kr = L.state
    iw2, jr2, kr2 = R.state
    assert np.array_equal(iw, iw2) and (jr, kr) == (jr2, kr2)
    xe, xo, it, rr = disconloops.fermprop_multi(srcs_e, srcs_o, kappa, co, bc, 1e-10, (40, 10), u)
    want = np.zeros((n[3], 16), dtype=np.complex128)
    for r in range(nrhs):
        eta = O.spinor_ref_to_canon(n, srcs_e[r], srcs_o[r])
        x = O.spinor_ref_to_canon(n, xe[r], xo[r])
        want += O.trace16(n, eta, x) / nrhs
    # scaled sources, traces divided by scale^2: the same estimate
    assert rel(a["traces"], want) < 1e-8


def test_noise_loop_lagfib_against_reference_loop(cfg1):
    # the reference's own loop on the config-1 field (golden fixture written by
    # the reference Fortran: initlagfib(0, 5), then per source z4source ->
    # fermprop (GMRES-DR to 1e-12) -> nsaverage, averaged): the batched loop on
    # the same stream gives the same zero-momentum operators (both sets of
    # solutions meet their residuals, so they agree to that order) and leaves
    # the stream exactly where the reference's loop left it
    from baylorlqcd_amd.lagfib import LagFib
    n, bc, U, u, co, g = cfg1
    gl = np.load(os.path.join(os.path.dirname(GOLD), "golden_noiseloop_8x8x8x16.npz"))
    iseed, nrhs = (int(v) for v in gl["case"])
    kappa = (float(gl["kappa"]), 0.0)
    L = LagFib(0, iseed)
    print(f"[multi] noise_loop_lagfib against the reference's loop, {nrhs} sources", flush=True)
    a = disconloops.noise_loop(nrhs, 0, kappa, (40, 10), tol=1e-11, coact=co, bc=bc, lagfib=L)
    assert np.all(a["res_full"] < 1e-11)
    assert rel(a["loops"][:, 0, :], gl["nsaverage_mean"]) < 1e-8
    iw, jr, kr = L.state
    assert np.array_equal(iw, gl["lagfib_after_iw"]) and (jr, kr) == tuple(int(v) for v in gl["lagfib_after_jrkr"])


@pytest.mark.parametrize("method,kappa", [(0, (0.150, 0.0)), (1, (0.156, 0.005)), (2, (0.150, 0.0))])
def test_propagator(cfg1, method, kappa):
    # all twelve spin-colour sources at the origin; the first is the golden
    # point source of the reference's GMRES-DR solve (Wilson kappa 0.150)
    n, bc, U, u, co, g = cfg1
    if method == 2:
        # untrained network: kappa^2 times the identity (point Jacobi) as the preconditioner
        learned.init((1, 8, 8, 8, 1), scale=kappa[0] ** 2, noise=0.0, seed=1)
    print(f"[multi] propagator method={method} kappa={kappa}", flush=True)
    out = quark.propagator(12, (1, 1, 1, 1), kappa, GMRES=(40, 16), tol=1e-10, op_prec=2, op_recon=12,
                           method=method, solutions=True)
    print("[multi] propagator", out["iters"], out["times"], out["stats"], flush=True)
    assert np.all(out["res_full"] < 1e-10)
    odd = O.parity_mask(n)
    V = int(np.prod(n))
    for r in (0, 4, 11):
        b = np.zeros((V, 4, 3), dtype=np.complex128)
        b[0, r // 3, r % 3] = 1.0
        x = O.spinor_ref_to_canon(n, out["xe"][r], out["xo"][r])
        res = full_rel_res(n, bc, U, kappa, b, x)
        assert res < 1e-10, (r, res)
        assert abs(res - out["res_full"][r]) < 1e-11
        if r == 0 and kappa == (0.150, 0.0):
            assert rel(x[~odd], g["gmresdr_k150_xe"]) < 1e-8
            assert rel(x[odd], g["gmresdr_k150_xo"]) < 1e-8
