"""The oracle's lagged-Fibonacci generator (oracle/lagfib.py) against the
reference's own outputs (tests/golden/golden_lagfib_8x8x8x16.npz, written by
tests/golden/make_golden.py lagfib from qqcd-test/cfgsprops/glue/lagfib.f90
and disconloops.f90's z4source / z2source): bit-identical states and draws,
the zero skip included.  Beside the fixture: the recurrence itself and the
Z4 value set."""
import hashlib
import os

import numpy as np
import pytest

from oracle import lagfib as LF

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N = (8, 8, 8, 16)


@pytest.fixture(scope="module")
def g():
    return np.load(os.path.join(ROOT, "tests", "golden", "golden_lagfib_8x8x8x16.npz"))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def st(g, key):
    jr, kr = g[key + "_jrkr"]
    return g[key + "_iw"], int(jr), int(kr)


def same(a, b):
    return np.array_equal(a[0], b[0]) and a[1:] == b[1:]


def test_init_states(g):
    for (my, sd), iw, (jr, kr) in zip(g["init_cases"], g["init_iw"], g["init_jrkr"]):
        assert same(LF.init(int(my), int(sd)), (iw, int(jr), int(kr)))


def test_draws_and_state(g):
    rn, s = LF.rnds(LF.init(0, 1), 1000)
    assert np.array_equal(rn, g["rn_1000"]) and same(s, st(g, "after_1000"))
    rn, s = LF.rnds(s, 200003)
    assert sha(rn) == str(g["rn_200003_sha256"]) and same(s, st(g, "after_200003"))


def test_zero_skip(g):
    s = st(g, "zero_in")
    assert s[0][s[2]] == s[0][s[1]]  # the next draw XORs to zero
    rn, s = LF.rnds(s, 100)
    assert np.array_equal(rn, g["rn_zero_100"]) and same(s, st(g, "zero_out"))
    assert np.all(rn > 0)


def test_z4_z2_sources(g):
    be, bo, s = LF.source(N, LF.init(0, 5), 0)
    assert sha(be) == str(g["z4_be_sha256"]) and sha(bo) == str(g["z4_bo_sha256"])
    vals = be.reshape(-1, 2)
    nz = np.abs(vals).sum(1) > 0
    assert set(map(tuple, vals[nz])) <= {(1, 0), (-1, 0), (0, 1), (0, -1)}
    be, bo, s = LF.source(N, s, 1)
    assert sha(be) == str(g["z2_be_sha256"]) and sha(bo) == str(g["z2_bo_sha256"])
    assert same(s, st(g, "after_z"))


def test_recurrence():
    # y_t = y_{t-521} xor y_{t-32} on the raw stream (no zero in this stretch)
    raw, _ = LF.draws(LF.init(0, 3), 5000)
    w = np.roll(LF.init(0, 3)[0].astype(np.int64), 0)
    y = np.concatenate([w, raw])
    assert np.array_equal(y[521:], y[:-521] ^ y[521 - 32:-32])


def test_noiseloop_fixture_stream():
    # the reference's noise-loop fixture consumed exactly nrhs z4source draws of
    # the stream of initlagfib(0, iseed): the oracle lands on the state it recorded
    gl = np.load(os.path.join(ROOT, "tests", "golden", "golden_noiseloop_8x8x8x16.npz"))
    iseed, nrhs = (int(v) for v in gl["case"])
    st = LF.init(0, iseed)
    for _ in range(nrhs):
        _, _, st = LF.source(N, st, 0)
    iw, jr, kr = st
    assert np.array_equal(np.asarray(iw, dtype=np.int64) & 0xFFFFFFFF,
                          np.asarray(gl["lagfib_after_iw"], dtype=np.int64) & 0xFFFFFFFF)
    assert (int(jr), int(kr)) == tuple(int(v) for v in gl["lagfib_after_jrkr"])
