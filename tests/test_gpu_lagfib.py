"""The reference's lagged-Fibonacci stream generated on the GPU (jump-ahead
over chunks, qqcd_lagfib.hip) and the Z4 / Z2 sources drawn from it, against
the reference's own draws (golden_lagfib_8x8x8x16.npz) and, at sizes the
fixture does not hold, against the oracle's sequential generator (pinned on
the same fixture).  Bar: bit-identical draws, sources and states.

Zero draws (the one non-linear step: rnds XORs again) are planted in the
start state at chosen positions -- inside a later chunk, on a chunk
boundary, on the last draw -- by solving the linear recurrence for one
window word on the CPU, so the device path that corrects a zero and
regenerates the rest of the stream runs."""
import hashlib
import os

import numpy as np
import pytest

from baylorlqcd_amd import lattice
from baylorlqcd_amd.lagfib import LagFib
from oracle import lagfib as LF

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N = (8, 8, 8, 16)


@pytest.fixture(scope="module")
def g():
    lattice.init(N, (1, 1, 1, -1))
    yield np.load(os.path.join(ROOT, "tests", "golden", "golden_lagfib_8x8x8x16.npz"))
    lattice.finalize()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def gst(g, key):
    jr, kr = g[key + "_jrkr"]
    return g[key + "_iw"], int(jr), int(kr)


def same(a, b):
    return np.array_equal(a[0], b[0]) and tuple(a[1:]) == tuple(b[1:])


def test_rnds_reference(g):
    print("[lagfib] rnds against the reference", flush=True)
    L = LagFib(0, 1)
    assert np.array_equal(L.rnds(1000), g["rn_1000"]) and same(L.state, gst(g, "after_1000"))
    assert sha(L.rnds(200003)) == str(g["rn_200003_sha256"]) and same(L.state, gst(g, "after_200003"))
    L = LagFib.from_state(*gst(g, "zero_in"))
    assert np.array_equal(L.rnds(100), g["rn_zero_100"]) and same(L.state, gst(g, "zero_out"))


def test_sources_reference(g):
    print("[lagfib] z4source / z2source against the reference", flush=True)
    L = LagFib(0, 5)
    slen = 6 * (O_ntotal()) * 64
    be, bo = np.zeros(slen), np.zeros(slen)
    L.z4source(be, bo)
    assert sha(be) == str(g["z4_be_sha256"]) and sha(bo) == str(g["z4_bo_sha256"])
    L.z2source(be, bo)
    assert sha(be) == str(g["z2_be_sha256"]) and sha(bo) == str(g["z2_bo_sha256"])
    assert same(L.state, gst(g, "after_z"))


def O_ntotal():
    nvh = int(np.prod(N)) // 32
    return nvh + max(2 * nvh // m for m in N)


def test_long_stream_many_chunks(g):
    # 12.5M draws: several hundred chunks of one launch
    print("[lagfib] 12.5M draws against the oracle", flush=True)
    st = LF.init(2, 11)
    L = LagFib.from_state(*st)
    m = 12_500_017
    got = L.rnds(m)
    want, wst = LF.rnds(st, m)
    assert np.array_equal(got, want) and same(L.state, wst)


# ---- planted zeros ----------------------------------------------------------

P = 521


def _mulx(a):
    a <<= 1
    if a >> P & 1:
        a ^= (1 << P) | 1 | (1 << 489)
    return a


def _mulmod(a, b):
    r = 0
    for i in range(P - 1, -1, -1):
        r = _mulx(r)
        if a >> i & 1:
            r ^= b
    return r


def _xpow(e):
    r, b = 1, 2
    while e:
        if e & 1:
            r = _mulmod(r, b)
        b = _mulmod(b, b)
        e >>= 1
    return r


def planted(T, seed):
    """A state (jr = 0) whose linear stream has y_T = 0: y_T = XOR of the
    window words selected by x^(T+521) mod P; flipping one selected word by
    y_T zeroes it."""
    iw, jr, kr = LF.init(0, seed)
    base = [int(iw[(jr + k) % P]) & 0xFFFFFFFF for k in range(P)]
    a = _xpow(T + P)
    sel = [k for k in range(P) if a >> k & 1]
    y = 0
    for k in sel:
        y ^= base[k]
    base[sel[0]] ^= y
    w = np.array(base, dtype=np.uint32).astype(np.int32)
    return w, 0, P - 32


@pytest.mark.parametrize("T,m", [(3_000_123, 5_000_000), (8192 * 7, 400_000), (399_999, 400_000), (0, 70_000)])
def test_planted_zero(g, T, m):
    print(f"[lagfib] planted zero at {T} of {m}", flush=True)
    st = planted(T, 13)
    raw, _ = LF.draws(st, min(m, T + 1))
    assert T == 0 or np.all(raw[:T] != 0)
    # the linear stream would give 0 at T; rnds gives y_{T-32} there instead
    L = LagFib.from_state(*st)
    got = L.rnds(m)
    want, wst = LF.rnds(st, m)
    assert np.array_equal(got, want) and same(L.state, wst)
    prev = want[T - 32] if T >= 32 else float(np.int32(st[0][(st[1] + P + T - 32) % P])) * LF.FNORM
    assert got[T] == prev


def test_bad_cursors_rejected(g):
    iw, jr, kr = LF.init(0, 1)
    with pytest.raises(Exception):
        LagFib.from_state(iw, 5, 17).rnds(10)
