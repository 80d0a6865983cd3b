"""qqcd_rnds_noalloc -- rnds through the library's static device workspace,
with no allocation -- against qqcd_rnds (which allocates), through the C
ABI: draws and stream states bit-identical on two calls in a row, the
second smaller than the first, in a fresh context each and in one context,
with a zero draw planted in the stream (the skip path regenerates from a
corrected window in the workspace); and the size limit is an error."""
import numpy as np
import pytest

from baylorlqcd_amd import _lib, lattice
from baylorlqcd_amd.lagfib import LagFib
from oracle import lagfib as LF
from test_gpu_lagfib import planted

pytestmark = pytest.mark.gpu

N = (8, 8, 8, 16)
MAXM = 1 << 20


def noalloc(L, m):
    rn = np.zeros(max(int(m), 1))
    _lib.check(_lib.lib().qqcd_rnds_noalloc(int(m), _lib.dp(rn), *L._args()))
    return rn[:int(m)]


def same(a, b):
    return np.array_equal(a[0], b[0]) and tuple(a[1:]) == tuple(b[1:])


@pytest.fixture
def ctx():
    lattice.init(N, (1, 1, 1, -1))
    yield
    lattice.finalize()


def test_rnds_noalloc_two_calls(ctx):
    print("[no-alloc] rnds against qqcd_rnds, 1048576 then 70001 draws", flush=True)
    a, b = LagFib(0, 7), LagFib(0, 7)
    for m in (MAXM, 70_001):
        want = a.rnds(m)
        got = noalloc(b, m)
        assert np.array_equal(got, want) and same(b.state, a.state)
    # the stream itself: the oracle's generator (pinned on the reference's draws)
    st = LF.init(0, 7)
    ref, rst = LF.rnds(st, MAXM + 70_001)
    assert np.array_equal(ref[MAXM:], got) and same(b.state, rst)


def test_rnds_noalloc_planted_zero(ctx):
    print("[no-alloc] planted zero at 400000, then a shorter call", flush=True)
    st = planted(400_000, 13)
    a, b = LagFib.from_state(*st), LagFib.from_state(*st)
    for m in (600_000, 5_000):
        want = a.rnds(m)
        got = noalloc(b, m)
        assert np.array_equal(got, want) and same(b.state, a.state)


def test_rnds_noalloc_new_context_smaller_lattice():
    # the workspace outlives a context: a second context on a smaller
    # lattice gets the same address and a shorter call
    print("[no-alloc] second context, smaller call", flush=True)
    for n, m in ((N, 300_000), ((4, 4, 4, 8), 1_000)):
        lattice.init(n, (1, 1, 1, -1))
        try:
            a, b = LagFib(1, 3), LagFib(1, 3)
            want = a.rnds(m)
            got = noalloc(b, m)
            assert np.array_equal(got, want) and same(b.state, a.state)
        finally:
            lattice.finalize()


def test_rnds_noalloc_limit(ctx):
    L = LagFib(0, 1)
    before = L.state
    rn = np.zeros(MAXM + 1)
    rc = _lib.lib().qqcd_rnds_noalloc(MAXM + 1, _lib.dp(rn), *L._args())
    assert rc != 0
    assert "1048576" in _lib.lib().qqcd_last_error().decode()
    assert same(L.state, before)
