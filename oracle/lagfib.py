"""Test infrastructure only (never imported by the product): a numpy
restatement of the reference's lagged-Fibonacci generator and of the Z4 / Z2
sources drawn from it.

  initlagfib = initbirep + bigdelay   qqcd-test/cfgsprops/glue/lagfib.f90:80-216
  rnds                                qqcd-test/cfgsprops/glue/lagfib.f90:219-259
  z4source, z2source                  qqcd-test/cfgsprops/quark/disconloops.f90:4518-4649

State: iw[0:521] (iwlagfib), jr, kr (jrlagfib, krlagfib).  A draw sets
iw[jr] ^= iw[kr], repeats that while the result is zero, returns iw[jr] *
fnorm (fnorm the single-precision literal 0.465661e-9 held in a double) and
advances both cursors modulo 521.  Pinned against the reference's own
outputs in tests/golden/golden_lagfib.npz (tests/test_oracle_lagfib.py).
"""
import numpy as np

P, Q = 521, 32
FNORM = float(np.float32(0.465661e-9))


def _u32(v):
    return v & 0xFFFFFFFF


def initbirep():
    """lagfib.f90:94-138: the process-independent start (iw, jr, kr)."""
    ib = [0] * P
    ix = 1
    for i in range(P):
        ix = _u32(ix * 69069)
        ib[i] = ix >> 31
    jr, kr = 0, P - Q
    iw = np.zeros(P, dtype=np.int64)
    for j in range(P):
        work = 0
        for _ in range(32):
            work = _u32(2 * work + ib[jr])
            ib[jr] ^= ib[kr]
            jr = 0 if jr + 1 == P else jr + 1
            kr = 0 if kr + 1 == P else kr + 1
        iw[j] = work >> 1
    return iw, jr, kr


def bigdelay(iw, ndelay):
    """lagfib.f90:142-216: the state of process ndelay from the common start."""
    mu = 40
    iwk = np.zeros(2 * P - 1, dtype=np.int64)
    iwk[:P] = iw
    for i in range(P, 2 * P - 1):
        iwk[i] = iwk[i - P] ^ iwk[i - Q]
    ib = [0] * (P + 32)
    m = ndelay
    nb = mu - 1
    while m > P - 1:
        nb += 1
        ib[nb] = m % 2
        m //= 2
    c = [0] * (2 * P)
    c[m] = 1
    for j in range(nb, -1, -1):
        for i in range(P - 1, -1, -1):
            c[2 * i + ib[j]] = c[i]
            c[2 * i + 1 - ib[j]] = 0
        for i in range(2 * P - 1, P - 1, -1):
            c[i - P] ^= c[i]
            c[i - Q] ^= c[i]
    sel = np.array(c[:P], dtype=bool)
    out = np.zeros(P, dtype=np.int64)
    for j in range(P):
        v = 0
        for x in iwk[j:j + P][sel]:
            v ^= int(x)
        out[j] = v
    return out


def init(myid, iseed):
    """initlagfib(myid, iseed): (iw int32[521], jr, kr)."""
    iw, jr, kr = initbirep()
    return bigdelay(iw, myid + iseed).astype(np.int32), jr, kr


def draws(state, m):
    """m raw draws (the integers iw[jr] after each update, int64) and the new
    state; a zero result is XOR-ed again with iw[kr], as rnds does."""
    iw, jr, kr = state
    if (jr + P - Q) % P != kr:
        raise ValueError("kr must equal jr - 32 modulo 521")
    # y[t] for t = -521 .. -1 is the window: y[-521 + k] = iw[(jr + k) mod 521]
    y = np.empty(P + m, dtype=np.int64)
    y[:P] = np.roll(np.asarray(iw, dtype=np.int64) & 0xFFFFFFFF, -jr)
    t = 0
    while t < m:
        e = min(t + Q, m)
        a, b = P + t, P + e
        blk = y[a - P:b - P] ^ y[a - Q:b - Q]
        z = np.flatnonzero(blk == 0)
        if z.size:
            e = t + int(z[0]) + 1
            b = P + e
            blk = blk[:e - t]
            blk[-1] = y[b - 1 - Q]  # iw[jr] = 0 ^ iw[kr]
            if blk[-1] == 0:
                raise ValueError("degenerate state: the zero skip never ends")
        y[a:b] = blk
        t = e
    out = y[P:]
    njr = (jr + m) % P
    niw = np.empty(P, dtype=np.int64)
    last = y[m:m + P]  # y[m-521 .. m-1]
    niw[(njr + np.arange(P)) % P] = last
    return out, (niw.astype(np.uint32).astype(np.int32), njr, (njr + P - Q) % P)


def rnds(state, m):
    """rnds(m, rn): m numbers (iw as a signed 32-bit integer times fnorm) and the new state."""
    raw, st = draws(state, m)
    return raw.astype(np.uint32).astype(np.int32).astype(np.float64) * FNORM, st


def _source_ref(n, rn, kind):
    """z4source (kind 0) / z2source (kind 1) body for one checkerboard: the
    (6, ntotal, 4, 2, 8) Fortran array as a flat vector, from 192 nvhalf numbers
    in the order (ic, isite, id, ieo, ibl), ibl fastest."""
    nvh = int(np.prod(n)) // 32
    nb = max(2 * nvh // m for m in n)
    ntot = nvh + nb
    a = np.zeros((8, 2, 4, ntot, 6))  # C-order view of the Fortran array
    r = rn.reshape(3, nvh, 4, 2, 8)
    for ic in range(3):
        v = r[ic].transpose(3, 2, 1, 0)  # (ibl, ieo, id, isite)
        re, im = np.zeros_like(v), np.zeros_like(v)
        if kind == 1:
            re = np.where(v > 0.5, 1.0, -1.0)
        else:
            c1 = v > 0.75
            c2 = ~c1 & (0.5 < v) & (v < 0.75)
            c3 = ~c1 & ~c2 & (0.25 < v) & (v < 0.5)
            c4 = ~c1 & ~c2 & ~c3
            re = np.where(c1, 1.0, np.where(c2, -1.0, 0.0))
            im = np.where(c3, 1.0, np.where(c4, -1.0, 0.0))
        a[:, :, :, :nvh, 2 * ic] = re
        a[:, :, :, :nvh, 2 * ic + 1] = im
    return a.ravel()


def source(n, state, kind):
    """z4source (kind 0) or z2source (kind 1): be, bo in the reference layout and the new state."""
    nvh = int(np.prod(n)) // 32
    rn, st = rnds(state, 384 * nvh)
    return _source_ref(n, rn[:192 * nvh], kind), _source_ref(n, rn[192 * nvh:], kind), st
