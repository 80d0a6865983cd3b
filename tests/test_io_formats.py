"""The reference's gauge-field and propagator files (GFwrite / GFread,
qqcd-test/cfgsprops/glue/gaugetools.f90:1259-1334; FPwrite / FPread,
qqcd-test/cfgsprops/quark/quark.f90:3379-3481) written and read by the
library's host code, against the reference Fortran's own files.

golden_io_8x8x8x16.json holds SHA-256 digests of the files the reference
wrote for seeded inputs and of the arrays its GFread / FPread returned
(tests/golden/make_golden.py io): equal digests mean byte-identical files and
bit-identical read-backs (fixsu3 included).  Host-only calls: no device.
"""
import hashlib
import json
import os
import struct

import numpy as np
import pytest

import oracle as O
from baylorlqcd_amd import gauge, gaugetools, quark

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "golden_io_8x8x8x16.json")


def sha(b):
    return hashlib.sha256(b).hexdigest()


@pytest.fixture(scope="module")
def gold():
    with open(GOLD) as f:
        g = json.load(f)
    n = tuple(g["lattice"])
    c = g["case"]
    u = O.links_canon_to_ref(n, gauge.haar(n, c["links_seed"]))
    rng = np.random.default_rng(c["spinor_seed"])
    V = int(np.prod(n))
    psi = rng.standard_normal((V, 4, 3)) + 1j * rng.standard_normal((V, 4, 3))
    xe, xo = O.spinor_canon_to_ref(n, psi)
    return n, g, c, u, xe, xo


def test_gauge_file_matches_reference(gold, tmp_path):
    n, g, c, u, xe, xo = gold
    path = tmp_path / "cfg.gf"
    uw = u.copy()
    gaugetools.gfwrite(n, uw, path, c["gaction"], c["beta"], c["alat"], c["tad"], c["icfg"])
    data = path.read_bytes()
    assert len(data) == g["gf_file_bytes"]
    assert sha(data) == g["gf_file_sha256"]
    assert sha(uw.tobytes()) == g["gfwrite_u_sha256"]  # fixsu3 applied in place, as the reference does
    # status "new": the file may not exist
    with pytest.raises(Exception):
        gaugetools.gfwrite(n, u.copy(), path)
    ur = np.zeros_like(u)
    meta = gaugetools.gfread(n, ur, path)
    assert sha(ur.tobytes()) == g["gfread_u_sha256"]
    assert list(meta["gaction"]) == c["gaction"] and meta["beta"] == c["beta"] and meta["icfgsave"] == c["icfg"]
    assert np.array_equal(meta["alat"], c["alat"]) and np.array_equal(meta["tad"], c["tad"])


def test_gauge_read_rebuilds_su3(gold, tmp_path):
    # semantics beside the digests: read-back links are unitary with det 1 and equal
    # the written (re-unitarised) ones
    n, g, c, u, xe, xo = gold
    path = tmp_path / "cfg.gf"
    uw = u.copy()
    gaugetools.gfwrite(n, uw, path)
    ur = np.zeros_like(u)
    gaugetools.gfread(n, ur, path)
    Ur = O.links_ref_to_canon(n, ur)
    eye = np.broadcast_to(np.eye(3), Ur.shape)
    assert np.allclose(Ur @ np.conj(np.swapaxes(Ur, -1, -2)), eye, atol=1e-14)
    assert np.allclose(np.linalg.det(Ur), 1.0, atol=1e-14)
    assert np.allclose(O.links_ref_to_canon(n, uw), Ur, rtol=0, atol=1e-15)


def test_propagator_file_matches_reference(gold, tmp_path):
    n, g, c, u, xe, xo = gold
    path = tmp_path / "prop.fp"
    quark.fpwrite(n, xe, xo, path, c["kappa"], c["tmqcd"], c["cSW"], c["icsrc"], c["idsrc"], c["icfg"])
    data = path.read_bytes()
    assert len(data) == g["fp_file_bytes"]
    assert sha(data) == g["fp_file_sha256"]
    ye, yo = np.zeros_like(xe), np.zeros_like(xo)
    meta = quark.fpread(n, ye, yo, path)
    assert sha(ye.tobytes()) == g["fpread_xe_sha256"] and sha(yo.tobytes()) == g["fpread_xo_sha256"]
    assert meta == {"kappa": c["kappa"], "tmqcd": c["tmqcd"], "cSW": c["cSW"], "icsrc": c["icsrc"],
                    "idsrc": c["idsrc"], "icfg": c["icfg"], "myid": 0}
    # the records: 64-byte payloads framed by their length; the trailer last
    assert struct.unpack("<i", data[:4])[0] == 64 and struct.unpack("<i", data[-4:])[0] == 40


def test_truncated_files_are_rejected(gold, tmp_path):
    n, g, c, u, xe, xo = gold
    path = tmp_path / "prop.fp"
    quark.fpwrite(n, xe, xo, path, 0.15)
    path.write_bytes(path.read_bytes()[:-10])
    with pytest.raises(Exception):
        quark.fpread(n, np.zeros_like(xe), np.zeros_like(xo), path)
    # a file of another lattice
    with pytest.raises(Exception):
        quark.fpread((8, 8, 8, 8), np.zeros_like(xe), np.zeros_like(xo), path)
