"""The SVF fold and its sky-vault tables against the reference's own fold pins.

``tests/golden/fold_pins.npz`` comes from the reference's
``native/cuda/data/fold_pins.json`` (``make_fold_pins.py``): the 153-patch
tables as float32 bits (captured from the reference's torch computation on
its authors' machine) and 10 cases of packed vegetation / vegetation-building
shadow bits with the fold's 11 output planes (``shadow.py:545-613``).

* ``skyvault.build_skyvault`` (ours, from this host's torch) reproduces ring,
  annulus counts, patch azimuths and E/S/W/N windows exactly, and the annulus
  weights exactly except one entry within 2 ulp: torch's float32 ``sin`` of a
  single element differs by an ulp between the two hosts (the golden SVF
  vectors written here pin this host's values);
* the C oracle's fold, fed the pinned tables, reproduces every output plane
  of every case bit for bit (vegetation and aveg families, the ``last``
  correction where vegdem2 == 0, the clamps), and SVFtotal follows from the
  pinned building SVF with the pinned transmissivity.
"""
from pathlib import Path

import numpy as np
import pytest

from solweig_project_amd import skyvault as SV

PINS = Path(__file__).parent / "golden" / "fold_pins.npz"


@pytest.fixture(scope="module")
def pins():
    with np.load(PINS) as z:
        return {k: z[k] for k in z.files}


def _ulps(a, b):
    return np.abs(a.astype(np.float32).view(np.int32).astype(np.int64) - b.astype(np.float32).view(np.int32))


def test_sky_vault_tables(pins):
    v = SV.build_skyvault()
    assert np.array_equal(np.asarray(v.ring), pins["ring"])
    assert np.array_equal(np.asarray(v.n_annuli), pins["na"])
    dm = np.asarray(v.dir_mask)
    assert np.array_equal(np.array([(dm >> k) & 1 for k in range(4)]), pins["dir"])
    assert np.array_equal(np.asarray(v.azi, np.float32).view(np.uint32), pins["azimuth"].view(np.uint32))
    for name in ("w_iso", "w_aniso"):
        u = _ulps(np.asarray(getattr(v, name), np.float32).ravel(), pins[name])
        assert int((u > 0).sum()) <= 1 and int(u.max()) <= 2, (name, np.nonzero(u)[0], u.max())
    last, trans, omt = pins["consts"]
    assert np.float32(3.0459e-4) == last and np.float32(0.03) == trans and np.float32(1.0) - trans == omt


def _unpack(b):
    return np.unpackbits(b, axis=2, bitorder="little", count=153).astype(np.float32)


def test_oracle_fold_matches_the_pinned_cases(pins):
    from oracle import driver as D
    L = D.lib()
    w_iso = np.ascontiguousarray(pins["w_iso"].reshape(8, 12), np.float32)
    w_aniso = np.ascontiguousarray(pins["w_aniso"].reshape(8, 12), np.float32)
    ring = np.ascontiguousarray(pins["ring"], np.int32)
    na = np.ascontiguousarray(pins["na"], np.int32)
    d = pins["dir"]
    dirm = np.ascontiguousarray(d[0] | (d[1] << 1) | (d[2] << 2) | (d[3] << 3), np.int32)
    trans = pins["consts"][1]
    for name in pins["cases"]:
        rows, cols = (int(x) for x in pins[f"{name}_shape"])
        n = rows * cols
        veg = np.ascontiguousarray(_unpack(pins[f"{name}_veg"]).reshape(n, 153).T)
        vb = np.ascontiguousarray(_unpack(pins[f"{name}_vbsh"]).reshape(n, 153).T)
        sh = np.zeros_like(veg)
        out = np.empty((16, n), np.float32)
        L.or_svf_fold(D._p(sh), D._p(veg), D._p(vb), n, D._p(w_iso), D._p(w_aniso), D._p(ring), D._p(na),
                      D._p(dirm), D._p(np.ascontiguousarray(pins[f"{name}_vegdem2"], np.float32)), D._p(out))
        want = pins[f"{name}_out"].reshape(11, n)
        got = np.concatenate([out[5:15], ((pins[f"{name}_svf"].ravel() - (np.float32(1) - out[5]) *
                                          (np.float32(1) - trans)).astype(np.float32))[None]])
        sel = np.ones(n, bool) if f"{name}_mask" not in pins else pins[f"{name}_mask"].ravel().astype(bool)
        assert np.array_equal(got[:, sel].view(np.uint32), want[:, sel].view(np.uint32)), str(name)


def test_a_planted_fold_mistake_fails(pins):
    """Dropping the ``last`` correction (S/W variants where vegdem2 == 0) fails a pinned case."""
    from oracle import driver as D
    L = D.lib()
    name = "special_planes"
    rows, cols = (int(x) for x in pins[f"{name}_shape"])
    n = rows * cols
    veg = np.ascontiguousarray(_unpack(pins[f"{name}_veg"]).reshape(n, 153).T)
    vb = np.ascontiguousarray(_unpack(pins[f"{name}_vbsh"]).reshape(n, 153).T)
    d = pins["dir"]
    out = np.empty((16, n), np.float32)
    vegdem2 = np.ones(n, np.float32)                   # planted: no cell takes the correction
    L.or_svf_fold(D._p(np.zeros_like(veg)), D._p(veg), D._p(vb), n,
                  D._p(np.ascontiguousarray(pins["w_iso"].reshape(8, 12))),
                  D._p(np.ascontiguousarray(pins["w_aniso"].reshape(8, 12))),
                  D._p(np.ascontiguousarray(pins["ring"])), D._p(np.ascontiguousarray(pins["na"])),
                  D._p(np.ascontiguousarray(d[0] | (d[1] << 1) | (d[2] << 2) | (d[3] << 3), np.int32)),
                  D._p(vegdem2), D._p(out))
    assert not np.array_equal(out[5:15], pins[f"{name}_out"].reshape(11, n)[:10])
