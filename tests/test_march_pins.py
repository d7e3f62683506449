"""The sun step tables and the sun march against the reference's own pins.

``tests/golden/march_site_pins.npz`` comes from the reference's
``native/cuda/data/march_site_pins.json`` (``make_march_pins.py``): a 500 x 500
site at 0.5 m, 13 daytime sun positions, the ``shadowingfunction_wallheight_23``
step tables the reference's torch stepping produced (``solweig.py:1079-1170``),
and the sha256 of the sh / vegsh / vbsh planes of its canonical CPU march.

* ``skyvault.sun_table`` (the product's table producer, also the oracle's)
  reproduces every table bit for bit (dx, dy, dz, dzprev);
* the C oracle's sun march on the pinned tables reproduces every plane's digest.
"""
import hashlib
from pathlib import Path

import numpy as np
import pytest

from solweig_project_amd import skyvault as SV

PINS = Path(__file__).parent / "golden" / "march_site_pins.npz"


@pytest.fixture(scope="module")
def pins():
    with np.load(PINS) as z:
        return {k: z[k] for k in z.files}


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


def test_step_tables_match_the_pins(pins):
    rows, cols, scale, amax = int(pins["rows"]), int(pins["cols"]), float(pins["scale"]), float(pins["amax"])
    for t in pins["steps"]:
        az, alt = (float(v) for v in pins[f"t{t}_azalt"])
        tab = SV.sun_table(az, alt, scale, amax, rows, cols)
        assert tab.count == pins[f"t{t}_dx"].size, t
        assert np.array_equal(tab.dx, pins[f"t{t}_dx"]) and np.array_equal(tab.dy, pins[f"t{t}_dy"]), t
        assert np.array_equal(np.asarray(tab.dz, np.float32).view(np.uint32), pins[f"t{t}_dz"].view(np.uint32)), t
        assert np.array_equal(np.asarray(tab.dzprev, np.float32).view(np.uint32),
                              pins[f"t{t}_dzprev"].view(np.uint32)), t


def test_oracle_sun_march_matches_the_pinned_digests(pins):
    from oracle import driver as D
    L = D.lib()
    rows, cols = int(pins["rows"]), int(pins["cols"])
    n = rows * cols
    a, veg, veg2 = (np.ascontiguousarray(pins[k], np.float32) for k in ("a", "vegdsm", "vegdsm2"))
    walls = np.zeros(n, np.float32)
    aspect = np.zeros(n, np.float32)
    out = np.empty(8 * n, np.float32)
    work = np.empty(25 * n, np.float32)
    for t in pins["steps"]:
        dx, dy = (np.ascontiguousarray(pins[f"t{t}_{k}"], np.int32) for k in ("dx", "dy"))
        dz, dzp = (np.ascontiguousarray(pins[f"t{t}_{k}"], np.float32) for k in ("dz", "dzprev"))
        az = float(pins[f"t{t}_azalt"][0])
        br, lo, hi = SV.sun_facesh_bounds(az)
        L.or_march_sun(D._p(a), D._p(veg), D._p(veg2), rows, cols, D._p(dx), D._p(dy), D._p(dz), D._p(dzp),
                       dx.size, D._p(walls), D._p(aspect), br, lo, hi, D._p(out), D._p(work))
        sh_want, veg_want, vb_want = (str(x) for x in pins[f"t{t}_sha"])
        assert digest(out[n:2 * n]) == sh_want, t
        assert digest(out[:n]) == veg_want, t
        assert digest(out[2 * n:3 * n]) == vb_want, t


def test_a_planted_table_error_fails(pins):
    """An off-by-one step index (each step with the next step's height) changes the digests."""
    from oracle import driver as D
    L = D.lib()
    rows, cols = int(pins["rows"]), int(pins["cols"])
    n = rows * cols
    a, veg, veg2 = (np.ascontiguousarray(pins[k], np.float32) for k in ("a", "vegdsm", "vegdsm2"))
    t = int(pins["steps"][0])
    dx, dy = (np.ascontiguousarray(pins[f"t{t}_{k}"], np.int32) for k in ("dx", "dy"))
    dz0 = np.ascontiguousarray(pins[f"t{t}_dz"], np.float32)
    dz = np.concatenate([dz0[1:], dz0[-1:]]).astype(np.float32)     # planted: dz of step k + 1 at step k
    dzp = np.concatenate([np.zeros(1, np.float32), dz[:-1]])
    out = np.empty(8 * n, np.float32)
    work = np.empty(25 * n, np.float32)
    z = np.zeros(n, np.float32)
    L.or_march_sun(D._p(a), D._p(veg), D._p(veg2), rows, cols, D._p(dx), D._p(dy), D._p(dz), D._p(dzp),
                   dx.size, D._p(z), D._p(z), -1, 0.0, 0.0, D._p(out), D._p(work))
    assert digest(out[n:2 * n]) != str(pins[f"t{t}_sha"][0]) or digest(out[:n]) != str(pins[f"t{t}_sha"][1])
