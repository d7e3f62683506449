"""WBGT chain against values written by the reference's own code.

``tests/golden/wbgt_pins.npz`` holds what ``solweig_gpu/calculate_wbgt.py``
returned (``make_wbgt_pins.py``): the liquid-phase isobaric wet-bulb
temperature of four forcing tables (``isobaric_wet_bulb_temperature_from_rh``,
``:691-715``) and the black-globe temperature and shade/sun WBGT blend on a
wind x Tmrt x Ta grid (``black_globe_temperature``, ``:717-761``;
``utci_process.py:1021-1033``).

Bars: the wet bulb bit for bit (same float32 operation order, the Lambert-W
branch in float64 rounded once as numba's float32 loop does); h_cg within one
float32 ulp (torch's vectorised ``pow`` against numpy's); the black globe
within 2e-3 K and WBGT within 1e-3 degC -- the closed form subtracts two
square roots of ~5e2 in float32, so one ulp of a cube root moves Tg by up to
~8e-4 K (both results sit about 5e-3 K from the quartic's float64 root,
through the closed form's truncated constants).
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

from oracle import wbgt as owbgt  # noqa: E402
from solweig_project_amd import met  # noqa: E402

PINS = np.load(ROOT / "tests" / "golden" / "wbgt_pins.npz")
F32 = np.float32


@pytest.mark.parametrize("table", ["day", "grid", "rand", "edge"])
def test_wet_bulb_equals_the_reference_bit_for_bit(table):
    p = PINS[f"wb_{table}_P"]
    ta = PINS[f"wb_{table}_Ta"]
    rh = PINS[f"wb_{table}_RH"]
    ref = PINS[f"wb_{table}_tw"]
    assert p.dtype == ta.dtype == rh.dtype == np.float32
    got = met.wet_bulb_from_rh(p * 1000.0, ta + 273.15, rh)
    assert got.tw_c.dtype == ref.dtype
    assert np.array_equal(got.tw_c, ref, equal_nan=True)


def test_edge_table_keeps_the_missing_row_missing():
    ref = PINS["wb_edge_tw"]
    assert np.isnan(ref[3]) and np.isfinite(np.delete(ref, 3)).all()
    # saturated rows: the wet bulb is the air temperature
    ta = PINS["wb_edge_Ta"]
    np.testing.assert_allclose(ref[:2], ta[:2], atol=2e-3)


def test_globe_coefficient_within_one_ulp():
    va = PINS["bg_va"]
    hcg_ref = PINS["bg_hcg"]
    hcg = F32(met.GLOBE_HCG_COEF) * np.power(va, F32(0.6))
    assert np.all(np.abs(hcg - hcg_ref) <= np.spacing(hcg_ref))
    # the product's scalar form (per-step constant of the flux epilogue)
    for v in np.unique(va):
        a_ref = F32(hcg_ref[va == v][0]) / (F32(0.95) * F32(5.670374419e-8))
        a = F32(met.globe_coefficient(float(v)))
        assert abs(float(a) - float(a_ref)) <= 2 * float(np.spacing(a_ref))


def test_black_globe_and_wbgt_within_the_reference():
    a = owbgt.globe_coefficient(PINS["bg_va"])
    tg = owbgt.black_globe(a, PINS["bg_tmrt"], PINS["bg_ta"])
    ref = PINS["bg_tg"]
    assert np.array_equal(np.isfinite(tg), np.isfinite(ref))
    assert np.max(np.abs(tg.astype(np.float64) - ref)) <= 2e-3
    w = owbgt.wbgt(PINS["bg_tmrt"], PINS["bg_shadow"], PINS["bg_ta"],
                   F32(0.7 * float(PINS["bg_wbt"])), a)
    assert np.max(np.abs(w.astype(np.float64) - PINS["bg_wbgt"])) <= 1e-3


def test_blend_branch_follows_the_shadow_plane():
    # shade (shadow < 0.1) carries 0.3 Tg, sun 0.2 Tg + 0.1 Ta: a swapped
    # condition moves WBGT by 0.1 (Tg - Ta), far beyond the 1e-3 bar here
    a = owbgt.globe_coefficient(PINS["bg_va"])
    swapped = owbgt.wbgt(PINS["bg_tmrt"], F32(1.0) - PINS["bg_shadow"], PINS["bg_ta"],
                         F32(0.7 * float(PINS["bg_wbt"])), a)
    assert np.max(np.abs(swapped.astype(np.float64) - PINS["bg_wbgt"])) > 0.5
