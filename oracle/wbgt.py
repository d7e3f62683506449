"""WBGT oracle -- TEST INFRASTRUCTURE ONLY (never imported by the product).

Per-pixel wet-bulb globe temperature of one time step in numpy float32, in
the reference's tensor operation order:

* black-globe temperature, ``black_globe_temperature``
  (``solweig_gpu/calculate_wbgt.py:717-761``): the closed-form root of
  Tg^4 + a Tg = (Tmrt + 273.15)^4 + a (Ta + 273.15), a = h_cg / (eps sigma);
* ``h_cg = (6.3 / 0.46821) * va^0.6`` and the sun/shade blend
  (``solweig_gpu/utci_process.py:1021-1033``): where ``shadow < 0.1``
  (shade, and every pixel at night) 0.7 Tw + 0.3 Tg, elsewhere
  0.7 Tw + 0.2 Tg + 0.1 Ta.

The wet-bulb temperature Tw is a per-step scalar of the host chain
(``solweig_project_amd.met.wet_bulb_from_rh``), pinned separately.

Pinned by (tests/test_wbgt_oracle.py): the quartic's real root from
``numpy.roots`` in float64 (the closed form's truncated constants
1.73205 / 3.4943 / 0.381571 stay within 0.01 K of it); Tg = Ta when
Tmrt = Ta; float32 overflow to -inf at large h_cg; NaN propagation; the
ISO 7243 weights (0.7/0.3 without solar load, 0.7/0.2/0.1 with it).
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
TINY = np.finfo(np.float32).tiny


def globe_coefficient(va):
    """a = h_cg / (emissivity * sigma), float32 (utci_process.py:1022-1025, calculate_wbgt.py:733-740)."""
    va = np.asarray(va, F32)
    hcg = F32(6.3 / 0.46821) * np.power(va, F32(0.6))
    return hcg / (F32(0.95) * F32(5.670374419e-8))


def black_globe(a, tmrt, ta):
    """Black-globe temperature (degC), float32 (calculate_wbgt.py:742-761)."""
    a = np.asarray(a, F32)
    tmrt = np.asarray(tmrt, F32)
    ta = np.asarray(ta, F32)
    with np.errstate(all="ignore"):
        b = np.power(tmrt + F32(273.15), F32(4)) + a * (ta + F32(273.15))
        m = F32(9.0) * (a * a)
        n = F32(27.0) * np.power(a, F32(4))
        p = F32(256.0) * ((b * b) * b)
        E = np.power(m + F32(1.73205) * np.sqrt(n + p), F32(1.0 / 3.0))
        Q = F32(3.4943) * b
        k = F32(0.381571) * E - Q / E
        ks = np.where(k < TINY, TINY, k).astype(F32)          # clamp(min) keeps NaN
        inner = F32(2.0) * a / np.sqrt(ks) - k
        inner = np.where(inner < 0, F32(0), inner).astype(F32)
        kz = np.where(k < 0, F32(0), k).astype(F32)
        return (F32(0.5) * np.sqrt(inner) - F32(0.5) * np.sqrt(kz) - F32(273.15)).astype(F32)


def wbgt(tmrt, shadow, ta, wbgt_w, a):
    """WBGT plane of one step: ``wbgt_w`` = f32(0.7 Tw), ``a`` the globe coefficient (utci_process.py:1031-1033)."""
    tg = black_globe(a, tmrt, ta)
    w = F32(wbgt_w)
    ta = np.asarray(ta, F32)
    with np.errstate(all="ignore"):
        no_solar = w + F32(0.3) * tg                    # shadow < 0.1: shaded, or night
        solar = (w + F32(0.2) * tg) + F32(0.1) * ta
    return np.where(np.asarray(shadow) < F32(0.1), no_solar, solar).astype(F32)
