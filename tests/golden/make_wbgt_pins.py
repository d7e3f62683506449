#!/usr/bin/env python3
"""Write WBGT pins from the reference's own wet-bulb and black-globe code.

Runs ONLY where the reference checkout exists (``/root/reference``); the
GPU host never runs it. It imports ``solweig_gpu.calculate_wbgt`` with the
import-time stand-ins of ``make_golden.py`` and one for ``numba.vectorize``
that keeps numba's typing of the reference's ``_lambertw`` signatures
(``float32(float32)``, ``float64(float64)``): the body is evaluated in
float64 and the result rounded to the argument's dtype once.

What it pins (reference file:line):
  * ``isobaric_wet_bulb_temperature_from_rh`` (``calculate_wbgt.py:691-715``,
    liquid phase, Romps' Newton iteration ``:596-679``) on forcing tables
    shaped like the reference's call (``utci_process.py:817-823``: float32
    ``P * 1000``, ``Ta + uhii + 273.15`` and RH columns), one table per call
    because the iteration stops on the table-wide largest update;
  * ``black_globe_temperature`` (``calculate_wbgt.py:717-761``) on torch
    float32 CPU tensors, with ``h_cg = (6.3 / 0.46821) va^0.6`` and the
    shade/sun blend of ``utci_process.py:1021-1033``.

Usage: ``python tests/golden/make_wbgt_pins.py`` (seconds).
"""
from __future__ import annotations

import sys
import types
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))

import make_golden  # noqa: E402  (stand-ins and the reference path)


def _numba_standin():
    """``numba.vectorize`` that computes in float64 and returns the
    argument's float dtype, as numba's ``float32(float32)`` loop does."""
    numba = types.ModuleType("numba")

    def vectorize(*a, **k):
        def deco(f):
            g = np.vectorize(f, otypes=[np.float64])

            def call(x):
                x = np.asarray(x)
                out = g(x.astype(np.float64))
                return out.astype(x.dtype if x.dtype.kind == "f" else np.float64)
            return call
        return deco
    numba.vectorize = vectorize
    numba.njit = lambda *a, **k: (lambda f: f)
    sys.modules["numba"] = numba


def _tables(rng):
    """Forcing tables (P kPa, Ta degC, RH %) as float32 columns."""
    hours = np.arange(24, dtype=np.float64)
    day = {
        "P": np.full(24, 101.2, np.float32),
        "Ta": (24.0 + 8.0 * np.sin((hours - 9.0) / 24.0 * 2 * np.pi)).astype(np.float32),
        "RH": (60.0 - 25.0 * np.sin((hours - 9.0) / 24.0 * 2 * np.pi)).astype(np.float32),
    }
    ta, rh, p = np.meshgrid(np.linspace(-5.0, 45.0, 11), np.linspace(5.0, 100.0, 8),
                            np.array([86.0, 95.5, 104.0]), indexing="ij")
    grid = {"P": p.ravel().astype(np.float32), "Ta": ta.ravel().astype(np.float32),
            "RH": rh.ravel().astype(np.float32)}
    n = 200
    rand = {"P": rng.uniform(80.0, 106.0, n).astype(np.float32),
            "Ta": rng.uniform(-10.0, 48.0, n).astype(np.float32),
            "RH": rng.uniform(1.0, 100.0, n).astype(np.float32)}
    # saturated and nearly dry rows plus a missing humidity reading in one table
    edge = {"P": np.array([101.325, 101.325, 90.0, 101.325, 99.0], np.float32),
            "Ta": np.array([30.0, 0.5, 40.0, 20.0, 10.0], np.float32),
            "RH": np.array([100.0, 100.0, 1.0, np.nan, 35.0], np.float32)}
    return {"day": day, "grid": grid, "rand": rand, "edge": edge}


def main():
    _numba_standin()
    make_golden._install_stubs()
    sys.path.insert(0, str(make_golden.REF))
    import torch
    import solweig_gpu.calculate_wbgt as cw  # noqa: E402

    rng = np.random.default_rng(20261204)
    out = {}
    for name, t in _tables(rng).items():
        tw = cw.isobaric_wet_bulb_temperature_from_rh(
            p=t["P"] * 1000.0, T=t["Ta"] + 273.15, rh=t["RH"],
            phase="liquid", method="Romps", limit=True)
        for k, v in t.items():
            out[f"wb_{name}_{k}"] = v
        out[f"wb_{name}_tw"] = np.asarray(tw)

    # black globe and blend on a grid of wind, Tmrt and Ta (float32 tensors)
    va, tmrt, ta = np.meshgrid(np.array([0.15, 0.5, 1.0, 2.5, 5.0, 10.0, 20.0]),
                               np.linspace(-20.0, 75.0, 20), np.linspace(-5.0, 45.0, 6),
                               indexing="ij")
    va_t = torch.tensor(va.ravel(), dtype=torch.float32)
    tmrt_t = torch.tensor(tmrt.ravel(), dtype=torch.float32)
    ta_t = torch.tensor(ta.ravel(), dtype=torch.float32)
    hcg = (6.3 / 0.46821) * torch.pow(va_t, 0.6)
    tg = cw.black_globe_temperature(hcg, tmrt_t, ta_t, emissivity=0.95)
    shadow = torch.tensor((np.arange(va.size) % 3 == 0).astype(np.float32))
    wbt = torch.tensor(np.float32(21.375))
    cond = shadow < 0.1
    wbgt = torch.where(cond, 0.7 * wbt + 0.3 * tg, 0.7 * wbt + 0.2 * tg + 0.1 * ta_t)
    out.update(bg_va=va_t.numpy(), bg_tmrt=tmrt_t.numpy(), bg_ta=ta_t.numpy(),
               bg_hcg=hcg.numpy(), bg_tg=tg.numpy(), bg_shadow=shadow.numpy(),
               bg_wbt=np.float32(21.375), bg_wbgt=wbgt.numpy())
    np.savez_compressed(HERE / "wbgt_pins.npz", **out)
    print("wrote", HERE / "wbgt_pins.npz", len(out), "arrays")


if __name__ == "__main__":
    main()
