"""Fixture from the reference's own fold pins (data, not code).

Reads ``native/cuda/data/fold_pins.json`` of the reference: the 153-patch fold
tables of its canonical CPU fold (``solweig_core/numba_cpu/svf_fold.py``,
restating ``shadow.py:527-613``) as float32 bits (iso/aniso annulus weights per
ring and slot, ring, annulus counts, patch azimuths, E/S/W/N windows, the
``last`` and transmissivity constants) and 10 cases: vegetation and
vegetation-building shadow bits packed 20 bytes per pixel (``np.packbits``,
little bit order, patch p in bit p % 8 of byte p // 8), vegdem2 and the
building SVF plane in, the 11 vegetation/aveg/total SVF planes out.

Writes ``fold_pins.npz`` beside this script. Run in a checkout that has the
reference at ``/root/reference``.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/native/cuda/data/fold_pins.json")
OUT = Path(__file__).resolve().parent / "fold_pins.npz"
OUTPUTS = ("svfveg", "svfEveg", "svfSveg", "svfWveg", "svfNveg", "svfaveg", "svfEaveg", "svfSaveg", "svfWaveg",
           "svfNaveg", "svftotal")


def f32(h) -> np.ndarray:
    return np.array([int(x, 16) for x in h], dtype=np.uint64).astype(np.uint32).view(np.float32)


def u8(h) -> np.ndarray:
    return np.array([int(x, 16) for x in h], dtype=np.uint8)


def main():
    d = json.loads(SRC.read_text())
    t = d["tables"]
    out = {"w_iso": f32(t["w_iso"]), "w_aniso": f32(t["w_aniso"]), "ring": np.array(t["ring"], np.int32),
           "na": np.array(t["na"], np.int32), "azimuth": f32(t["azimuth"]),
           "dir": np.array([t["dir_e"], t["dir_s"], t["dir_w"], t["dir_n"]], np.int32),
           "consts": f32([t["last_bits"], t["trans_bits"], t["one_minus_trans_bits"]])}
    names = []
    for c in d["cases"]:
        name, rows, cols = c["name"], int(c["rows"]), int(c["cols"])
        names.append(name)
        out[f"{name}_shape"] = np.array([rows, cols], np.int32)
        out[f"{name}_veg"] = u8(c["veg_bytes"]).reshape(rows, cols, -1)
        out[f"{name}_vbsh"] = u8(c["vbsh_bytes"]).reshape(rows, cols, -1)
        out[f"{name}_vegdem2"] = f32(c["vegdem2"]).reshape(rows, cols)
        out[f"{name}_svf"] = f32(c["svf_building"]).reshape(rows, cols)
        out[f"{name}_out"] = np.stack([f32(c[k]).reshape(rows, cols) for k in OUTPUTS])
        if "mask" in c:
            out[f"{name}_mask"] = np.array(c["mask"], np.uint8).reshape(rows, cols)
    out["cases"] = np.array(names)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size / 1e6:.2f} MB), {len(names)} cases")


if __name__ == "__main__":
    main()
