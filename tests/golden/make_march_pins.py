"""Fixture from the reference's own march pins (data, not code).

Reads ``native/cuda/data/march_site_pins.json`` of the reference: the site_500
anchor its canonical CPU march (``solweig_core/numba_cpu/march.py``) was pinned
on. The file holds the composed heights (a, vegdsm, vegdsm2) of a 500 x 500
site at 0.5 m, the effective amax, and for each of the 13 daytime steps the
sun azimuth/altitude (float32 bits), the ``shadowingfunction_wallheight_23``
step table (dx, dy as decimal integers, dz, dzprev as float32 bits) and the sha256 of the float32 bytes of the
sh / vegsh / vbsh planes.

Writes ``march_site_pins.npz`` beside this script (heights as float32 planes,
tables as arrays, digests as strings). Run in a checkout that has the
reference at ``/root/reference``.
"""
from __future__ import annotations

import json
import struct
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/native/cuda/data/march_site_pins.json")
OUT = Path(__file__).resolve().parent / "march_site_pins.npz"


def f32_from_hex(h) -> np.ndarray:
    u = np.array([int(x, 16) for x in h], dtype=np.uint64).astype(np.uint32)
    return u.view(np.float32)


def i32_from_dec(h) -> np.ndarray:
    """Integer columns are written as decimal strings (floats as hex bits)."""
    return np.array([int(x) for x in h], dtype=np.int32)


def bits_to_f32(b: str) -> np.float32:
    return np.frombuffer(struct.pack("<I", int(b, 16)), np.float32)[0]


def main():
    d = json.loads(SRC.read_text())
    s = d["site"]
    rows, cols = int(s["rows"]), int(s["cols"])
    out = dict(rows=np.int32(rows), cols=np.int32(cols), scale=np.float64(float(s["scale_repr"])),
               amax=bits_to_f32(s["amax_bits"]))
    for k in ("a", "vegdsm", "vegdsm2"):
        out[k] = f32_from_hex(s[k]).reshape(rows, cols)
    steps = sorted(s["per_t"], key=int)
    out["steps"] = np.array([int(t) for t in steps], np.int32)
    for t in steps:
        e = s["per_t"][t]
        out[f"t{t}_azalt"] = np.array([bits_to_f32(e["az_bits"]), bits_to_f32(e["alt_bits"])], np.float32)
        out[f"t{t}_dx"] = i32_from_dec(e["dx"])
        out[f"t{t}_dy"] = i32_from_dec(e["dy"])
        out[f"t{t}_dz"] = f32_from_hex(e["dz"])
        out[f"t{t}_dzprev"] = f32_from_hex(e["dzprev"])
        out[f"t{t}_sha"] = np.array([e["sh_sha256"], e["vegsh_sha256"], e["vbsh_sha256"]])
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size / 1e6:.2f} MB), {len(steps)} steps")


if __name__ == "__main__":
    main()
