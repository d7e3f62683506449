"""Fixture from the reference's own radiation pins (data, not code).

Reads ``native/cuda/data/rad_pins.json`` of the reference: one 48 x 48 crop
(rows 120:168, cols 200:248) of a captured 0.5 m site at steps 1, 5, 23 (night)
and 12 (day). Each step holds the flat inputs of the reference's fused
radiation step (``solweig_core/numba_cpu/radiation.py`` ``_fused_day_kernel``
/ ``_fused_night_kernel``, restating ``solweig.py:75-2311``) and the planes its
canonical CPU evaluation wrote:

* float32 planes and per-patch tables as float32 bits, packed patch cubes as
  bytes (20 bytes per pixel, patch p in bit p % 8 of byte p // 8; the sunlit /
  shaded decisions 29 bytes per pixel: 153 Kside bits, then one bit per
  guarded longwave patch at ``kside_n + dp_rank``), integer tables as ints,
  float32 scalars as bits, float64 scalars as big-endian bits;
* outputs (Tmrt, K/L fluxes, Kside terms, TgOut, cardinal planes) and, for
  the day step, the six thermal-delay state planes after the step.

Writes ``rad_pins.npz`` beside this script, keys ``t{step}_{name}``. Run in a
checkout that has the reference at ``/root/reference``.
"""
from __future__ import annotations

import json
import struct
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/native/cuda/data/rad_pins.json")
OUT = Path(__file__).resolve().parent / "rad_pins.npz"
BYTE_KEYS = ("sh_pb", "veg_pb", "vbsh_pb", "sun_pb", "shd_pb")


def convert(key: str, v):
    if isinstance(v, list):
        if key in BYTE_KEYS:
            return np.array([int(x, 16) for x in v], dtype=np.uint8)
        if v and isinstance(v[0], str):
            return np.array([int(x, 16) for x in v], dtype=np.uint64).astype(np.uint32).view(np.float32)
        return np.array(v, dtype=np.int32)
    if isinstance(v, str) and len(v) == 16:
        return np.float64(struct.unpack(">d", bytes.fromhex(v))[0])
    if isinstance(v, str):
        return np.array([int(v, 16)], dtype=np.uint32).view(np.float32)[0]
    return np.int32(v)


def main():
    d = json.loads(SRC.read_text())
    out = {"crop": np.array([d["crop"][k] for k in ("r0", "r1", "c0", "c1")], np.int32)}
    steps = []
    for ts in d["timesteps"]:
        t = int(ts["t"])
        steps.append(t)
        out[f"t{t}_is_day"] = np.int32(bool(ts["is_day"]))
        for k, v in ts["inputs"].items():
            out[f"t{t}_{k}"] = convert(k, v)
        for k, v in ts["outputs"].items():
            out[f"t{t}_out_{k}"] = convert(k, v)
        for k, v in ts["next"].items():
            out[f"t{t}_next_{k}"] = convert(k, v)
    out["steps"] = np.array(steps, np.int32)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size / 1e6:.2f} MB), steps {steps}")


if __name__ == "__main__":
    main()
