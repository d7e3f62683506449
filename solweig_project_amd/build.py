"""Build libsolweig_hip.so (gfx950) in-tree with hipcc.

Flags: -O3, wave64 gfx950 code object, -ffp-contract=off and no fast-math
(one rounding per arithmetic operation), packed-FP32 instruction selection
disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libsolweig_hip.so"
SOURCES = [CSRC / "sw_hip.hip", CSRC / "sw_runtime.hip", CSRC / "sw_walls.hip", CSRC / "sw_cnn.hip"]
# compiled on its own, with machine LICM off: the persistent day kernel holds the
# day-step bodies inside its phase loops (see the comment at the top of the file)
DAY_SOURCE = CSRC / "sw_day.hip"
DAY_FLAGS = ["-mllvm", "-disable-machine-licm"]
ARCH = os.environ.get("SW_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-shared", "-std=c++17",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libsolweig_hip.so)")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed:\n{res.stdout}\n{res.stderr}")


def build(verbose: bool = False) -> Path:
    subprocess.run([sys.executable, str(CSRC / "gen_utci_poly.py")], check=True,
                   stdout=subprocess.DEVNULL)
    tmp = OUT.with_suffix(".so.tmp")
    day_obj = PKG / "sw_day.o"
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    _run([hipcc(), *compile_flags, *DAY_FLAGS, "-c", "-o", str(day_obj), str(DAY_SOURCE)], verbose)
    try:
        _run([hipcc(), *HIPCC_FLAGS, "-o", str(tmp), *map(str, SOURCES), "-x", "none", str(day_obj)], verbose)
    finally:
        day_obj.unlink(missing_ok=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
