"""Write ``sun_tabs_day.npz``: the day-loop sun-march tables of the two golden scenes.

The host scalar chain (``solweig_project_amd.skyvault.sun_table`` through ``solver.plan_day``)
evaluates float32 transcendentals with torch, whose CPU kernels go through MKL's vector math.
MKL takes a different code path on AMD hosts than on the Intel host the golden vectors were
recorded on, so a handful of dz values differ by an ulp between machines (the reference itself
shares this property). This script records the tables as computed on the recording host, where
the oracle reproduces every plane of the reference's wall tail bit for bit with them
(``tests/test_oracle_golden.py``). The GPU test feeds them to the device kernel so its planes can
be compared with the reference's bit for bit on any host.

Per scene: ``{name}_sun_tab`` int32 (S, 4) rows {dx, dy, dz bits, dzprev bits},
``{name}_sun_off`` / ``{name}_sun_count`` int32 (T,) per step, ``{name}_is_day`` int32 (T,).
Run from the repository root: ``python tests/golden/make_sun_tabs.py``.
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from solweig_project_amd import _lib as L  # noqa: E402
from solweig_project_amd import met as M  # noqa: E402
from solweig_project_amd.solver import compose_scene, plan_day  # noqa: E402

HERE = Path(__file__).resolve().parent


def main():
    out = {}
    for name in ("a48", "b40x56"):
        g = np.load(HERE / f"svf_{name}.npz")
        d = np.load(HERE / f"day_{name}.npz")
        scene = compose_scene(g["dsm"], g["dem"], g["cdsm"], g["walls"], g["aspect"], g["landcover"],
                              scale=float(g["scale"]))
        sun = M.SunTable(d["altitude"][0], d["azimuth"][0], d["zen"][0], d["jday"][0], d["dectime"],
                         d["altmax"][0])
        plan = plan_day(scene, d["met"], sun, float(d["latitude"]))
        rec = np.frombuffer(np.ascontiguousarray(plan.steps).tobytes(), dtype=np.dtype(L.StepStruct))
        out[f"{name}_sun_tab"] = np.ascontiguousarray(plan.sun_tab, np.int32)
        out[f"{name}_sun_off"] = rec["sun_off"].astype(np.int32)
        out[f"{name}_sun_count"] = rec["sun_count"].astype(np.int32)
        out[f"{name}_is_day"] = rec["is_day"].astype(np.int32)
    np.savez_compressed(HERE / "sun_tabs_day.npz", **out)


if __name__ == "__main__":
    main()
