"""A large tile (8192 x 8192 = 67 Mpx, 24 h) against a crop of itself.

Every output of a pixel depends only on heights within the reach of its rays:
the SVF march ends where dz exceeds amax (amax * scale / tan 6 deg cells away),
the sun march likewise at the day's lowest sun, the ground walk within 22
cells. A crop that keeps a halo wider than the longest reach and has the same
amax therefore gives the interior pixels bit-identical masks, SVF fields and
day outputs. The tile repeats one generated 1024 x 1024 city 8 x 8 times; any
crop wider than 1024 cells holds a full period, so its amax equals the tile's.

This checks the full solve at a size where every plane offset exceeds 2^31
bytes and the day's output volume exceeds 2^31 floats.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _bits(x):
    return x.contiguous().view(torch.int32)


def test_large_tile_equals_a_crop_with_halo():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")
    from solweig_project_amd import sunpos, synth
    from solweig_project_amd.solver import DeviceScene, HipSolver, compose_scene, upload_inputs
    print("[gpu-test] 8192x8192 tile vs crop with halo", flush=True)
    base = synth.make_site(1024, 1024, seed=91)
    reps = 8
    big = {k: np.tile(np.asarray(getattr(base, k), np.float32), (reps, reps))
           for k in ("dsm", "dem", "cdsm", "walls", "aspect", "landcover")}
    rows = cols = 1024 * reps
    sun = sunpos.day_sun_table(base.latitude, base.longitude, base.utc_offset, 2021, 180)
    met = synth.make_met(sun.altitude, seed=92)
    amax = float(max(big["dsm"].max(), (big["cdsm"] + big["dem"]).max()))
    low = float(sun.altitude[sun.altitude > 0].min())
    reach = math.ceil(amax * base.scale / math.tan(math.radians(min(low, 6.0)))) + 24
    halo = int(math.ceil(reach / 64.0) * 64)
    m = 1024
    r0 = c0 = 3 * 1024 + 512
    assert r0 - halo >= 0 and r0 + m + halo <= rows and m + 2 * halo > 1024
    solver = HipSolver("cuda:0", tune_march=False)
    dev = solver.device

    def solve(arrs):
        tile = compose_scene(arrs["dsm"], arrs["dem"], arrs["cdsm"], arrs["walls"], arrs["aspect"],
                             arrs["landcover"], scale=base.scale)
        ds = DeviceScene(tile, dev, upload_inputs(tile, dev))
        fields, out = solver.solve_device(ds, met, sun, base.latitude)
        torch.cuda.synchronize()
        return ds, fields, out

    ds_b, f_b, o_b = solve(big)
    print(f"[gpu-test] big tile solved; crop {m}+2x{halo}", flush=True)
    crop = {k: np.ascontiguousarray(v[r0 - halo:r0 + m + halo, c0 - halo:c0 + m + halo]) for k, v in big.items()}
    ds_c, f_c, o_c = solve(crop)
    w = m + 2 * halo
    fb = f_b.view(-1, rows, cols)[:, r0:r0 + m, c0:c0 + m]
    fc = f_c.view(-1, w, w)[:, halo:halo + m, halo:halo + m]
    assert torch.equal(_bits(fb), _bits(fc))
    mb = ds_b.masks.view(-1, rows, cols)[:, r0:r0 + m, c0:c0 + m]
    mc = ds_c.masks.view(-1, w, w)[:, halo:halo + m, halo:halo + m]
    assert torch.equal(mb.contiguous(), mc.contiguous())
    ob = o_b[:, :, r0:r0 + m, c0:c0 + m]
    oc = o_c[:, :, halo:halo + m, halo:halo + m]
    assert torch.equal(_bits(ob), _bits(oc))
    # the whole tile: Tmrt finite everywhere, UTCI NaN exactly off the ground
    assert bool(torch.isfinite(o_b[:, 0]).all())
    ground = (ds_b.buildings.view(rows, cols) == 1.0)
    assert bool((torch.isnan(o_b[:, 1]) == ~ground).all())
