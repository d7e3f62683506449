"""Run-API mirrors on the GPU against the reference's golden vectors and each other."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

SCENES = ["a48", "b40x56"]


def _say(msg):
    print(f"[gpu-test] {msg}", flush=True)


@pytest.fixture(scope="module")
def api():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")
    from solweig_project_amd import api as A
    return A


def _golden(golden_dir, name):
    return np.load(golden_dir / f"svf_{name}.npz"), np.load(golden_dir / f"day_{name}.npz")


def _ref_tuple(g):
    from solweig_project_amd.api import REF_ORDER
    return tuple(g[k].astype(np.float32) for k in REF_ORDER)


def _window_kw(g, d, **over):
    kw = dict(a=g["dsm"], temp1=g["cdsm"], temp2=g["dem"], walls=g["walls"], dirwalls=g["aspect"],
              met_file=d["met"], altitude=d["altitude"], azimuth=d["azimuth"], zen=d["zen"], jday=d["jday"],
              dectime=d["dectime"], altmax=d["altmax"],
              location={"latitude": float(d["latitude"]), "longitude": float(d["longitude"]), "altitude": 0.0},
              scale=float(g["scale"]), landcover_grid=g["landcover"],
              requested_variables=("utci", "tmrt", "shadow", "kdown", "kup", "ldown", "lup", "ta", "wind"))
    kw.update(over)
    return kw


@pytest.mark.parametrize("name", SCENES)
def test_svf_calculator_matches_reference(api, golden_dir, name):
    _say(f"svf_calculator {name}")
    g, _ = _golden(golden_dir, name)
    b = api.svf_calculator(2, amaxvalue=g["amax"], a=g["dsm"], vegdem=g["vegdsm"], vegdem2=g["vegdsm2"],
                           bush=np.zeros_like(g["dsm"]), scale=float(g["scale"]))
    got = [x.cpu().numpy() for x in b]
    for k, (x, nm) in enumerate(zip(got, api.REF_ORDER)):
        ref = g[nm].astype(np.float32)
        if nm in ("vegshmat", "vbshvegshmat", "shmat"):
            assert np.array_equal(x, ref), nm
        else:
            assert np.abs(x - ref).max() <= 1e-5, nm


@pytest.mark.parametrize("name", SCENES)
@pytest.mark.parametrize("bundle", ["reference_tuple", "device_bundle"])
def test_run_utci_window_matches_reference(api, golden_dir, name, bundle):
    _say(f"run_utci_window {name} {bundle}")
    g, d = _golden(golden_dir, name)
    if bundle == "reference_tuple":
        svf = _ref_tuple(g)
    else:
        svf = api.svf_calculator(2, amaxvalue=g["amax"], a=g["dsm"], vegdem=g["vegdsm"], vegdem2=g["vegdsm2"],
                                 bush=np.zeros_like(g["dsm"]), scale=float(g["scale"]))
    out = api.run_utci_window(**_window_kw(g, d, svf_bundle=svf))
    assert np.array_equal(out["shadow"], d["o_shadow"])
    assert np.abs(out["tmrt"] - d["o_tmrt"]).max() <= 0.01
    m = ~np.isnan(d["o_utci"])
    assert np.array_equal(np.isnan(out["utci"]), ~m)
    assert np.abs(out["utci"][m] - d["o_utci"][m]).max() <= 0.01
    for k in ("kdown", "kup", "ldown", "lup"):
        assert np.abs(out[k] - d["o_" + k]).max() <= 0.01, k
    met = d["met"]
    ta = (np.float32(met[:, 11]) + np.float32(met[:, 24])).astype(np.float32)
    assert np.array_equal(out["ta"][:, 0, 0], ta)
    assert np.array_equal(out["wind"][:, 0, 0], np.maximum(np.float32(met[:, 9]), np.float32(0.15)))
    assert all(v.dtype == np.float32 and v.shape == (24,) + g["dsm"].shape for v in out.values())


def test_warm_start_continues_the_cold_run_bitwise(api, golden_dir):
    _say("warm start at step 10")
    g, d = _golden(golden_dir, "a48")
    svf = _ref_tuple(g)
    cold = api.run_utci_window(**_window_kw(g, d, svf_bundle=svf))
    head, state = api.run_utci_window(**_window_kw(g, d, svf_bundle=svf, time_stop=10), return_final_state=True)
    assert state["next_step"] == 10
    tail = api.run_utci_window(**_window_kw(g, d, svf_bundle=svf, time_start=10, initial_state=state))
    for k in cold:
        assert np.array_equal(head[k], cold[k][:10], equal_nan=True), k
        assert np.array_equal(tail[k], cold[k][10:], equal_nan=True), k


def test_out_window_is_the_crop_of_the_full_run(api, golden_dir):
    _say("out_window crop, cropped cubes")
    g, d = _golden(golden_dir, "b40x56")
    svf = _ref_tuple(g)
    full = api.run_utci_window(**_window_kw(g, d, svf_bundle=svf))
    win = (5, 31, 7, 44)
    part = api.run_utci_window(**_window_kw(g, d, svf_bundle=svf, out_window=win))
    r0, r1, c0, c1 = win
    cropped = tuple(x[r0:r1, c0:c1] if x.ndim == 3 else x for x in svf)
    cropped = tuple(svf[i] if api.REF_ORDER[i] not in ("vegshmat", "vbshvegshmat", "shmat") else cropped[i]
                    for i in range(19))
    part2 = api.run_utci_window(**_window_kw(g, d, svf_bundle=cropped, out_window=win,
                                             svf_bundle_cubes_cropped=True))
    for k in full:
        assert np.array_equal(part[k], full[k][:, r0:r1, c0:c1], equal_nan=True), k
        assert np.array_equal(part2[k], full[k][:, r0:r1, c0:c1], equal_nan=True), k


def test_device_refusals(api, golden_dir):
    from solweig_project_amd.solver import UnsupportedInput
    g, d = _golden(golden_dir, "a48")
    svf = _ref_tuple(g)
    with pytest.raises(UnsupportedInput, match="windcoeff"):
        api.run_utci_window(**_window_kw(g, d, svf_bundle=svf, windcoeff=np.full(g["dsm"].shape, 2.0)))
    bad = list(svf)
    bad[api.REF_ORDER.index("shmat")] = bad[api.REF_ORDER.index("shmat")] * 0.5
    with pytest.raises(UnsupportedInput, match="cubes hold values"):
        api.run_utci_window(**_window_kw(g, d, svf_bundle=tuple(bad)))
    lc = np.array([[1, 0.18, 0.95, 0.58, -9.78, 15.0], [99, 0.2, 0.9, 0.58, -3.41, 15.0]])
    with pytest.raises(UnsupportedInput, match="missing from lc_class"):
        api.run_utci_window(**_window_kw(g, d, svf_bundle=svf, lc_class=lc))
    z = np.ones_like(g["dsm"])
    with pytest.raises(UnsupportedInput, match="bush"):
        api.svf_calculator(2, amaxvalue=g["amax"], a=g["dsm"], vegdem=g["vegdsm"], vegdem2=g["vegdsm2"], bush=z,
                           scale=1.0)


def test_core_adapter_solves_a_write_window(api, golden_dir):
    """SolveRequestView -> dispatch -> HIP backend; outputs are the write-window crop with global origins."""
    _say("core adapter write window")
    from solweig_project_amd import met as M
    from solweig_project_amd.adapter import solve_with_core_adapter
    from solweig_project_amd.solweig_core import ArrayView, PhysicalWindow, SolveRequestView, TimeCoverage
    g, d = _golden(golden_dir, "a48")
    dom = "golden/a48"
    req = SolveRequestView(logical_domain_id=dom, rows=48, cols=48, origin_x_m=0.0, origin_y_m=0.0,
                           pixel_size_m=1.0 / float(g["scale"]), read_window=PhysicalWindow(0, 48, 0, 48),
                           write_window=PhysicalWindow(6, 40, 3, 45), time=TimeCoverage(0, 24, 24),
                           requested_variables=("utci", "tmrt", "shadow"), profile_id="mi355x_v1", device="cuda",
                           dtype_plan={"building_dsm": "<f4", "dem": "<f4", "veg_canopy": "<f4", "walls": "<f4",
                                       "wall_aspect": "<f4", "landcover": "<f4"})
    names = {"building_dsm": "dsm", "dem": "dem", "veg_canopy": "cdsm", "walls": "walls", "wall_aspect": "aspect",
             "landcover": "landcover"}
    planes = {k: ArrayView.from_numpy(np.ascontiguousarray(g[v]), logical_domain_id=dom, freeze=True)
              for k, v in names.items()}
    sun = M.SunTable(d["altitude"][0], d["azimuth"][0], d["zen"][0], d["jday"][0], d["dectime"], d["altmax"][0])
    views = {}
    out = solve_with_core_adapter(req, planes, met=d["met"], sun=sun, latitude=float(d["latitude"]),
                                  output_views=views)
    assert set(out) == {"utci", "tmrt", "shadow"}
    sl = (slice(None), slice(6, 40), slice(3, 45))
    assert np.array_equal(out["shadow"], d["o_shadow"][sl])
    assert np.abs(out["tmrt"] - d["o_tmrt"][sl]).max() <= 0.01
    m = ~np.isnan(d["o_utci"][sl])
    assert np.abs(out["utci"][m] - d["o_utci"][sl][m]).max() <= 0.01
    assert views["tmrt"].global_origin == (0, 6, 3) and views["tmrt"].frozen
    assert not out["tmrt"].flags.writeable


def test_svf_cache_files_round_trip_through_run_utci_window(api, golden_dir, tmp_path):
    """svf_calculator -> the reference's cache files (svfs zip, SkyViewFactor tif, shadowmats npz)
    -> load -> run_utci_window: identical outputs to the run on the in-memory bundle."""
    from solweig_project_amd import svf_cache as SC
    _say("SVF cache files round trip")
    g, d = _golden(golden_dir, "a48")
    b = api.svf_calculator(2, amaxvalue=g["amax"], a=g["dsm"], vegdem=g["vegdsm"], vegdem2=g["vegdsm2"],
                           bush=np.zeros_like(g["dsm"]), scale=float(g["scale"]))
    host = tuple(x.cpu().numpy() if hasattr(x, "cpu") else np.asarray(x) for x in b)
    gt = (1000.0, 1.0 / float(g["scale"]), 0.0, 2000.0, 0.0, -1.0 / float(g["scale"]))
    SC.save_svf_cache(str(tmp_path), "0_0", host, gt)
    loaded = SC.load_svf_cache(str(tmp_path), "0_0", g["dsm"].shape, gt)
    for x, y in zip(loaded, host):
        assert np.array_equal(x, y)
    ref = api.run_utci_window(**_window_kw(g, d, svf_bundle=b))
    got = api.run_utci_window(**_window_kw(g, d, svf_bundle=loaded))
    for k in ref:
        assert np.array_equal(got[k], ref[k], equal_nan=True), k
