"""Runtime entry points of the C ABI on the GPU: memory, copies, events, the
composite SVF sweep and day loop, and graph capture/replay of a day.

Outputs of the one-call paths are compared bitwise with the per-launch paths
that the parity tests check against the reference.
"""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-test] {msg}", flush=True)


@pytest.fixture(scope="module")
def L():
    from solweig_project_amd import _lib
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")
    _lib.load()
    return _lib


@pytest.fixture(scope="module")
def day_case(L):
    from solweig_project_amd import met as M
    from solweig_project_amd import synth
    from solweig_project_amd.solver import HipSolver, TileInputs
    site = synth.make_site(120, 96, seed=31)
    from solweig_project_amd import sunpos
    sun = sunpos.day_sun_table(site.latitude, site.longitude, site.utc_offset, 2021, 180)
    met = synth.make_met(sun.altitude, seed=32)
    solver = HipSolver("cuda:0")
    tile = TileInputs.from_site(site)
    return dict(solver=solver, tile=tile, met=met, sun=sun, lat=site.latitude, M=M)


def test_device_name_and_mem_info(L):
    buf = ctypes.create_string_buffer(256)
    L.call("sw_device_name", 0, buf, 256)
    assert b"gfx950" in buf.value
    free, total = ctypes.c_int64(), ctypes.c_int64()
    L.call("sw_mem_info", ctypes.byref(free), ctypes.byref(total))
    assert 0 < free.value <= total.value


def test_alloc_counters_and_copies(L):
    live0, na0, nf0 = (ctypes.c_int64() for _ in range(3))
    L.call("sw_alloc_stats", ctypes.byref(live0), ctypes.byref(na0), ctypes.byref(nf0))
    n = 1 << 16
    d = ctypes.c_void_p()
    L.call("sw_alloc", ctypes.byref(d), n * 4)
    live, na, nf = (ctypes.c_int64() for _ in range(3))
    L.call("sw_alloc_stats", ctypes.byref(live), ctypes.byref(na), ctypes.byref(nf))
    assert live.value - live0.value == n * 4 and na.value == na0.value + 1
    h = ctypes.c_void_p()
    L.call("sw_host_alloc", ctypes.byref(h), n * 4)
    src = np.arange(n, dtype=np.float32)
    L.call("sw_copy", d, src.ctypes.data, n * 4)
    L.call("sw_fill_f32", ctypes.c_void_p(d.value + 4 * 16), ctypes.c_float(-2.5), 8, None)
    L.call("sw_copy_async", h, d, n * 4, None)
    L.call("sw_sync", None)
    got = np.ctypeslib.as_array((ctypes.c_float * n).from_address(h.value)).copy()
    exp = src.copy()
    exp[16:24] = -2.5
    assert np.array_equal(got, exp)
    # 2D crop: rows 3..9, cols 5..20 of a 256 x 256 image
    out = np.zeros((7, 16), np.float32)
    img = src.reshape(256, 256)
    L.call("sw_copy2d_async", out.ctypes.data, 16 * 4, ctypes.c_void_p(d.value + (3 * 256 + 5) * 4), 256 * 4,
           16 * 4, 7, None)
    L.call("sw_sync", None)
    ref = img[3:10, 5:21].copy()
    ref[(ref >= 16) & (ref < 24)] = -2.5
    assert np.array_equal(out, ref)
    L.call("sw_host_free", h)
    L.call("sw_free", d)
    L.call("sw_alloc_stats", ctypes.byref(live), ctypes.byref(na), ctypes.byref(nf))
    assert live.value == live0.value and nf.value == nf0.value + 1
    with pytest.raises(L.HipError, match="not allocated by sw_alloc"):
        L.call("sw_free", ctypes.c_void_p(d.value + 4096))


def test_streams_and_events(L):
    s, e0, e1 = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    L.call("sw_stream_create", ctypes.byref(s))
    L.call("sw_event_create", ctypes.byref(e0))
    L.call("sw_event_create", ctypes.byref(e1))
    x = torch.empty(1 << 22, device="cuda:0")
    L.call("sw_event_record", e0, s)
    L.call("sw_fill_f32", L.ptr(x), ctypes.c_float(1.0), x.numel(), s)
    L.call("sw_event_record", e1, s)
    L.call("sw_event_sync", e1)
    ms = ctypes.c_float()
    L.call("sw_event_elapsed", e0, e1, ctypes.byref(ms))
    assert ms.value > 0
    L.call("sw_stream_wait_event", None, e1)
    assert float(x.sum()) == float(x.numel())
    for h in (e0, e1):
        L.call("sw_event_destroy", h)
    L.call("sw_stream_destroy", s)


def _prepared(day_case):
    from solweig_project_amd.solver import plan_day
    solver, tile = day_case["solver"], day_case["tile"]
    ds = solver.upload(tile)
    solver.svf(ds)
    plan = plan_day(tile, day_case["met"], day_case["sun"], day_case["lat"])
    dplan = solver.upload_plan(plan)
    solver.tile_static(ds, dplan)
    return ds, plan, dplan


def test_svf_sweep_one_call_matches_split(L, day_case):
    _say("sw_svf_sweep vs march+fold")
    solver = day_case["solver"]
    ds, _, _ = _prepared(day_case)
    torch.cuda.synchronize()
    ref_fields = ds.fields.clone()
    ref_masks = ds.masks.clone()
    ds.fields.zero_()
    ds.masks.zero_()
    offs, tab = solver.svf_tables(ds.rows, ds.cols, ds.tile.scale)
    c = solver._fold_consts
    P = L.ptr
    L.call("sw_svf_sweep", P(ds.hpack), P(ds.stats), P(offs), P(tab), P(ds.vegdsm2), P(c["w_iso"]),
           P(c["w_aniso"]), P(c["ring"]), P(c["nann"]), P(c["dirm"]), P(ds.counts), P(ds.hpad), P(ds.masks),
           P(ds.fields), ds.rows, ds.cols, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert torch.equal(ds.masks, ref_masks)
    assert torch.equal(ds.fields, ref_fields)


def _per_launch_day(day_case, ds, plan, dplan):
    solver = day_case["solver"]
    out = solver.run_day(ds, plan, dplan, ("tmrt", "utci", "shadow", "kdown", "lup"), per_launch=True)
    torch.cuda.synchronize()
    return out.clone()


def test_timed_day_matches_and_reports_every_launch(L, day_case):
    _say("sw_solve_day_timed")
    from solweig_project_amd.solver import KernelTimer
    solver = day_case["solver"]
    ds, plan, dplan = _prepared(day_case)
    ref = _per_launch_day(day_case, ds, plan, dplan)
    old = solver.timer
    solver.timer = KernelTimer(reserve=3 * plan.n_steps)
    try:
        out = solver.run_day(ds, plan, dplan, ("tmrt", "utci", "shadow", "kdown", "lup"))
        torch.cuda.synchronize()
        summ = solver.timer.summary()
    finally:
        solver.timer.close()
        solver.timer = old
    assert torch.equal(torch.nan_to_num(out, nan=-1e30), torch.nan_to_num(ref, nan=-1e30))
    nday = sum(plan.is_day)
    assert summ["sun_step"][0] == nday and summ["flux_day"][0] == nday
    assert summ["flux_night"][0] == plan.n_steps - nday
    assert all(t > 0 for _, t in summ.values())


def test_solve_day_one_call_matches_per_launch(L, day_case):
    _say("sw_solve_day vs per-launch day")
    solver = day_case["solver"]
    ds, plan, dplan = _prepared(day_case)
    ref = _per_launch_day(day_case, ds, plan, dplan)
    out = solver.run_day(ds, plan, dplan, ("tmrt", "utci", "shadow", "kdown", "lup"))
    torch.cuda.synchronize()
    assert torch.equal(torch.nan_to_num(out, nan=-1e30), torch.nan_to_num(ref, nan=-1e30))


def test_day_graph_replays_match_eager(L, day_case):
    _say("captured day graph, 3 replays")
    ds, plan, dplan = _prepared(day_case)
    ref = _per_launch_day(day_case, ds, plan, dplan)
    T = plan.n_steps
    mask = 1 | 2 | 4 | 8 | 64
    out = torch.full_like(ref, 7.0)
    s = ctypes.c_void_p()
    L.call("sw_stream_create", ctypes.byref(s))
    is_day = (ctypes.c_int32 * T)(*[int(x) for x in plan.is_day])
    P = L.ptr
    tile = ds.tile_struct()
    L.call("sw_graph_begin", s)
    rc_day = L.load().sw_solve_day(ctypes.byref(tile), P(dplan["steps"]), P(dplan["model"]),
                                   P(dplan["sun_tab"]), is_day, 0, T, P(out), mask, plan.walk_halo, 1, s)
    g = ctypes.c_void_p()
    L.call("sw_graph_end", s, ctypes.byref(g))
    assert rc_day == 0
    try:
        for rep in range(3):
            out.fill_(float(rep))               # eager work between replays
            torch.cuda.synchronize()
            L.call("sw_graph_launch", g, s)
            L.call("sw_sync", s)
            assert torch.equal(torch.nan_to_num(out, nan=-1e30), torch.nan_to_num(ref, nan=-1e30)), rep
    finally:
        L.call("sw_graph_destroy", g)
        L.call("sw_stream_destroy", s)


def test_solve_day_rejects_bad_arguments(L, day_case):
    ds, plan, dplan = _prepared(day_case)
    P = L.ptr
    is_day = (ctypes.c_int32 * 1)(0)
    with pytest.raises(L.HipError, match="step range"):
        L.call("sw_solve_day", ctypes.byref(ds.tile_struct()), P(dplan["steps"]), P(dplan["model"]),
               P(dplan["sun_tab"]), is_day, 5, 2, P(ds.state), 1, plan.walk_halo, 0, None)
    with pytest.raises(L.HipError, match="out_mask"):
        L.call("sw_solve_day", ctypes.byref(ds.tile_struct()), P(dplan["steps"]), P(dplan["model"]),
               P(dplan["sun_tab"]), is_day, 0, 1, P(ds.state), 0, plan.walk_halo, 0, None)


class _InlinePlanner:
    """solve_many's planner interface, planning in this process (no child process after GPU init)."""

    class _Done:
        def __init__(self, v):
            self.v = v

        def get(self, timeout=None):
            return self.v

    def submit(self, tile, job):
        from solweig_project_amd.solver import plan_day
        return self._Done(plan_day(tile, job.met, job.sun, job.latitude, time_start=job.time_start,
                                   time_stop=job.time_stop))


def test_pipelined_solves_match_solve_device(L, day_case):
    _say("solve_many vs solve_device")
    from solweig_project_amd.pipeline import Job, solve_many
    solver, tile = day_case["solver"], day_case["tile"]
    ds = solver.upload(tile)
    _, ref = solver.solve_device(ds, day_case["met"], day_case["sun"], day_case["lat"])
    ref = ref.clone()
    job = Job(day_case["met"], day_case["sun"], day_case["lat"])
    out = solve_many(solver, ds, [job, job], _InlinePlanner())
    torch.cuda.synchronize()
    assert torch.equal(torch.nan_to_num(out, nan=-1e30), torch.nan_to_num(ref, nan=-1e30))


def test_copy_kernel_upload_from_pinned_memory():
    """sw_copy_to_device_cu (a copy kernel reading pinned host memory) delivers every byte, for sizes with
    and without a tail past the last 16-byte word."""
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")
    from solweig_project_amd import _lib as L
    print("[gpu-test] copy-kernel upload", flush=True)
    rng = np.random.default_rng(5)
    for n in (1, 15, 16, 17, 4103, 3 * (1 << 20) + 5):
        host = torch.from_numpy(rng.integers(0, 256, n, dtype=np.uint8)).pin_memory()
        dev = torch.full((n + 64,), 7, dtype=torch.uint8, device="cuda:0")
        s = torch.cuda.Stream("cuda:0")
        L.call("sw_copy_to_device_cu", dev.data_ptr(), host.data_ptr(), n, s.cuda_stream)
        s.synchronize()
        assert torch.equal(dev[:n].cpu(), host), n
        assert bool((dev[n:] == 7).all()), n          # nothing written past the end
