"""The request loop (serving.Server): requests served concurrently (overlapping
slots, edits interleaved with in-flight met-only and full re-solves) give the
same results as each request served alone on the site version it was admitted
against. Results are compared through their checksums (NaN count, sum and a
position-weighted sum), which the deterministic kernels reproduce exactly.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-test] {msg}", flush=True)


class _InlinePlanner:
    """The planner interface without a child process (planned on submit)."""

    class _Done:
        def __init__(self, v):
            self.v = v

        def ready(self):
            return True

        def get(self, timeout=None):
            return self.v

    def submit(self, tile, job):
        from solweig_project_amd.solver import plan_day
        return self._Done(plan_day(tile, job.met, job.sun, job.latitude))


def _site():
    from solweig_project_amd import synth, sunpos
    from solweig_project_amd import walls as WG
    from solweig_project_amd.solver import TileInputs
    base = synth.make_site(288, 256, seed=91, with_aspect=False)
    base.dsm = (base.dsm * np.float32(0.3)).astype(np.float32)
    w, a = WG.walls_and_aspect(base.dsm, base.scale)
    tile = TileInputs.from_site(synth.site_with(base, base.dsm, base.cdsm, base.landcover, walls=w, aspect=a))
    sun = sunpos.day_sun_table(base.latitude, base.longitude, base.utc_offset, 2021, 180)
    mets = [synth.make_met(sun.altitude, seed=92 + k) for k in range(3)]
    return tile, sun, mets, base.latitude


def test_day_in_chunks_equals_the_whole_day():
    """A day enqueued in step ranges (as serving slots feed it) equals the day in one call, bit for bit."""
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")
    from solweig_project_amd.solver import HipSolver, _owned_plan, plan_day
    solver = HipSolver("cuda:0")
    tile, sun, mets, lat = _site()
    ds = solver.upload(tile)
    solver.svf(ds)
    plan = plan_day(tile, mets[1], sun, lat)
    dplan = _owned_plan(plan, solver.device)
    solver.tile_static(ds, dplan)
    outs = ("tmrt", "utci", "shadow", "kdown")
    whole = solver.run_day(ds, plan, dplan, outs).clone()
    part = torch.full_like(whole, -1.0)
    for t0 in range(0, plan.n_steps, 5):
        solver.run_day(ds, plan, dplan, outs, out=part, steps=(t0, min(plan.n_steps, t0 + 5)))
    torch.cuda.synchronize()
    assert sum(plan.is_day[5:]) > 0
    assert np.array_equal(whole.cpu().numpy(), part.cpu().numpy(), equal_nan=True)
    assert not torch.isnan(whole[:, 0]).any()          # Tmrt is defined everywhere, day and night


def test_requests_in_the_mix_equal_requests_alone():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")
    from solweig_project_amd import incremental as I
    from solweig_project_amd import serving as SV
    from solweig_project_amd.solver import HipSolver
    solver = HipSolver("cuda:0")
    tile, sun, mets, lat = _site()
    st = I.solve_site(solver, tile, mets[0], sun, lat)
    reqs = SV.poisson_workload(14, rate=400.0, mix={"edit": 0.5, "met": 0.35, "full": 0.15}, seed=5,
                               site_tile=st.tile, met_pool=mets)
    _say(f"serving {len(reqs)} requests: {[r.kind for r in reqs]}")
    server = SV.Server(solver, st, mets, sun, lat, _InlinePlanner(), slots=3)
    done = server.run(reqs)
    assert all(r.finish > 0 for r in done)
    overlapped = any(a.start < b.finish and b.start < a.finish for a in done for b in done
                     if a is not b and a.kind != "edit" and b.kind != "edit")
    _say(f"slots overlapped: {overlapped}")
    # alone: a fresh site, every request by itself, edits in order, others at their admission version
    tile2, _, _, _ = _site()
    st2 = I.solve_site(solver, tile2, mets[0], sun, lat)
    alone = SV.Server(solver, st2, mets, sun, lat, _InlinePlanner(), slots=1)
    order = sorted(done, key=lambda r: (r.payload["version"], 0 if r.kind == "edit" else 1))
    for r in order:
        twin = SV.Request(r.kind, 0.0, {k: v for k, v in r.payload.items() if k != "version"})
        alone.run([twin])
        assert twin.result == r.result, (r.kind, r.payload, twin.result, r.result)
