"""Consecutive solves alternating over two scene copies on their own streams (``solve_many``
lanes, the bench's default) give every job the outputs it gets when solved alone."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


class _InlinePlanner:
    """The planner interface without a child process (planned on submit)."""

    class _Done:
        def __init__(self, v):
            self.v = v

        def get(self, timeout=None):
            return self.v

    def submit(self, tile, job):
        from solweig_project_amd.solver import plan_day
        return self._Done(plan_day(tile, job.met, job.sun, job.latitude))


def test_two_lanes_equal_solves_alone():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")
    from solweig_project_amd import sunpos, synth
    from solweig_project_amd.pipeline import Job, solve_many
    from solweig_project_amd.solver import DeviceScene, HipSolver, TileInputs, upload_inputs
    print("[gpu-test] two lanes vs alone", flush=True)
    site = synth.make_site(160, 144, seed=61)
    tile = TileInputs.from_site(site)
    sun = sunpos.day_sun_table(site.latitude, site.longitude, site.utc_offset, 2021, 180)
    jobs = [Job(synth.make_met(sun.altitude, seed=62 + k), sun, site.latitude) for k in range(4)]
    solver = HipSolver("cuda:0")
    dev = solver.device
    T = jobs[0].met.shape[0]
    lanes = []
    for _ in range(2):
        ds = DeviceScene(tile, dev, upload_inputs(tile, dev))
        lanes.append((ds, torch.cuda.Stream(dev), torch.full((T, 3, 160, 144), -1.0, device=dev)))
    solve_many(solver, lanes[0][0], jobs, _InlinePlanner(), lanes=lanes)
    torch.cuda.synchronize()
    # lane k ends with job 2 + k
    for k in range(2):
        ds = DeviceScene(tile, dev, upload_inputs(tile, dev))
        alone = solve_many(solver, ds, [jobs[2 + k]], _InlinePlanner(), out=torch.empty_like(lanes[k][2]))
        torch.cuda.synchronize()
        assert np.array_equal(lanes[k][2].cpu().numpy(), alone.cpu().numpy(), equal_nan=True), k
        assert torch.equal(lanes[k][0].fields, ds.fields)


@pytest.mark.parametrize("fields_last", [True, False])
def test_host_to_host_solves_deliver_what_the_device_computes(fields_last):
    """The host-to-host pipeline (pinned inputs uploaded per job, day in chunks, each chunk copied to
    the job's pinned host buffers on the copy engines, SVF fields copied before or after the day)
    delivers every job's fields and outputs as the job solved alone computes them, over three lanes
    and a ring of two device output buffers (so a buffer is reused while earlier copies run)."""
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")
    from solweig_project_amd import sunpos, synth
    from solweig_project_amd.pipeline import HostBuffers, Job, pinned_inputs, solve_many, solve_many_host
    from solweig_project_amd.solver import DeviceScene, HipSolver, TileInputs, upload_inputs
    print(f"[gpu-test] host-to-host lanes, fields_last={fields_last}", flush=True)
    site = synth.make_site(160, 144, seed=71)
    tile = TileInputs.from_site(site)
    sun = sunpos.day_sun_table(site.latitude, site.longitude, site.utc_offset, 2021, 180)
    jobs = [Job(synth.make_met(sun.altitude, seed=72 + k), sun, site.latitude) for k in range(5)]
    solver = HipSolver("cuda:0")
    dev = solver.device
    T = jobs[0].met.shape[0]
    lanes = [(DeviceScene(tile, dev, upload_inputs(tile, dev)), torch.cuda.Stream(dev), None) for _ in range(3)]
    host_in = pinned_inputs(tile)
    hb = [HostBuffers(torch.full((16, 160, 144), -1.0).pin_memory(), torch.full((T, 3, 160, 144), -1.0).pin_memory())
          for _ in jobs]
    dev_outs = [torch.empty((T, 3, 160, 144), device=dev) for _ in range(2)]
    done = solve_many_host(solver, lanes, jobs, _InlinePlanner(), host_in, hb, dev_outs, chunk_steps=5,
                           fields_last=fields_last)
    assert len(done) == len(jobs)
    torch.cuda.synchronize()
    for k, job in enumerate(jobs):
        ds = DeviceScene(tile, dev, upload_inputs(tile, dev))
        alone = solve_many(solver, ds, [job], _InlinePlanner(), out=torch.empty_like(dev_outs[0]))
        torch.cuda.synchronize()
        assert np.array_equal(hb[k].out.numpy(), alone.cpu().numpy(), equal_nan=True), k
        assert np.array_equal(hb[k].fields.numpy().reshape(16, -1), ds.fields.cpu().numpy()), k


def test_threaded_host_to_host_solves_deliver_what_the_device_computes():
    """The copy-thread pipeline (downloads issued by a host thread on a stream of their own after each
    chunk's event, device rings of two day-output and two field buffers, so a ring slot is reused
    while earlier downloads may still run) delivers every job's fields and outputs as the job solved
    alone computes them, over two lanes."""
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")
    from solweig_project_amd import sunpos, synth
    from solweig_project_amd.pipeline import (HostBuffers, Job, pinned_inputs, solve_many,
                                              solve_many_host_threaded)
    from solweig_project_amd.solver import DeviceScene, HipSolver, TileInputs, upload_inputs
    print("[gpu-test] threaded host-to-host lanes", flush=True)
    site = synth.make_site(150, 170, seed=73)
    tile = TileInputs.from_site(site)
    sun = sunpos.day_sun_table(site.latitude, site.longitude, site.utc_offset, 2021, 180)
    jobs = [Job(synth.make_met(sun.altitude, seed=74 + k), sun, site.latitude) for k in range(6)]
    solver = HipSolver("cuda:0")
    dev = solver.device
    T = jobs[0].met.shape[0]
    lanes = [(DeviceScene(tile, dev, upload_inputs(tile, dev)), torch.cuda.Stream(dev)) for _ in range(2)]
    host_in = pinned_inputs(tile)
    hb = [HostBuffers(torch.full((16, 150, 170), -1.0).pin_memory(), torch.full((T, 3, 150, 170), -1.0).pin_memory())
          for _ in jobs]
    dev_outs = [torch.empty((T, 3, 150, 170), device=dev) for _ in range(2)]
    fields_ring = [torch.empty_like(lanes[0][0].fields) for _ in range(2)]
    done = solve_many_host_threaded(solver, lanes, jobs, _InlinePlanner(), host_in, hb, dev_outs, fields_ring,
                                    chunk_steps=7)
    assert len(done) == len(jobs)
    torch.cuda.synchronize()
    for k, job in enumerate(jobs):
        ds = DeviceScene(tile, dev, upload_inputs(tile, dev))
        alone = solve_many(solver, ds, [job], _InlinePlanner(), out=torch.empty_like(dev_outs[0]))
        torch.cuda.synchronize()
        assert np.array_equal(hb[k].out.numpy(), alone.cpu().numpy(), equal_nan=True), k
        assert np.array_equal(hb[k].fields.numpy().reshape(16, -1), ds.fields.cpu().numpy()), k


def test_threaded_host_to_host_reports_a_failed_download():
    """A download the library refuses (a missing host buffer: sw_copy_to_host returns SW_E_ARG before any
    copy) is raised in the caller; the copy thread keeps releasing the caller's waits, so nothing hangs."""
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")
    from solweig_project_amd import sunpos, synth
    from solweig_project_amd.pipeline import HostBuffers, Job, pinned_inputs, solve_many_host_threaded
    from solweig_project_amd.solver import DeviceScene, HipSolver, TileInputs, upload_inputs
    print("[gpu-test] threaded host-to-host with a refused download", flush=True)
    site = synth.make_site(64, 80, seed=75)
    tile = TileInputs.from_site(site)
    sun = sunpos.day_sun_table(site.latitude, site.longitude, site.utc_offset, 2021, 180)
    jobs = [Job(synth.make_met(sun.altitude, seed=76 + k), sun, site.latitude) for k in range(4)]
    solver = HipSolver("cuda:0")
    dev = solver.device
    T = jobs[0].met.shape[0]
    lanes = [(DeviceScene(tile, dev, upload_inputs(tile, dev)), torch.cuda.Stream(dev)) for _ in range(2)]
    hb = [HostBuffers(torch.empty((16, 64, 80)).pin_memory(), torch.empty((T, 3, 64, 80)).pin_memory())
          for _ in jobs]
    hb[1] = HostBuffers(hb[1].fields, torch.empty((T, 3, 0, 80)))    # no room for job 1's day outputs
    dev_outs = [torch.empty((T, 3, 64, 80), device=dev) for _ in range(2)]
    fields_ring = [torch.empty_like(lanes[0][0].fields) for _ in range(2)]
    with pytest.raises(RuntimeError, match="download failed"):
        solve_many_host_threaded(solver, lanes, jobs, _InlinePlanner(), pinned_inputs(tile), hb, dev_outs,
                                 fields_ring, chunk_steps=8)
    torch.cuda.synchronize()
