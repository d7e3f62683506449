"""Pipelined solves: the host scalar chain of the next day runs in a planner
process while the GPU runs the current one.

The scalar chain (sun tables, clearness index, Perez luminance, sky tables,
per-step records) is Python on the host and costs about as much as the GPU
work of a 1000 x 1000 day. :class:`PlanPipeline` computes each job's
:class:`~solweig_project_amd.solver.DayPlan` in a child process (started
with ``spawn``, CPU only: it never touches the GPU) and
:meth:`solve_many` keeps one job planned ahead, so the GPU is fed without
waiting for the host. Every job is planned; the plans are byte-identical to
``plan_day`` in the calling process (same code, same inputs).
"""
from __future__ import annotations

import multiprocessing as mp
from dataclasses import dataclass

import numpy as np

from . import met as M


def _init_worker():
    # the chain is thousands of scalar-sized ops: one intra-op thread, no spinning pool
    import torch
    torch.set_num_threads(1)


def _plan_job(spec: tuple, met: np.ndarray, sun: tuple, latitude: float, time_start: int, time_stop):
    """Planner-process side: plan one day and return it without the per-step objects."""
    import time
    from .solver import DayPlan, TileSpec, plan_day
    t0 = time.perf_counter()
    rows, cols, scale, has_lc, lc_table = spec
    tile = TileSpec(rows, cols, scale, has_lc, lc_table)
    p = plan_day(tile, met, M.SunTable(*sun), latitude, time_start=time_start, time_stop=time_stop)
    out = DayPlan(steps=p.steps, sun_tab=p.sun_tab, model=np.array(p.model), n_steps=p.n_steps,
                  walk_halo=p.walk_halo, is_day=list(p.is_day), params=[])
    out.plan_ms = (time.perf_counter() - t0) * 1e3
    return out


@dataclass
class Job:
    """One day to solve on a device-resident scene."""

    met: np.ndarray
    sun: M.SunTable
    latitude: float
    time_start: int = 0
    time_stop: int | None = None


class PlanPipeline:
    """A pool of planner processes (CPU only)."""

    def __init__(self, workers: int = 1):
        ctx = mp.get_context("spawn")
        self.workers = workers
        self.pool = ctx.Pool(processes=workers, initializer=_init_worker)

    def warm(self, tile, job: Job):
        """Plan one job on every worker (first-call caches, imports), outside any timed region."""
        for f in [self.submit(tile, job) for _ in range(2 * self.workers)]:
            f.get()

    def submit(self, tile, job: Job):
        spec = (int(tile.shape[0]), int(tile.shape[1]), float(tile.scale), bool(tile.has_landcover),
                tile.lc_table)
        s = job.sun
        sun = (s.altitude, s.azimuth, s.zen, s.jday, s.dectime, s.altmax)
        return self.pool.apply_async(_plan_job, (spec, np.asarray(job.met, np.float64), sun, float(job.latitude),
                                                 int(job.time_start), job.time_stop))

    def close(self):
        self.pool.close()
        self.pool.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


def solve_many(solver, ds, jobs, pipeline: PlanPipeline, outputs=("tmrt", "utci", "shadow"), out=None,
               stream=None, lanes=None):
    """Solve every job on the scene ``ds`` (SVF sweep + day), planning as many jobs ahead as the
    pipeline has planner processes (a plan takes longer on the host than a solve on the device).

    ``out`` receives each job's outputs in turn (the last job's remain);
    returns the device output tensor. The host never waits for the GPU.

    ``lanes``: a list of (scene, stream, out) triples; job k runs on lane k mod len(lanes), so
    consecutive independent solves overlap on the device (each lane's buffers, plan staging and
    outputs are its own; ``ds``, ``stream`` and ``out`` are then unused). The SVF sweeps run one
    lane ahead of the days: job k + len(lanes) - 1's sweep is enqueued before job k waits for its
    plan, so the device has sweeps to run while the first plans are computed. Returns the last
    lane's out.
    """
    import time
    if not jobs:
        return out
    lane_of = (lambda k: lanes[k % len(lanes)]) if lanes else (lambda k: (ds, stream, out))
    depth = len(lanes) - 1 if lanes else 0
    ahead = max(1, int(getattr(pipeline, "workers", 1)))
    pending = [pipeline.submit(lane_of(k)[0].tile, j) for k, j in enumerate(jobs[:ahead])]
    sweep_ms = [0.0] * len(jobs)

    def sweep(k):
        t = time.perf_counter()
        sc, st, _ = lane_of(k)
        solver.svf(sc, st)
        sweep_ms[k] = time.perf_counter() - t

    for k in range(min(depth, len(jobs))):
        sweep(k)
    marks = []
    for k, job in enumerate(jobs):
        sc, st, o = lane_of(k)
        if not lanes:
            o = out                      # one lane: each job reuses the previous job's output buffer
        if k + depth < len(jobs):
            sweep(k + depth)
        t1 = time.perf_counter()
        if k + ahead < len(jobs):
            pending.append(pipeline.submit(sc.tile, jobs[k + ahead]))
        plan = pending.pop(0).get()
        t2 = time.perf_counter()
        if hasattr(solver, "plan_ms") and hasattr(plan, "plan_ms"):
            solver.plan_ms.append(plan.plan_ms)
        dplan = solver.upload_plan(plan, st)
        t3 = time.perf_counter()
        solver.tile_static(sc, dplan, st)
        out = solver.run_day(sc, plan, dplan, outputs, out=o, stream=st)
        t4 = time.perf_counter()
        marks.append((sweep_ms[k + depth] if k + depth < len(jobs) else 0.0, t2 - t1, t3 - t2, t4 - t3))
    solver.pipeline_marks = marks
    return out


@dataclass
class HostBuffers:
    """Pinned host buffers of one solve: input rasters in, SVF fields and day outputs out."""

    fields: "object"          # torch uint8/float32 pinned (16, rows, cols)
    out: "object"             # torch float32 pinned (T, k, rows, cols)


def pinned_inputs(tile):
    """The tile's input rasters as pinned host tensors (what a host-to-host solve uploads)."""
    import torch
    names = ["dsm", "dem", "cdsm", "walls", "aspect_deg"] + (["landcover"] if tile.landcover is not None else [])
    return {k: torch.from_numpy(np.ascontiguousarray(getattr(tile, k), np.float32)).pin_memory() for k in names}


def solve_many_host(solver, lanes, jobs, pipeline: PlanPipeline, host_inputs, host_outs, dev_outs,
                    outputs=("tmrt", "utci", "shadow"), chunk_steps: int = 8, fields_last: bool = True):
    """Host-to-host solves: for job k, upload the pinned input rasters to lane k mod len(lanes),
    run the SVF sweep and the day, and deliver the 16 SVF fields and every step's output planes
    into ``host_outs[k]`` (pinned :class:`HostBuffers`).

    ``lanes``: (scene, stream, copy stream) triples; a copy stream of None copies on the lane's own
    stream (in order, no events: the lane waits for its copies while the other lanes compute, and
    no stream shares a hardware queue when lanes + 1 <= GPU_MAX_HW_QUEUES). The day of job k writes
    into device buffer
    ``dev_outs[k mod len(dev_outs)]`` in chunks of ``chunk_steps`` steps. Each chunk's planes go
    out on the lane's copy stream as soon as the chunk is computed, through ``sw_copy_to_host``
    (the DMA engines: no compute units, so the downloads overlap the solves at full PCIe rate);
    the SVF fields go out right after the sweep. Ordering is
    stream-ordered with events, never a host wait: a lane's next sweep waits for the copy of its
    previous fields, and a day waits for the copies out of the device buffer it is about to reuse
    (``len(dev_outs)`` jobs back). Plans come from the planner processes, ``pipeline.workers`` jobs
    ahead. ``fields_last`` (own-stream copies only): the SVF fields go out after the day rather than
    right after the sweep, so the day starts without waiting for them on the copy engine (the
    lane's next sweep follows the copy in stream order). Returns each job's
    completion event (its last copy), in job order."""
    import torch

    from . import _lib as L
    ahead = max(1, int(getattr(pipeline, "workers", 1)))
    pending = [pipeline.submit(lanes[0][0].tile, j) for j in jobs[:ahead]]
    fields_out = [None] * len(lanes)        # event: the copy of a lane's SVF fields is done
    buf_free = [None] * len(dev_outs)       # event: the copies out of a device output buffer are done
    done = []

    def copy_out(dst, src, cstream):
        L.call("sw_copy_to_host", dst.data_ptr(), src.data_ptr(), src.numel() * src.element_size(), cstream.cuda_stream)

    import time

    def sweep(k):
        """Upload job k's inputs and sweep its SVF on its lane (and copy the fields out unless late)."""
        t = time.perf_counter()
        li = k % len(lanes)
        ds, stream, cstream = lanes[li]
        own = cstream is None
        cstream = stream if own else cstream
        with torch.cuda.stream(stream):
            if fields_out[li] is not None:
                stream.wait_event(fields_out[li])
            for name, tin in host_inputs.items():
                ds.inputs[name].copy_(tin, non_blocking=True)
            solver.svf(ds, stream)
            ev_fields = torch.cuda.Event()
            ev_fields.record(stream)
        if not own:
            cstream.wait_event(ev_fields)
        if not (fields_last and own):
            copy_out(host_outs[k].fields, ds.fields, cstream)
        if not own:
            fields_out[li] = torch.cuda.Event()
            fields_out[li].record(cstream)
        sweep_s[k] = time.perf_counter() - t

    # the sweeps run len(lanes) - 1 jobs ahead of the days: job k + depth's sweep is enqueued (on its
    # lane, after that lane's previous day) before job k waits for its plan
    depth = len(lanes) - 1
    sweep_s = [0.0] * len(jobs)
    for k in range(min(depth, len(jobs))):
        sweep(k)
    marks = []
    for k, job in enumerate(jobs):
        li, bi = k % len(lanes), k % len(dev_outs)
        ds, stream, cstream = lanes[li]
        own = cstream is None
        cstream = stream if own else cstream
        out = dev_outs[bi]
        hb = host_outs[k]
        if k + depth < len(jobs):
            sweep(k + depth)
        tm = [time.perf_counter()]
        if k + ahead < len(jobs):
            pending.append(pipeline.submit(ds.tile, jobs[k + ahead]))
        plan = pending.pop(0).get()
        tm.append(time.perf_counter())
        dplan = solver.upload_plan(plan, stream)
        tm.append(time.perf_counter())
        solver.tile_static(ds, dplan, stream)
        if buf_free[bi] is not None:
            stream.wait_event(buf_free[bi])
        T = plan.n_steps
        for t0 in range(0, T, chunk_steps):
            t1 = min(T, t0 + chunk_steps)
            solver.run_day(ds, plan, dplan, outputs, out=out, stream=stream, steps=(t0, t1))
            if not own:
                ev = torch.cuda.Event()
                ev.record(stream)
                cstream.wait_event(ev)
            copy_out(hb.out[t0:t1], out[t0:t1], cstream)
        if fields_last and own:
            copy_out(hb.fields, ds.fields, cstream)
        fin = torch.cuda.Event(enable_timing=True)
        fin.record(cstream)
        buf_free[bi] = fin
        done.append(fin)
        tm.append(time.perf_counter())
        marks.append([sweep_s[k + depth] if k + depth < len(jobs) else 0.0] + [b - a for a, b in zip(tm, tm[1:])])
    solver.pipeline_marks = marks
    return done


def solve_many_host_threaded(solver, lanes, jobs, pipeline: PlanPipeline, host_inputs, host_outs, dev_outs,
                             fields_ring, outputs=("tmrt", "utci", "shadow"), chunk_steps: int = 8,
                             stagger: bool = True, copy_stream=None):
    """Host-to-host solves with the downloads issued by a copy thread on a stream of their own.

    The lanes' streams never wait for a download: each day chunk writes into device buffer
    ``dev_outs[k mod len(dev_outs)]`` and each job's SVF fields are copied on the device into
    ``fields_ring[k mod len(fields_ring)]``; the copy thread polls the event that ends each chunk
    and issues the chunk's download on the copy stream, which therefore holds nothing but copies, as
    soon as the chunk is done (per lane in order, the lanes independently of each other). A lane
    reuses a device buffer only after the download out of it (``len(dev_outs)`` jobs back) has
    completed, which the host checks before it enqueues the day that reuses it. ``lanes``: (scene,
    stream) pairs. Returns each job's completion event (its last download), in job order.

    ``copy_stream``: the stream of the downloads (default: a new one). The runtime maps streams onto
    its hardware queues (GPU_MAX_HW_QUEUES, 4 by default) in creation order, round robin, and a copy
    stream that lands on a lane's queue serialises that lane behind the downloads (14-16 against
    10.5 ms per 1000 x 1000 x 24 h solve). Create the lanes' streams and the copy stream one after
    another (``torch.cuda.Stream`` hands out pool streams in order) and pass the copy stream here, so
    that the three streams sit on three different queues."""
    import collections
    import queue
    import threading
    import time

    import torch

    from . import _lib as L
    dev = lanes[0][1].device
    cs = copy_stream if copy_stream is not None else torch.cuda.Stream(dev)
    work = queue.Queue()
    failure = []                             # the copy thread's exception, re-raised by the caller

    def issue(item):
        _, ready, copies, fin, issued = item
        try:
            if not failure:
                for dst, src in copies:
                    L.call("sw_copy_to_host", dst.data_ptr(), src.data_ptr(), src.numel() * src.element_size(),
                           cs.cuda_stream)
                fin.record(cs)
        except BaseException as e:           # noqa: BLE001 - handed to the caller
            failure.append(e)
        finally:
            issued.set()                     # never leave the caller waiting, also after a failure

    def copier():
        # Downloads are issued per lane in order, each once its chunk's event has completed; the lanes
        # are polled (0.2 ms apart while some download waits), so a lane whose chunk is done is never
        # held behind another lane's chunk that is still computing.
        pend = [collections.deque() for _ in lanes]
        closing = False
        while True:
            try:
                item = work.get(timeout=2e-4) if any(pend) else work.get()
                while True:
                    if item is None:
                        closing = True
                    else:
                        pend[item[0]].append(item)
                    item = work.get_nowait()
            except queue.Empty:
                pass
            for q in pend:
                while q and (failure or q[0][1].query()):
                    issue(q.popleft())
            if closing and not any(pend):
                return
    def wait_issued(issued):
        issued.wait()
        if failure:
            raise RuntimeError("a host-to-host download failed") from failure[0]

    th = threading.Thread(target=copier, daemon=True)
    th.start()
    ahead = max(1, int(getattr(pipeline, "workers", 1)))
    pending = [pipeline.submit(lanes[0][0].tile, j) for j in jobs[:ahead]]
    R = len(dev_outs)
    last_copy = [None] * R                  # (completion event, issued flag) of the last download per ring slot
    done = []
    depth = len(lanes) - 1
    sweep_s = [0.0] * len(jobs)

    first_sweep = [None]

    def sweep(k):
        t = time.perf_counter()
        ds, stream = lanes[k % len(lanes)]
        with torch.cuda.stream(stream):
            if stagger and 0 < k < len(lanes) and first_sweep[0] is not None:
                # each lane's first sweep starts after the previous lane's: the lanes run out of phase
                # (one sweeps while another runs its day), not sweep against sweep
                stream.wait_event(first_sweep[0])
            for name, tin in host_inputs.items():
                # a copy kernel, not the DMA engines, which are busy with the downloads
                L.call("sw_copy_to_device_cu", ds.inputs[name].data_ptr(), tin.data_ptr(),
                       tin.numel() * tin.element_size(), stream.cuda_stream)
            solver.svf(ds, stream)
            if stagger and k < len(lanes) - 1:
                first_sweep[0] = torch.cuda.Event()
                first_sweep[0].record(stream)
        sweep_s[k] = time.perf_counter() - t

    try:
        for k in range(min(depth, len(jobs))):
            sweep(k)
        marks = []
        for k, job in enumerate(jobs):
            ds, stream = lanes[k % len(lanes)]
            b = k % R
            out, fr, hb = dev_outs[b], fields_ring[b], host_outs[k]
            if k + depth < len(jobs):
                sweep(k + depth)
            tm = [time.perf_counter()]
            if k + ahead < len(jobs):
                pending.append(pipeline.submit(ds.tile, jobs[k + ahead]))
            plan = pending.pop(0).get()
            tm.append(time.perf_counter())
            dplan = solver.upload_plan(plan, stream)
            tm.append(time.perf_counter())
            solver.tile_static(ds, dplan, stream)
            if last_copy[b] is not None:
                # the ring slot's previous downloads must be complete before this day overwrites it: a
                # host wait (normally already true, the slot was filled len(dev_outs) jobs back), not a
                # stream wait on the copy stream's event, which stalled the lanes for milliseconds
                fin, issued = last_copy[b]
                wait_issued(issued)
                fin.synchronize()
            T = plan.n_steps
            fin = None
            for t0 in range(0, T, chunk_steps):
                t1 = min(T, t0 + chunk_steps)
                solver.run_day(ds, plan, dplan, outputs, out=out, stream=stream, steps=(t0, t1))
                ready = torch.cuda.Event()
                ready.record(stream)
                fin, issued = torch.cuda.Event(enable_timing=True), threading.Event()
                work.put((k % len(lanes), ready, [(hb.out[t0:t1], out[t0:t1])], fin, issued))
            with torch.cuda.stream(stream):
                fr.copy_(ds.fields, non_blocking=True)
            ready = torch.cuda.Event()
            ready.record(stream)
            fin, issued = torch.cuda.Event(enable_timing=True), threading.Event()
            work.put((k % len(lanes), ready, [(hb.fields, fr)], fin, issued))
            last_copy[b] = (fin, issued)
            done.append((fin, issued))
            tm.append(time.perf_counter())
            marks.append([sweep_s[k + depth] if k + depth < len(jobs) else 0.0] + [b_ - a_ for a_, b_ in zip(tm, tm[1:])])
        solver.pipeline_marks = marks
    finally:
        work.put(None)
        th.join()
    for _, issued in done:
        wait_issued(issued)
    return [fin for fin, _ in done]
