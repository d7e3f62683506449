"""Device-resident request loop over one site (edits, met-only re-runs, full re-solves).

Requests arrive on a clock (Poisson arrivals in the benchmark) and are
served with continuous batching: a request is admitted as soon as an
execution slot is free and its HBM footprint fits, and the work of admitted
requests overlaps on the device (each slot has its own HIP stream).

* ``edit`` - a single-window change of DSM / CDSM / land cover, applied in
  arrival order to the master site (``incremental.apply_edit_window``); the
  site's outputs are the edited site's full-solve outputs.
* ``met`` - the day re-run over the site with another forcing table: the
  slot copies the site's planes at admission (device copies on its stream,
  so later edits do not change the request's result), gets the day plan
  from the planner processes and runs the day on its own state and outputs.
* ``full`` - a full re-solve of the site's current inputs (compose, SVF
  sweep, static planes, day) in a slot.

Edits go first: queued edits are applied before slot requests are admitted
or advanced, and slots enqueue their day a few steps at a time (one chunk in
flight per slot), so an edit shares the GPU with at most one chunk per slot.

Every request's result equals the same request served alone on the site
version it was admitted against; ``checksum`` fingerprints results for that
comparison.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import incremental as I
from .pipeline import Job, PlanPipeline
from .solver import DeviceScene, HipSolver, _owned_plan


@dataclass
class Request:
    kind: str                 # "edit" | "met" | "full"
    arrival: float            # seconds from the start of the run
    payload: dict = field(default_factory=dict)
    start: float = 0.0
    finish: float = 0.0
    result: object = None

    @property
    def latency_ms(self) -> float:
        return (self.finish - self.arrival) * 1e3


def poisson_workload(n: int, rate: float, mix: dict, seed: int, site_tile, met_pool: list) -> list:
    """``n`` requests with exponential inter-arrival times at ``rate`` per second; ``mix`` maps kind
    to probability. Edits add a tree crown of radius 2-6 px on open ground or clear the canopy within that
    radius around a canopy cell."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kinds = list(mix)
    p = np.array([mix[k] for k in kinds], np.float64)
    p /= p.sum()
    rows, cols = site_tile.shape
    ground = np.flatnonzero((site_tile.dsm == site_tile.dem).ravel())
    canopy = np.flatnonzero((site_tile.cdsm > 0).ravel())
    t = 0.0
    out = []
    for _ in range(n):
        t += float(rng.exponential(1.0 / rate))
        kind = kinds[int(rng.choice(len(kinds), p=p))]
        payload = {}
        if kind == "edit":
            remove = bool(rng.random() < 0.5) and canopy.size > 0
            cells = canopy if remove else ground          # removals clear an existing crown
            ci, cj = divmod(int(cells[rng.integers(0, cells.size)]), cols)
            r = int(rng.integers(2, 7))
            payload = dict(center=(ci, cj), radius=r, height=float(rng.uniform(3.0, 15.0)), remove=remove)
        elif kind == "met":
            payload = dict(met=int(rng.integers(0, len(met_pool))))
        out.append(Request(kind, t, payload))
    return out


def tree_window(tile, center, radius, height, remove):
    """(box, new CDSM window values) of a tree crown added or cleared at ``center``."""
    rows, cols = tile.shape
    ci, cj = center
    r0, r1, c0, c1 = max(0, ci - radius), min(rows, ci + radius + 1), max(0, cj - radius), min(cols, cj + radius + 1)
    cd = tile.cdsm[r0:r1, c0:c1].copy()
    ii, jj = np.mgrid[r0:r1, c0:c1]
    d2 = (ii - ci) ** 2 + (jj - cj) ** 2
    disc = d2 <= radius * radius
    if remove:
        cd[disc] = 0
    else:
        crown = (height * np.exp(-d2 / (2.0 * (radius / 2.0) ** 2))).astype(np.float32)
        crown[~disc | (tile.dsm[r0:r1, c0:c1] != tile.dem[r0:r1, c0:c1])] = 0
        cd = np.maximum(cd, crown)
    return (r0, r1, c0, c1), cd


def checksum(t: torch.Tensor) -> tuple:
    """Order-independent fingerprint of a device tensor (NaN counted, then excluded)."""
    x = t.float()
    nan = torch.isnan(x)
    y = torch.where(nan, torch.zeros_like(x), x).double()
    w = torch.arange(1, y.numel() + 1, device=y.device, dtype=torch.float64).view_as(y)
    return int(nan.sum().item()), float(y.sum().item()), float((y * torch.remainder(w, 9973.0)).sum().item())


class Slot:
    """An execution slot: a stream and a scene copy with its own state, walk sources and outputs."""

    def __init__(self, solver: HipSolver, st: I.SiteState):
        self.stream = torch.cuda.Stream(solver.device)
        with torch.cuda.stream(self.stream):
            self.scene = DeviceScene(st.tile, solver.device, {k: (None if v is None else v.clone())
                                                              for k, v in st.scene.inputs.items()})
            self.out = torch.empty_like(st.out)
        self.busy = None                  # the request in flight
        self.event = torch.cuda.Event()
        self.pending_plan = None
        self.plan = None
        self.dplan = None
        self.next_step = 0                # next day step to enqueue

    def footprint(self) -> int:
        s = self.scene
        return sum(t.numel() * t.element_size() for t in (s.hpack, s.vegdsm2, s.buildings, s.cls, s.aspect,
                                                          s.masks, s.fields, s.svfbuveg, s.tan_asvf, s.tan_sa,
                                                          s.walkface, s.src, s.state, s.hpad, self.out)) + sum(
            v.numel() * v.element_size() for v in s.inputs.values() if v is not None)


class Server:
    """One site, a FIFO of arrived requests, ``slots`` execution slots (admitted while HBM allows)."""

    def __init__(self, solver: HipSolver, st: I.SiteState, met_pool: list, sun, latitude: float,
                 pipeline: PlanPipeline, slots: int = 4, hbm_reserve: float = 0.1, chunk_steps: int = 4):
        self.solver, self.st, self.met_pool, self.sun, self.lat = solver, st, met_pool, sun, latitude
        self.pipeline = pipeline
        self.chunk_steps = max(1, int(chunk_steps))
        self.slots = []
        free, total = torch.cuda.mem_get_info(solver.device)
        for _ in range(slots):
            s = Slot(solver, st)
            if torch.cuda.mem_get_info(solver.device)[0] < hbm_reserve * total:
                del s                   # admission by HBM footprint: stop adding slots
                break
            self.slots.append(s)
        self.master = torch.cuda.Stream(solver.device)

    # -- request execution -----------------------------------------------------
    def _snapshot(self, slot: Slot):
        """Copy the site's planes into the slot: on the slot's stream after the master's queued edits,
        and the master's later edits wait for the copies."""
        ev = torch.cuda.Event()
        ev.record(self.master)
        slot.stream.wait_event(ev)
        src, dst = self.st.scene, slot.scene
        with torch.cuda.stream(slot.stream):
            for k, v in src.inputs.items():
                if v is not None:
                    dst.inputs[k].copy_(v, non_blocking=True)
            for name in ("hpack", "vegdsm2", "buildings", "cls", "aspect", "stats", "masks", "fields", "svfbuveg",
                         "tan_asvf", "tan_sa", "walkface"):
                getattr(dst, name).copy_(getattr(src, name), non_blocking=True)
        done = torch.cuda.Event()
        done.record(slot.stream)
        self.master.wait_event(done)
        dst.tile = src.tile

    def _start(self, slot: Slot, req: Request, now: float):
        req.start = now
        req.payload["version"] = self.version
        slot.busy = req
        if req.kind == "met":
            self._snapshot(slot)
            job = Job(self.met_pool[req.payload["met"]], self.sun, self.lat)
            slot.pending_plan = self.pipeline.submit(self.st.tile, job)
        else:                               # full re-solve of the current inputs
            self._snapshot(slot)
            with torch.cuda.stream(slot.stream):
                self.solver.svf(slot.scene, slot.stream)
            slot.pending_plan = self.pipeline.submit(self.st.tile, Job(self.st_met(), self.sun, self.lat))

    def st_met(self):
        return self.met_pool[0]

    def _advance(self, slot: Slot, now: float):
        """Move a slot's request on; returns True when it finished.

        The day is enqueued a few steps at a time (one chunk in flight per slot), so an edit that
        arrives shares the GPU with at most one chunk per slot."""
        req = slot.busy
        if slot.pending_plan is not None:
            if not slot.pending_plan.ready():
                return False
            plan = slot.pending_plan.get()
            slot.pending_plan = None
            slot.plan, slot.next_step = plan, 0
            with torch.cuda.stream(slot.stream):
                slot.dplan = _owned_plan(plan, self.solver.device)
                if req.kind == "full":
                    self.solver.tile_static(slot.scene, slot.dplan, slot.stream)
            self._feed(slot)
            return False
        if not slot.event.query():
            return False
        if slot.plan is not None and slot.next_step < slot.plan.n_steps:
            self._feed(slot)
            return False
        req.finish = time.perf_counter() - self.t0
        if self.checksums:
            req.result = checksum(slot.out)
        slot.busy = None
        slot.plan = None
        return True

    def _feed(self, slot: Slot):
        """Enqueue the slot's next chunk of day steps and mark its end."""
        t0 = slot.next_step
        t1 = min(slot.plan.n_steps, t0 + self.chunk_steps)
        with torch.cuda.stream(slot.stream):
            self.solver.run_day(slot.scene, slot.plan, slot.dplan, self.st.outputs, out=slot.out,
                                stream=slot.stream, steps=(t0, t1))
            slot.event.record(slot.stream)
        slot.next_step = t1

    def _edit(self, req: Request):
        p = req.payload
        box, cd = tree_window(self.st.tile, p["center"], p["radius"], p["height"], p["remove"])
        with torch.cuda.stream(self.master):
            info = I.apply_edit_window(self.solver, self.st, box, cdsm=cd, stream=self.master)
            self.version += 1
            req.payload["version"] = self.version
            req.payload["mode"] = info.get("mode")
            if self.checksums:
                req.result = checksum(self.st.out)
        return info

    # -- the loop --------------------------------------------------------------
    def run(self, requests: list, checksums: bool = True) -> list:
        """Serve ``requests`` (sorted by arrival) against the wall clock; returns them with timings."""
        self.t0 = time.perf_counter()
        self.checksums = checksums
        self.version = 0
        torch.cuda.synchronize()
        todo = deque(sorted(requests, key=lambda r: r.arrival))
        queue = deque()
        done = []
        while todo or queue or any(s.busy for s in self.slots):
            now = time.perf_counter() - self.t0
            while todo and todo[0].arrival <= now:
                queue.append(todo.popleft())
            progressed = False
            # edits first, in arrival order among themselves, on the master stream; requests waiting
            # for a slot do not hold them back (each is served on the site version it is admitted at)
            edits = [r for r in queue if r.kind == "edit"]
            for req in edits:
                queue.remove(req)
                req.start = time.perf_counter() - self.t0
                self._edit(req)
                req.finish = time.perf_counter() - self.t0
                done.append(req)
                progressed = True
            if edits:
                continue                    # new arrivals may be edits too
            for s in self.slots:
                req = s.busy
                if req is not None and self._advance(s, now):
                    done.append(req)
            # slot work is admitted as slots free up, in arrival order
            while queue:
                free = next((s for s in self.slots if s.busy is None), None)
                if free is None:
                    break
                self._start(free, queue.popleft(), now)
                progressed = True
            if not progressed and not any(s.busy for s in self.slots) and todo and not queue:
                time.sleep(max(0.0, min(0.002, todo[0].arrival - now)))
            elif not progressed:
                time.sleep(0.0002)
        torch.cuda.synchronize()
        return [r for r in requests]


def percentiles(requests, kinds=None):
    lat = np.array([r.latency_ms for r in requests if kinds is None or r.kind in kinds])
    if lat.size == 0:
        return None
    return {"n": int(lat.size), "p50_ms": round(float(np.percentile(lat, 50)), 2),
            "p99_ms": round(float(np.percentile(lat, 99)), 2), "max_ms": round(float(lat.max()), 2),
            "mean_ms": round(float(lat.mean()), 2)}
