"""Full SOLWEIG solve on one MI355X through libsolweig_hip.

Host work per tile: scene compose and validation, the 153 sky-patch march
tables, the per-step scalar chain (``met.DayLoop``) for the whole day, and
the walk tables. All of it is O(patches x steps) and is uploaded once; the
day is then a sequence of two kernels per step (sun march + walk sources,
flux -> Tmrt -> UTCI) with no host synchronisation between steps.

Device memory per tile (n pixels): heights 16 B/px, masks 80 B/px, SVF
fields 64 B/px, static planes ~30 B/px, walk sources 8 B/px, thermal state
24 B/px, outputs 4 B/px per requested plane per step.

The reference entry points this replaces: ``svf_calculator``
(``solweig_gpu/shadow.py:409-625``) and ``run_utci_window``
(``solweig_gpu/utci_process.py:477-1090``).
"""
from __future__ import annotations

import ctypes
from contextlib import nullcontext as _nullcontext
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L
from . import met as M
from . import skyvault as SV

F32 = np.float32

#: output plane bits of sw_flux_step, in packing order
OUT_BITS = {"tmrt": 1, "utci": 2, "shadow": 4, "kdown": 8, "kup": 16, "ldown": 32, "lup": 64, "wbgt": 128}
OUT_ORDER = ("tmrt", "utci", "shadow", "kdown", "kup", "ldown", "lup", "wbgt")
#: planes of HipSolver.sun_planes, in the order shadowingfunction_wallheight_23 returns them
SUN_PLANES = ("vegsh", "sh", "vbsh", "wallsh", "wallsun", "wallshve", "facesh", "facesun")


class UnsupportedInput(ValueError):
    """The input lies outside the domain the HIP backend reproduces exactly."""


@dataclass
class TileInputs:
    """Host-resident inputs of one tile (the rasters the reference loads).

    dsm: building DSM incl. ground; dem: ground; cdsm: tree canopy height
    above ground ("temp1"); walls: wall heights; aspect_deg: wall aspect in
    degrees; landcover: SOLWEIG land-cover raster or None. scale is pixels
    per metre (1 / pixel size).
    """

    dsm: np.ndarray
    dem: np.ndarray
    cdsm: np.ndarray
    walls: np.ndarray
    aspect_deg: np.ndarray
    landcover: np.ndarray | None
    scale: float
    lc_table: np.ndarray | None = None   # [class, albedo, emissivity, TgK, Tstart, TmaxLST] rows

    @property
    def shape(self):
        return self.dsm.shape

    @property
    def has_landcover(self) -> bool:
        return self.landcover is not None

    def spec(self) -> "TileSpec":
        r, c = self.shape
        return TileSpec(r, c, self.scale, self.has_landcover, self.lc_table)

    @classmethod
    def from_site(cls, site):
        return cls(dsm=site.dsm, dem=site.dem, cdsm=site.cdsm, walls=site.walls, aspect_deg=site.aspect,
                   landcover=site.landcover, scale=site.scale)


@dataclass
class TileSpec:
    """Shape and site constants of a tile whose rasters are already in HBM."""

    rows: int
    cols: int
    scale: float
    has_landcover: bool
    lc_table: np.ndarray | None = None

    @property
    def shape(self):
        return (self.rows, self.cols)


def compose_scene(dsm, dem, cdsm, walls, aspect_deg, landcover=None, *, scale: float) -> TileInputs:
    """Validate shapes and wrap the host rasters (compose itself runs on the device)."""
    arrs = [np.ascontiguousarray(x, F32) for x in (dsm, dem, cdsm, walls, aspect_deg)]
    if any(x.shape != arrs[0].shape or x.ndim != 2 for x in arrs):
        raise ValueError("input rasters must be 2-D and of one shape")
    lc = None if landcover is None else np.ascontiguousarray(landcover, F32)
    if lc is not None and lc.shape != arrs[0].shape:
        raise ValueError("land cover shape differs from the DSM")
    return TileInputs(*arrs, landcover=lc, scale=float(scale))


# ---------------------------------------------------------------------------
# sunlit / shaded patch thresholds
# ---------------------------------------------------------------------------
_RAD2DEG = np.float32(180.0 / math.pi)


def _g(x: float) -> np.float32:
    return np.float32(np.float32(math.atan(x)) * _RAD2DEG)


def _key_to_f32(k: int) -> float:
    u = (k - 0x80000000) if k >= 0x80000000 else (0xFFFFFFFF - k)
    return float(np.array([u], dtype=np.uint32).view(np.float32)[0])


def atan_thresholds(alt: float):
    """(T_lt, T_gt) with  atan(x)*180/pi < alt <=> x < T_lt  and  > alt <=> x > T_gt.

    ``shaded_or_sunlit`` (``solweig.py:541-561``) compares
    ``float32(atan(x)) * float32(180/pi)`` with the patch altitude; the map
    x -> that product is monotone in x, so each comparison is a threshold on
    x. Found by bisection over the ordered float32 values.
    """
    alt = np.float32(alt)
    neg_inf = 0xFFFFFFFF - 0xFF800000      # key of -inf
    pos_inf = 0x7F800000 + 0x80000000      # key of +inf

    def first_key(pred):
        lo, hi = neg_inf, pos_inf + 1
        while lo < hi:
            mid = (lo + hi) // 2
            if pred(_g(_key_to_f32(mid))):
                hi = mid
            else:
                lo = mid + 1
        return lo

    k_ge = first_key(lambda v: v >= alt)    # smallest x with g(x) >= alt
    t_lt = _key_to_f32(k_ge) if k_ge <= pos_inf else float("inf")
    k_gt = first_key(lambda v: v > alt)     # smallest x with g(x) > alt
    t_gt = _key_to_f32(k_gt - 1) if k_gt - 1 >= neg_inf else float("-inf")
    return np.float32(t_lt), np.float32(t_gt)


_THR_CACHE: dict = {}


def patch_thresholds(alt: np.ndarray):
    key = alt.tobytes()
    if key not in _THR_CACHE:
        lt = np.empty(alt.shape, F32)
        gt = np.empty(alt.shape, F32)
        memo = {}
        for i, a in enumerate(alt):
            if float(a) not in memo:
                memo[float(a)] = atan_thresholds(float(a))
            lt[i], gt[i] = memo[float(a)]
        _THR_CACHE[key] = (lt, gt)
    return _THR_CACHE[key]


# ---------------------------------------------------------------------------
# per-step / model structs
# ---------------------------------------------------------------------------
def _fill(arr, vals):
    arr[:] = [float(v) for v in vals]


STEP_DTYPE = np.dtype(L.StepStruct)
_SCALARS = ("ta", "ta_mat", "rh", "va", "tak4", "radI", "radD", "radG", "sinalt", "cosalt_k",
            "tan_zen", "one_m_psi", "tw_w", "tw_omw", "sunlit_k", "shaded_k", "sunlit_l",
            "shaded_l", "veg_l", "tgwall", "lwall", "wbgt_w", "wbgt_a")


def step_records(params, sun_offs) -> np.ndarray:
    """``sw_step_t`` records (numpy structured array, one per step) of a day."""
    rec = np.zeros(len(params), STEP_DTYPE)
    geo_memo = {}
    for i, (sp, off) in enumerate(zip(params, sun_offs)):
        r = rec[i]
        r["is_day"] = int(sp.is_day)
        r["sun_off"] = off
        r["sun_count"] = sp.table.count if (sp.is_day and sp.table is not None) else 0
        r["facesh_branch"] = int(sp.facesh_branch)
        r["azilow"], r["azihigh"] = sp.azilow, sp.azihigh
        for k in _SCALARS:
            r[k] = getattr(sp, k)
        r["tw_first"] = int(sp.tw_first)
        r["tw_update"] = int(sp.tw_update)
        r["tg_class"] = sp.tg_class
        g = sp.geom
        if id(g) not in geo_memo:
            lt, gt = patch_thresholds(np.asarray(g.alt, F32))
            geo_memo[id(g)] = (lt, gt)
        lt, gt = geo_memo[id(g)]
        ster, calt, salt = (np.asarray(x, F32) for x in (g.ster, g.cos_alt, g.sin_alt))
        pk, pl = r["pk"], r["pl"]
        pk["lv"], pk["lumchi"], pk["angl"], pk["ster"] = sp.lv, sp.lumchi, g.angl_inc, ster
        # step x patch products in float32, in the flux kernel's operation order
        pk["k_sh"] = (F32(r["shaded_k"]) * ster) * calt
        pk["k_sun"] = (F32(r["sunlit_k"]) * ster) * calt
        pk["yi_k"], pk["sun_lt"], pk["sun_gt"] = sp.yi_k, lt, gt
        pk["lsky_down"], pk["lsky_side"] = sp.lsky_down, sp.lsky_side
        vbase = F32(r["veg_l"]) * ster
        pk["lv_side"], pk["lv_down"] = vbase * calt, vbase * salt
        pk["cos_alt"], pk["sin_alt"], pk["l_sun"] = calt, salt, sp.l_sun_side
        sbase, hbase = F32(r["sunlit_l"]) * ster, F32(r["shaded_l"]) * ster
        pl["yi_l"] = sp.yi_l
        pl["ls_side"], pl["ls_down"] = sbase * calt, sbase * salt
        pl["lh_side"], pl["lh_down"] = hbase * calt, hbase * salt
    return rec


def step_struct(sp: M.StepParams, sun_off: int) -> L.StepStruct:
    """One ``sw_step_t`` as a ctypes structure."""
    return L.StepStruct.from_buffer_copy(step_records([sp], [sun_off]).tobytes())


def model_struct(scale: float, lcc: M.LandCover) -> L.ModelStruct:
    wt = SV.walk_tables(scale)
    nsec = wt.dx.shape[1]
    if nsec > L.WALK_MAX:
        raise UnsupportedInput(f"ground-view walk of {nsec} steps exceeds {L.WALK_MAX} (pixel too small)")
    if (wt.branch < 0).any():
        raise UnsupportedInput("walk azimuth without a face-shading branch")
    m = L.ModelStruct()
    m.absK, m.absL, m.Fside, m.Fup, m.Fcyl = M.ABSK, M.ABSL, M.FSIDE, M.FUP, M.FCYL
    m.albedo_b, m.ewall, m.sbc = M.ALBEDO_B, M.EWALL, M.SBC_PY
    m.first, m.second, m.nsec = wt.first, wt.second, nsec
    m.walk_halo = int(max(np.abs(wt.dx).max(), np.abs(wt.dy).max()))
    if m.walk_halo > 48:
        raise UnsupportedInput(f"ground-view walk halo {m.walk_halo} px exceeds 48 (pixel too small)")
    _fill(m.alb_class, lcc.alb)
    _fill(m.emis_class, lcc.emis)
    dx = np.zeros((L.WALK_AZ, L.WALK_MAX), np.int32)
    dy = np.zeros((L.WALK_AZ, L.WALK_MAX), np.int32)
    dx[:, :nsec] = wt.dx
    dy[:, :nsec] = wt.dy
    m.walk_dx[:] = dx.ravel().tolist()
    m.walk_dy[:] = dy.ravel().tolist()
    m.walk_branch[:] = wt.branch.tolist()
    m.walk_dir[:] = wt.dir_mask.tolist()
    _fill(m.walk_lo, wt.azilow)
    _fill(m.walk_hi, wt.azihigh)
    return m


_MODEL_CACHE: dict = {}


def _model_struct_cached(scale: float, lcc: M.LandCover) -> L.ModelStruct:
    key = (float(scale), lcc.key())
    if key not in _MODEL_CACHE:
        _MODEL_CACHE[key] = model_struct(scale, lcc)
    return _MODEL_CACHE[key]


@dataclass
class DayPlan:
    """All host-side inputs of one day, ready to upload."""

    steps: np.ndarray          # uint8 bytes of sw_step_t[T]
    sun_tab: np.ndarray        # int32 (S, 4)
    model: np.ndarray          # uint8 bytes of sw_model_t
    n_steps: int
    walk_halo: int
    is_day: list
    params: list = field(default_factory=list)
    wet_bulb: object = None    # met.WetBulb of the whole forcing table (WBGT)


def plan_day(tile: TileInputs, met: np.ndarray, sun: M.SunTable, latitude: float,
             state: M.DayState | None = None, time_start: int = 0, time_stop: int | None = None) -> DayPlan:
    """Host scalar chain for the whole day; the sun tables stop only at the grid
    edge (the kernels apply the amax stop test on the device)."""
    lcc = M.landcover_constants(tile.has_landcover, tile.lc_table)
    rows, cols = tile.shape
    loop = M.DayLoop(met, sun, latitude=latitude, landcover=lcc, scale=tile.scale, amax=None,
                     rows=rows, cols=cols, state=state)
    stop = met.shape[0] if time_stop is None else time_stop
    tabs, params, offs = [], [], []
    off = 0
    for i in range(time_start, stop):
        sp = loop.step(i)
        params.append(sp)
        if sp.is_day:
            if sp.facesh_branch < 0:
                raise UnsupportedInput(f"step {i}: sun azimuth without a face-shading branch")
            t = sp.table
            tab = np.stack([t.dx, t.dy, t.dz.view(np.int32), t.dzprev.view(np.int32)], axis=1).astype(np.int32)
            tabs.append(tab)
            offs.append(off)
            off += tab.shape[0]
        else:
            offs.append(0)
    steps = step_records(params, offs).view(np.uint8)
    sun_tab = np.concatenate(tabs) if tabs else np.zeros((1, 4), np.int32)
    mstruct = _model_struct_cached(tile.scale, lcc)
    return DayPlan(steps=steps, sun_tab=np.ascontiguousarray(sun_tab), model=L.struct_bytes(mstruct),
                   n_steps=len(params), walk_halo=int(mstruct.walk_halo), is_day=[p.is_day for p in params],
                   params=params, wet_bulb=loop.wet_bulb)


def plan_for_shape(plan: DayPlan, rows: int, cols: int, scale: float) -> DayPlan:
    """``plan`` for a tile of another shape (same site, day, forcing and pixel size).

    Only the sun-march tables depend on the shape (the march stops at the
    grid edge); they are rebuilt and the step records' table offsets and
    counts rewritten. Equal to ``plan_day`` on a tile of that shape.
    """
    T = plan.n_steps
    steps = np.array(plan.steps, dtype=np.uint8, copy=True)
    words = steps.view(np.int32).reshape(T, -1)
    off_w = L.StepStruct.sun_off.offset // 4
    cnt_w = L.StepStruct.sun_count.offset // 4
    tabs, off = [], 0
    for i, sp in enumerate(plan.params):
        if not sp.is_day:
            continue
        t = SV.sun_table(sp.sun_azalt[0], sp.sun_azalt[1], scale, None, rows, cols)
        tab = np.stack([t.dx, t.dy, t.dz.view(np.int32), t.dzprev.view(np.int32)], axis=1).astype(np.int32)
        words[i, off_w] = off
        words[i, cnt_w] = tab.shape[0]
        tabs.append(tab)
        off += tab.shape[0]
    sun_tab = np.concatenate(tabs) if tabs else np.zeros((1, 4), np.int32)
    return DayPlan(steps=steps, sun_tab=np.ascontiguousarray(sun_tab), model=plan.model, n_steps=T,
                   walk_halo=plan.walk_halo, is_day=list(plan.is_day), params=plan.params)


# ---------------------------------------------------------------------------
# device solver
# ---------------------------------------------------------------------------
def _stream_ptr(stream) -> int:
    return stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream


FLAG_TEXT = {1: "non-finite surface heights", 4: "bush layer present (DEM = -0.25 * trees somewhere)",
             8: "non-integer land-cover codes"}


class DeviceScene:
    """Device-resident tile: inputs, composed planes, SVF masks and fields."""

    def __init__(self, tile: TileInputs, device, inputs=None):
        self.tile = tile
        self.device = device
        rows, cols = tile.shape
        self.rows, self.cols, self.n = rows, cols, rows * cols
        dev = device
        if inputs is None:
            inputs = upload_inputs(tile, device)
        self.inputs = inputs
        n = self.n
        self.hpack = torch.empty((n, 4), dtype=torch.float32, device=dev)
        self.vegdsm2 = torch.empty(n, dtype=torch.float32, device=dev)
        self.buildings = torch.empty(n, dtype=torch.float32, device=dev)
        self.cls = torch.empty(n, dtype=torch.uint8, device=dev)
        self.aspect = torch.empty(n, dtype=torch.float32, device=dev)
        self.stats = torch.empty(4, dtype=torch.int32, device=dev)
        self.counts = torch.empty(L.NPATCH, dtype=torch.int32, device=dev)
        self.hpad = torch.empty(L.svf_workspace_floats(rows, cols), dtype=torch.float32,
                                device=dev)
        self.masks = torch.empty((L.MASK_PLANES * L.MASK_WORDS, n), dtype=torch.int32, device=dev)
        self.fields = torch.empty((L.NFIELDS, n), dtype=torch.float32, device=dev)
        self.svfbuveg = torch.empty(n, dtype=torch.float32, device=dev)
        self.tan_asvf = torch.empty(n, dtype=torch.float32, device=dev)
        self.tan_sa = torch.empty(n, dtype=torch.float32, device=dev)
        self.walkface = torch.empty(n, dtype=torch.int32, device=dev)
        self.src = torch.empty((n, 2), dtype=torch.int32, device=dev)
        self.state = torch.zeros((6, n), dtype=torch.float32, device=dev)

    @property
    def walls(self):
        return self.inputs["walls"]

    def tile_struct(self) -> L.TileStruct:
        """``sw_tile_t`` of this scene's device buffers."""
        P = L.ptr
        return L.TileStruct(self.rows, self.cols, P(self.hpack), P(self.walls), P(self.aspect), P(self.buildings),
                            P(self.cls), P(self.stats), P(self.masks), P(self.svfbuveg), P(self.tan_asvf),
                            P(self.tan_sa), P(self.walkface), P(self.src), P(self.state))

    def check_flags(self):
        """Refuse the solve if compose flagged inputs outside the supported domain (syncs)."""
        flags = int(self.stats[2].item()) & 0xFFFFFFFF
        if flags:
            msgs = [t for b, t in FLAG_TEXT.items() if flags & b]
            raise UnsupportedInput("; ".join(msgs) or f"input flags {flags:#x}")

    def amax_hmax(self):
        k = self.stats[:2].cpu().numpy().view(np.uint32)
        vals = []
        for key in k:
            key = int(key)
            u = (key & 0x7FFFFFFF) if key & 0x80000000 else (~key & 0xFFFFFFFF)
            vals.append(float(np.array([u], np.uint32).view(np.float32)[0]))
        return tuple(vals)


def upload_inputs(tile: TileInputs, device):
    """Host -> HBM copy of the input rasters (outside the device-resident timed region)."""
    d = {k: torch.from_numpy(np.ascontiguousarray(getattr(tile, k), F32)).to(device)
         for k in ("dsm", "dem", "cdsm", "walls", "aspect_deg")}
    d["landcover"] = None if tile.landcover is None else torch.from_numpy(
        np.ascontiguousarray(tile.landcover, F32)).to(device)
    return d


class MarchTuner:
    """Autotuner of the SVF march launch (``sw_march_config_t``): the gather chunk length (steps per
    wave between exit tests), the wave's pixel shape in the 16 x 16 tile and the blocks per 32-patch
    mask word. On the first sweep of a tile shape it times every candidate on the scene itself
    (HIP events on the sweep's stream, best of ``reps``) and keeps the fastest for that
    (rows, cols, scale); every candidate writes the same masks. ``enabled`` False (or the
    environment variable SW_MARCH_TUNE=0) keeps the defaults (chunk 8, 8 x 8 squares, split 1);
    ``fixed`` = (chunk, shape, split) pins one candidate for every tile without timing."""

    CANDIDATES = tuple((c, sh, sp) for c in (8, 4, 12) for sh in (0, 1, 2) for sp in (1, 2))
    DEFAULT = (8, 0, 1)

    def __init__(self, enabled: bool = True, reps: int = 2, fixed: tuple | None = None):
        import os
        if fixed is not None and tuple(fixed) not in self.CANDIDATES:
            raise ValueError(f"march config {fixed} is not one of {self.CANDIDATES}")
        self.fixed = None if fixed is None else tuple(fixed)
        self.enabled = enabled and self.fixed is None and os.environ.get("SW_MARCH_TUNE", "1") != "0"
        self.reps = reps
        self.best = {}        # (rows, cols, scale) -> (chunk, shape, split)
        self.timings = {}     # (rows, cols, scale) -> {candidate: ms}

    @staticmethod
    def struct(c) -> L.MarchConfig:
        return L.MarchConfig(chunk=c[0], shape=c[1], split=c[2])

    def config(self, solver, ds, offs, tab, stream=None) -> L.MarchConfig:
        if self.fixed is not None:
            return self.struct(self.fixed)
        if not self.enabled:
            return self.struct(self.DEFAULT)
        key = (ds.rows, ds.cols, float(ds.tile.scale))
        if key not in self.best:
            self.best[key], self.timings[key] = self._tune(solver, ds, offs, tab, stream)
        return self.struct(self.best[key])

    def _tune(self, solver, ds, offs, tab, stream):
        P = L.ptr
        sp = _stream_ptr(stream)
        s = torch.cuda.current_stream(solver.device) if stream is None else stream
        args = (P(ds.hpack), P(offs), P(tab), P(ds.counts), P(ds.hpad), P(ds.masks), P(ds.stats), ds.rows, ds.cols)
        L.call("sw_svf_march_cfg", *args, ctypes.byref(self.struct(self.DEFAULT)), sp)        # warm
        times = {}
        for c in self.CANDIDATES:
            cs = self.struct(c)
            best = float("inf")
            for _ in range(self.reps):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(s)
                L.call("sw_svf_march_cfg", *args, ctypes.byref(cs), sp)
                e1.record(s)
                e1.synchronize()
                best = min(best, e0.elapsed_time(e1))
            times[c] = round(best, 4)
        return min(times, key=times.get), times


class KernelTimer:
    """Per-launch kernel times from HIP events recorded on the launch stream.

    Events come from a pool created up front (reserve), so timing adds no
    event creation between launches; a day loop records its events inside
    the library (sw_solve_day_timed).
    """

    def __init__(self, reserve: int = 0):
        self.records = []          # (name, event_before, event_after)
        self._pool = []
        self._next = 0
        self.reserve(reserve)

    def reserve(self, n: int):
        while len(self._pool) < n:
            e = ctypes.c_void_p()
            L.call("sw_event_create", ctypes.byref(e))
            self._pool.append(e.value)

    def _take(self, k: int):
        self.reserve(self._next + k)
        evs = self._pool[self._next:self._next + k]
        self._next += k
        return evs

    def span(self, name, stream=None):
        timer = self
        sp = _stream_ptr(stream)

        class _Span:
            def __enter__(self_):
                self_.e0, self_.e1 = timer._take(2)
                L.call("sw_event_record", self_.e0, sp)
                return self_

            def __exit__(self_, *exc):
                L.call("sw_event_record", self_.e1, sp)
                timer.records.append((name, self_.e0, self_.e1))
                return False
        return _Span()

    def day_events(self, is_day):
        """ctypes array of 3 events per step for sw_solve_day_timed, registered for the summary."""
        evs = self._take(3 * len(is_day))
        for i, d in enumerate(is_day):
            e0, e1, e2 = evs[3 * i:3 * i + 3]
            if d:
                self.records.append(("sun_step", e0, e1))
            self.records.append(("flux_day" if d else "flux_night", e1, e2))
        return (ctypes.c_void_p * len(evs))(*evs)

    def summary(self):
        """{name: (launches, total_ms)} (call after synchronising)."""
        out = {}
        ms = ctypes.c_float()
        for name, e0, e1 in self.records:
            L.call("sw_event_elapsed", e0, e1, ctypes.byref(ms))
            n, t = out.get(name, (0, 0.0))
            out[name] = (n + 1, t + float(ms.value))
        return out

    def close(self):
        for e in self._pool:
            L.call("sw_event_destroy", e)
        self._pool = []
        self._next = 0


class _NoTimer:
    def span(self, name, stream=None):
        import contextlib
        return contextlib.nullcontext()


class HipSolver:
    """Drives the HIP kernels for one tile at a time on one GPU."""

    def __init__(self, device: str | torch.device = "cuda", tune_march: bool = True,
                 march_config: tuple | None = None):
        if not torch.cuda.is_available():
            raise RuntimeError("HipSolver needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = torch.device(device)
        self.march_tuner = MarchTuner(enabled=tune_march, fixed=march_config)
        self.lib = L.load()
        self.vault = SV.build_skyvault()
        v = self.vault
        self._fold_consts = {k: torch.from_numpy(np.ascontiguousarray(x)).to(self.device) for k, x in (
            ("w_iso", v.w_iso), ("w_aniso", v.w_aniso), ("ring", v.ring), ("nann", v.n_annuli),
            ("dirm", v.dir_mask))}
        self._tables = {}
        self.timer = _NoTimer()
        self.plan_ms: list[float] = []      # host planning time of each solve
        self._staging = {}            # plan staging (pinned host + device), STAGING_DEPTH per stream
        self._staging_turn = {}

    # -- SVF ---------------------------------------------------------------
    def upload(self, tile: TileInputs) -> DeviceScene:
        return DeviceScene(tile, self.device)

    def svf_tables(self, rows: int, cols: int, scale: float):
        """Geometric march tables of the 153 patches (host, torch float32 scalar chain)."""
        key = (rows, cols, scale)
        if key not in self._tables:
            offs, dx, dy, dz = SV.svf_tables(self.vault, scale, None, rows, cols)
            # the padded-buffer march needs steps of at most one cell
            for p in range(L.NPATCH):
                a, b = int(offs[p]), int(offs[p + 1])
                if b > a and (max(abs(int(dx[a])), abs(int(dy[a]))) > 1 or (b - a > 1 and max(
                        np.abs(np.diff(dx[a:b])).max(), np.abs(np.diff(dy[a:b])).max()) > 1)):
                    raise UnsupportedInput(f"patch {p}: march step longer than one cell")
                # the fast march bounds a chunk of steps by its first and last step: monotone x and y,
                # non-decreasing dz (a table with negative dz, the zenith patch, takes the checked march)
                if b - a > 1 and dz[a] >= 0 and not ((np.all(np.diff(dx[a:b]) >= 0) or np.all(np.diff(dx[a:b]) <= 0)) and (
                        np.all(np.diff(dy[a:b]) >= 0) or np.all(np.diff(dy[a:b]) <= 0)) and np.all(
                        np.diff(dz[a:b]) >= 0)):
                    raise UnsupportedInput(f"patch {p}: march steps not monotone")
            pitch = cols + 2 * L.SVF_PAD
            tab = np.stack([dx, dy, dz.view(np.int32), (dx * pitch + dy) * 12], axis=1).astype(np.int32)
            self._tables[key] = (torch.from_numpy(np.ascontiguousarray(offs)).to(self.device),
                                 torch.from_numpy(np.ascontiguousarray(tab)).to(self.device))
        return self._tables[key]

    def compose(self, ds: DeviceScene, stream=None):
        sp = _stream_ptr(stream)
        P = L.ptr
        i = ds.inputs
        lc = 0 if i["landcover"] is None else P(i["landcover"])
        with self.timer.span("compose", stream):
            L.call("sw_compose", P(i["dsm"]), P(i["dem"]), P(i["cdsm"]), P(i["aspect_deg"]), lc, P(ds.hpack),
                   P(ds.vegdsm2), P(ds.buildings), P(ds.cls), P(ds.aspect), P(ds.stats), ds.rows, ds.cols, sp)

    def svf(self, ds: DeviceScene, stream=None, tables=None, compose=True):
        """Compose (unless ``compose`` is False) and the SVF sweep; returns the field planes.

        The march runs with the configuration the autotuner chose for this tile shape (tuned on the
        first sweep of the shape when ``self.march_tuner`` is enabled)."""
        sp = _stream_ptr(stream)
        P = L.ptr
        if compose:
            self.compose(ds, stream)
        offs, tab = tables or self.svf_tables(ds.rows, ds.cols, ds.tile.scale)
        cfg = self.march_tuner.config(self, ds, offs, tab, stream)
        with self.timer.span("svf_march", stream):
            L.call("sw_svf_march_cfg", P(ds.hpack), P(offs), P(tab), P(ds.counts), P(ds.hpad), P(ds.masks),
                   P(ds.stats), ds.rows, ds.cols, ctypes.byref(cfg), sp)
        c = self._fold_consts
        with self.timer.span("svf_fold", stream):
            L.call("sw_svf_fold", P(ds.masks), P(ds.vegdsm2), P(c["w_iso"]), P(c["w_aniso"]), P(c["ring"]),
                   P(c["nann"]), P(c["dirm"]), P(ds.fields), ds.rows, ds.cols, sp)
        return ds.fields

    def unpack_masks(self, ds: DeviceScene, stream=None):
        """(shmat, vegshmat, vbshvegshmat) float32 cubes (rows, cols, 153)."""
        cubes = [torch.empty((ds.rows, ds.cols, L.NPATCH), dtype=torch.float32, device=self.device) for _ in range(3)]
        L.call("sw_unpack_masks", L.ptr(ds.masks), *[L.ptr(c) for c in cubes], ds.rows, ds.cols, _stream_ptr(stream))
        return cubes

    # -- day ---------------------------------------------------------------
    STAGING_DEPTH = 4      # plan staging buffers per stream

    def upload_plan(self, plan: DayPlan, stream=None):
        """Copy the day's tables to HBM from persistent pinned staging, without a host sync.

        Each stream rotates over :attr:`STAGING_DEPTH` staging buffers (pinned host + device), so
        the host can enqueue that many days ahead. A staging buffer is rewritten only after the
        event of its previous copy has completed, and its device half only after every kernel
        queued before that copy (stream order), so a day never reads a plan rewritten under it
        unless it is more than STAGING_DEPTH days behind on its stream, which the event wait rules
        out."""
        parts = [np.ascontiguousarray(plan.steps).view(np.uint8),
                 np.ascontiguousarray(plan.sun_tab).view(np.uint8).reshape(-1),
                 np.ascontiguousarray(plan.model).view(np.uint8).reshape(-1)]
        sizes = [p.nbytes for p in parts]
        offs = np.concatenate([[0], np.cumsum([(z + 255) // 256 * 256 for z in sizes])]).astype(int)
        total = int(offs[-1])
        s = torch.cuda.current_stream(self.device) if stream is None else stream
        turn = self._staging_turn.get(s.cuda_stream, 0)
        self._staging_turn[s.cuda_stream] = turn + 1
        key = (s.cuda_stream, turn % self.STAGING_DEPTH)
        st = self._staging.get(key)
        if st is None or st["host"].numel() < total:
            if st is not None:
                st["event"].synchronize()
            cap = max(total, 1 << 20)
            st = dict(host=torch.empty(cap, dtype=torch.uint8, pin_memory=True),
                      dev=torch.empty(cap, dtype=torch.uint8, device=self.device), event=torch.cuda.Event())
            st["event"].record(s)
            self._staging[key] = st
        st["event"].synchronize()          # the previous copy out of the staging buffer has run
        host = st["host"].numpy()
        for p, o in zip(parts, offs):
            host[o:o + p.nbytes] = p
        # the upload runs on the compute units (a copy kernel reading the pinned staging buffer), so it
        # never queues behind output downloads on the DMA engines
        L.call("sw_copy_to_device_cu", st["dev"].data_ptr(), st["host"].data_ptr(), total, s.cuda_stream)
        st["event"].record(s)
        d = st["dev"]
        return dict(steps=d[offs[0]:offs[0] + sizes[0]], sun_tab=d[offs[1]:offs[1] + sizes[1]],
                    model=d[offs[2]:offs[2] + sizes[2]])

    def tile_static(self, ds: DeviceScene, dplan, stream=None):
        P = L.ptr
        L.call("sw_tile_static", P(ds.fields), P(ds.walls), P(ds.aspect), P(dplan["model"]), P(ds.svfbuveg),
               P(ds.tan_asvf), P(ds.tan_sa), P(ds.walkface), ds.rows, ds.cols, _stream_ptr(stream))

    def sun_planes(self, ds: DeviceScene, plan: DayPlan, dplan, step: int, stream=None):
        """The day loop's sun march and wall tail of daytime ``step`` with the reference's eight
        output planes (``shadowingfunction_wallheight_23``'s return tuple, ``solweig.py:1187-1224``):
        a (8, rows, cols) device tensor in the order :data:`SUN_PLANES`. Also rewrites the scene's
        walk sources as the day loop's sun step does."""
        if not plan.is_day[step]:
            raise ValueError(f"step {step} is a night step: no sun march")
        planes = torch.empty((len(SUN_PLANES), ds.rows, ds.cols), dtype=torch.float32, device=self.device)
        P = L.ptr
        L.call("sw_sun_step_planes", P(ds.hpack), P(ds.walls), P(ds.aspect), P(ds.cls), P(ds.buildings),
               P(dplan["steps"]), P(dplan["model"]), P(dplan["sun_tab"]), P(ds.stats), P(ds.src), P(planes),
               int(step), ds.rows, ds.cols, _stream_ptr(stream))
        return planes

    def run_day(self, ds: DeviceScene, plan: DayPlan, dplan, outputs=("tmrt", "utci", "shadow"),
                out=None, stream=None, reset_state=True, per_launch=False, steps=None):
        """Enqueue the time loop; returns the device output tensor (T, k, rows, cols).

        ``steps`` = (t0, t1) enqueues only those steps (their planes of ``out``); the state carries
        over from the previous range, and is cleared only for a range starting at 0."""
        mask = 0
        for name in outputs:
            mask |= OUT_BITS[name]
        k = bin(mask).count("1")
        T = plan.n_steps
        if out is None:
            out = torch.empty((T, k, ds.rows, ds.cols), dtype=torch.float32, device=self.device)
        sp = _stream_ptr(stream)
        P = L.ptr
        if steps is not None:
            t0, t1 = (int(v) for v in steps)
            if not 0 <= t0 < t1 <= T or per_launch:
                raise ValueError(f"step range {steps} of a {T}-step day")
            # the day flags of the range: sw_solve_day reads is_day[i - t0]; the state is cleared by the
            # library's fill kernel on the day's own stream (a torch fill would run on torch's current
            # stream and could land inside another stream's day)
            is_day = (ctypes.c_int32 * (t1 - t0))(*[int(x) for x in plan.is_day[t0:t1]])
            base = P(out) + t0 * k * ds.n * 4
            L.call("sw_solve_day", ctypes.byref(ds.tile_struct()), P(dplan["steps"]), P(dplan["model"]),
                   P(dplan["sun_tab"]), is_day, t0, t1, base, mask, plan.walk_halo, int(reset_state and t0 == 0), sp)
            return out
        is_day = (ctypes.c_int32 * T)(*[int(x) for x in plan.is_day])
        if not per_launch:
            # one library call enqueues the whole day (state cleared on the day's stream)
            if isinstance(self.timer, KernelTimer):
                evs = self.timer.day_events(plan.is_day)
                L.call("sw_solve_day_timed", ctypes.byref(ds.tile_struct()), P(dplan["steps"]), P(dplan["model"]),
                       P(dplan["sun_tab"]), is_day, evs, 0, T, P(out), mask, plan.walk_halo, int(bool(reset_state)), sp)
            else:
                L.call("sw_solve_day", ctypes.byref(ds.tile_struct()), P(dplan["steps"]), P(dplan["model"]),
                       P(dplan["sun_tab"]), is_day, 0, T, P(out), mask, plan.walk_halo, int(bool(reset_state)), sp)
            return out
        if reset_state:
            L.call("sw_fill_f32", P(ds.state), 0.0, 6 * ds.n, sp)
        # per-launch path (kernel timing spans around each launch)
        stride = k * ds.n * 4
        base = P(out)
        for i in range(T):
            if plan.is_day[i]:
                with self.timer.span("sun_step", stream):
                    L.call("sw_sun_step", P(ds.hpack), P(ds.walls), P(ds.aspect), P(ds.cls), P(ds.buildings),
                           P(dplan["steps"]), P(dplan["model"]), P(dplan["sun_tab"]), P(ds.stats), P(ds.src), i,
                           ds.rows, ds.cols, sp)
            with self.timer.span("flux_day" if plan.is_day[i] else "flux_night", stream):
                L.call("sw_flux_step", P(ds.masks), P(ds.src), P(ds.svfbuveg), P(ds.tan_asvf), P(ds.tan_sa),
                       P(ds.walkface), P(ds.buildings), P(ds.cls), P(dplan["steps"]), P(dplan["model"]),
                       P(ds.state), base + i * stride, mask, i, plan.walk_halo if plan.is_day[i] else -1,
                       ds.rows, ds.cols, sp)
        return out

    def output_names(self, outputs):
        mask = 0
        for name in outputs:
            mask |= OUT_BITS[name]
        return [n for n in OUT_ORDER if mask & OUT_BITS[n]]

    def solve_device(self, ds: DeviceScene, met, sun: M.SunTable, latitude: float,
                     outputs=("tmrt", "utci", "shadow"), out=None, stream=None):
        """Enqueue a full solve of a device-resident tile; returns (fields, out) device tensors.

        The SVF sweep is enqueued first; the day's host scalar chain runs
        while the GPU marches.
        """
        self.svf(ds, stream)
        t0 = time.perf_counter()
        with torch.inference_mode():
            plan = plan_day(ds.tile, met, sun, latitude)
        self.plan_ms.append((time.perf_counter() - t0) * 1e3)
        dplan = self.upload_plan(plan)
        self.tile_static(ds, dplan, stream)
        out = self.run_day(ds, plan, dplan, outputs, out=out, stream=stream)
        return ds.fields, out

    def solve_batch(self, scenes, met, sun: M.SunTable, latitude: float, outputs=("tmrt", "utci", "shadow"),
                    outs=None, stream=None) -> "DayBatch":
        """Enqueue full solves of several device-resident scenes sharing one forcing.

        Per-item SVF sweeps, then one batched day loop (``sw_solve_batch``);
        returns the :class:`DayBatch` whose ``outs`` hold each item's
        (T, k, rows, cols) outputs.
        """
        batch = DayBatch(self, scenes, met, sun, latitude, outputs, outs)
        batch.prepare(stream)
        batch.run(stream)
        return batch

    def solve(self, tile: TileInputs, met, sun: M.SunTable, latitude: float,
              outputs=("tmrt", "utci", "shadow")):
        """Full solve of one tile; returns (fields (16, rows, cols), {name: (T, rows, cols)}, DeviceScene)."""
        ds = self.upload(tile)
        fields, out = self.solve_device(ds, met, sun, latitude, outputs)
        torch.cuda.synchronize()
        ds.check_flags()
        names = self.output_names(outputs)
        host = out.cpu().numpy()
        res = {nm: host[:, j] for j, nm in enumerate(names)}
        return fields.reshape(L.NFIELDS, ds.rows, ds.cols).cpu().numpy(), res, ds


# ---------------------------------------------------------------------------
# batched day loop (many tiles / edit scenarios, one launch per kernel and step)
# ---------------------------------------------------------------------------
def _owned_plan(plan: DayPlan, device):
    """A day plan copied to device buffers of its own (pinned staging, non-blocking)."""
    out = {}
    for key, arr in (("steps", plan.steps), ("sun_tab", plan.sun_tab), ("model", plan.model)):
        raw = np.ascontiguousarray(arr).view(np.uint8).reshape(-1)
        host = torch.from_numpy(raw.copy()).pin_memory()
        out[key] = host.to(device, non_blocking=True)
    return out


class DayBatch:
    """Items of one ``sw_solve_batch``: scenes, their per-shape plans and outputs.

    Items share the model constants and day/night pattern (same forcing,
    sun, latitude, pixel size and land-cover table); each distinct tile
    shape gets its own plan because the sun march ends at the tile edge.
    ``items_dev`` is the device copy of the laid-out item table.
    """

    def __init__(self, solver: "HipSolver", scenes, met, sun: M.SunTable, latitude: float,
                 outputs=("tmrt", "utci", "shadow"), outs=None):
        if not scenes:
            raise ValueError("empty batch")
        t0 = scenes[0].tile
        for ds in scenes[1:]:
            t = ds.tile
            if t.scale != t0.scale or t.has_landcover != t0.has_landcover or not np.array_equal(
                    np.asarray(t.lc_table), np.asarray(t0.lc_table)):
                raise UnsupportedInput("a batch needs one pixel size and one land-cover table")
        self.solver = solver
        self.scenes = list(scenes)
        self.mask = 0
        for name in outputs:
            self.mask |= OUT_BITS[name]
        self.k = bin(self.mask).count("1")
        self.plans = {}
        with torch.inference_mode():
            base = plan_day(t0, met, sun, latitude)       # the scalar chain once
            self.plans[t0.shape] = base
            for ds in self.scenes:
                if ds.tile.shape not in self.plans:
                    self.plans[ds.tile.shape] = plan_for_shape(base, ds.rows, ds.cols, t0.scale)
        first = next(iter(self.plans.values()))
        for p in self.plans.values():
            if list(p.is_day) != list(first.is_day) or p.walk_halo != first.walk_halo:
                raise UnsupportedInput("batch items disagree on the day/night pattern")
        self.n_steps = first.n_steps
        self.is_day = list(first.is_day)
        self.walk_halo = first.walk_halo
        dev = solver.device
        self.dplans = {shape: _owned_plan(p, dev) for shape, p in self.plans.items()}
        self.model = self.dplans[t0.shape]["model"]
        T = self.n_steps
        if outs is None:
            outs = [torch.empty((T, self.k, ds.rows, ds.cols), dtype=torch.float32, device=dev) for ds in scenes]
        self.outs = outs
        items = (L.BatchItemStruct * len(scenes))()
        for k, ds in enumerate(self.scenes):
            dp = self.dplans[ds.tile.shape]
            items[k].tile = ds.tile_struct()
            items[k].steps = L.ptr(dp["steps"])
            items[k].sun_tab = L.ptr(dp["sun_tab"])
            items[k].out = L.ptr(outs[k])
        sb, fb = ctypes.c_int32(), ctypes.c_int32()
        L.call("sw_batch_layout", items, len(scenes), ctypes.byref(sb), ctypes.byref(fb))
        self.sun_blocks, self.flux_blocks = sb.value, fb.value
        self.items_host = items
        raw = np.frombuffer(bytes(items), dtype=np.uint8)
        self.items_dev = torch.from_numpy(raw.copy()).pin_memory().to(dev, non_blocking=True)
        self._is_day_c = (ctypes.c_int32 * T)(*[int(x) for x in self.is_day])

    def prepare(self, stream=None):
        """SVF sweep and per-day static planes of every item (enqueued)."""
        for ds in self.scenes:
            self.solver.svf(ds, stream)
            self.solver.tile_static(ds, dict(model=self.model), stream)

    def run(self, stream=None, reset_state=True, persistent=False, timeout_s=10.0):
        """Enqueue the batched day; outputs land in ``self.outs``.

        ``persistent``: the whole day in one persistent launch
        (``sw_solve_batch_persistent``) instead of one launch per phase and
        step; the outputs are the same bit for bit. Its barrier workspace is
        ``self.sync``; :meth:`check` reads its error flag after the stream."""
        if persistent:
            if getattr(self, "sync", None) is None:
                self.sync = torch.zeros(L.DAY_SYNC_WORDS, dtype=torch.int32, device=self.solver.device)
            L.call("sw_solve_batch_persistent", L.ptr(self.items_dev), self.items_host, len(self.scenes),
                   L.ptr(self.model), self._is_day_c, 0, self.n_steps, self.mask, self.walk_halo,
                   1 if reset_state else 0, L.ptr(self.sync), float(timeout_s), _stream_ptr(stream))
        else:
            L.call("sw_solve_batch", L.ptr(self.items_dev), self.items_host, len(self.scenes), L.ptr(self.model),
                   self._is_day_c, 0, self.n_steps, self.mask, self.walk_halo, 1 if reset_state else 0,
                   _stream_ptr(stream))
        return self.outs

    def check(self):
        """Raise if the last persistent launch stopped at a barrier time-out (call after its stream)."""
        sync = getattr(self, "sync", None)
        if sync is not None and int(sync[L.DAY_SYNC_ERR].item()) != 0:
            raise RuntimeError("sw_solve_batch_persistent: a grid barrier timed out; the day is incomplete")
