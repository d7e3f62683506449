"""Incremental edits: recompute only the tiles an edit can reach, in place.

The reference recomputes edited regions by halo/corridor closure
(``solweig_gpu/incremental/solver.py:595-957``) and sparse work lists
(``sparse_work.py:633``). Here the closure is computed on the 16 x 16 tile
grid from the step tables themselves, and the kernels run on the full
tile's buffers over the listed tiles only (``sw_svf_sweep_tiles``,
``sw_solve_day_tiles``).

An edit changes heights (DSM, CDSM) inside a box E and possibly land
cover. With o ranging over the march offsets a kernel executes (every step
up to and including the first with dz > amax, at the tile's amax):

* walls change within E + 1 and the wall aspect within E' = E + (filter
  half width + 2) (filter window plus gradient); E' also holds the land-cover
  and building-mask changes;
* the SVF masks of p can change only if p + o hits E for an SVF patch
  offset o: A_svf = union over o of (E - o);
* the walk sources of p at day step i can change only if p is in E' or
  p + o hits E for a sun offset o of step i: S_i;
* the outputs of p can change (at any step, through the state) only in
  U = A_svf + E' + (union S_i) dilated by the walk reach; their flux steps
  read walk sources within the walk reach, so the sun march runs on
  V = U dilated by the walk reach.

Every other tile keeps its masks, fields and outputs, which a full solve of
the edited tile reproduces unchanged; the listed tiles are recomputed by the
same kernels with the same tables and amax, so the edited site equals a
full solve. An edit that changes the tile's amax changes the executed steps
everywhere and is solved in full.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L
from . import met as M
from . import skyvault as SV
from . import walls as WG
from .solver import OUT_BITS, DayPlan, DeviceScene, HipSolver, TileInputs, _owned_plan, _stream_ptr, plan_day

F32 = np.float32
TILE = 16


def _sync(stream):
    """Wait for the edit's own stream only (other streams keep running)."""
    (stream if stream is not None else torch.cuda.current_stream()).synchronize()


def tile_amax(dsm: np.ndarray, dem: np.ndarray, cdsm: np.ndarray) -> np.float32:
    """max(a, trees + DEM) as the compose step forms it (float32, trees clamped at 0)."""
    t1 = np.maximum(np.asarray(cdsm, F32), F32(0))
    vegdem = (t1 + np.asarray(dem, F32)).astype(F32)
    return F32(max(F32(np.max(dsm)), F32(np.max(vegdem))))


def executed(dz: np.ndarray, amax) -> int:
    """Steps a march executes: up to and including the first step with dz > amax (the loop
    tests the previous step's dz), or the whole table."""
    stop = np.nonzero(~(np.float32(amax) >= np.asarray(dz, np.float32)))[0]
    return int(stop[0]) + 1 if stop.size else len(dz)


@dataclass
class Offsets:
    """Executed march offsets (row, col) of a tile at its amax."""

    svf: np.ndarray                   # (K, 2) unique over the 153 patches
    sun: list                         # per day step: (K_i, 2); None at night
    walk: int                         # walk reach in pixels
    aspect: int                       # aspect margin in pixels


def march_offsets(solver: HipSolver, tile: TileInputs, plan: DayPlan, amax) -> Offsets:
    rows, cols = tile.shape
    offs, dx, dy, dz = SV.svf_tables(solver.vault, tile.scale, None, rows, cols)
    parts = [np.zeros((1, 2), np.int32)]
    for p in range(L.NPATCH):
        a, b = int(offs[p]), int(offs[p + 1])
        if b > a:
            k = executed(dz[a:b], amax)
            parts.append(np.stack([dx[a:a + k], dy[a:a + k]], axis=1))
    svf = np.unique(np.concatenate(parts).astype(np.int32), axis=0)
    tab = np.asarray(plan.sun_tab, np.int32)
    rec = np.asarray(plan.steps).view(np.int32).reshape(plan.n_steps, -1)
    o_w, c_w = L.StepStruct.sun_off.offset // 4, L.StepStruct.sun_count.offset // 4
    sun = []
    for i in range(plan.n_steps):
        if not plan.is_day[i]:
            sun.append(None)
            continue
        t = tab[int(rec[i, o_w]):int(rec[i, o_w]) + int(rec[i, c_w])]
        k = executed(t[:, 2].view(np.float32), amax) if len(t) else 0
        sun.append(np.unique(np.concatenate([np.zeros((1, 2), np.int32), t[:k, :2]]), axis=0))
    return Offsets(svf=svf, sun=sun, walk=int(plan.walk_halo), aspect=WG.filter_size(tile.scale) // 2 + 2)


def changed_boxes(a: TileInputs, dsm, cdsm, landcover):
    """Bounding boxes (r0, r1, c0, c1) of changed heights and of changed land cover (None if unchanged)."""
    def box(m):
        if not m.any():
            return None
        rr = np.nonzero(m.any(axis=1))[0]
        cc = np.nonzero(m.any(axis=0))[0]
        return int(rr[0]), int(rr[-1]) + 1, int(cc[0]), int(cc[-1]) + 1
    h = (np.asarray(dsm) != a.dsm) | (np.asarray(cdsm) != a.cdsm)
    lc = None
    if a.landcover is not None and landcover is not None:
        lc = box(np.asarray(landcover) != a.landcover)
    return box(h), lc


def grow(box, d, rows, cols):
    r0, r1, c0, c1 = box
    return max(0, r0 - d), min(rows, r1 + d), max(0, c0 - d), min(cols, c1 + d)


def union(a, b):
    if a is None:
        return b
    if b is None:
        return a
    return min(a[0], b[0]), max(a[1], b[1]), min(a[2], b[2]), max(a[3], b[3])


class TileGrid:
    """Boolean sets of the 16 x 16 tiles of a rows x cols tile."""

    def __init__(self, rows, cols):
        self.rows, self.cols = rows, cols
        self.ty, self.tx = -(-rows // TILE), -(-cols // TILE)

    def empty(self):
        return np.zeros((self.ty, self.tx), bool)

    def boxes(self, box, offsets):
        """Tiles touched by the pixel boxes ``box - o`` for every offset o (rows of ``offsets``)."""
        r0, r1, c0, c1 = box
        o = np.asarray(offsets, np.int64).reshape(-1, 2)
        pr0 = np.clip(r0 - o[:, 0], 0, self.rows)
        pr1 = np.clip(r1 - o[:, 0], 0, self.rows)
        pc0 = np.clip(c0 - o[:, 1], 0, self.cols)
        pc1 = np.clip(c1 - o[:, 1], 0, self.cols)
        keep = (pr1 > pr0) & (pc1 > pc0)
        if not keep.any():
            return self.empty()
        tr0, tr1 = pr0[keep] // TILE, -(-pr1[keep] // TILE)
        tc0, tc1 = pc0[keep] // TILE, -(-pc1[keep] // TILE)
        d = np.zeros((self.ty + 1, self.tx + 1), np.int32)
        np.add.at(d, (tr0, tc0), 1)
        np.add.at(d, (tr0, tc1), -1)
        np.add.at(d, (tr1, tc0), -1)
        np.add.at(d, (tr1, tc1), 1)
        return d.cumsum(0).cumsum(1)[:self.ty, :self.tx] > 0

    def dilate(self, m, pixels):
        """Tiles within ``pixels`` of a tile of ``m`` (ceil(pixels / 16) tile rings)."""
        k = -(-pixels // TILE)
        out = m.copy()
        if k == 0 or not m.any():
            return out
        ys, xs = np.nonzero(m)
        for dy in range(-k, k + 1):
            yy = np.clip(ys + dy, 0, self.ty - 1)
            for dx in range(-k, k + 1):
                out[yy, np.clip(xs + dx, 0, self.tx - 1)] = True
        return out

    @staticmethod
    def ids(m):
        return np.flatnonzero(m.ravel()).astype(np.int32)


@dataclass
class SiteState:
    """A device-resident solved tile that accepts edits (one forcing for every edit)."""

    tile: TileInputs
    scene: DeviceScene
    plan: DayPlan
    dplan: dict
    out: torch.Tensor                 # (T, k, rows, cols)
    amax: np.float32
    offsets: Offsets
    mask: int
    outputs: tuple
    log: list = field(default_factory=list)

    @property
    def fields(self):
        return self.scene.fields


def solve_site(solver: HipSolver, tile: TileInputs, met, sun: M.SunTable, latitude: float,
               outputs=("tmrt", "utci", "shadow")) -> SiteState:
    """Full solve of a tile kept on the device for later edits (the site keeps copies of the planes)."""
    tile = TileInputs(dsm=np.array(tile.dsm, F32), dem=np.array(tile.dem, F32), cdsm=np.array(tile.cdsm, F32),
                      walls=np.array(tile.walls, F32), aspect_deg=np.array(tile.aspect_deg, F32),
                      landcover=None if tile.landcover is None else np.array(tile.landcover, F32),
                      scale=tile.scale, lc_table=tile.lc_table)
    ds = solver.upload(tile)
    solver.svf(ds)
    with torch.inference_mode():
        plan = plan_day(tile, met, sun, latitude)
    dplan = _owned_plan(plan, solver.device)
    solver.tile_static(ds, dplan)
    out = solver.run_day(ds, plan, dplan, outputs)
    torch.cuda.synchronize()
    ds.check_flags()
    amax = tile_amax(tile.dsm, tile.dem, tile.cdsm)
    mask = 0
    for name in outputs:
        mask |= OUT_BITS[name]
    return SiteState(tile=tile, scene=ds, plan=plan, dplan=dplan, out=out, amax=amax,
                     offsets=march_offsets(solver, tile, plan, amax), mask=mask, outputs=tuple(outputs))


def _walls_aspect_window(tile: TileInputs, box, margin: int):
    """Walls and aspect of the (already edited) DSM over E' = box + margin, on the device.

    The device crop extends E' by the margin again, so the filter windows,
    the filter's interior bounds and the gradients of E' cells see the same
    cells as on the full tile (or the same tile border). Writes E' into the
    tile's wall and aspect planes and returns E'."""
    rows, cols = tile.shape
    inner = grow(box, margin, rows, cols)
    r0, r1, c0, c1 = grow(inner, margin, rows, cols)
    w, a = WG.walls_and_aspect(np.ascontiguousarray(tile.dsm[r0:r1, c0:c1]), tile.scale)
    i0, i1, j0, j1 = inner
    tile.walls[i0:i1, j0:j1] = w[i0 - r0:i1 - r0, j0 - c0:j1 - c0]
    tile.aspect_deg[i0:i1, j0:j1] = a[i0 - r0:i1 - r0, j0 - c0:j1 - c0]
    return inner


def edit_tiles(st: SiteState, hbox, ebox):
    """(SVF tiles, sun-march tiles, flux tiles) as boolean tile sets for an edit with changed-height
    box ``hbox`` and changed-input box ``ebox`` (heights, land cover, walls, aspect)."""
    rows, cols = st.tile.shape
    g = TileGrid(rows, cols)
    o = st.offsets
    a_svf = g.boxes(hbox, o.svf) if hbox is not None else g.empty()
    s_any = g.boxes(ebox, np.zeros((1, 2), np.int32))
    if hbox is not None:
        for sun in o.sun:
            if sun is not None:
                s_any |= g.boxes(hbox, sun)
    u = a_svf | g.dilate(s_any, o.walk)
    v = g.dilate(u, o.walk)
    return a_svf, v, u


def apply_edit(solver: HipSolver, st: SiteState, dsm, cdsm, landcover=None, stream=None) -> dict:
    """Apply an edit given as new full DSM / CDSM / land-cover planes (see :func:`apply_edit_window`)."""
    hbox, lbox = changed_boxes(st.tile, dsm, cdsm, landcover)
    box = union(hbox, lbox)
    if box is None:
        return {"mode": "none", "ms": 0.0}
    r0, r1, c0, c1 = box
    win = lambda x: None if x is None else np.asarray(x, F32)[r0:r1, c0:c1]  # noqa: E731
    return apply_edit_window(solver, st, box, win(dsm), win(cdsm), win(landcover), stream=stream)


def apply_edit_window(solver: HipSolver, st: SiteState, box, dsm=None, cdsm=None, landcover=None,
                      stream=None) -> dict:
    """Apply an edit of the window ``box`` = (r0, r1, c0, c1) and recompute the tiles it reaches.

    ``dsm`` / ``cdsm`` / ``landcover`` are the window's new values (None:
    unchanged). The site state becomes the edited one: its host planes
    (updated in place), the device scene and the day outputs, all equal to a
    full solve of the edited tile. Returns the tile counts and the host wall
    time (the call returns after the device work has finished).
    """
    t0 = time.perf_counter()
    tile = st.tile
    rows, cols = tile.shape
    r0, r1, c0, c1 = box
    if not (0 <= r0 < r1 <= rows and 0 <= c0 < c1 <= cols):
        raise ValueError(f"edit window {box} outside the {rows} x {cols} tile")
    sl = (slice(r0, r1), slice(c0, c1))
    shape = (r1 - r0, c1 - c0)

    def changed(new, old):
        if new is None:
            return None
        new = np.asarray(new, F32)
        if new.shape != shape:
            raise ValueError(f"edit window values of shape {new.shape}, expected {shape}")
        return new if not np.array_equal(new, old) else None
    dsm = changed(dsm, tile.dsm[sl])
    cdsm = changed(cdsm, tile.cdsm[sl])
    lc = changed(landcover, tile.landcover[sl]) if tile.landcover is not None else None
    info = {"box": box}
    if dsm is None and cdsm is None and lc is None:
        info.update(mode="none", ms=(time.perf_counter() - t0) * 1e3)
        return info
    # changed cells only (tighter boxes than the window)
    def tight(new, old):
        m = new != old
        rr, cc = np.nonzero(m.any(axis=1))[0], np.nonzero(m.any(axis=0))[0]
        return r0 + int(rr[0]), r0 + int(rr[-1]) + 1, c0 + int(cc[0]), c0 + int(cc[-1]) + 1
    hbox = None
    for new, old in ((dsm, tile.dsm[sl]), (cdsm, tile.cdsm[sl])):
        if new is not None:
            hbox = union(hbox, tight(new, old))
    lbox = tight(lc, tile.landcover[sl]) if lc is not None else None
    # amax of the edited tile: the window can raise it; lowering it needs the window to have held it
    old_win = tile_amax(tile.dsm[sl], tile.dem[sl], tile.cdsm[sl])
    if dsm is not None:
        tile.dsm[sl] = dsm
    if cdsm is not None:
        tile.cdsm[sl] = cdsm
    if lc is not None:
        tile.landcover[sl] = lc
    new_win = tile_amax(tile.dsm[sl], tile.dem[sl], tile.cdsm[sl])
    amax = st.amax
    if new_win > amax:
        amax = new_win
    elif old_win == st.amax and new_win < old_win:
        amax = tile_amax(tile.dsm, tile.dem, tile.cdsm)
    ebox = lbox
    if hbox is not None:
        ebox = union(_walls_aspect_window(tile, hbox, st.offsets.aspect), lbox)
    ds = st.scene
    # changed input windows to the device; compose then runs over the whole tile (per pixel, cheap)
    i0, i1, j0, j1 = ebox
    for key, plane in (("dsm", tile.dsm), ("cdsm", tile.cdsm), ("walls", tile.walls), ("aspect_deg", tile.aspect_deg),
                       ("landcover", tile.landcover)):
        dst = ds.inputs.get(key)
        if dst is not None and plane is not None:
            dst.view(rows, cols)[i0:i1, j0:j1] = torch.from_numpy(np.ascontiguousarray(plane[i0:i1, j0:j1])).to(
                dst.device)
    if amax != st.amax:
        # every march executes a different number of steps: solve the tile in full
        solver.svf(ds, stream)
        solver.tile_static(ds, st.dplan, stream)
        solver.run_day(ds, st.plan, st.dplan, st.outputs, out=st.out, stream=stream)
        _sync(stream)
        st.amax = amax
        st.offsets = march_offsets(solver, tile, st.plan, amax)
        info.update(mode="full", out_box=(0, rows, 0, cols), ms=(time.perf_counter() - t0) * 1e3)
        st.log.append(info)
        return info
    svf_t, sun_t, flux_t = edit_tiles(st, hbox, ebox)
    ids = [TileGrid.ids(m) for m in (svf_t, sun_t, flux_t)]
    dev_ids = [torch.from_numpy(x).to(solver.device) for x in ids]
    sp = _stream_ptr(stream)
    P = L.ptr
    solver.compose(ds, stream)
    if ids[0].size:
        offs, tab = solver.svf_tables(rows, cols, tile.scale)
        c = solver._fold_consts
        L.call("sw_svf_sweep_tiles", P(ds.hpack), P(ds.stats), P(offs), P(tab), P(ds.vegdsm2), P(c["w_iso"]),
               P(c["w_aniso"]), P(c["ring"]), P(c["nann"]), P(c["dirm"]), P(ds.counts), P(ds.hpad), P(ds.masks),
               P(ds.fields), P(dev_ids[0]), int(ids[0].size), rows, cols, sp)
    solver.tile_static(ds, st.dplan, stream)
    T = st.plan.n_steps
    is_day = (ctypes.c_int32 * T)(*[int(x) for x in st.plan.is_day])
    L.call("sw_solve_day_tiles", ctypes.byref(ds.tile_struct()), P(st.dplan["steps"]), P(st.dplan["model"]),
           P(st.dplan["sun_tab"]), is_day, 0, T, P(st.out), st.mask, st.plan.walk_halo, 1, P(dev_ids[1]),
           int(ids[1].size), P(dev_ids[2]), int(ids[2].size), sp)
    _sync(stream)
    g = TileGrid(rows, cols)
    # the pixel box of every tile whose outputs were recomputed (an edit's result window)
    ty, tx = np.nonzero(flux_t)
    out_box = None
    if ty.size:
        out_box = (int(ty.min()) * TILE, min(rows, (int(ty.max()) + 1) * TILE),
                   int(tx.min()) * TILE, min(cols, (int(tx.max()) + 1) * TILE))
    info.update(mode="tiles", svf_tiles=int(ids[0].size), sun_tiles=int(ids[1].size), flux_tiles=int(ids[2].size),
                grid_tiles=g.ty * g.tx, out_box=out_box, ms=(time.perf_counter() - t0) * 1e3)
    st.log.append(info)
    return info
