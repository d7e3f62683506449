"""Seeded synthetic city tiles and meteorological forcing.

The generator produces the host-resident inputs of one SOLWEIG full solve:
DSM, DEM, CDSM (trees above ground), TDSM rule, walls, wall aspect, land
cover and a 25-column UMEP-layout forcing table. It replaces the GeoTIFF /
ERA5 preprocessing chain of the reference (``preprocessor.py``,
``create_inputs.py``) which needs network data.

Walls follow ``findwalls`` (``solweig_gpu/walls_aspect.py:27-58``): the
4-neighbour maximum minus the cell, zeroed below the wall limit and on the
border. The wall aspect follows the rotating-filter method of
``filter1Goodwin_as_aspect_v3`` (``walls_aspect.py:99-166``) with the
gradient fallback of ``get_ders`` (``walls_aspect.py:78-97``), evaluated
with whole-plane correlations instead of a per-wall-cell loop.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy import ndimage

F32 = np.float32

#: Land-cover codes of the SOLWEIG class table (``landcoverclasses_2016a.txt``).
LC_PAVED = 1
LC_BUILDING = 2
LC_GRASS = 5
LC_WATER = 7


@dataclass
class Site:
    """One tile of host-resident inputs (all planes row-major float32)."""

    dsm: np.ndarray          # building DSM incl. ground (m)
    dem: np.ndarray          # ground elevation (m)
    cdsm: np.ndarray         # tree canopy height above ground (m); "temp1"
    walls: np.ndarray        # wall height (m)
    aspect: np.ndarray       # wall aspect (deg, 0 = north, clockwise)
    landcover: np.ndarray    # land-cover class codes (float32 like a raster)
    pixel_size: float = 1.0  # metres per pixel
    latitude: float = 30.0
    longitude: float = 0.0
    utc_offset: float = 0.0
    seed: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def shape(self):
        return self.dsm.shape

    @property
    def scale(self) -> float:
        """Pixels per metre, ``1 / geotransform[1]`` in the reference."""
        return 1.0 / self.pixel_size


def findwalls(dsm: np.ndarray, walllimit: float = 3.0) -> np.ndarray:
    """Wall heights from a DSM (``walls_aspect.py:27-58``).

    ``max(N, S, E, W neighbours) - cell``; values below ``walllimit`` are
    zeroed and the one-pixel border is zero. Computed in float64 like the
    reference (``np.zeros`` planes), returned as float32 (the raster dtype).
    """
    a = dsm.astype(np.float64)
    walls = np.zeros_like(a)
    nb = np.maximum.reduce([a[:-2, 1:-1], a[2:, 1:-1], a[1:-1, :-2], a[1:-1, 2:]])
    walls[1:-1, 1:-1] = nb
    walls = walls - a
    walls[walls < walllimit] = 0.0
    walls[:, 0] = 0.0
    walls[:, -1] = 0.0
    walls[0, :] = 0.0
    walls[-1, :] = 0.0
    return walls.astype(F32)


def _get_ders(dsm: np.ndarray, scale: float):
    """Slope aspect in radians (``walls_aspect.py:78-97``)."""
    dx = 1.0 / scale
    fy, fx = np.gradient(dsm.astype(np.float64), dx, dx)
    asp = np.arctan2(fx, fy)
    asp = -asp
    asp[asp < 0] += 2 * np.pi
    return asp


def wall_aspect(walls: np.ndarray, scale: float, dsm: np.ndarray) -> np.ndarray:
    """Wall aspect in degrees by the rotating line-filter method.

    For every rotation h in 0..179 a 9-pixel line filter (rotated with
    bilinear interpolation, rounded) is correlated with the binary wall map;
    the rotation with the strictly largest count wins (first maximum kept),
    and the side of the filter with the larger DSM sum decides between
    ``270 - h`` and ``270 - h - 180``. Cells left at 0 fall back to the DSM
    gradient aspect. Mirrors ``walls_aspect.py:99-166``.
    """
    rows, cols = dsm.shape
    filtersize = int(np.floor((scale + 1e-10) * 9))
    if filtersize <= 2:
        filtersize = 3
    elif filtersize != 9 and filtersize % 2 == 0:
        filtersize += 1
    n = filtersize - 1
    fceil = int(np.ceil(filtersize / 2.0))
    ffloor = int(np.floor(filtersize / 2.0))

    filt = np.zeros((filtersize, filtersize))
    build = np.zeros((filtersize, filtersize))
    filt[:, fceil - 1] = 1
    build[fceil - 1, :ffloor] = 1
    build[fceil - 1, fceil:] = 2

    wallb = (walls > 0).astype(np.float64)
    a64 = dsm.astype(np.float64)
    y = np.zeros((rows, cols))
    z = np.zeros((rows, cols))
    x = np.zeros((rows, cols))
    # interior window of the reference loop: i in [fceil-1, rows-fceil-1)
    interior = np.zeros((rows, cols), dtype=bool)
    interior[fceil - 1: max(fceil - 1, rows - fceil - 1), fceil - 1: max(fceil - 1, cols - fceil - 1)] = True
    active = interior & (wallb == 1)
    for h in range(180):
        f1 = np.round(ndimage.rotate(filt, h, order=1, reshape=False, mode="nearest"))
        fb = np.round(ndimage.rotate(build, h, order=0, reshape=False, mode="nearest"))
        index = 270 - h
        if h in (150, 30):
            fb[:, n] = 0
        if index == 225:
            f1[0, 0] = f1[n, n] = 1
        if index == 135:
            f1[0, n] = f1[n, 0] = 1
        s = ndimage.correlate(wallb, f1, mode="constant", cval=0.0)
        s1 = ndimage.correlate(a64, (fb == 1).astype(np.float64), mode="constant", cval=0.0)
        s2 = ndimage.correlate(a64, (fb == 2).astype(np.float64), mode="constant", cval=0.0)
        upd = active & (z < s)
        z[upd] = s[upd]
        x[upd] = np.where(s1[upd] > s2[upd], 1, 2)
        y[upd] = index
    y[x == 1] -= 180
    y[y < 0] += 360
    asp = _get_ders(dsm, scale)
    y += ((wallb == 1) & (y == 0)) * (asp / (math.pi / 180.0))
    return y.astype(F32)


def _place_buildings(rng, rows, cols, cover, dsm):
    occupied = np.zeros((rows, cols), dtype=bool)
    lo, hi = 8, 40
    target = cover * rows * cols
    guard = 0
    count = 0                                   # occupied.sum(), kept incrementally
    while count < target and guard < 100000:
        guard += 1
        h = int(rng.integers(lo, hi + 1))
        w = int(rng.integers(lo, hi + 1))
        h = min(h, rows)
        w = min(w, cols)
        r0 = int(rng.integers(0, rows - h + 1))
        c0 = int(rng.integers(0, cols - w + 1))
        height = F32(rng.uniform(5.0, 40.0))
        sl = (slice(r0, r0 + h), slice(c0, c0 + w))
        dsm[sl] = np.maximum(dsm[sl], height)
        count += h * w - int(occupied[sl].sum())
        occupied[sl] = True
    return occupied


def _place_trees(rng, rows, cols, cover, building):
    cdsm = np.zeros((rows, cols), dtype=F32)
    target = cover * rows * cols
    ground = np.flatnonzero(~building.ravel())
    if ground.size == 0:
        return cdsm
    guard = 0
    count = 0                                   # (cdsm > 0).sum(), kept incrementally
    while count < target and guard < 100000:
        guard += 1
        k = int(ground[rng.integers(0, ground.size)])
        ci, cj = divmod(k, cols)
        r = int(rng.integers(2, 7))
        h = float(rng.uniform(3.0, 15.0))
        sigma = r / 2.0
        i0, i1 = max(0, ci - r), min(rows, ci + r + 1)
        j0, j1 = max(0, cj - r), min(cols, cj + r + 1)
        ii, jj = np.mgrid[i0:i1, j0:j1]
        d2 = (ii - ci) ** 2 + (jj - cj) ** 2
        crown = (h * np.exp(-d2 / (2.0 * sigma * sigma))).astype(F32)
        crown[d2 > r * r] = 0
        crown[building[i0:i1, j0:j1]] = 0
        before = int((cdsm[i0:i1, j0:j1] > 0).sum())
        cdsm[i0:i1, j0:j1] = np.maximum(cdsm[i0:i1, j0:j1], crown)
        count += int((cdsm[i0:i1, j0:j1] > 0).sum()) - before
    return cdsm


def mixed_dsm(rows: int, cols: int, seed: int = 7, *, cover: float = 0.25, slope: float = 0.05) -> np.ndarray:
    """DSM with walls of every orientation: rotated rectangles and round towers on sloped ground.

    Ground rises ``slope`` m per pixel along a random direction; roofs are
    flat at U(4, 30) m above the ground under the footprint centre, with
    fractional heights, so the wall aspect and its gradient fallback both see
    non-trivial float32 sums.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    ii, jj = np.mgrid[0:rows, 0:cols].astype(np.float64)
    th = rng.uniform(0.0, 2.0 * np.pi)
    ground = slope * (np.cos(th) * ii + np.sin(th) * jj)
    ground -= ground.min()
    dsm = ground.astype(F32)
    roof = np.zeros((rows, cols), bool)
    guard = 0
    while roof.sum() < cover * rows * cols and guard < 1000:
        guard += 1
        ci, cj = rng.uniform(0, rows), rng.uniform(0, cols)
        if rng.random() < 0.35:
            r = rng.uniform(2.5, max(3.0, min(rows, cols) / 6))
            fp = (ii - ci) ** 2 + (jj - cj) ** 2 <= r * r
        else:
            a = rng.uniform(0.0, np.pi)
            hl, hw = rng.uniform(2.5, max(3.0, rows / 4)), rng.uniform(2.5, max(3.0, cols / 5))
            u = (ii - ci) * np.cos(a) + (jj - cj) * np.sin(a)
            v = -(ii - ci) * np.sin(a) + (jj - cj) * np.cos(a)
            fp = (np.abs(u) <= hl) & (np.abs(v) <= hw)
        if not fp.any():
            continue
        base = ground[min(rows - 1, int(ci)), min(cols - 1, int(cj))]
        top = F32(base + rng.uniform(4.0, 30.0))
        dsm[fp] = np.maximum(dsm[fp], top)
        roof |= fp
    return dsm


def make_site(rows: int, cols: int, seed: int = 20261203, *, pixel_size: float = 1.0,
              building_cover: float = 0.30, canopy_cover: float = 0.05,
              with_aspect: bool = True, latitude: float = 30.0,
              longitude: float = 0.0, utc_offset: float = 0.0) -> Site:
    """Seeded city tile: flat ground, rectangular buildings, Gaussian crowns.

    Buildings: axis-aligned rectangles with sides U{8..40} px and heights
    U(5, 40) m placed (maximum on overlap) until ``building_cover`` of the
    cells are roofs. Trees: crowns ``h exp(-d^2 / (2 sigma^2))`` with
    h ~ U(3, 15) m, r ~ U{2..6} px, sigma = r/2, cut at d > r, on non-roof
    cells, combined by maximum until ``canopy_cover`` of the cells carry
    canopy. Land cover: roofs are class 2, other cells random 8x8 blocks of
    paved (1), grass (5) and water (7).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    dem = np.zeros((rows, cols), dtype=F32)
    dsm = np.zeros((rows, cols), dtype=F32)
    building = _place_buildings(rng, rows, cols, building_cover, dsm)
    rng_t = np.random.Generator(np.random.PCG64(seed + 1000))
    cdsm = _place_trees(rng_t, rows, cols, canopy_cover, building)
    rng_l = np.random.Generator(np.random.PCG64(seed + 2000))
    br, bc = (rows + 7) // 8, (cols + 7) // 8
    blocks = rng_l.choice(np.array([LC_PAVED, LC_GRASS, LC_WATER], dtype=F32), size=(br, bc))
    lc = np.kron(blocks, np.ones((8, 8), dtype=F32))[:rows, :cols].astype(F32)
    lc[building] = LC_BUILDING
    walls = findwalls(dsm, 3.0)
    scale = 1.0 / pixel_size
    if with_aspect:
        aspect = wall_aspect(walls, scale, dsm)
    else:
        aspect = np.zeros_like(walls)
    return Site(dsm=dsm, dem=dem, cdsm=cdsm, walls=walls, aspect=aspect,
                landcover=np.ascontiguousarray(lc), pixel_size=pixel_size,
                latitude=latitude, longitude=longitude, utc_offset=utc_offset,
                seed=seed)


def make_met(altitude_deg: np.ndarray, *, year: int = 2021, doy: int = 180,
             hours=None, seed: int = 20261204, pressure_kpa: float = 101.3) -> np.ndarray:
    """25-column UMEP forcing table for one synthetic day.

    Columns used by the solver: 0 year, 1 DOY, 2 hour, 3 minute, 9 wind
    speed, 10 RH, 11 Ta, 12 pressure (kPa), 14 global radiation, 21 diffuse
    and 22 direct radiation (-999: recomputed from global), 23 wind
    direction, 24 urban heat island increment. Ta = 29 + 7 sin(2 pi (h-9)/24)
    (22-36 C), RH = 50 - 20 sin(...) (30-70 %), wind ~ U(0.5, 4) m/s,
    clear-sky global radiation max(0, 0.75 * 1367 * sin(alt)).
    """
    if hours is None:
        hours = np.arange(24)
    hours = np.asarray(hours, dtype=np.float64)
    n = hours.size
    rng = np.random.Generator(np.random.PCG64(seed))
    met = np.full((n, 25), -999.0)
    met[:, 0] = year
    met[:, 1] = doy
    met[:, 2] = hours
    met[:, 3] = 0.0
    ph = 2.0 * np.pi * (hours - 9.0) / 24.0
    met[:, 11] = 29.0 + 7.0 * np.sin(ph)
    met[:, 10] = 50.0 - 20.0 * np.sin(ph)
    met[:, 9] = rng.uniform(0.5, 4.0, n)
    met[:, 12] = pressure_kpa
    alt = np.asarray(altitude_deg, dtype=np.float64).reshape(-1)
    met[:, 14] = np.maximum(0.0, 0.75 * 1367.0 * np.sin(np.radians(alt)))
    met[:, 21] = -999.0
    met[:, 22] = -999.0
    met[:, 23] = rng.uniform(0.0, 360.0, n)
    met[:, 24] = 0.0
    for c in (4, 5, 6, 7, 8, 13, 15, 16, 17, 18, 19, 20):
        met[:, c] = -999.0
    return met


@dataclass
class Edit:
    kind: str                # "add_tree", "remove_tree", "add_building", "remove_building"
    rows: tuple              # (r0, r1) bounding rows of the footprint
    cols: tuple              # (c0, c1)
    pixels: int              # footprint in pixels


def edit_scenarios(base: Site, n: int, seed: int = 20261223, *, max_edits: int = 20):
    """``n`` edited copies of ``base``: each adds or removes 1..``max_edits`` trees or buildings.

    Footprints are 10..400 px: trees are discs of radius 2..11 px with crowns
    ``h exp(-d^2 / (2 sigma^2))``, h ~ U(3, 15) m, on non-roof cells;
    buildings are rectangles of 10..400 px at U(5, 40) m (class 2, no canopy
    under them). Removing a tree clears the canopy in a disc around a canopy
    cell; removing a building lowers a rectangle around a roof cell to the
    ground (class 1). Returns ``[(dsm, cdsm, landcover, [Edit, ...]), ...]``
    as float32 planes; walls and aspect are left to the caller (they follow
    from the DSM).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    rows, cols = base.dsm.shape
    out = []
    for _ in range(n):
        dsm = base.dsm.copy()
        cdsm = base.cdsm.copy()
        lc = base.landcover.copy()
        edits = []
        for _ in range(int(rng.integers(1, max_edits + 1))):
            kind = ("add_tree", "remove_tree", "add_building", "remove_building")[int(rng.integers(0, 4))]
            roof = dsm > base.dem + 2.0
            if kind in ("add_tree", "remove_tree"):
                r = int(rng.integers(2, 12))
                if kind == "add_tree":
                    ci, cj = int(rng.integers(0, rows)), int(rng.integers(0, cols))
                else:
                    cells = np.flatnonzero(cdsm.ravel() > 0)
                    if cells.size == 0:
                        continue
                    ci, cj = divmod(int(cells[rng.integers(0, cells.size)]), cols)
                i0, i1, j0, j1 = max(0, ci - r), min(rows, ci + r + 1), max(0, cj - r), min(cols, cj + r + 1)
                ii, jj = np.mgrid[i0:i1, j0:j1]
                d2 = (ii - ci) ** 2 + (jj - cj) ** 2
                disc = d2 <= r * r
                if kind == "add_tree":
                    h = float(rng.uniform(3.0, 15.0))
                    sigma = r / 2.0
                    crown = (h * np.exp(-d2 / (2.0 * sigma * sigma))).astype(F32)
                    crown[~disc | roof[i0:i1, j0:j1]] = 0
                    cdsm[i0:i1, j0:j1] = np.maximum(cdsm[i0:i1, j0:j1], crown)
                else:
                    cdsm[i0:i1, j0:j1][disc] = 0
                edits.append(Edit(kind, (i0, i1), (j0, j1), int(disc.sum())))
            else:
                area = int(rng.integers(10, 401))
                h = int(rng.integers(max(2, -(-area // 20)), min(20, area // 2) + 1))
                w = max(1, min(cols, area // h))
                h = min(h, rows)
                if kind == "add_building":
                    r0, c0 = int(rng.integers(0, rows - h + 1)), int(rng.integers(0, cols - w + 1))
                else:
                    cells = np.flatnonzero(roof.ravel())
                    if cells.size == 0:
                        continue
                    ci, cj = divmod(int(cells[rng.integers(0, cells.size)]), cols)
                    r0, c0 = max(0, min(rows - h, ci - h // 2)), max(0, min(cols - w, cj - w // 2))
                sl = (slice(r0, r0 + h), slice(c0, c0 + w))
                if kind == "add_building":
                    dsm[sl] = np.maximum(dsm[sl], base.dem[sl] + F32(rng.uniform(5.0, 40.0)))
                    lc[sl] = LC_BUILDING
                    cdsm[sl] = 0
                else:
                    dsm[sl] = base.dem[sl]
                    lc[sl] = np.where(lc[sl] == LC_BUILDING, F32(LC_PAVED), lc[sl])
                edits.append(Edit(kind, (r0, r0 + h), (c0, c0 + w), h * w))
        out.append((dsm.astype(F32), cdsm.astype(F32), lc.astype(F32), edits))
    return out


def site_with(base: Site, dsm: np.ndarray, cdsm: np.ndarray, landcover: np.ndarray, walls=None,
              aspect=None) -> Site:
    """A copy of ``base`` with new surfaces; walls/aspect recomputed on the CPU unless given."""
    if walls is None:
        walls = findwalls(dsm, 3.0)
    if aspect is None:
        aspect = wall_aspect(walls, base.scale, dsm)
    return Site(dsm=dsm, dem=base.dem, cdsm=cdsm, walls=np.asarray(walls, F32), aspect=np.asarray(aspect, F32),
                landcover=landcover, pixel_size=base.pixel_size, latitude=base.latitude,
                longitude=base.longitude, utc_offset=base.utc_offset, seed=base.seed)
