"""The per-tile file driver ``compute_utci`` on the GPU path.

Mirrors ``solweig_gpu/utci_process.py:1199-1663``: same positional and
keyword arguments, same files read and written, same validation errors.
For one tile it

1. reads the Building DSM, tree (canopy) DSM, DEM, wall height and wall
   aspect rasters, and the optional land-cover and wind-coefficient rasters
   (a single raster, or a dict of the 12 directional rasters 0-330 by 30,
   each of the DSM's shape);
2. cleans the land-cover grid (codes outside 1-7 become 6, codes 3 and 4
   become 5, with the reference's warnings; class constants of
   ``landcoverclasses_2016a``);
3. locates the tile: the raster centre converted to WGS 84 (:mod:`.geo`),
   altitude 3 m when the DEM's lower median is positive (else that median),
   and the UTC offset of the tile's time zone on ``selected_date_str``;
4. derives the sun table from the forcing rows;
5. reuses the SVF cache in ``<base>/SVF`` (``base`` = the parent of the
   Building DSM's folder) when all three files exist and match the DSM's
   size and geotransform, else computes SVF on the GPU and writes the cache;
6. runs the time loop (:func:`~.api.run_utci_window`) for UTCI and the
   requested outputs;
7. writes ``UTCI_<number>.tif`` and each requested ``<Name>_<number>.tif``
   as multi-band Float32 GeoTIFFs (one band per forcing row, band metadata
   ``Time`` = the row's local time on the selected date), and
   ``SVF_<number>.tif`` when ``save_svf`` and the cache was not reused.

Differences a caller can observe:

* the time zone: the reference looks it up from the tile's coordinates with
  ``timezonefinder``, which this package does not carry. Pass ``timezone``
  (an IANA name) or ``utc_offset`` (hours); without either the call is
  refused before any work is done;
* projections: WGS 84 UTM and EPSG:4326 rasters are located here; for others
  pass ``location`` (``{'latitude', 'longitude', 'altitude'}``);
* the return value is a summary dictionary (files written, location, UTC
  offset, whether the SVF cache was reused) instead of None;
* what the time loop refuses (:mod:`.api`: direction-dependent wind, a
  non-uniform wind coefficient, a non-zero bush layer) is refused here too.
"""
from __future__ import annotations

import datetime
import os
import time

import numpy as np

from . import geo
from . import met as M
from . import svf_cache as C
from . import sunpos

F32 = np.float32

#: output flag -> (requested variable, file prefix), in the reference's write order
OUTPUTS = (("save_tmrt", "tmrt", "TMRT"), ("save_kup", "kup", "Kup"), ("save_kdown", "kdown", "Kdown"),
           ("save_lup", "lup", "Lup"), ("save_ldown", "ldown", "Ldown"), ("save_shadow", "shadow", "Shadow"),
           ("save_wbgt", "wbgt", "WBGT"), ("save_ta", "ta", "Ta"), ("save_wind", "wind", "Wind"))
WIND_DIRECTIONS = tuple(range(0, 360, 30))


def svf_cache_paths(building_dsm_path: str, number: str):
    """(base_path, svf_dir, SkyViewFactor path, svfs zip path, shadowmats path) of a tile
    (``utci_process.py:226-247``)."""
    base = os.path.dirname(os.path.dirname(building_dsm_path))
    svf_dir = os.path.join(base, "SVF")
    tot, zpath, npz = C.cache_paths(svf_dir, number)
    return base, svf_dir, tot, zpath, npz


def svf_cache_usable(building_dsm_path: str, number: str) -> bool:
    """All three cache files exist and match the DSM tile's size and geotransform; an unreadable DSM or
    damaged cache file counts as unusable (``utci_process.py:259-313``)."""
    try:
        dsm = C.read_geotiff_full(building_dsm_path)
    except (OSError, ValueError, KeyError, IndexError):
        return False
    _, svf_dir, _, _, _ = svf_cache_paths(building_dsm_path, number)
    return C.cache_matches(svf_dir, number, dsm.shape, dsm.geotransform)


def lower_median(x: np.ndarray) -> float:
    """The lower of the two middle values for an even count (what ``torch.median`` returns)."""
    v = np.asarray(x, np.float64).ravel()
    if v.size == 0:
        raise ValueError("median of an empty raster")
    if np.isnan(v).any():
        return float("nan")
    k = (v.size - 1) // 2
    return float(np.partition(v, k)[k])


def tile_location(dsm: C.GeoRaster, dem: np.ndarray) -> dict:
    """WGS 84 location of the tile centre and the driver's altitude rule (``utci_process.py:1330-1370``)."""
    rows, cols = dsm.shape
    x, y = geo.tile_centre(dsm.geotransform, cols, rows)
    lon, lat = geo.to_lonlat(x, y, dsm.epsg)
    alt = lower_median(dem)
    if alt > 0:
        alt = 3.0
    return {"longitude": lon, "latitude": lat, "altitude": alt}


def utc_offset_hours(timezone_name: str, date: datetime.datetime) -> float:
    """UTC offset (hours) of ``timezone_name`` at local ``date`` (pytz ``localize``, as the reference)."""
    import pytz
    tz = pytz.timezone(timezone_name)
    return tz.localize(date).utcoffset().total_seconds() / 3600


def _load(path, what: str) -> C.GeoRaster:
    if path is None or not os.path.isfile(path):
        raise FileNotFoundError(f"Could not open raster: {path}")
    try:
        return C.read_geotiff_full(path)
    except ValueError as e:
        raise ValueError(f"{what} raster {path}: {e}") from None


def _band(r: C.GeoRaster) -> np.ndarray:
    return np.ascontiguousarray(r.data[0], F32)


def load_windcoeff(windcoeff_path, number: str, shape):
    """(windcoeff, windcoeff_by_dir) of the driver's three forms (``utci_process.py:1248-1286``)."""
    if windcoeff_path is None:
        return None, None
    if isinstance(windcoeff_path, dict):
        missing = [d for d in WIND_DIRECTIONS if d not in windcoeff_path]
        if missing:
            raise FileNotFoundError(f"Tile {number} is missing directional wind coefficient files: "
                                    + ", ".join(f"WindCoeff_dir{d:03d}_{number}.tif" for d in missing))
        by_dir = {}
        for d in WIND_DIRECTIONS:
            arr = _band(_load(windcoeff_path[d], "wind coefficient"))
            if arr.shape != tuple(shape):
                raise ValueError(f"Wind coefficient raster shape {arr.shape} for direction {d:03d} does not match "
                                 f"Building DSM shape {tuple(shape)}")
            by_dir[d] = arr
        return None, by_dir
    arr = _band(_load(windcoeff_path, "wind coefficient"))
    if arr.shape != tuple(shape):
        raise ValueError(f"Wind coefficient raster shape {arr.shape} does not match Building DSM shape {tuple(shape)}")
    return arr, None


def clean_landcover_with_warnings(lc: np.ndarray, log=print) -> np.ndarray:
    """Land-cover cleanup of the driver with its two warnings (``utci_process.py:1296-1313``)."""
    g = np.array(lc, dtype=F32, copy=True)
    bad = (g < 1) | (g > 7)
    if bad.any():
        log("Warning: land-cover grid contains values outside 1-7. Invalid cells are set to 6 (bare soil). ")
        g[bad] = 6
    veg = (g == 3) | (g == 4)
    if veg.any():
        log("Attention! The land cover grid includes values (deciduous and/or conifer) not appropriate for the "
            "SOLWEIG-formatted land cover grid (should not include 3 or 4). Land cover under the vegetation is "
            "required. Setting the invalid landcover types to grass.")
        g[veg] = 5
    return g


def band_times(met: np.ndarray, base_date: datetime.datetime) -> list:
    """Per-band ``Time`` metadata: the row's hour and minute on the selected date."""
    return [{"Time": base_date.replace(hour=int(r[2]), minute=int(r[3])).isoformat()} for r in met]


def compute_utci(building_dsm_path, tree_path, dem_path, walls_path, aspect_path, landcover_path, windcoeff_path,
                 met_file, output_path, number, selected_date_str, save_tmrt=False, save_svf=False, save_kup=False,
                 save_kdown=False, save_lup=False, save_ldown=False, save_shadow=False, save_wbgt=False,
                 save_ta=False, save_wind=False, *, timezone: str | None = None, utc_offset: float | None = None,
                 location: dict | None = None, device="cuda", log=print):
    """UTCI and the requested outputs of one tile, written as GeoTIFFs into ``output_path``."""
    from . import api
    t_start = time.time()
    dsm = _load(building_dsm_path, "Building DSM")
    a = _band(dsm)
    temp1 = _band(_load(tree_path, "tree"))
    temp2 = _band(_load(dem_path, "DEM"))
    walls = _band(_load(walls_path, "walls"))
    dirwalls = _band(_load(aspect_path, "aspect"))
    for name, arr in (("tree", temp1), ("DEM", temp2), ("walls", walls), ("aspect", dirwalls)):
        if arr.shape != a.shape:
            raise ValueError(f"{name} raster shape {arr.shape} does not match Building DSM shape {a.shape}")
    windcoeff, windcoeff_by_dir = load_windcoeff(windcoeff_path, number, a.shape)
    lcgrid = None
    if landcover_path is not None:
        lcgrid = clean_landcover_with_warnings(_band(_load(landcover_path, "land cover")), log)
        if lcgrid.shape != a.shape:
            raise ValueError(f"land-cover raster shape {lcgrid.shape} does not match Building DSM shape {a.shape}")
    met = np.asarray(met_file, dtype=np.float64)
    if met.ndim != 2 or met.shape[1] < 4 or met.shape[0] < 1:
        raise ValueError("met_file must be a forcing table (rows x columns, year/doy/hour/minute first)")
    base_date = datetime.datetime.strptime(selected_date_str, "%Y-%m-%d")
    rows, cols = a.shape
    gt = dsm.geotransform
    if gt is None:
        raise ValueError(f"{building_dsm_path} carries no geotransform")
    scale = 1 / gt[1]
    if location is None:
        location = tile_location(dsm, temp2)
    else:
        location = {"longitude": float(location["longitude"]), "latitude": float(location["latitude"]),
                    "altitude": float(location.get("altitude", 0.0))}
    if utc_offset is None:
        if timezone is None:
            raise ValueError("the tile's time zone is needed (the reference looks it up with timezonefinder, "
                             "which this package does not carry): pass timezone='<IANA name>' or utc_offset=<hours>")
        utc = utc_offset_hours(timezone, base_date)
        log(f"[INFO] Timezone: {timezone}, UTC offset: {utc} hours")
    else:
        utc = float(utc_offset)
    sun = sunpos.sun_table(met, location["latitude"], location["longitude"], utc, location["altitude"])

    # scene compose of the driver (utci_process.py:1370-1377)
    t1 = np.where(temp1 < 0, F32(0), temp1).astype(F32)
    vegdem = (t1 + temp2).astype(F32)
    vegdem2 = (t1 * F32(0.25) + temp2).astype(F32)
    bush = (np.logical_not(vegdem2 * vegdem) * vegdem).astype(F32)
    vegdsm = (t1 + a).astype(F32)
    vegdsm[vegdsm == a] = 0
    vegdsm2 = (t1 * F32(0.25) + a).astype(F32)
    vegdsm2[vegdsm2 == a] = 0
    amaxvalue = float(max(a.max(), vegdem.max()))

    _, svf_dir, _, _, _ = svf_cache_paths(building_dsm_path, number)
    cached = C.cache_matches(svf_dir, number, (rows, cols), gt)
    if cached:
        log(f"[INFO] Loading cached SVF outputs for tile {number} from {svf_dir}")
        bundle = C.load_svf_cache(svf_dir, number, (rows, cols), gt)
    else:
        log(f"[INFO] Cached SVF outputs not found for tile {number}. Calculating SVF.")
        bundle = api.svf_calculator(2, amaxvalue, a, vegdsm, vegdsm2, bush, scale, device=device)
        C.save_svf_cache(svf_dir, number, bundle, gt, dsm.epsg)

    flags = {"save_tmrt": save_tmrt, "save_kup": save_kup, "save_kdown": save_kdown, "save_lup": save_lup,
             "save_ldown": save_ldown, "save_shadow": save_shadow, "save_wbgt": save_wbgt, "save_ta": save_ta,
             "save_wind": save_wind}
    requested = ["utci"] + [var for flag, var, _ in OUTPUTS if flags[flag] and var != "wbgt"]
    outputs = api.run_utci_window(
        a=a, temp1=t1, temp2=temp2, walls=walls, dirwalls=dirwalls, svf_bundle=bundle, met_file=met,
        altitude=sun.altitude, azimuth=sun.azimuth, zen=sun.zen, jday=sun.jday, dectime=sun.dectime,
        altmax=sun.altmax, location=location, scale=scale, amaxvalue=amaxvalue, landcover_grid=lcgrid,
        lc_class=None, windcoeff=windcoeff, windcoeff_by_dir=windcoeff_by_dir,
        requested_variables=tuple(requested), save_wbgt=bool(save_wbgt), device=device)

    os.makedirs(output_path, exist_ok=True)
    times = band_times(met, base_date)
    written = []

    def put(prefix, cube):
        path = os.path.join(output_path, f"{prefix}_{number}.tif")
        C.write_geotiff(path, cube, gt, dsm.epsg, band_metadata=times)
        written.append(path)
    put("UTCI", outputs["utci"])
    if save_tmrt:
        put("TMRT", outputs["tmrt"])
    if save_svf and not cached:
        path = os.path.join(output_path, f"SVF_{number}.tif")
        C.write_geotiff(path, np.asarray(bundle[18] if not hasattr(bundle, "field") else
                                         bundle.field("svftotal").detach().cpu().numpy(), F32).reshape(rows, cols),
                        gt, dsm.epsg)
        written.append(path)
    elif save_svf:
        log(f"[INFO] SVF cache already exists for tile {number}; skipping duplicate SVF_{number}.tif write.")
    for flag, var, prefix in OUTPUTS[1:]:
        if flags[flag]:
            put(prefix, outputs[var])
    took = time.time() - t_start
    log(f"Time taken to execute tile {number}: {took:.2f} seconds")
    return {"files": written, "location": location, "utc_offset": utc, "svf_cache_reused": cached,
            "svf_dir": svf_dir, "seconds": took}
