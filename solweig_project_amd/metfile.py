"""Forcing preprocessing: UMEP 25-column met files from gridded forcing.

The array-level core of the reference's preprocessor (``solweig_gpu/``):

* ``era5_fields`` - ERA5 single-level fields to the forcing variables
  (``preprocessor.py:541-566``): Magnus relative humidity from 2 m
  temperature and dewpoint, wind speed and meteorological direction from
  u10/v10, hourly ssrd accumulations to W m-2;
* ``uhi_cycle`` - the nighttime urban heat island cycle
  (``preprocessor.py:72-189``): daily amplitude (S_down/rho_cp * DTR^3 /
  wind)^(1/4) shaped by a half sine over each night;
* ``sample_tile`` - one tile's met rows (``preprocessor.py:1081-1352``):
  the UTC records of the tile's local day, each variable averaged over the
  grid points inside the tile's lon/lat box (nearest grid point when the
  grid is coarser than the tile or no point falls inside), unit
  conversions K -> C and Pa -> kPa, -999 for missing values;
* ``write_metfile`` / ``read_metfile`` - the 25-column text format with its
  header line and per-column precision (``preprocessor.py:1336-1352``,
  read as ``solweig_gpu.py:472`` does).

File I/O of NetCDF (netCDF4 / xarray), raster CRS transforms (GDAL) and the
timezone lookup (timezonefinder) are not part of this module: callers pass
arrays, the tile's lon/lat box and an IANA timezone name.
"""
from __future__ import annotations

import datetime as _dt
import math
from dataclasses import dataclass

import numpy as np

#: column order of the met file (preprocessor.py:1135-1148)
COLUMNS = ("iy", "id", "it", "imin", "Q*", "QH", "QE", "Qs", "Qf", "Wind", "RH", "Td", "press", "rain", "Kdn",
           "snow", "ldown", "fcld", "wuh", "xsmd", "lai_hr", "Kdiff", "Kdir", "Wd", "uhii")
#: printf precision per column after the four integer time columns
_DECIMALS = (2, 2, 2, 2, 2, 5, 2, 2, 2) + (2,) * 12
#: gridded variable sampled into each met column (preprocessor.py:1095-1103)
GRID_VARS = {"Wind": "WIND", "RH": "RH2", "Td": "T2", "press": "PSFC", "Kdn": "SWDOWN", "Wd": "WDIR",
             "uhii": "UHI_CYCLE"}
MISSING = -999.0
RHO_CP = 1.225 * 1005.0          # air density * cp (preprocessor.py:37)


# ---------------------------------------------------------------------------
# met file text format
# ---------------------------------------------------------------------------
def format_row(row) -> str:
    """One met-file line: four integers, then 21 fixed-point values."""
    if len(row) != len(COLUMNS):
        raise ValueError(f"a met row has {len(COLUMNS)} values, got {len(row)}")
    head = " ".join(str(int(v)) for v in row[:4])
    tail = " ".join(f"{float(v):.{d}f}" for v, d in zip(row[4:], _DECIMALS))
    return f"{head} {tail}"


def write_metfile(path, rows) -> None:
    """Write ``rows`` (N x 25) with the header line of column names."""
    with open(path, "w") as f:
        f.write(" ".join(COLUMNS) + "\n")
        for r in rows:
            f.write(format_row(r) + "\n")


def read_metfile(path) -> np.ndarray:
    """The forcing table of a met file: float64 (N, ncols), header skipped, single-space separated."""
    a = np.loadtxt(path, skiprows=1, delimiter=" ", ndmin=2)
    if a.shape[1] < 23:
        raise ValueError(f"{path}: a met file has at least 23 columns, found {a.shape[1]}")
    return a


# ---------------------------------------------------------------------------
# ERA5 single-level fields -> forcing variables
# ---------------------------------------------------------------------------
def magnus_svp_hpa(tc):
    """Magnus saturation vapour pressure in hPa, T in C (preprocessor.py:541-553)."""
    return 6.112 * np.exp((17.67 * tc) / (tc + 243.5))


def era5_fields(t2m, d2m, sp, u10, v10, ssrd) -> dict:
    """Gridded forcing (time, lat, lon) float32 from ERA5 arrays: T2 (K), PSFC (Pa), RH2 (%),
    WIND (m/s), WDIR (degrees the wind blows from; 0 below 0.01 m/s), SWDOWN (W m-2)."""
    t2m, d2m = np.asarray(t2m), np.asarray(d2m)
    u10, v10 = np.asarray(u10), np.asarray(v10)
    speed = np.sqrt(u10 ** 2 + v10 ** 2)
    wdir = (270.0 - np.degrees(np.arctan2(v10, u10))) % 360.0
    wdir = np.where(speed > 0.01, wdir, 0.0)
    rh = np.clip(100.0 * (magnus_svp_hpa(d2m - 273.15) / magnus_svp_hpa(t2m - 273.15)), 0, 100)
    sw = (np.asarray(ssrd) / 3600.0).astype(np.float32)
    f = np.float32
    return {"T2": t2m.astype(f), "PSFC": np.asarray(sp).astype(f), "RH2": rh.astype(f), "WIND": speed.astype(f),
            "WDIR": wdir.astype(f), "SWDOWN": sw}


# ---------------------------------------------------------------------------
# nighttime UHI cycle
# ---------------------------------------------------------------------------
def night_sine_profile(sw, amp, extend_before: int = 0, extend_after: int = 1) -> np.ndarray:
    """Nighttime profile of one series (preprocessor.py:72-124): each run of valid steps with
    SW <= 5 W m-2, extended by the given counts, gets mean(amp) * sin(linspace(0, pi, L)) when
    it spans two or more steps and its mean amplitude is positive."""
    sw = np.asarray(sw, dtype=float)
    amp = np.asarray(amp, dtype=float)
    n = sw.size
    out = np.zeros(n)
    if n == 0 or amp.shape != sw.shape:
        return out
    night = np.isfinite(sw) & np.isfinite(amp) & (sw <= 5.0)
    idx = np.flatnonzero(night)
    if idx.size == 0:
        return out
    for run in np.split(idx, np.flatnonzero(np.diff(idx) > 1) + 1):
        s = max(0, int(run[0]) - extend_before)
        e = min(n, int(run[-1]) + 1 + extend_after)
        if e - s < 2:
            continue
        a = amp[s:e]
        a = a[np.isfinite(a)]
        if a.size == 0:
            continue
        A = float(np.mean(a))
        if A <= 0:
            continue
        out[s:e] = A * np.maximum(np.sin(np.linspace(0, np.pi, e - s)), 0.0)
    return out


def _day_index(times) -> np.ndarray:
    t = np.asarray(times, dtype="datetime64[ns]")
    return t.astype("datetime64[D]")


def uhi_cycle(times, t2_k, sw, u, v) -> np.ndarray:
    """UHI_CYCLE (time, lat, lon) float32 in K (preprocessor.py:126-189).

    Per calendar day of ``times``: mean S_down / rho_cp, mean wind speed, daily
    temperature range; amplitude (max(0, S DTR^3) / max(wind, 1))^(1/4) (0 where
    undefined), held over the day's records, then shaped per grid point by
    ``night_sine_profile`` with no extension."""
    t2 = np.asarray(t2_k, dtype=float)
    swa = np.asarray(sw, dtype=float)
    wd = np.sqrt(np.asarray(u, dtype=float) ** 2 + np.asarray(v, dtype=float) ** 2)
    days = _day_index(times)
    uniq, inv = np.unique(days, return_inverse=True)
    amp_day = np.zeros((uniq.size,) + t2.shape[1:])
    with np.errstate(all="ignore"):
        for k in range(uniq.size):
            sel = inv == k
            s_mean = np.nanmean(swa[sel] / RHO_CP, axis=0)
            w_mean = np.nanmean(wd[sel], axis=0)
            dtr = np.nanmax(t2[sel], axis=0) - np.nanmin(t2[sel], axis=0)
            num = s_mean * dtr ** 3
            num = np.where(np.isfinite(num), num, 0.0)
            num = np.maximum(num, 0.0)
            w_safe = np.where(w_mean > 1.0, w_mean, 1.0)
            a = (num / w_safe) ** 0.25
            amp_day[k] = np.where(np.isnan(a), 0.0, a)
    amp = amp_day[inv]
    out = np.zeros_like(t2)
    flat_sw = swa.reshape(swa.shape[0], -1)
    flat_amp = amp.reshape(amp.shape[0], -1)
    flat_out = out.reshape(out.shape[0], -1)
    for j in range(flat_sw.shape[1]):
        flat_out[:, j] = night_sine_profile(flat_sw[:, j], flat_amp[:, j], 0, 0)
    return out.astype(np.float32)


# ---------------------------------------------------------------------------
# per-tile sampling
# ---------------------------------------------------------------------------
def haversine_m(lat1, lon1, lat2, lon2) -> float:
    """Great-circle distance on a 6371 km sphere (preprocessor.py:1012-1030)."""
    p1, p2 = math.radians(lat1), math.radians(lat2)
    a = math.sin((p2 - p1) / 2) ** 2 + math.cos(p1) * math.cos(p2) * math.sin(math.radians(lon2 - lon1) / 2) ** 2
    return 2 * 6371000.0 * math.asin(math.sqrt(a))


def grid_cell_size_m(lon2d, lat2d, cx, cy, tree) -> tuple:
    """Median east-west / north-south spacing around the grid point nearest (cx, cy)
    (preprocessor.py:1033-1062)."""
    ny, nx = lat2d.shape
    _, idx = tree.query([cx, cy], k=1)
    i, j = np.unravel_index(idx, (ny, nx))

    def d(i2, j2):
        return haversine_m(lat2d[i, j], lon2d[i, j], lat2d[i2, j2], lon2d[i2, j2])

    ew = [d(i, jj) for jj in (j - 1, j + 1) if 0 <= jj < nx]
    ns = [d(ii, j) for ii in (i - 1, i + 1) if 0 <= ii < ny]
    both = ew + ns
    if not both:
        return 1e30, 1e30
    return (float(np.median(ew if ew else both)), float(np.median(ns if ns else both)))


@dataclass
class TileBox:
    """A tile's extent in degrees (the lon/lat bounding box of its corners)."""
    min_lon: float
    min_lat: float
    max_lon: float
    max_lat: float

    @property
    def center(self):
        return (self.min_lon + self.max_lon) / 2.0, (self.min_lat + self.max_lat) / 2.0

    def size_m(self):
        """(width, height) through the centre (preprocessor.py:1065-1078)."""
        cx, cy = self.center
        return (haversine_m(cy, self.min_lon, cy, self.max_lon), haversine_m(self.min_lat, cx, self.max_lat, cx))

    def ring(self) -> np.ndarray:
        """Closed counter-clockwise ring starting at (max_lon, min_lat)."""
        return np.array([(self.max_lon, self.min_lat), (self.max_lon, self.max_lat), (self.min_lon, self.max_lat),
                         (self.min_lon, self.min_lat), (self.max_lon, self.min_lat)])


def _utc_datetimes(times):
    out = []
    for t in times:
        if isinstance(t, _dt.datetime):
            out.append(t.replace(tzinfo=None) if t.tzinfo is None else
                       t.astimezone(_dt.timezone.utc).replace(tzinfo=None))
        else:
            out.append(np.datetime64(t, "s").astype(_dt.datetime))
    return out


def sample_tile(fields: dict, lat2d, lon2d, times_utc, box: TileBox, timezone: str, local_date,
                use_uhi: bool = True) -> np.ndarray:
    """The met rows (N x 25, float64) of one tile for one local calendar day.

    ``fields`` maps gridded variable names (``GRID_VARS`` values) to
    (time, ny, nx) arrays; ``times_utc`` are the records' UTC times;
    ``timezone`` an IANA name (the reference looks it up from the tile centre).
    Records between local 00:00 and 23:59 of ``local_date`` are used."""
    import pytz
    from matplotlib.path import Path
    from scipy.spatial import cKDTree

    lat2d = np.asarray(lat2d, dtype=float)
    lon2d = np.asarray(lon2d, dtype=float)
    ny, nx = lat2d.shape
    tree = cKDTree(np.column_stack([lon2d.ravel(), lat2d.ravel()]))
    tz = pytz.timezone(timezone)
    if isinstance(local_date, str):
        local_date = _dt.datetime.strptime(local_date, "%Y-%m-%d").date()
    start = tz.localize(_dt.datetime.combine(local_date, _dt.time(0, 0))).astimezone(pytz.utc)
    end = tz.localize(_dt.datetime.combine(local_date, _dt.time(23, 59))).astimezone(pytz.utc)
    utc = [pytz.utc.localize(t) for t in _utc_datetimes(times_utc)]
    steps = [k for k, t in enumerate(utc) if start <= t <= end]
    if not steps:
        raise ValueError(f"no forcing records fall on local date {local_date} ({timezone})")
    cx, cy = box.center
    tw, th = box.size_m()
    cw, ch = grid_cell_size_m(lon2d, lat2d, cx, cy, tree)
    use_nn = cw > tw and ch > th
    inside = None
    if not use_nn:
        inside = Path(box.ring()).contains_points(np.column_stack([lon2d.ravel(), lat2d.ravel()])).reshape(ny, nx)
        use_nn = not inside.any()
    _, nn = tree.query([cx, cy], k=1)
    ni, nj = np.unravel_index(nn, (ny, nx))

    def sample(name, t):
        a = np.asarray(fields[name][t], dtype=float)
        if np.ma.isMaskedArray(fields[name]):
            a = np.where(np.ma.getmaskarray(fields[name][t]), np.nan, a)
        if use_nn:
            return float(a[ni, nj])
        m = np.where(inside, a, np.nan)
        v = float(np.nanmean(m)) if np.any(~np.isnan(m)) else np.nan
        return v if np.isfinite(v) else float(a[ni, nj])

    rows = []
    for t in steps:
        lt = utc[t].astimezone(tz)
        row = {"iy": lt.year, "id": lt.timetuple().tm_yday, "it": lt.hour, "imin": lt.minute}
        for key in ("Q*", "QH", "QE", "Qs", "Qf", "snow", "ldown", "fcld", "wuh", "xsmd", "lai_hr", "Kdiff", "Kdir"):
            row[key] = MISSING
        row["rain"] = 0.0
        for key in ("Wind", "RH", "Td", "press", "Kdn"):
            name = GRID_VARS[key]
            v = sample(name, t) if name in fields else np.nan
            if key == "Td" and np.isfinite(v):
                v -= 273.15
            if key == "press" and np.isfinite(v):
                v /= 1000.0
            row[key] = v if np.isfinite(v) else MISSING
        wd = sample("WDIR", t) if "WDIR" in fields else np.nan
        row["Wd"] = wd % 360.0 if np.isfinite(wd) else MISSING
        uh = sample("UHI_CYCLE", t) if (use_uhi and "UHI_CYCLE" in fields) else 0.0
        row["uhii"] = uh if np.isfinite(uh) else 0.0
        rows.append([row[c] for c in COLUMNS])
    return np.array(rows, dtype=np.float64)
