"""Fixture: the reference's recorded default-path outputs of ``run_utci_window`` with their inputs.

The reference holds ``tests/fixtures/params_physics_prechange_default.npz``:
Tmrt, UTCI, Kdown, Kup, Ldown, Lup and shadow (4, 64, 64) of
``run_utci_window`` (``utci_process.py:477-1090``) on a prepared site, written
by its torch-CPU code. The file holds outputs only; the site is built by the
reference's tests from a recipe, restated here as data:

* 64 x 64 rasters at 2 m, origin (1000 m, 2000 m) of EPSG:32616
  (``test_incremental_params_physics.py:90-93``);
* flat DEM; building DSM 12 m over rows 5:12 x cols 3:12 and 4 m over rows
  39:48 x cols 38:48; land cover 1 with blocks of 5, 7 and 6; walls and aspect
  zero (``test_incremental_worker.py:187-203, 206-243``);
* one tree of 8 m with a 3 m crown at (1037 m, 1907 m): the cells whose
  centres lie within the crown, rows 45:48 x cols 17:20
  (``test_incremental_params_physics.py:100-106``,
  ``solweig_gpu/incremental/trees.py:103-153``);
* met rows for 10:00-13:00 of 2024 DOY 172 (``test_incremental_worker.py:131-149``);
* location: the raster centre (1064 m, 1936 m) of UTM zone 16N taken to
  WGS84 as the reference does with GDAL (``incremental/solver.py:221-289``),
  here by the Krueger series of the transverse Mercator projection to sixth
  order in n (Karney 2011, J. Geodesy 85:475), checked by its own forward
  series; altitude 0 (median DEM);
* UTC offset -6 h: the point (0.0175 N, 91.49 W) lies beside Isabela Island,
  Galapagos; the time-zone lookup the reference uses returns
  Pacific/Galapagos or the ocean zone Etc/GMT+6 there, both UTC-6 with no
  daylight saving (``incremental/solver.py:291-313``).

Writes ``params_physics_site.npz`` beside this script. Run in a checkout that
has the reference at ``/root/reference``.
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/tests/fixtures/params_physics_prechange_default.npz")
OUT = Path(__file__).resolve().parent / "params_physics_site.npz"
ROWS = COLS = 64
PIXEL = 2.0
ORIGIN = (1000.0, 2000.0)
HOURS = (10, 11, 12, 13)

# WGS84 and UTM zone 16N
A_WGS84 = 6378137.0
F_WGS84 = 1.0 / 298.257223563
K0 = 0.9996
LON0 = -87.0
FALSE_EASTING = 500000.0


def _series():
    n = F_WGS84 / (2.0 - F_WGS84)
    big_a = A_WGS84 / (1.0 + n) * (1.0 + n ** 2 / 4.0 + n ** 4 / 64.0 + n ** 6 / 256.0)
    alpha = [n / 2 - 2 * n ** 2 / 3 + 5 * n ** 3 / 16 + 41 * n ** 4 / 180 - 127 * n ** 5 / 288 + 7891 * n ** 6 / 37800,
             13 * n ** 2 / 48 - 3 * n ** 3 / 5 + 557 * n ** 4 / 1440 + 281 * n ** 5 / 630 - 1983433 * n ** 6 / 1935360,
             61 * n ** 3 / 240 - 103 * n ** 4 / 140 + 15061 * n ** 5 / 26880 + 167603 * n ** 6 / 181440,
             49561 * n ** 4 / 161280 - 179 * n ** 5 / 168 + 6601661 * n ** 6 / 7257600,
             34729 * n ** 5 / 80640 - 3418889 * n ** 6 / 1995840,
             212378941 * n ** 6 / 319334400]
    beta = [n / 2 - 2 * n ** 2 / 3 + 37 * n ** 3 / 96 - n ** 4 / 360 - 81 * n ** 5 / 512 + 96199 * n ** 6 / 604800,
            n ** 2 / 48 + n ** 3 / 15 - 437 * n ** 4 / 1440 + 46 * n ** 5 / 105 - 1118711 * n ** 6 / 3870720,
            17 * n ** 3 / 480 - 37 * n ** 4 / 840 - 209 * n ** 5 / 4480 + 5569 * n ** 6 / 90720,
            4397 * n ** 4 / 161280 - 11 * n ** 5 / 504 - 830251 * n ** 6 / 7257600,
            4583 * n ** 5 / 161280 - 108847 * n ** 6 / 3991680,
            20648693 * n ** 6 / 638668800]
    delta = [2 * n - 2 * n ** 2 / 3 - 2 * n ** 3 + 116 * n ** 4 / 45 + 26 * n ** 5 / 45 - 2854 * n ** 6 / 675,
             7 * n ** 2 / 3 - 8 * n ** 3 / 5 - 227 * n ** 4 / 45 + 2704 * n ** 5 / 315 + 2323 * n ** 6 / 945,
             56 * n ** 3 / 15 - 136 * n ** 4 / 35 - 1262 * n ** 5 / 105 + 73814 * n ** 6 / 2835,
             4279 * n ** 4 / 630 - 332 * n ** 5 / 35 - 399572 * n ** 6 / 14175,
             4174 * n ** 5 / 315 - 144838 * n ** 6 / 6237,
             601676 * n ** 6 / 22275]
    return big_a, alpha, beta, delta


def utm_inverse(easting: float, northing: float) -> tuple[float, float]:
    """(lon, lat) in degrees of a UTM 16N point (Krueger series)."""
    big_a, _, beta, delta = _series()
    xi = northing / (K0 * big_a)
    eta = (easting - FALSE_EASTING) / (K0 * big_a)
    xi1 = xi - sum(b * math.sin(2 * j * xi) * math.cosh(2 * j * eta) for j, b in enumerate(beta, 1))
    eta1 = eta - sum(b * math.cos(2 * j * xi) * math.sinh(2 * j * eta) for j, b in enumerate(beta, 1))
    chi = math.asin(math.sin(xi1) / math.cosh(eta1))
    lat = chi + sum(d * math.sin(2 * j * chi) for j, d in enumerate(delta, 1))
    lon = math.atan2(math.sinh(eta1), math.cos(xi1))
    return LON0 + math.degrees(lon), math.degrees(lat)


def utm_forward(lon: float, lat: float) -> tuple[float, float]:
    """(easting, northing) of a point in UTM 16N (Krueger series, exact conformal latitude)."""
    big_a, alpha, _, _ = _series()
    e = math.sqrt(F_WGS84 * (2.0 - F_WGS84))
    phi = math.radians(lat)
    lam = math.radians(lon - LON0)
    t = math.sinh(math.atanh(math.sin(phi)) - e * math.atanh(e * math.sin(phi)))
    xi1 = math.atan2(t, math.cos(lam))
    eta1 = math.asinh(math.sin(lam) / math.hypot(t, math.cos(lam)))
    xi = xi1 + sum(a * math.sin(2 * j * xi1) * math.cosh(2 * j * eta1) for j, a in enumerate(alpha, 1))
    eta = eta1 + sum(a * math.cos(2 * j * xi1) * math.sinh(2 * j * eta1) for j, a in enumerate(alpha, 1))
    return FALSE_EASTING + K0 * big_a * eta, K0 * big_a * xi


def met_row(hour: int) -> list[float]:
    daytime = 6 <= hour <= 19
    ramp = math.sin(math.pi * (hour - 6) / 13.0) if daytime else -0.3
    radg = 850.0 * max(0.0, ramp) if daytime else 0.0
    ta = 22.0 + 8.0 * ramp
    rh = 60.0 - 20.0 * ramp
    row = [2024.0, 172.0, float(hour), 0.0] + [0.0] * 5 + [2.0, rh, ta, 1013.0, 0.0, radg] + [0.0] * 6 \
        + [150.0, 600.0, 270.0, 0.0]
    assert len(row) == 25
    return row


def site():
    dem = np.zeros((ROWS, COLS), np.float32)
    dsm = dem.copy()
    dsm[int(ROWS * 0.08):int(ROWS * 0.20), int(COLS * 0.06):int(COLS * 0.20)] = 12.0
    dsm[int(ROWS * 0.62):int(ROWS * 0.75), int(COLS * 0.60):int(COLS * 0.75)] = 4.0
    lc = np.ones((ROWS, COLS), np.float32)
    lc[int(ROWS * 0.3):int(ROWS * 0.4), int(COLS * 0.3):int(COLS * 0.4)] = 5
    lc[int(ROWS * 0.85):int(ROWS * 0.92), int(COLS * 0.05):int(COLS * 0.15)] = 7
    lc[int(ROWS * 0.85):int(ROWS * 0.92), int(COLS * 0.8):int(COLS * 0.9)] = 6
    # tree: cells whose centres lie within the crown disc
    tx, ty, height, radius = ORIGIN[0] + 18.5 * PIXEL, ORIGIN[1] - 46.5 * PIXEL, 8.0, 3.0
    cx = ORIGIN[0] + (np.arange(COLS) + 0.5) * PIXEL
    cy = ORIGIN[1] - (np.arange(ROWS) + 0.5) * PIXEL
    inside = (cx[None, :] - tx) ** 2 + (cy[:, None] - ty) ** 2 <= radius ** 2
    trees = np.where(inside, np.float32(height), np.float32(0.0)).astype(np.float32)
    return dem, dsm, trees, lc


def main():
    e0, n0 = ORIGIN[0] + PIXEL * COLS / 2.0, ORIGIN[1] - PIXEL * ROWS / 2.0
    lon, lat = utm_inverse(e0, n0)
    e1, n1 = utm_forward(lon, lat)
    assert abs(e1 - e0) < 1e-6 and abs(n1 - n0) < 1e-6, (e1 - e0, n1 - n0)
    assert utm_inverse(FALSE_EASTING, 0.0) == (LON0, 0.0)
    dem, dsm, trees, lc = site()
    out = {"dem": dem, "dsm": dsm, "trees": trees, "landcover": lc,
           "walls": np.zeros_like(dem), "aspect": np.zeros_like(dem),
           "met": np.array([met_row(h) for h in HOURS], np.float64),
           "lon": np.float64(lon), "lat": np.float64(lat), "utc": np.float64(-6.0), "elev": np.float64(0.0),
           "pixel": np.float64(PIXEL)}
    with np.load(SRC) as z:
        for k in z.files:
            out[f"out_{k}"] = z[k]
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size / 1e3:.0f} kB); lon {lon:.9f} lat {lat:.9f}; "
          f"trees {int((trees > 0).sum())} cells")


if __name__ == "__main__":
    main()
