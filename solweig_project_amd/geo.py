"""Map coordinates of a tile: the projected centre to WGS 84 longitude/latitude.

The reference's tile driver (``solweig_gpu/utci_process.py:1330-1365``) takes
the raster centre ``(gt[0] + gt[1] * width / 2, gt[3] + gt[5] * height / 2)``
and converts it from the raster's projection to WGS 84 with GDAL/OSR. GDAL is
not a dependency here, so the projections a SOLWEIG tile comes in are
evaluated directly:

* UTM on WGS 84 (EPSG 32601-32660 north, 32701-32760 south): the ellipsoidal
  transverse Mercator of Snyder, *Map Projections: A Working Manual* (USGS
  Professional Paper 1395, 1987), eqs. 3-21 and 8-9 to 8-25 (forward and
  inverse series, k0 = 0.9996, false easting 500 km, false northing 10 000 km
  south of the equator); the series are accurate to well under a millimetre
  within a UTM zone;
* geographic WGS 84 (EPSG 4326): the centre is already longitude/latitude.

Any other projection is refused (pass the location explicitly).
"""
from __future__ import annotations

import math
from dataclasses import dataclass


class UnsupportedProjection(ValueError):
    """The raster's projection is not one this module evaluates."""


@dataclass(frozen=True)
class Ellipsoid:
    a: float          # semi-major axis (m)
    e2: float         # first eccentricity squared


WGS84 = Ellipsoid(6378137.0, 0.00669437999014132)     # 1/f = 298.257223563
UTM_K0 = 0.9996
UTM_FALSE_EASTING = 500000.0
UTM_FALSE_NORTHING_SOUTH = 10000000.0


def _meridian_arc(phi: float, el: Ellipsoid) -> float:
    """Distance along the meridian from the equator to latitude ``phi`` (Snyder 3-21)."""
    e2 = el.e2
    e4, e6 = e2 * e2, e2 * e2 * e2
    return el.a * ((1 - e2 / 4 - 3 * e4 / 64 - 5 * e6 / 256) * phi
                   - (3 * e2 / 8 + 3 * e4 / 32 + 45 * e6 / 1024) * math.sin(2 * phi)
                   + (15 * e4 / 256 + 45 * e6 / 1024) * math.sin(4 * phi)
                   - (35 * e6 / 3072) * math.sin(6 * phi))


def tm_forward(lat_deg: float, lon_deg: float, lon0_deg: float, k0: float, el: Ellipsoid = WGS84):
    """Transverse Mercator (x, y) in metres of a point, origin at (0, lon0), no false offsets
    (Snyder 8-9 to 8-13)."""
    phi = math.radians(lat_deg)
    ep2 = el.e2 / (1 - el.e2)
    s, c = math.sin(phi), math.cos(phi)
    n = el.a / math.sqrt(1 - el.e2 * s * s)
    t = math.tan(phi) ** 2
    cc = ep2 * c * c
    a = math.radians(lon_deg - lon0_deg) * c
    m = _meridian_arc(phi, el)
    x = k0 * n * (a + (1 - t + cc) * a ** 3 / 6 + (5 - 18 * t + t * t + 72 * cc - 58 * ep2) * a ** 5 / 120)
    y = k0 * (m + n * math.tan(phi) * (a * a / 2 + (5 - t + 9 * cc + 4 * cc * cc) * a ** 4 / 24
                                       + (61 - 58 * t + t * t + 600 * cc - 330 * ep2) * a ** 6 / 720))
    return x, y


def tm_inverse(x: float, y: float, lon0_deg: float, k0: float, el: Ellipsoid = WGS84):
    """(latitude, longitude) in degrees of transverse Mercator (x, y), origin at (0, lon0), no false
    offsets (Snyder 8-18 to 8-25, footpoint latitude by the series 3-26)."""
    e2 = el.e2
    ep2 = e2 / (1 - e2)
    m = y / k0
    mu = m / (el.a * (1 - e2 / 4 - 3 * e2 * e2 / 64 - 5 * e2 ** 3 / 256))
    r = math.sqrt(1 - e2)
    e1 = (1 - r) / (1 + r)
    phi1 = (mu + (3 * e1 / 2 - 27 * e1 ** 3 / 32) * math.sin(2 * mu)
            + (21 * e1 ** 2 / 16 - 55 * e1 ** 4 / 32) * math.sin(4 * mu)
            + (151 * e1 ** 3 / 96) * math.sin(6 * mu)
            + (1097 * e1 ** 4 / 512) * math.sin(8 * mu))
    s, c = math.sin(phi1), math.cos(phi1)
    c1 = ep2 * c * c
    t1 = math.tan(phi1) ** 2
    w = 1 - e2 * s * s
    n1 = el.a / math.sqrt(w)
    r1 = el.a * (1 - e2) / w ** 1.5
    d = x / (n1 * k0)
    phi = phi1 - (n1 * math.tan(phi1) / r1) * (
        d * d / 2 - (5 + 3 * t1 + 10 * c1 - 4 * c1 * c1 - 9 * ep2) * d ** 4 / 24
        + (61 + 90 * t1 + 298 * c1 + 45 * t1 * t1 - 252 * ep2 - 3 * c1 * c1) * d ** 6 / 720)
    lam = (d - (1 + 2 * t1 + c1) * d ** 3 / 6
           + (5 - 2 * c1 + 28 * t1 - 3 * c1 * c1 + 8 * ep2 + 24 * t1 * t1) * d ** 5 / 120) / c
    return math.degrees(phi), lon0_deg + math.degrees(lam)


def utm_zone(epsg: int):
    """(zone, north) of a WGS 84 UTM EPSG code, or None."""
    if 32601 <= epsg <= 32660:
        return epsg - 32600, True
    if 32701 <= epsg <= 32760:
        return epsg - 32700, False
    return None


def to_lonlat(x: float, y: float, epsg: int | None):
    """WGS 84 (longitude, latitude) in degrees of map coordinates (x, y) in the projection ``epsg``."""
    if epsg == 4326:
        return float(x), float(y)
    z = utm_zone(int(epsg)) if epsg is not None else None
    if z is None:
        raise UnsupportedProjection(
            f"projection EPSG:{epsg} is not evaluated here (WGS 84 UTM zones and EPSG:4326 are); "
            "pass the tile's location explicitly")
    zone, north = z
    lon0 = -183.0 + 6.0 * zone
    yy = y if north else y - UTM_FALSE_NORTHING_SOUTH
    lat, lon = tm_inverse(x - UTM_FALSE_EASTING, yy, lon0, UTM_K0)
    return lon, lat


def tile_centre(geotransform, width: int, height: int):
    """Map coordinates of the raster centre (``utci_process.py:1360-1361``)."""
    gt = geotransform
    return gt[0] + gt[1] * width / 2.0, gt[3] + gt[5] * height / 2.0
