"""Solar position (NREL SPA, Reda & Andreas 2004) and the per-row sun table.

Restates the reference's ``sun_position`` (``solweig_gpu/sun_position.py:26-118``)
and ``Solweig_2015a_metdata_noload`` (``sun_position.py:905-1010``) in float64
numpy, with the same choices: delta_t = 0, refraction correction always
applied, angles folded into [0, 360) after each stage (also the
geocentric latitude), zenith 89..90 clamped to 89, sun evaluated half a
time step before each forcing row, and the daily maximum altitude found by
stepping 15 minutes forward from 10:00 until the altitude decreases.

The periodic-term tables are data (``data/spa_terms.json``) in the
reference's transcription, which differs from the published tables in a
few digits; using them keeps the sun table, and through it the march
tables, identical to the reference's.
"""
from __future__ import annotations

import datetime
import json
from functools import lru_cache
from pathlib import Path

import numpy as np

from .met import SunTable

_DATA = Path(__file__).resolve().parent / "data" / "spa_terms.json"


@lru_cache(maxsize=1)
def _terms():
    raw = json.loads(_DATA.read_text())["terms"]
    return {k: np.atleast_2d(np.array(v, dtype=np.float64)) for k, v in raw.items()}


def _wrap(v, hi=360.0):
    v = v - hi * np.floor(v / hi)
    if v < 0:
        v = v + hi
    return v


def _series(tab, jme):
    return np.sum(tab[:, 0] * np.cos(tab[:, 1] + (tab[:, 2] * jme)))


def _julian(t):
    if t["month"] in (1, 2):
        Y, Mo = t["year"] - 1, t["month"] + 12
    else:
        Y, Mo = t["year"], t["month"]
    ut = ((t["hour"] - t["UTC"]) / 24) + (t["min"] / (60 * 24)) + (t["sec"] / (60 * 60 * 24))
    D = t["day"] + ut
    if t["year"] < 1582 or (t["year"] == 1582 and (t["month"] < 10 or (t["month"] == 10 and t["day"] <= 4))):
        B = 0
    else:
        A = np.floor(Y / 100)
        B = 2 - A + np.floor(A / 4)
    jd = D + B + np.floor(365.25 * (Y + 4716)) + np.floor(30.6001 * (Mo + 1)) - 1524.5
    jde = jd + (0 / 86400)
    jc = (jd - 2451545) / 36525
    jce = (jde - 2451545) / 36525
    return jd, jc, jce, jce / 10


def sun_position(t: dict, lat: float, lon: float, elev: float):
    """(zenith deg, azimuth deg) of the sun for a local time dict (year..sec, UTC)."""
    T = _terms()
    jd, jc, jce, jme = _julian(t)
    d2r = np.pi / 180
    # 2. heliocentric longitude / latitude / radius
    L = [_series(T[f"L{k}_terms"], jme) for k in range(5)] + [T["L5_terms"][:, 0] * np.cos(T["L5_terms"][:, 1] + (T["L5_terms"][:, 2] * jme))]
    hl = (L[0] + (L[1] * jme) + (L[2] * np.power(jme, 2)) + (L[3] * np.power(jme, 3))
          + (L[4] * np.power(jme, 4)) + (L[5] * np.power(jme, 5))) / 1e8
    hl = _wrap(hl * 180 / np.pi)
    Bt = [_series(T["B0_terms"], jme), _series(T["B1_terms"], jme)]
    hb = _wrap(((Bt[0] + (Bt[1] * jme)) / 1e8) * 180 / np.pi)
    R = [_series(T[f"R{k}_terms"], jme) for k in range(4)] + [T["R4_terms"][:, 0] * np.cos(T["R4_terms"][:, 1] + (T["R4_terms"][:, 2] * jme))]
    radius = (R[0] + (R[1] * jme) + (R[2] * np.power(jme, 2)) + (R[3] * np.power(jme, 3))
              + (R[4] * np.power(jme, 4))) / 1e8
    # 3. geocentric
    glon = _wrap(hl + 180)
    glat = _wrap(-hb)
    # 4. nutation
    X = [(1 / 189474) * np.power(jce, 3) + -0.0019142 * np.power(jce, 2) + 445267.11148 * jce + 297.85036,
         -(1 / 300000) * np.power(jce, 3) + -0.0001603 * np.power(jce, 2) + 35999.05034 * jce + 357.52772,
         (1 / 56250) * np.power(jce, 3) + 0.0086972 * np.power(jce, 2) + 477198.867398 * jce + 134.96298,
         (1 / 327270) * np.power(jce, 3) + -0.0036825 * np.power(jce, 2) + 483202.017538 * jce + 93.27191,
         (1 / 450000) * np.power(jce, 3) + 0.0020708 * np.power(jce, 2) + -1934.136261 * jce + 125.04452]
    arg = T["Y_terms"].dot(np.transpose(np.array(X))) * (np.pi / 180)
    nt = T["nutation_terms"]
    dpsi = np.sum((nt[:, 0] + (nt[:, 1] * jce)) * np.sin(arg)) / 36000000
    deps = np.sum((nt[:, 2] + (nt[:, 3] * jce)) * np.cos(arg)) / 36000000
    # 5. true obliquity
    U = jme / 10
    p = [2.45, 5.79, 27.87, 7.12, -39.05, -249.67, -51.38, 1999.25, -1.55, -4680.93, 84381.448]
    eps0 = p[0] * np.power(U, 10) + p[1] * np.power(U, 9) + p[2] * np.power(U, 8) + p[3] * np.power(U, 7) + \
        p[4] * np.power(U, 6) + p[5] * np.power(U, 5) + p[6] * np.power(U, 4) + p[7] * np.power(U, 3) + \
        p[8] * np.power(U, 2) + p[9] * U + p[10]
    eps = (eps0 / 3600) + deps
    # 6-7. aberration, apparent longitude
    lam = glon + dpsi + (-20.4898 / (3600 * radius))
    # 8. apparent sidereal time at Greenwich
    nu0 = 280.46061837 + (360.98564736629 * (jd - 2451545)) + (0.000387933 * np.power(jc, 2)) - (np.power(jc, 3) / 38710000)
    nu = _wrap(nu0) + (dpsi * np.cos(eps * np.pi / 180))
    # 9-10. right ascension, declination
    num = (np.sin(lam * np.pi / 180) * np.cos(eps * np.pi / 180)) - (np.tan(glat * np.pi / 180) * np.sin(eps * np.pi / 180))
    alpha = _wrap(np.arctan2(num, np.cos(lam * np.pi / 180)) * 180 / np.pi)
    delta = np.arcsin((np.sin(glat * np.pi / 180) * np.cos(eps * np.pi / 180))
                      + (np.cos(glat * np.pi / 180) * np.sin(eps * np.pi / 180) * np.sin(lam * np.pi / 180))) * 180 / np.pi
    # 11. observer local hour angle
    H = _wrap(nu + lon - alpha)
    # 12. topocentric
    xi = 8.794 / (3600 * radius)
    u = np.arctan(0.99664719 * np.tan(lat * np.pi / 180))
    x = np.cos(u) + ((elev / 6378140) * np.cos(lat * np.pi / 180))
    y = (0.99664719 * np.sin(u)) + ((elev / 6378140) * np.sin(lat * np.pi / 180))
    dalpha = np.arctan2(-x * np.sin(xi * np.pi / 180) * np.sin(H * np.pi / 180),
                        np.cos(delta * np.pi / 180) - (x * np.sin(xi * np.pi / 180) * np.cos(H * np.pi / 180)))
    num = (np.sin(delta * np.pi / 180) - (y * np.sin(xi * np.pi / 180))) * np.cos(dalpha)
    den = np.cos(delta * np.pi / 180) - (y * np.sin(xi * np.pi / 180)) * np.cos(H * np.pi / 180)
    delta_t = np.arctan2(num, den) * 180 / np.pi
    Ht = H - dalpha * 180 / np.pi
    # 14. zenith (with refraction) and azimuth
    e0 = np.arcsin((np.sin(lat * np.pi / 180) * np.sin(delta_t * np.pi / 180))
                   + (np.cos(lat * np.pi / 180) * np.cos(delta_t * np.pi / 180) * np.cos(Ht * np.pi / 180))) * 180 / np.pi
    refr = 1.02 / (60 * np.tan((e0 + (10.3 / (e0 + 5.11))) * np.pi / 180))
    zenith = 90 - (e0 + refr)
    az = (np.arctan2(np.sin(Ht * np.pi / 180),
                     (np.cos(Ht * np.pi / 180) * np.sin(lat * np.pi / 180)) - (np.tan(delta_t * np.pi / 180) * np.cos(lat * np.pi / 180)))
          * 180 / np.pi) + 180
    return float(np.asarray(zenith).squeeze()), float(np.asarray(_wrap(az)).squeeze())


def _tdict(dt: datetime.datetime, utc: float):
    return {"year": dt.year, "month": dt.month, "day": dt.day, "hour": dt.hour, "min": dt.minute, "sec": 0, "UTC": utc}


def sun_table(met: np.ndarray, lat: float, lon: float, utc: float, elev: float = 0.0) -> SunTable:
    """Per-row sun geometry of a UMEP forcing table (``Solweig_2015a_metdata_noload``)."""
    met = np.asarray(met, dtype=np.float64)
    n = met.shape[0]
    dectime = met[:, 1] + met[:, 2] / 24 + met[:, 3] / (60 * 24.)
    half = 0 if n == 1 else (dectime[1] - dectime[0]) / 2.
    alt = np.empty(n)
    azi = np.empty(n)
    zen = np.empty(n)
    jday = np.empty(n)
    altmax = np.empty(n)
    sunmaximum = 0.0
    for i in range(n):
        YMD = datetime.datetime(int(met[i, 0]), 1, 1) + datetime.timedelta(int(met[i, 1]) - 1)
        if i == 0 or np.mod(dectime[i], np.floor(dectime[i])) == 0:
            fifteen = 0.
            sunmaximum = -90.
            zmax = 90.
            while sunmaximum <= 90. - zmax:
                sunmaximum = 90. - zmax
                fifteen = fifteen + 15. / 1440.
                t = YMD + datetime.timedelta(days=(60 * 10) / 1440.0 + fifteen)
                zmax, _ = sun_position(_tdict(t, utc), lat, lon, elev)
        altmax[i] = sunmaximum
        t = YMD + datetime.timedelta(hours=met[i, 2]) + datetime.timedelta(minutes=met[i, 3]) \
            - datetime.timedelta(days=half)
        z, a = sun_position(_tdict(t, utc), lat, lon, elev)
        if 89.0 < z <= 90.0:
            z = 89.0
        alt[i] = 90. - z
        zen[i] = z * (np.pi / 180.)
        azi[i] = a
        jday[i] = YMD.timetuple().tm_yday
    return SunTable(altitude=alt, azimuth=azi, zen=zen, jday=jday, dectime=dectime, altmax=altmax)


def day_sun_table(lat: float, lon: float, utc: float, year: int, doy: int, hours=None, elev: float = 0.0) -> SunTable:
    """Sun table of an hourly day (hours 0..23 by default)."""
    hours = np.arange(24) if hours is None else np.asarray(hours)
    met = np.zeros((hours.size, 4))
    met[:, 0] = year
    met[:, 1] = doy
    met[:, 2] = hours
    return sun_table(met, lat, lon, utc, elev)
