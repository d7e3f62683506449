"""Per-timestep host scalar chain of the SOLWEIG day loop.

One call per time step turns the forcing row, the sun position and the
carried day-cycle scalars (CI, firstdaytime, timeadd, Twater) into a flat
bundle of float32 constants that the per-pixel kernels consume: clearness
index and Reindl diffuse split, Perez sky luminance per patch, sky
emissivity and patch longwave tables, ground/wall temperature amplitudes
per land-cover class, thermal-delay weights, and the sun march table.

The reference evaluates this chain with 0-dim torch tensors whose dtypes
mix float32 (``torch.tensor(python_float)``) and float64 (values coming
from numpy forcing arrays); the rounding points follow from that. This
module evaluates the same expressions with the same tensor dtypes, and
hands the kernels the float32 value the reference would broadcast into
each per-pixel expression.

References (``solweig_gpu/``):
  * driver scalars ``utci_process.py:796-866, 917-934``;
  * ``daylen`` ``solweig.py:47-73``; vapour pressure / Prata sky
    emissivity ``solweig.py:2041-2046``;
  * ``clearnessindex_2013b`` ``solweig.py:851-926``; ``diffusefraction``
    ``solweig.py:928-974``; ``Perez_v3`` ``solweig.py:1227-1361``;
  * ground temperature wave ``solweig.py:2101-2122``;
  * ``TsWaveDelay_2015a`` ``solweig.py:451-488``;
  * ``Kside_veg_v2022a`` scalars ``solweig.py:598-724``;
  * ``Lcyl_v2022a``/``model2``/``define_patch_characteristics`` scalars
    ``solweig.py:1394-1414, 1505-1625, 1651-1760``; CI<0.95 sky blend
    ``solweig.py:2247-2252``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import skyvault

_CPU = torch.device("cpu")
F32 = np.float32

# Model constants (utci_process.py:49-69, 774-795).
ALBEDO_B = 0.2
ALBEDO_G = 0.15
EWALL = 0.9
EGROUND = 0.95
ABSK = 0.7
ABSL = 0.95
FSIDE = 0.22
FUP = 0.06
FCYL = 0.28
TRANS_VEG = 3.0 / 100.0
SBC_PY = 5.67051e-8

#: SOLWEIG land-cover class table (landcoverclasses_2016a.txt):
#: code -> (albedo, emissivity, TgK, Tstart, TmaxLST); code 99 = walls.
LC_CLASSES = {
    2: (0.18, 0.95, 0.58, -9.78, 15.0),
    1: (0.18, 0.95, 0.58, -9.78, 15.0),
    0: (0.20, 0.95, 0.37, -3.41, 15.0),
    7: (0.05, 0.98, 0.00, 0.00, 12.0),
    5: (0.16, 0.94, 0.21, -3.38, 14.0),
    6: (0.25, 0.94, 0.33, -3.01, 14.0),
    99: (0.20, 0.90, 0.58, -3.41, 15.0),
}
N_CLASS = 8  # class index space 0..7 used on the device


def _f(x) -> float:
    return float(x)


@dataclass
class LandCover:
    """Per-class surface constants (index = cleaned land-cover code)."""

    present: bool
    alb: np.ndarray      # (8,) f32
    emis: np.ndarray     # (8,) f32
    tgk: np.ndarray      # (8,) f32
    tstart: np.ndarray   # (8,) f32
    tmaxlst: np.ndarray  # (8,) f32
    tgk_wall: float
    tstart_wall: float
    tmaxlst_wall: float
    classes: frozenset = frozenset(range(8))   # class codes the table defines

    def key(self):
        """Hashable identity of the constants (for caching derived tables)."""
        return (self.present, self.alb.tobytes(), self.emis.tobytes(), self.tgk.tobytes(), self.tstart.tobytes(),
                self.tmaxlst.tobytes(), self.tgk_wall, self.tstart_wall, self.tmaxlst_wall)


def clean_landcover(lc: np.ndarray | None) -> np.ndarray | None:
    """Class cleanup of the driver (``utci_process.py:738-744``)."""
    if lc is None:
        return None
    g = np.array(lc, dtype=np.float32, copy=True)
    g[(g < 1) | (g > 7)] = 6
    g[(g == 3) | (g == 4)] = 5
    return g


def landcover_constants(present: bool, table=None) -> LandCover:
    """Per-class constants; ``table`` rows are [class, albedo, emissivity, TgK,
    Tstart, TmaxLST] with a wall row 99 (the land-cover class file read by
    ``utci_process.py:1318-1322`` and applied by ``Tgmaps_v1.py:15-53``);
    None selects the default SOLWEIG table."""
    alb = np.zeros(N_CLASS, F32)
    emis = np.zeros(N_CLASS, F32)
    tgk = np.zeros(N_CLASS, F32)
    tst = np.zeros(N_CLASS, F32)
    tmx = np.zeros(N_CLASS, F32)
    if present:
        classes = LC_CLASSES
        if table is not None:
            t = np.asarray(table, dtype=np.float64)
            if t.ndim != 2 or t.shape[1] != 6:
                raise ValueError("land-cover class table must have 6 columns")
            classes = {int(r[0]): tuple(float(x) for x in r[1:]) for r in t}
            if 99 not in classes:
                raise ValueError("land-cover class table has no wall row (class 99)")
        for code in range(N_CLASS):
            row = classes.get(code)
            if row is None:
                continue
            alb[code], emis[code], tgk[code], tst[code], tmx[code] = row
        w = classes[99]
        lc = LandCover(True, alb, emis, tgk, tst, tmx, w[2], w[3], w[4])
        lc.classes = frozenset(c for c in classes if 0 <= c < N_CLASS)
        return lc
    # no land cover: uniform defaults (utci_process.py:757-766)
    alb[:] = F32(0.0) + F32(ALBEDO_G)
    emis[:] = F32(0.0) + F32(EGROUND)
    tgk[:] = F32(0.0) + F32(0.37)
    tst[:] = F32(0.0) - F32(3.41)
    tmx[:] = 15.0
    return LandCover(False, alb, emis, tgk, tst, tmx, 0.37, -3.41, 15.0)


@dataclass
class SunTable:
    """Per-row solar geometry (``Solweig_2015a_metdata_noload`` outputs)."""

    altitude: np.ndarray  # deg, f64
    azimuth: np.ndarray   # deg, f64
    zen: np.ndarray       # rad, f64
    jday: np.ndarray      # f64
    dectime: np.ndarray   # f64
    altmax: np.ndarray    # f64


@dataclass
class DayState:
    """Carried day-cycle scalars (``utci_process.py:404-411``)."""

    CI: object = 1.0
    firstdaytime: float = 1.0
    timeadd: object = 0.0
    Twater: object = None


@dataclass
class StepParams:
    """Everything one time step's kernels need, as float32 constants."""

    i: int
    is_day: bool
    # forcing at the pixel
    ta: float = 0.0            # f32(Ta)
    ta_mat: float = 0.0        # f32(f32(Ta) + uhii): UTCI air temperature
    rh: float = 0.0            # f32(RH)
    va: float = 0.0            # clamp(f32(1)*f32(Ws), 0.15)
    tak4: float = 0.0          # f32((Ta + 273.15)^4 in float64)
    # shortwave
    radI: float = 0.0
    radD: float = 0.0
    radG: float = 0.0          # f32(radG)
    sinalt: float = 0.0        # f32(sin(alt*pi/180)) in float64
    cosalt_k: float = 0.0      # cos(f32(alt) * f32(pi/180)) (Kside)
    psi: float = 0.0           # f32 leaf transmissivity
    one_m_psi: float = 0.0     # f32(1 - psi)
    tan_zen: float = 0.0       # tan(f32(zen)) for the cylinder wedge
    # ground / wall temperature
    tg_class: np.ndarray = field(default_factory=lambda: np.zeros(N_CLASS, F32))
    tgwall: float = 0.0
    lwall: float = 0.0
    # thermal delay
    tw_first: bool = False     # firstdaytime == 1 -> Tgmap1 := Tgmap0
    tw_update: bool = False    # timeadd >= 59/1440 -> state blended
    tw_w: float = 0.0          # f32(weight1)
    tw_omw: float = 0.0        # f32(1 - weight1)
    # WBGT
    wbgt_w: float = 0.0        # f32(0.7 * wet-bulb temperature in degC, float64)
    wbgt_a: float = 0.0        # black-globe h_cg / (emissivity * sigma), float32
    # sky patches (153)
    lv: np.ndarray = field(default_factory=lambda: np.zeros(153, F32))
    lumchi: np.ndarray = field(default_factory=lambda: np.zeros(153, F32))
    yi_k: np.ndarray = field(default_factory=lambda: np.zeros(153, F32))
    yi_l: np.ndarray = field(default_factory=lambda: np.zeros(153, F32))
    lsky_down: np.ndarray = field(default_factory=lambda: np.zeros(153, F32))
    lsky_side: np.ndarray = field(default_factory=lambda: np.zeros(153, F32))
    l_sun_side: np.ndarray = field(default_factory=lambda: np.zeros(153, np.int32))
    sunlit_k: float = 0.0      # Kside sunlit surface radiance
    shaded_k: float = 0.0      # Kside shaded surface radiance
    veg_l: float = 0.0         # vegetation surface longwave (f64 in reference)
    sunlit_l: float = 0.0      # sunlit wall longwave
    shaded_l: float = 0.0      # shaded wall longwave (f64 in reference)
    # sun march
    table: skyvault.StepTable | None = None
    facesh_branch: int = 0
    azilow: float = 0.0
    azihigh: float = 0.0
    geom: object = None
    # scalar diagnostics (for checks against recorded reference values)
    diag: dict = field(default_factory=dict)


def _t(x):
    return torch.tensor(x, device=_CPU)


_CONST = {}


def _c(x):
    """``_t(x)`` for a constant: one tensor per value, created on first use (the chain only reads
    these, so sharing them changes no result and saves a tensor construction per use)."""
    key = (type(x), x)
    t = _CONST.get(key)
    if t is None:
        t = _CONST[key] = torch.tensor(x, device=_CPU)
    return t


def daylen_snup(jday: float, lat: float):
    DOY = _t(jday)
    XLAT = _t(lat)
    RAD = torch.pi / 180.0
    DEC = -23.45 * torch.cos(2.0 * torch.pi * (DOY + 10.0) / 365.0)
    SOC = torch.tan(RAD * DEC) * torch.tan(RAD * XLAT)
    SOC = torch.clamp(SOC, -1.0, 1.0)
    DAYL = 12.0 + 24.0 * torch.arcsin(SOC) / torch.pi
    return 12.0 - DAYL / 2.0


def clearness_index(zen, jday, Ta, RH, radG, latitude: float, P):
    """(I0, CI, Kt) of ``clearnessindex_2013b`` (float32 tensors)."""
    if P == -999.0:
        p = 1013.0
    else:
        p = P * 10.0
    Itoa = 1370.0
    b = 2.0 * torch.pi * jday / 365.0
    D = torch.sqrt(1.00011 + 0.034221 * torch.cos(b) + 0.001280 * torch.sin(b)
                   + 0.000719 * torch.cos(2.0 * b) + 0.000077 * torch.sin(2.0 * b))
    m = 35.0 * torch.cos(zen) * ((1224.0 * (torch.cos(zen) ** 2) + 1.0) ** (-1 / 2.0))
    Trpg = 1.021 - 0.084 * (m * (0.000949 * p + 0.051)) ** 0.5
    if latitude < 10.0:
        G = [3.37, 2.85, 2.80, 2.64]
    elif latitude < 20.0:
        G = [2.99, 3.02, 2.70, 2.93]
    elif latitude < 30.0:
        G = [3.60, 3.00, 2.98, 2.93]
    elif latitude < 40.0:
        G = [3.04, 3.11, 2.92, 2.94]
    elif latitude < 50.0:
        G = [2.70, 2.95, 2.77, 2.71]
    elif latitude < 60.0:
        G = [2.52, 3.07, 2.67, 2.93]
    elif latitude < 70.0:
        G = [1.76, 2.69, 2.61, 2.61]
    elif latitude < 80.0:
        G = [1.60, 1.67, 2.24, 2.63]
    else:
        G = [1.11, 1.44, 1.94, 2.02]
    if jday > 335 or jday <= 60:
        G = G[0]
    elif 60 < jday <= 152:
        G = G[1]
    elif 152 < jday <= 244:
        G = G[2]
    else:
        G = G[3]
    a2 = _c(17.27)
    b2 = _c(237.7)
    Td = (b2 * (((a2 * Ta) / (b2 + Ta)) + torch.log(RH))) / (a2 - (((a2 * Ta) / (b2 + Ta)) + torch.log(RH)))
    Td = (Td * 1.8) + 32
    u = torch.exp(0.1133 - torch.log(_c(G + 1.0)) + 0.0393 * Td)
    Tw = 1 - 0.077 * ((u * m) ** 0.3)
    Tar = 0.935 ** m
    I0 = Itoa * torch.cos(zen) * Trpg * Tw * D * Tar
    I0 = torch.where(torch.abs(zen) > torch.pi / 2, _c(0.0), I0)
    I0 = torch.where(torch.isnan(I0), _c(0.0), I0)
    corr = 0.1473 * torch.log(90 - (zen / torch.pi * 180)) + 0.3454
    CI = radG / I0 + (1 - corr)
    I0et = Itoa * torch.cos(zen) * D
    Kt = radG / I0et
    return I0, CI, Kt


def diffuse_fraction(radG, altitude, Kt, Ta, RH):
    """(radI, radD) of ``diffusefraction`` (Reindl et al. 1990)."""
    Ta = Ta if isinstance(Ta, torch.Tensor) else _t(Ta)
    RH = RH if isinstance(RH, torch.Tensor) else _t(RH)
    alfa = altitude * (torch.pi / 180.0)
    alfa = alfa if isinstance(alfa, torch.Tensor) else _t(alfa)
    if Ta <= -999.00 or RH <= -999.00 or torch.isnan(Ta) or torch.isnan(RH):
        if Kt <= 0.3:
            radD = radG * (1.020 - 0.248 * Kt)
        elif 0.3 < Kt < 0.78:
            radD = radG * (1.45 - 1.67 * Kt)
        else:
            radD = radG * 0.147
    else:
        RH = RH / 100
        if Kt <= 0.3:
            radD = radG * (1 - 0.232 * Kt + 0.0239 * torch.sin(alfa) - 0.000682 * Ta + 0.0195 * RH)
        elif 0.3 < Kt < 0.78:
            radD = radG * (1.329 - 1.716 * Kt + 0.267 * torch.sin(alfa) - 0.00357 * Ta + 0.106 * RH)
        else:
            radD = radG * (0.426 * Kt - 0.256 * torch.sin(alfa) + 0.00349 * Ta + 0.0734 * RH)
    radI = (radG - radD) / torch.sin(alfa)
    radI = torch.where(radI < 0, _c(0.0), radI)
    radI = torch.where((altitude < 1) & (radI > radG), radG, radI) if isinstance(altitude, torch.Tensor) \
        else (radG if (altitude < 1 and bool(radI > radG)) else radI)
    radD = torch.where(radD > radG, radG, radD)
    return radI, radD


_PEREZ = np.array([
    # a1..a4, b1..b4, c1..c4, d1..d4, e1..e4 for the 8 clearness bins
    [1.3525, -0.2576, -0.2690, -1.4366, -0.7670, 0.0007, 1.2734, -0.1233, 2.8000, 0.6004, 1.2375, 1.0000, 1.8734, 0.6297, 0.9738, 0.2809, 0.0356, -0.1246, -0.5718, 0.9938],
    [-1.2219, -0.7730, 1.4148, 1.1016, -0.2054, 0.0367, -3.9128, 0.9156, 6.9750, 0.1774, 6.4477, -0.1239, -1.5798, -0.5081, -1.7812, 0.1080, 0.2624, 0.0672, -0.2190, -0.4285],
    [-1.1000, -0.2515, 0.8952, 0.0156, 0.2782, -0.1812, -4.5000, 1.1766, 24.7219, -13.0812, -37.7000, 34.8438, -5.0000, 1.5218, 3.9229, -2.6204, -0.0156, 0.1597, 0.4199, -0.5562],
    [-0.5484, -0.6654, -0.2672, 0.7117, 0.7234, -0.6219, -5.6812, 2.6297, 33.3389, -18.3000, -62.2500, 52.0781, -3.5000, 0.0016, 1.1477, 0.1062, 0.4659, -0.3296, -0.0876, -0.0329],
    [-0.6000, -0.3566, -2.5000, 2.3250, 0.2937, 0.0496, -5.6812, 1.8415, 21.0000, -4.7656, -21.5906, 7.2492, -3.5000, -0.1554, 1.4062, 0.3988, 0.0032, 0.0766, -0.0656, -0.1294],
    [-1.0156, -0.3670, 1.0078, 1.4051, 0.2875, -0.5328, -3.8500, 3.3750, 14.0000, -0.9999, -7.1406, 7.5469, -3.4000, -0.1078, -1.0750, 1.5702, -0.0672, 0.4016, 0.3017, -0.4844],
    [-1.0000, 0.0211, 0.5025, -0.5119, -0.3000, 0.1922, 0.7023, -1.6317, 19.0000, -5.0000, 1.2438, -1.9094, -4.0000, 0.0250, 0.3844, 0.2656, 1.0468, -0.3788, -2.4517, 1.4656],
    [-1.0500, 0.0289, 0.4260, 0.3590, -0.3250, 0.1156, 0.7781, 0.0025, 31.0625, -14.5000, -46.1148, 55.3750, -7.2312, 0.4050, 13.3500, 0.6234, 1.5000, -0.6426, 1.8564, 0.5636],
], dtype=np.float64)


_PEREZ_CONSTS = None


def _perez_consts():
    """Coefficient columns and the step-invariant patch trigonometry of ``Perez_v3``."""
    global _PEREZ_CONSTS
    if _PEREZ_CONSTS is None:
        coef = torch.tensor(_PEREZ.astype(np.float32))
        cols = tuple(coef[:, 4 * k: 4 * k + 4] for k in range(5))
        deg2rad = torch.tensor(np.pi / 180, device=_CPU)
        skyalt, skyazi = skyvault.create_patches()
        skyzen = (90 - skyalt) * deg2rad
        skyalt = skyalt * deg2rad
        sky = dict(cos_zen=torch.cos(skyzen), sin_alt=torch.sin(skyalt), cos_alt=torch.cos(skyalt),
                   azi=skyazi * deg2rad)
        _PEREZ_CONSTS = cols + (sky,)
    return _PEREZ_CONSTS


def perez_luminance(zen_deg: float, azimuth_deg: float, radD, radI, jday: float):
    """Relative Perez luminance of the 153 patches (``Perez_v3``, patchchoice 1).

    Coefficient table: Perez et al. (1993), Solar Energy 50(3) 235-245.
    Returns a (153,) float32 tensor normalised to sum 1.
    """
    acoeff, bcoeff, ccoeff, dcoeff, ecoeff, sky = _perez_consts()
    deg2rad = _c(np.pi / 180)
    altitude = 90 - zen_deg
    zen = torch.tensor(zen_deg) * deg2rad
    azimuth = torch.tensor(azimuth_deg) * deg2rad
    altitude = torch.tensor(altitude) * deg2rad
    Idh = radD
    Ibn = radI
    PerezClearness = ((Idh + Ibn) / (Idh + 1.041 * torch.pow(zen, 3))) / (1 + 1.041 * torch.pow(zen, 3))
    day_angle = jday * 2 * torch.pi / 365
    I0 = 1367 * (1.00011 + 0.034221 * torch.cos(_t(day_angle)) + 0.00128 * torch.sin(_t(day_angle))
                 + 0.000719 * torch.cos(2 * _t(day_angle)) + 0.000077 * torch.sin(2 * _t(day_angle)))
    if altitude >= 10 * deg2rad:
        AirMass = 1 / torch.sin(altitude)
    elif altitude < 0:
        AirMass = 1 / torch.sin(altitude) + 0.50572 * torch.pow(
            180 * torch.complex(altitude, _t(0.0)) / torch.pi + 6.07995, -1.6364)
    else:
        AirMass = 1 / torch.sin(altitude) + 0.50572 * torch.pow(180 * altitude / torch.pi + 6.07995, -1.6364)
    PerezBrightness = (AirMass * Idh) / I0
    if Idh <= 10:
        PerezBrightness = _c(0.0)
    bounds = (1.065, 1.230, 1.500, 1.950, 2.800, 4.500, 6.200)
    k = 7
    for j, bnd in enumerate(bounds):
        if PerezClearness < bnd:
            k = j
            break
    m_a = acoeff[k, 0] + acoeff[k, 1] * zen + PerezBrightness * (acoeff[k, 2] + acoeff[k, 3] * zen)
    m_b = bcoeff[k, 0] + bcoeff[k, 1] * zen + PerezBrightness * (bcoeff[k, 2] + bcoeff[k, 3] * zen)
    m_e = ecoeff[k, 0] + ecoeff[k, 1] * zen + PerezBrightness * (ecoeff[k, 2] + ecoeff[k, 3] * zen)
    if k > 0:
        m_c = ccoeff[k, 0] + ccoeff[k, 1] * zen + PerezBrightness * (ccoeff[k, 2] + ccoeff[k, 3] * zen)
        m_d = dcoeff[k, 0] + dcoeff[k, 1] * zen + PerezBrightness * (dcoeff[k, 2] + dcoeff[k, 3] * zen)
    else:
        m_c = torch.exp(torch.pow(PerezBrightness * (ccoeff[k, 0] + ccoeff[k, 1] * zen), ccoeff[k, 2])) - 1
        m_d = -torch.exp(PerezBrightness * (dcoeff[k, 0] + dcoeff[k, 1] * zen)) + dcoeff[k, 2] + \
            PerezBrightness * dcoeff[k, 3] * PerezBrightness
    cosSky = sky["sin_alt"] * torch.sin(altitude) + \
        torch.cos(altitude) * sky["cos_alt"] * torch.cos(torch.abs(sky["azi"] - azimuth))
    lv = (1 + m_a * torch.exp(m_b / sky["cos_zen"])) * \
        ((1 + m_c * torch.exp(m_d * torch.arccos(cosSky)) + m_e * cosSky * cosSky))
    lv = lv / torch.sum(lv)
    return lv


def _patches(round_trip: bool):
    """Patch (alt, azi) as float32 tensors.

    Night steps use the ``create_patches`` values; day steps use the
    columns of the Perez table, which went degrees -> radians -> degrees in
    float32 (``solweig.py:1344-1360``) and can differ by an ulp.
    """
    alt, azi = skyvault.create_patches()
    if round_trip:
        deg2rad = torch.tensor(np.pi / 180, device=_CPU)
        rad2deg = torch.tensor(180 / np.pi, device=_CPU)
        alt = (alt * deg2rad) * rad2deg
        azi = (azi * deg2rad) * rad2deg
    return alt, azi


def patch_steradians(round_trip: bool = False):
    """Solid angle of each patch (``Lcyl_v2022a`` ``solweig.py:1708-1718``)."""
    alt, _ = _patches(round_trip)
    deg2rad = torch.tensor(np.pi / 180, device=_CPU)
    alt_np = alt.numpy()
    skyalt, skyalt_c = np.unique(alt_np, return_counts=True)
    skyalt_t = torch.tensor(skyalt)
    skyalt_ct = torch.tensor(skyalt_c)
    ster = torch.zeros(alt.shape[0])
    for i in range(alt.shape[0]):
        cnt = skyalt_ct[skyalt_t == alt[i]]
        if cnt > 1:
            ster[i] = ((360 / cnt) * deg2rad) * (torch.sin((alt[i] + alt[0]) * deg2rad)
                                                 - torch.sin((alt[i] - alt[0]) * deg2rad))
        else:
            ster[i] = ((360 / cnt) * deg2rad) * (torch.sin((alt[i]) * deg2rad)
                                                 - torch.sin((alt[i - 1] + alt[0]) * deg2rad))
    return ster


@dataclass
class PatchGeometry:
    """Step-invariant per-patch factors (float32)."""

    alt: np.ndarray
    azi: np.ndarray
    ster: np.ndarray
    cos_alt: np.ndarray      # cos(alt * deg2rad)
    sin_alt: np.ndarray      # sin(alt * deg2rad)
    angl_inc: np.ndarray     # cos(alt * deg2rad) * cos(0)
    card_e: np.ndarray       # cos((90 - azi) * deg2rad) (define_patch cardinals)
    card_s: np.ndarray
    card_w: np.ndarray
    card_n: np.ndarray
    dir_e: np.ndarray        # bool membership for the L cardinal sums
    dir_s: np.ndarray
    dir_w: np.ndarray
    dir_n: np.ndarray
    # step-invariant tensors of the per-step chain (same dtypes as there)
    t_ster: torch.Tensor = None        # f32 (153,)
    t_alt: torch.Tensor = None         # f32 (153,)
    t_azi: torch.Tensor = None         # f32 (153,)
    t_azi64: torch.Tensor = None       # f64 (153,)
    sin_alt_k: torch.Tensor = None     # sin(alt * f32(pi/180)), f32 (radTot terms)
    band_idx: np.ndarray = None        # ring index of each patch (unique altitudes)
    band_zterm: torch.Tensor = None    # 1.7 - 1/cos((90 - band alt) * pi/180), f32 (8,)
    band_sin: torch.Tensor = None      # f32(sin(band alt * pi/180) in float64), (8,)
    band_cos: torch.Tensor = None      # f32(cos(band alt * pi/180) in float64), (8,)


_GEOM_CACHE: dict = {}


def patch_geometry(round_trip: bool = False) -> PatchGeometry:
    if round_trip in _GEOM_CACHE:
        return _GEOM_CACHE[round_trip]
    g = _patch_geometry(round_trip)
    _GEOM_CACHE[round_trip] = g
    return g


def _patch_geometry(round_trip: bool) -> PatchGeometry:
    alt, azi = _patches(round_trip)
    deg2rad = torch.tensor(np.pi / 180, device=_CPU)
    ster = patch_steradians(round_trip)
    cos_alt = torch.cos(alt * deg2rad)
    sin_alt = torch.sin(alt * deg2rad)
    angl = torch.stack([torch.cos(alt[i] * deg2rad) * torch.cos(torch.tensor(0.)) for i in range(alt.shape[0])])
    ce = torch.stack([torch.cos((90 - azi[i]) * deg2rad) for i in range(alt.shape[0])])
    cs = torch.stack([torch.cos((180 - azi[i]) * deg2rad) for i in range(alt.shape[0])])
    cw = torch.stack([torch.cos((270 - azi[i]) * deg2rad) for i in range(alt.shape[0])])
    cn = torch.stack([torch.cos((0 - azi[i]) * deg2rad) for i in range(alt.shape[0])])
    a = azi.numpy()
    bands = np.unique(alt.numpy())
    skyzen = 90 - torch.tensor(bands)
    band_sin = torch.stack([torch.sin(torch.tensor(b) * deg2rad) for b in bands]).float()
    band_cos = torch.stack([torch.cos(torch.tensor(b) * deg2rad) for b in bands]).float()
    t_alt = alt.clone()
    return PatchGeometry(
        t_ster=ster.clone(), t_alt=t_alt, t_azi=azi.clone(), t_azi64=azi.double(),
        sin_alt_k=torch.sin(t_alt * torch.tensor(torch.pi / 180.0)),
        band_idx=np.searchsorted(bands, alt.numpy()),
        band_zterm=1.7 - (1 / torch.cos(skyzen * deg2rad)),
        band_sin=band_sin, band_cos=band_cos,
        alt=alt.numpy().astype(F32), azi=a.astype(F32), ster=ster.numpy().astype(F32),
        cos_alt=cos_alt.numpy().astype(F32), sin_alt=sin_alt.numpy().astype(F32),
        angl_inc=angl.numpy().astype(F32),
        card_e=ce.numpy().astype(F32), card_s=cs.numpy().astype(F32),
        card_w=cw.numpy().astype(F32), card_n=cn.numpy().astype(F32),
        dir_e=(a > 360) | (a < 180), dir_s=(a > 90) & (a < 270),
        dir_w=(a > 180) & (a < 360), dir_n=(a > 270) | (a < 90))


def _tgwave(amp, tmaxlst, frac_minus_snup, snup24):
    """``amp * sin(((frac - SNUP/24) / (TmaxLST/24 - SNUP/24)) * pi / 2)``."""
    den = tmaxlst / 24 - snup24
    return amp * torch.sin((frac_minus_snup / den) * np.pi / 2)


# ---------------------------------------------------------------------------
# WBGT forcing scalars: isobaric wet-bulb temperature (liquid phase, Romps
# Newton iteration; calculate_wbgt.py:17-63, 156-295, 412-511, 512-715) and
# the black-globe convection coefficient (utci_process.py:1021-1030,
# calculate_wbgt.py:717-745).
# ---------------------------------------------------------------------------
WB_RD, WB_RV = 287.0, 461.5          # gas constants of dry air / water vapour (J/kg/K)
WB_EPS = WB_RD / WB_RV
WB_CPD, WB_CPV, WB_CPL = 1005.0, 2040.0, 4220.0   # isobaric specific heats (J/kg/K)
WB_T0, WB_ES0, WB_LV0 = 273.16, 611.657, 2.501e6  # triple point (K, Pa) and latent heat there (J/kg)
WB_PRECISION, WB_MAX_ITER = 0.001, 20
GLOBE_EMISSIVITY, GLOBE_SIGMA = 0.95, 5.670374419e-8
GLOBE_HCG_COEF = 6.3 / 0.46821     # h_cg = coef * va^0.6


def latent_heat_liquid(T):
    """Lv(T) = Lv0 + (cpv - cpl) (T - T0) (calculate_wbgt.py:93-107)."""
    return WB_LV0 + (WB_CPV - WB_CPL) * (T - WB_T0)


def svp_liquid(T):
    """Saturation vapour pressure over liquid water in Pa (Ambaum 2020 eq. 13,
    calculate_wbgt.py:196-205): es0 (T0/T)^((cpl-cpv)/Rv) exp(Lv0/(Rv T0) - Lv(T)/(Rv T))."""
    lv = latent_heat_liquid(T)
    return WB_ES0 * np.power(WB_T0 / T, (WB_CPL - WB_CPV) / WB_RV) * np.exp(
        WB_LV0 / (WB_RV * WB_T0) - lv / (WB_RV * T))


def q_from_vapour_pressure(p, e):
    """Specific humidity from vapour pressure (inverse of calculate_wbgt.py:171-174)."""
    return WB_EPS * e / (p - (1 - WB_EPS) * e)


# PSEM rational approximations of the lower Lambert-W branch (Vazquez-Leal et
# al. 2019), on [-1/e, -0.34), [-0.34, -0.1) and [-0.1, 0): numerator and
# denominator coefficients a1.., b1.. (calculate_wbgt.py:314-349).
_W1_A = (-7.874564067684664, -63.11879948166995, -168.6110850408981, -150.1089086912451)
_W1_B = (15.97679839497612, 98.26612857148953, 293.9558944644677, 430.4471947824411, 247.8576700279611)
_W2_A = (-1362.78381643109, -1386.04132570149, 11892.1649836015, 16904.0507511421)
_W2_B = (251.440197724561, -1264.99554712435, -5687.63429510978, -2639.24130979048)
_W3_A = (1.01999365162218, -12.6917365519443, -45.1506015092455)
_W3_B = (-22.9809693297808, -104.692066099727, -95.2085341727207)


def _horner(x, coefs):
    """x * (c1 + x * (c2 + ... x * cn))."""
    acc = coefs[-1]
    for c in coefs[-2::-1]:
        acc = c + x * acc
    return x * acc


def lambertw_lower(x):
    """W_{-1}(x) for x in [-1/e, 0), NaN elsewhere: the PSEM approximation of
    the interval containing x plus one Fritsch-Shafer-Crowley step
    (calculate_wbgt.py:297-361). Evaluated in float64."""
    x = np.asarray(x, np.float64)
    y = np.full(x.shape, np.nan)
    with np.errstate(all="ignore"):
        s1 = (x >= -0.3678794411714423) & (x < -0.34)
        y1 = (_horner(x, _W1_A) / (1.0 + _horner(x, _W1_B))) * (x + math.exp(-1.0)) - 1.0
        s2 = (x >= -0.34) & (x < -0.1)
        y2 = _horner(x, _W2_A) / (1.0 + _horner(x, _W2_B))
        s3 = (x >= -0.1) & (x < 0.0)
        l1 = np.log(-x)
        y3 = _horner(x, _W3_A) / (1.0 + _horner(x, _W3_B)) + (l1 - np.log(-l1) + np.log(-l1) / l1)
        y = np.where(s1, y1, np.where(s2, y2, np.where(s3, y3, y)))
        z = np.log(x / y) - y
        t = 2.0 * (1.0 + y) * (1.0 + y + (2.0 / 3.0) * z)
        return y * (1.0 + (z / (1.0 + y)) * (t - z) / (t - 2.0 * z))


def dewpoint_liquid(p, T, q):
    """Dewpoint over liquid water (Romps 2021 closed form through W_{-1},
    calculate_wbgt.py:419-428), limited to T (:469-470)."""
    e = p * q / (WB_EPS * (1 - q) + q)
    rh = e / svp_liquid(T)
    beta = (WB_CPL - WB_CPV) / WB_RV
    alpha = -(1 / beta) * (WB_LV0 + (WB_CPL - WB_CPV) * WB_T0) / WB_RV
    with np.errstate(all="ignore"):
        arg = np.power(rh, 1 / beta) * (alpha / T) * np.exp(alpha / T)
    return np.minimum(alpha / lambertw_lower(arg).astype(np.result_type(arg)), T)


@dataclass
class WetBulb:
    """Isobaric wet-bulb temperature of every forcing row (degC) and the
    Newton iterations the whole table took."""
    tw_c: np.ndarray
    iterations: int
    converged: bool


def wet_bulb_from_rh(p_pa, t_k, rh_pct) -> WetBulb:
    """Liquid-phase isobaric wet-bulb temperature by Romps' Newton iteration
    (calculate_wbgt.py:476-511, 596-679, 691-715).

    The iteration runs over the whole array and stops when the largest
    update (NaN ignored) is below 0.001 K, or after 20 updates; the result is
    limited to T. Arithmetic is in the inputs' dtype (float64 for forcing
    tables)."""
    p = np.asarray(p_pa)
    T = np.asarray(t_k)
    es = svp_liquid(T)
    q = q_from_vapour_pressure(p, (np.asarray(rh_pct) / 100.0) * es)
    tw = T - (1 / 3) * (T - dewpoint_liquid(p, T, q))
    cpm = (1 - q) * WB_CPD + q * WB_CPV
    dlv = WB_CPV - WB_CPL
    count, converged = 0, False
    with np.errstate(all="ignore"):
        while True:
            prev = tw
            esw = svp_liquid(tw)
            qs = WB_EPS * esw / (p - (1 - WB_EPS) * esw)
            lv = latent_heat_liquid(tw)
            dqs = qs * (1 - qs + qs / WB_EPS) * lv / (WB_RV * tw ** 2)
            f = cpm * (T - tw) * (1 - qs) - (qs - q) * lv
            fp = -(cpm * (T - tw) + lv) * dqs - cpm * (1 - qs) - (qs - q) * dlv
            tw = tw - f / fp
            d = np.abs(tw - prev)
            if d.size and not np.all(np.isnan(d)) and np.nanmax(d) < WB_PRECISION:
                converged = True
                break
            count += 1
            if count == WB_MAX_ITER:
                break
    return WetBulb(np.minimum(T, tw) - 273.15, count + 1, converged)


def globe_coefficient(va: float) -> float:
    """a = h_cg / (emissivity * sigma) with h_cg = coef * va^0.6, in float32
    like the reference's tensors (utci_process.py:1022-1025,
    calculate_wbgt.py:733-740)."""
    hcg = F32(GLOBE_HCG_COEF) * np.power(F32(va), F32(0.6))
    return float(hcg / (F32(GLOBE_EMISSIVITY) * F32(GLOBE_SIGMA)))


class DayLoop:
    """Iterates the host chain over the forcing rows of one tile/day."""

    def __init__(self, met: np.ndarray, sun: SunTable, *, latitude: float, landcover: LandCover,
                 scale: float, amax: float, rows: int, cols: int, state: DayState | None = None,
                 psi_override=None):
        self.met = np.asarray(met, dtype=np.float64)
        self.sun = sun
        self.lat = float(latitude)
        self.lc = landcover
        self.scale = float(scale)
        self.amax = None if amax is None else float(np.float32(amax))
        self.rows = rows
        self.cols = cols
        self.state = state or DayState()
        self.geom_night = patch_geometry(False)
        self.geom_day = patch_geometry(True)
        n = self.met.shape[0]
        self.Ta = torch.tensor(self.met[:, 11])
        self.RH = torch.tensor(self.met[:, 10])
        self.radG = torch.tensor(self.met[:, 14])
        self.P = torch.tensor(self.met[:, 12])
        self.Ws = torch.tensor(self.met[:, 9])
        self.uhii = torch.tensor(self.met[:, 24]) if self.met.shape[1] > 24 else torch.zeros(n)
        DOY = torch.tensor(self.met[:, 1])
        leafon = torch.zeros((1, n))
        leafon[0, (DOY > 97) & (DOY < 300)] = 1
        psi = leafon * TRANS_VEG
        psi[leafon == 0] = 0.5
        self.psi = psi[0]
        self.timestepdec = 0 if n == 1 else (sun.dectime[1] - sun.dectime[0])
        self.dectime = np.asarray(sun.dectime, dtype=np.float64)
        self.jday_row = np.asarray(sun.jday, dtype=np.float64)
        # the wet-bulb iteration spans every forcing row at once (utci_process.py:817-823)
        # per-day constants of the chain: read-only tensors shared by the steps
        self._snup = {}
        self._lc_t = (torch.tensor(self.lc.tgk), torch.tensor(self.lc.tstart), torch.tensor(self.lc.tmaxlst))
        self._lcw_t = tuple(torch.tensor([[v]], dtype=torch.float64).float()
                            for v in (self.lc.tgk_wall, self.lc.tstart_wall, self.lc.tmaxlst_wall))
        self.wet_bulb = wet_bulb_from_rh(self.P.numpy() * 1000.0, (self.Ta + self.uhii).numpy() + 273.15,
                                         self.RH.numpy())

    # -- midnight bookkeeping (utci_process.py:917-934) --------------------
    def _midnight(self, i):
        st = self.state
        d = self.dectime
        if self.lc.present and (((d[i] - np.floor(d[i])) == 0) or (i == 0)):
            Ta_ = self.Ta.numpy()
            st.Twater = np.mean(Ta_[self.jday_row == np.floor(d[i])])
        if (d[i] - np.floor(d[i])) == 0:
            # np.where returns a 1-tuple for 1-D input, so the sunrise branch
            # of the reference is never taken and CI resets to 1 at midnight.
            st.CI = 1.0

    def step(self, i: int, uhii=None) -> StepParams:
        self._midnight(i)
        st = self.state
        sun = self.sun
        sp = StepParams(i=i, is_day=bool(sun.altitude[i] > 0))
        altitude = torch.tensor(np.float64(sun.altitude[i]))
        azimuth = torch.tensor(np.float64(sun.azimuth[i]))
        zen = torch.tensor(np.float64(sun.zen[i]))
        jday = float(sun.jday[i])
        dectime = torch.tensor(np.float64(self.dectime[i]))
        altmax = torch.tensor(np.float64(sun.altmax[i]))
        Ta = self.Ta[i]
        RH = self.RH[i]
        radG = self.radG[i]
        P = self.P[i]
        SBC = _c(5.67051e-8)
        SNUP = self._snup.get(jday)
        if SNUP is None:
            SNUP = self._snup[jday] = daylen_snup(jday, self.lat)
        ea = 6.107 * 10 ** ((7.5 * Ta) / (237.3 + Ta)) * (RH / 100.0)
        msteg = 46.5 * (ea / (Ta + 273.15))
        esky = (1 - (1 + msteg) * torch.exp(-((1.2 + 3.0 * msteg) ** 0.5))) + 0
        psi = self.psi[i]
        sp.psi = _f(psi)
        sp.one_m_psi = _f(1 - psi)
        sp.ta = float(np.float32(Ta.item()))
        uh = self.uhii[i]
        ta_mat = (torch.zeros(1) + Ta) + uh
        sp.ta_mat = float(ta_mat[0])
        sp.rh = float((torch.zeros(1) + RH)[0])
        va = torch.clamp(torch.ones(1) * self.Ws[i], min=0.15)
        sp.va = float(va[0])
        sp.wbgt_w = float(F32(0.7 * float(self.wet_bulb.tw_c[i])))
        sp.wbgt_a = globe_coefficient(sp.va)
        tak4_64 = (Ta + 273.15) ** 4
        sp.tak4 = float(np.float32(tak4_64.item()))
        deg2rad_t = _c(np.pi / 180)
        geom = self.geom_day if sp.is_day else self.geom_night
        sp.geom = geom
        d = sp.diag
        d["esky0"] = float(esky)
        d["ea"] = float(ea)
        CI = st.CI
        if sp.is_day:
            zt = _t(zen.item())
            jt = _t(jday)
            # (the reference evaluates this twice; the first result is overwritten)
            I0, CI, Kt = clearness_index(zt, jt, Ta.item(), _t(RH.item()) / 100.0, _t(radG.item()), self.lat, _t(P.item()))
            CI = min(CI, 1.0)
            radI, radD = diffuse_fraction(_t(radG.item()), _t(altitude.item()), Kt, _t(Ta.item()), _t(RH.item()))
            zenDeg = zen * (180 / np.pi)
            lv = perez_luminance(zenDeg.item(), azimuth.item(), radD, radI, jday)
            sp.lv = lv.numpy().astype(F32)
            sp.radI = _f(radI)
            sp.radD = _f(radD)
            sp.radG = float(np.float32(radG.item()))
            sp.sinalt = float(np.float32(torch.sin(altitude * (np.pi / 180)).item()))
            # sun march table + face-shading window
            sp.sun_azalt = (azimuth.item(), altitude.item())
            tab = skyvault.sun_table(azimuth.item(), altitude.item(), self.scale, self.amax, self.rows, self.cols)
            sp.table = tab
            sp.facesh_branch, sp.azilow, sp.azihigh = skyvault.sun_facesh_bounds(azimuth.item())
            # ground / wall temperature wave
            frac = (dectime - torch.floor(dectime)) - SNUP / 24
            snup24 = SNUP / 24
            tgk, tst, tmx = self._lc_t
            Tgamp = tgk * altmax + tst
            Tg = _tgwave(Tgamp, tmx, frac, snup24)
            if self.lc.present:
                tgkw, tstw, tmxw = self._lcw_t
                Tgampwall = tgkw * altmax + tstw
                Tgwall = _tgwave(Tgampwall, tmxw, frac, snup24)
            else:
                Tgampwall = self.lc.tgk_wall * altmax + self.lc.tstart_wall
                Tgwall = Tgampwall * torch.sin((((dectime - torch.floor(dectime)) - SNUP / 24)
                                                / (self.lc.tmaxlst_wall / 24 - SNUP / 24)) * np.pi / 2)
            Tgwall = torch.maximum(Tgwall, _c(0))
            radI0, radD0 = diffuse_fraction(I0, altitude.item(), 1.0, Ta.item(), RH.item())
            corr = 0.1473 * torch.log(90 - (zen / np.pi * 180)) + 0.3454
            CI_Tg = (radG / radI0) + (1 - corr)
            CI_Tg = min(CI_Tg, 1.0)
            radG0 = radI0 * (torch.sin(altitude * (np.pi / 180))) + radD0
            CI_TgG = (radG / radG0) + (1 - corr)
            CI_TgG = min(CI_TgG, 1.0)
            Tg = Tg * CI_TgG
            Tgwall = Tgwall * CI_TgG
            if self.lc.present:
                Tg = torch.maximum(Tg, _c(0))
            sp.tg_class = Tg.float().numpy().astype(F32).reshape(-1)
            sp.tgwall = float(torch.as_tensor(Tgwall).float().reshape(-1)[0])
            ewall_t = _c(EWALL)
            lwall = SBC * ewall_t * (Tgwall + Ta + 273.15) ** 4 - SBC * ewall_t * (Ta + 273.15) ** 4
            sp.lwall = float(torch.as_tensor(lwall).float().reshape(-1)[0])
            # thermal delay (TsWaveDelay_2015a, all six planes share it)
            sp.tw_first = (st.firstdaytime == 1)
            ta_in = st.timeadd
            if ta_in >= (59 / 1440):
                w = torch.exp(-33.27 * torch.tensor(ta_in))
                sp.tw_update = True
                new_timeadd = self.timestepdec if self.timestepdec > (59 / 1440) else 0
            else:
                ta_in = ta_in + self.timestepdec
                w = torch.exp(-33.27 * torch.tensor(ta_in))
                sp.tw_update = False
                new_timeadd = ta_in
            sp.tw_w = float(np.float32(w.item()))
            sp.tw_omw = float(np.float32((1 - w).item()))
            st.timeadd = new_timeadd
            st.firstdaytime = 0
            # cylinder wedge: beta = f32(zen)
            sp.tan_zen = float(torch.tan(torch.tensor(zen.item(), dtype=torch.float32)))
            # Kside scalars (alt/azi as Python floats -> f32 tensors)
            alt_k = torch.tensor(altitude.item())
            deg2rad_k = _c(torch.pi / 180.0)
            sp.cosalt_k = _f(torch.cos(alt_k * deg2rad_k))
            albedo = ALBEDO_B
            sp.sunlit_k = _f((albedo * (radI * torch.cos(alt_k * deg2rad_k)) + (radD * 0.5)) / torch.pi)
            sp.shaded_k = _f((albedo * radD * 0.5) / torch.pi)
            # radTot accumulates in float32, patch by patch (solweig.py:696-704);
            # add.accumulate is that sequential float32 sum
            terms = (lv * geom.t_ster * geom.sin_alt_k).numpy()
            radTot = np.add.accumulate(terms, dtype=np.float32)[-1]
            lumChi = (lv * radD) / torch.tensor([radTot])
            sp.lumchi = lumChi.numpy().astype(F32)
            az_py = azimuth.item()
            p2s = torch.abs(az_py - geom.t_azi)
            xi = torch.cos(p2s * (torch.pi / 180.0))
            yi = 2 * xi * torch.tan(alt_k * (torch.pi / 180.0))
            sp.yi_k = torch.where(yi > 0, 0.0, yi).numpy().astype(F32)
            d.update(I0=_f(I0), CI=float(CI), radI=sp.radI, radD=sp.radD,
                     CI_Tg=float(CI_Tg), CI_TgG=float(CI_TgG))
        else:
            # night (solweig.py:2183-2227)
            sp.tw_first = False
            sp.tw_update = False
            st.timeadd = 0
            st.firstdaytime = 1.0
            sp.radG = float(np.float32(radG.item()))
        # ---- longwave (both branches) ----
        if CI < 0.95:
            esky = CI * esky + (1 - CI) * 1.0
        st.CI = CI
        sp.diag["esky"] = float(esky)
        ewall_t = _c(EWALL)
        pi_t = _c(np.pi)
        esky_band = 1 - (1 - esky) * torch.exp(0.308 * geom.band_zterm)
        # per ring: (esky_band * SBC) in float32, then * (Ta + 273.15)^4 / pi in float64;
        # per patch: float32(base) * ster * float32(sin or cos of the ring altitude)
        base = ((esky_band * SBC).double() * ((Ta + 273.15) ** 4)) / pi_t
        b32 = base.float()[geom.band_idx]
        sp.lsky_down = ((b32 * geom.t_ster) * geom.band_sin[geom.band_idx]).numpy().astype(F32)
        sp.lsky_side = ((b32 * geom.t_ster) * geom.band_cos[geom.band_idx]).numpy().astype(F32)
        veg_surf = (ewall_t * SBC * ((Ta + 273.15) ** 4)) / pi_t
        sp.veg_l = float(veg_surf)
        sp.shaded_l = float((ewall_t * SBC * ((Ta + 273.15) ** 4)) / pi_t)
        if sp.is_day:
            tgw = torch.tensor([[sp.tgwall]], dtype=torch.float32)
            sp.sunlit_l = float(((ewall_t * SBC * ((Ta + tgw + 273.15) ** 4)) / pi_t).reshape(-1)[0])
            # float64 like the reference: 0-dim float64 azimuth minus a float32 patch azimuth
            azd = torch.abs(azimuth - geom.t_azi64)
            lsun = ((azd > 90) & (azd < 270) & (altitude > 0)).numpy()
            xi = torch.cos(azd * (torch.pi / 180.0))
            yi = 2 * xi * torch.tan(altitude * (torch.pi / 180.0))
            yl = torch.where(yi > 0, 0.0, yi).float().numpy()
            sp.yi_l = np.where(lsun, yl, F32(0)).astype(F32)
            sp.l_sun_side = lsun.astype(np.int32)
        else:
            sp.sunlit_l = 0.0
            sp.l_sun_side = np.zeros(153, np.int32)
        return sp
