"""Sky vault patches, annulus weights and march step tables (host side).

Everything here is O(patches x steps) scalar work done once per request on
the host. The per-pixel work runs in the HIP kernels, which consume these
tables from device memory.

The tables must carry exactly the float32 values the reference derives,
because the shadow masks are compared bit for bit: one flipped ``dz`` bit
moves a comparison ``a[src] - dz > a[t]``. The reference derives them with
float32 0-dim torch tensors (sin/cos/tan/round on CPU), so this module
evaluates the same expressions with the same tensor dtypes, step by step.
torch is used here only as the scalar arithmetic that defines those bits.

References:
  * ``create_patches`` ``solweig_gpu/shadow.py:352-406`` (patch_option 2);
  * sky-patch azimuths ``shadow.py:527-543``; annulus weights
    ``shadow.py:324-350``; fold direction windows ``shadow.py:556-584``;
  * SVF march tables ``shadow.py:194-301`` (azimuth 0 -> 1e-12, index
    from 1, stop test on the previous dz, so the overshoot step is kept);
  * sun march tables ``solweig.py:1088-1185`` (index from 0, ``dzprev``);
  * ground-view walk offsets ``solweig.py:113-231`` and azimuths ``:323``.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

_CPU = torch.device("cpu")

# patch_option 2 ("153 patches") of create_patches.
ANNULINO = (0, 12, 24, 36, 48, 60, 72, 84, 90)
RING_ALT = (6, 18, 30, 42, 54, 66, 78, 90)
RING_AZISTART = (0, 4, 2, 5, 8, 0, 10, 0)
RING_COUNT = (31, 30, 28, 24, 19, 13, 7, 1)
N_PATCHES = 153

#: Ground-view-factor search azimuths (deg): 5, 25, ..., 345.
N_WALK_AZ = 18


@dataclass(frozen=True)
class SkyVault:
    """The 153 sky patches and their SVF fold constants.

    ``alt``/``azi`` are the float32 patch altitude/azimuth in degrees
    (``create_patches`` values, identical to the ``iazimuth`` the SVF sweep
    uses). ``w_iso[p]``/``w_aniso[p]`` hold the ring's annulus weights in
    summation order; ``dir_mask[p]`` is a 4-bit E/S/W/N membership set.
    """

    alt: np.ndarray          # (153,) float32, degrees
    azi: np.ndarray          # (153,) float32, degrees
    ring: np.ndarray         # (153,) int32
    w_iso: np.ndarray        # (8, 12) float32, ring annulus weights (iso)
    w_aniso: np.ndarray      # (8, 12) float32, ring annulus weights (aniso)
    n_annuli: np.ndarray     # (8,) int32 (12 or 6)
    dir_mask: np.ndarray     # (153,) int32 bits: 1=E, 2=S, 4=W, 8=N


def _patch_tensors():
    annulino = torch.tensor(list(ANNULINO), device=_CPU)
    alt_int = torch.tensor(list(RING_ALT), device=_CPU)
    azistart = torch.tensor(list(RING_AZISTART), device=_CPU)
    count = torch.tensor(list(RING_COUNT), device=_CPU)
    return annulino, alt_int, azistart, count


_PATCHES = None


def create_patches():
    """Float32 patch altitudes/azimuths exactly as ``create_patches(2)`` (cached copies)."""
    global _PATCHES
    if _PATCHES is None:
        _PATCHES = _create_patches()
    return _PATCHES[0].clone(), _PATCHES[1].clone()


def _create_patches():
    annulino, alt_int, azistart, count = _patch_tensors()
    aziint = 360 / count
    alts = torch.tensor([], device=_CPU)
    azis = torch.tensor([], device=_CPU)
    for j in range(alt_int.shape[0]):
        for k in range(count[j]):
            alts = torch.cat((alts, torch.tensor([alt_int[j]], device=_CPU)))
            azis = torch.cat((azis, torch.tensor([k * aziint[j] + azistart[j]], device=_CPU)))
    return alts, azis


def sweep_azimuths():
    """Per-patch azimuths (float32 tensor) in SVF-sweep order (``shadow.py:535-543``)."""
    annulino, alt_int, azistart, count = _patch_tensors()
    aziint = torch.tensor([360 / c for c in count], device=_CPU)
    iaz = torch.zeros((1, int(torch.sum(count).item())), device=_CPU)
    index = 0
    for j in range(alt_int.shape[0]):
        for k in range(int(360 / aziint[j])):
            iaz[0, index] = k * aziint[j] + azistart[j]
            if iaz[0, index] > 360.0:
                iaz[0, index] = iaz[0, index] - 360.0
            index += 1
    return iaz[0]


def annulus_weight(altitude: int, aziinterval) -> torch.Tensor:
    """Solid-angle weight of one 1-degree annulus (``shadow.py:324-350``)."""
    n = torch.tensor(90.0, device=_CPU)
    alt = torch.tensor(altitude, device=_CPU)
    steprad = (360.0 / aziinterval) * (torch.pi / 180.0)
    annulus = 91.0 - alt
    w = (1.0 / (2.0 * torch.pi)) * torch.sin(torch.pi / (2.0 * n)) * \
        torch.sin((torch.pi * (2.0 * annulus - 1.0)) / (2.0 * n))
    return steprad * w


def build_skyvault() -> SkyVault:
    annulino, alt_int, azistart, count = _patch_tensors()
    alts, _ = create_patches()
    iaz = sweep_azimuths()
    aniso = torch.ceil(count / 2.0)
    w_iso = np.zeros((8, 12), dtype=np.float32)
    w_aniso = np.zeros((8, 12), dtype=np.float32)
    n_ann = np.zeros(8, dtype=np.int32)
    for i in range(8):
        ks = list(range(int(annulino[i]) + 1, int(annulino[i + 1]) + 1))
        n_ann[i] = len(ks)
        for m, k in enumerate(ks):
            w_iso[i, m] = annulus_weight(k, count[i]).item()
            w_aniso[i, m] = annulus_weight(k, aniso[i]).item()
    ring = np.repeat(np.arange(8, dtype=np.int32), RING_COUNT)
    az = iaz.numpy().astype(np.float32)
    dmask = np.zeros(N_PATCHES, dtype=np.int32)
    for p in range(N_PATCHES):
        a = iaz[p]
        if 0 <= a < 180:
            dmask[p] |= 1
        if 90 <= a < 270:
            dmask[p] |= 2
        if 180 <= a < 360:
            dmask[p] |= 4
        if a >= 270 or a < 90:
            dmask[p] |= 8
    return SkyVault(alt=alts.numpy().astype(np.float32), azi=az, ring=ring,
                    w_iso=w_iso, w_aniso=w_aniso, n_annuli=n_ann, dir_mask=dmask)


# ---------------------------------------------------------------------------
# March step tables
# ---------------------------------------------------------------------------


@dataclass
class StepTable:
    """One march: executed steps in order (overshoot step included)."""

    dx: np.ndarray      # int32 row offsets (source row = target row + dx)
    dy: np.ndarray      # int32 col offsets (source col = target col + dy)
    dz: np.ndarray      # float32 height drop at each step
    dzprev: np.ndarray  # float32 dz of the previous step (sun march only)

    @property
    def count(self) -> int:
        return int(self.dx.shape[0])


def _as_f32_tensor(x):
    if isinstance(x, torch.Tensor):
        return x
    return torch.tensor(x, device=_CPU)


def _steps(ssin, scos, tanaz, ds_sin, ds_cos, tanalt, branch: bool, amax, rows: int, cols: int,
           start: int, neg_one):
    """Vectorised step loop shared by both marches.

    Evaluates every candidate index at once with the same float32 element
    operations as the reference's scalar loop, then keeps the executed
    prefix: a step runs while the previous step still satisfied
    ``amax >= dz and |dx| < rows and |dy| < cols`` (so the first failing
    step, the overshoot, is kept). With ``amax=None`` only the geometric
    part is applied; the device kernels then apply the ``amax`` test with
    the tile maximum they computed themselves.
    """
    kmax = max(rows, cols) + 2
    index = torch.arange(start, start + kmax, dtype=torch.float32)
    if branch:
        dy = ssin * index
        dx = (neg_one * scos) * torch.abs(torch.round(index / tanaz))
        ds = ds_sin
    else:
        dy = ssin * torch.abs(torch.round(index * tanaz))
        dx = (neg_one * scos) * index
        ds = ds_cos
    dz = (ds * index) * tanalt
    ok = (torch.abs(dx) < rows) & (torch.abs(dy) < cols)
    if amax is not None:
        ok = ok & (amax >= dz)
    bad = torch.nonzero(~ok)
    count = int(bad[0, 0]) + 1 if bad.numel() else kmax
    return (dx[:count].numpy().astype(np.int32), dy[:count].numpy().astype(np.int32),
            dz[:count].numpy().astype(np.float32))


def svf_patch_table(azimuth, altitude, scale: float, amaxvalue, rows: int, cols: int) -> StepTable:
    """Step table of one SVF patch march (``shadow()``, ``shadow.py:194-301``).

    ``azimuth`` is the float32 patch azimuth (tensor or float, degrees),
    ``altitude`` the integer ring altitude, ``amaxvalue`` the float32
    maximum height used by the stop test. Index starts at 1.
    """
    degrees = torch.pi / 180.0
    if azimuth == 0.0:
        azimuth = 1e-12
    azimuth = _as_f32_tensor(azimuth)
    altitude = _as_f32_tensor(altitude)
    amax = None if amaxvalue is None else _as_f32_tensor(amaxvalue)
    azimuth = azimuth * degrees
    altitude = altitude * degrees
    pibyfour = torch.pi / 4.0
    branch = bool(pibyfour <= azimuth < 3.0 * pibyfour or 5.0 * pibyfour <= azimuth < 7.0 * pibyfour)
    sinaz = torch.sin(azimuth)
    cosaz = torch.cos(azimuth)
    tanaz = torch.tan(azimuth)
    dx, dy, dz = _steps(torch.sign(sinaz), torch.sign(cosaz), tanaz, torch.abs(1.0 / sinaz),
                        torch.abs(1.0 / cosaz), torch.tan(altitude) / scale, branch, amax, rows, cols,
                        1, -1.0)
    return StepTable(dx, dy, dz, np.zeros_like(dz))


def sun_table(azimuth_deg: float, altitude_deg: float, scale: float, amaxvalue: float,
              rows: int, cols: int) -> StepTable:
    """Step table of the sun march (``shadowingfunction_wallheight_23``).

    ``azimuth_deg``/``altitude_deg`` are the Python floats the reference
    passes (``azimuth.item()``); they become float32 tensors before the
    degree conversion (``solweig.py:1090-1092``). Index starts at 0 and each
    step also carries the previous step's dz.
    """
    degrees = torch.tensor(np.pi / 180.0, device=_CPU)
    azimuth = torch.tensor(azimuth_deg, device=_CPU) * degrees
    altitude = torch.tensor(altitude_deg, device=_CPU) * degrees
    pibyfour = torch.tensor(np.pi / 4.0, device=_CPU)
    branch = bool(((pibyfour <= azimuth) and (azimuth < 3 * pibyfour))
                  or ((5 * pibyfour <= azimuth) and (azimuth < 7 * pibyfour)))
    sinaz = torch.sin(azimuth)
    cosaz = torch.cos(azimuth)
    tanaz = torch.tan(azimuth)
    dx, dy, dz = _steps(torch.sign(sinaz), torch.sign(cosaz), tanaz, torch.abs(1 / sinaz),
                        torch.abs(1 / cosaz), torch.tan(altitude) / scale, branch, amaxvalue, rows, cols,
                        0, -1)
    dzprev = np.concatenate([np.zeros(1, np.float32), dz[:-1]]).astype(np.float32)
    return StepTable(dx, dy, dz, dzprev)


def sun_facesh_bounds(azimuth_deg: float):
    """The float32 azimuth window that decides wall face shading.

    Returns (branch, azilow, azihigh) as in ``solweig.py:1187-1196``: branch
    0 -> ``facesh = (aspect < lo | aspect >= hi) - wallbol + 1``; branch 1/2
    -> ``facesh = 1 - (aspect > lo | aspect <= hi)``; branch -1 -> no branch
    matched (facesh undefined in the reference; the solver refuses).
    """
    degrees = torch.tensor(np.pi / 180.0, device=_CPU)
    azimuth = torch.tensor(azimuth_deg, device=_CPU) * degrees
    azilow = azimuth - torch.pi / 2
    azihigh = azimuth + torch.pi / 2
    if azilow >= 0 and azihigh < 2 * torch.pi:
        return 0, float(azilow), float(azihigh)
    if azilow < 0 and azihigh <= 2 * torch.pi:
        azilow = azilow + 2 * torch.pi
        return 1, float(azilow), float(azihigh)
    if azilow > 0 and azihigh >= 2 * torch.pi:
        azihigh = azihigh - 2 * torch.pi
        return 2, float(azilow), float(azihigh)
    return -1, float(azilow), float(azihigh)


@dataclass
class WalkTables:
    """Ground-view-factor walk: 18 azimuths x ``second`` offsets."""

    dx: np.ndarray          # (18, second) int32
    dy: np.ndarray          # (18, second) int32
    branch: np.ndarray      # (18,) int32 facesh branch per walk azimuth
    azilow: np.ndarray      # (18,) float32
    azihigh: np.ndarray     # (18,) float32
    dir_mask: np.ndarray    # (18,) int32 E/S/W/N membership bits
    first: float            # rounded first-height steps (>= 1)
    second: float           # rounded second-height steps


_WALK_CACHE: dict = {}


def walk_tables(scale: float, height: float = 1.1) -> WalkTables:
    key = (float(scale), float(height))
    if key not in _WALK_CACHE:
        _WALK_CACHE[key] = _walk_tables(scale, height)
    return _WALK_CACHE[key]


def _walk_tables(scale: float, height: float) -> WalkTables:
    """Offsets of ``sunonsurface_2018a`` for the 18 GVF azimuths.

    ``first = round(height)``, ``second = round(height*20)`` as in the
    driver (``utci_process.py:847-853``), then both are rescaled by
    ``round(x * scale)`` inside the walk (``solweig.py:143-146``).
    """
    h = torch.tensor(height, device=_CPU)
    first = torch.round(h.clone())
    if first == 0.0:
        first = torch.tensor(1.0, device=_CPU)
    second = torch.round(h * 20.0)
    sc = torch.tensor(scale, device=_CPU)
    first = torch.round(first * sc)
    if first < 1:
        first = torch.tensor(1.0)
    second = torch.round(second * sc)
    nsec = int(second.item())
    az_all = torch.arange(5, 359, 20, dtype=torch.float32, device=_CPU)
    dxs = np.zeros((N_WALK_AZ, nsec), np.int32)
    dys = np.zeros((N_WALK_AZ, nsec), np.int32)
    br = np.zeros(N_WALK_AZ, np.int32)
    lo = np.zeros(N_WALK_AZ, np.float32)
    hi = np.zeros(N_WALK_AZ, np.float32)
    dmask = np.zeros(N_WALK_AZ, np.int32)
    for j in range(N_WALK_AZ):
        azA = az_all[j]
        azimuth = azA * (torch.pi / 180)
        pibyfour = torch.pi / 4
        three = 3 * pibyfour
        five = 5 * pibyfour
        seven = 7 * pibyfour
        sinaz = torch.sin(azimuth)
        cosaz = torch.cos(azimuth)
        tanaz = torch.tan(azimuth)
        ssin = torch.sign(sinaz)
        scos = torch.sign(cosaz)
        index = 0
        for n in range(nsec):
            if (pibyfour <= azimuth and azimuth < three) or (five <= azimuth and azimuth < seven):
                dy = ssin * index
                dx = -1 * scos * torch.abs(torch.round(index / tanaz))
            else:
                dy = ssin * torch.abs(torch.round(index * tanaz))
                dx = -1 * scos * index
            dxs[j, n] = int(dx)
            dys[j, n] = int(dy)
            index += 1
        azilow = azimuth - torch.pi / 2
        azihigh = azimuth + torch.pi / 2
        if azilow >= 0 and azihigh < 2 * torch.pi:
            br[j] = 0
        elif azilow < 0 and azihigh <= 2 * torch.pi:
            azilow = azilow + 2 * torch.pi
            br[j] = 1
        elif azilow > 0 and azihigh >= 2 * torch.pi:
            azihigh = azihigh - 2 * torch.pi
            br[j] = 2
        else:
            br[j] = -1
        lo[j] = float(azilow)
        hi[j] = float(azihigh)
        if 0 <= azA < 180:
            dmask[j] |= 1
        if 90 <= azA < 270:
            dmask[j] |= 2
        if 180 <= azA < 360:
            dmask[j] |= 4
        if 270 <= azA or azA < 90:
            dmask[j] |= 8
    return WalkTables(dxs, dys, br, lo, hi, dmask, float(first), float(second))


def svf_tables(vault: SkyVault, scale: float, amaxvalue, rows: int, cols: int):
    """All 153 SVF patch tables concatenated: (offsets, dx, dy, dz).

    ``amaxvalue=None`` keeps every geometrically valid step (device form)."""
    iaz = sweep_azimuths()
    annulino, alt_int, azistart, count = _patch_tensors()
    tabs = []
    p = 0
    for i in range(8):
        for _ in range(int(count[i])):
            tabs.append(svf_patch_table(iaz[p], alt_int[i], scale, amaxvalue, rows, cols))
            p += 1
    offs = np.zeros(len(tabs) + 1, np.int32)
    for i, t in enumerate(tabs):
        offs[i + 1] = offs[i] + t.count
    dx = np.concatenate([t.dx for t in tabs]).astype(np.int32)
    dy = np.concatenate([t.dy for t in tabs]).astype(np.int32)
    dz = np.concatenate([t.dz for t in tabs]).astype(np.float32)
    return offs, dx, dy, dz
