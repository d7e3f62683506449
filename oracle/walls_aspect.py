"""CPU oracle of wall heights and wall aspect -- TEST INFRASTRUCTURE ONLY.

Restates ``findwalls`` (``solweig_gpu/walls_aspect.py:27-58``), ``get_ders``
(``:78-97``) and ``filter1Goodwin_as_aspect_v3`` (``:99-166``) with the
reference's dtypes:

* walls: the 4-neighbour maximum of the float32 DSM minus the cell, in
  float64 (``np.zeros`` planes), zeroed below the limit and on the border;
* aspect: for h = 0..179 the line filter rotated by ``ndimage.rotate``
  (order 1, rounded) counts wall cells; the strictly larger count wins
  (first maximum kept) and the two halves of the rotated building filter
  (order 0) decide the side by comparing float32 sums of the DSM in the
  window, each summed the way ``np.sum`` does it (:func:`np_sum_f32_rows`);
* the gradient fallback (wall cells left at 0) uses ``np.gradient`` of the
  float32 DSM, float32 ``arctan2`` and float32 degree conversion.

The per-wall-cell window loop of the reference is evaluated for all wall
cells of one rotation at once; every comparison and sum is the same.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import ndimage

F32 = np.float32


def findwalls(dsm: np.ndarray, walllimit: float = 3.0) -> np.ndarray:
    """Wall heights (float64), ``walls_aspect.py:27-58``."""
    a = np.asarray(dsm)
    walls = np.zeros(a.shape)
    if a.shape[0] >= 3 and a.shape[1] >= 3:
        nb = np.stack([a[:-2, 1:-1], a[2:, 1:-1], a[1:-1, :-2], a[1:-1, 2:]])
        walls[1:-1, 1:-1] = np.max(nb, axis=0)          # np.max propagates NaN like the reference
    walls = walls - a
    walls[walls < walllimit] = 0
    walls[:, 0] = 0
    walls[:, -1] = 0
    walls[0, :] = 0
    walls[-1, :] = 0
    return walls


def filter_size(scale: float) -> int:
    """Line-filter size (``walls_aspect.py:114-118``)."""
    fs = int(np.floor((scale + 1e-10) * 9))
    if fs <= 2:
        fs = 3
    elif fs != 9 and fs % 2 == 0:
        fs += 1
    return fs


def rotated_filters(fs: int):
    """The 180 (line, building) filter pairs of ``walls_aspect.py:121-142``."""
    n = fs - 1
    fceil = int(np.ceil(fs / 2.0))
    ffloor = int(np.floor(fs / 2.0))
    filt = np.zeros((fs, fs))
    build = np.zeros((fs, fs))
    filt[:, fceil - 1] = 1
    build[fceil - 1, :ffloor] = 1
    build[fceil - 1, fceil:] = 2
    out = []
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
        out.append((f1, fb))
    return out


def np_sum_f32_rows(v: np.ndarray) -> np.ndarray:
    """Row sums of a float32 matrix in ``np.sum``'s order (pairwise, 8 accumulators).

    The reduction starts from the identity +0.0 and adds the pairwise sum:
    rows shorter than 8 are one running sum from -0.0; rows of 8..128 are
    eight running sums over blocks of 8, combined as
    ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)), then the remainder added in order.
    Longer rows are not needed here.
    """
    v = np.asarray(v, F32)
    m, n = v.shape
    if n > 128:
        raise ValueError("rows longer than 128 are not restated")
    if n < 8:
        r = np.full(m, -0.0, F32)
        for k in range(n):
            r = (r + v[:, k]).astype(F32)
        return (F32(0.0) + r).astype(F32)
    acc = [v[:, k].copy() for k in range(8)]
    i = 8
    lim = n - (n % 8)
    while i < lim:
        for k in range(8):
            acc[k] = (acc[k] + v[:, i + k]).astype(F32)
        i += 8
    res = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]))
    for k in range(i, n):
        res = (res + v[:, k]).astype(F32)
    return (F32(0.0) + res).astype(F32)


def gradient_aspect_deg(dsm: np.ndarray, scale: float) -> np.ndarray:
    """``get_ders`` aspect in degrees as float32 (``walls_aspect.py:78-97, 163``)."""
    dx = 1 / scale
    fy, fx = np.gradient(np.asarray(dsm, F32), dx, dx)
    asp = np.arctan2(fx, fy)
    asp = -asp
    asp[asp < 0] += 2 * np.pi
    return asp / (math.pi / 180.)


def wall_aspect(walls: np.ndarray, scale: float, dsm: np.ndarray, parts: bool = False):
    """Wall aspect in degrees (float64 like the reference), ``walls_aspect.py:99-166``.

    With ``parts`` also returns the mask of wall cells that took the
    gradient fallback (cells still at 0 after the rotating filter).
    """
    a = np.asarray(dsm, F32)
    rows, cols = a.shape
    fs = filter_size(scale)
    fceil = int(np.ceil(fs / 2.0))
    ffloor = int(np.floor(fs / 2.0))
    wallb = (np.asarray(walls) > 0).astype(np.uint8)
    y = np.zeros((rows, cols))
    x = np.zeros((rows, cols))
    act = np.zeros((rows, cols), bool)
    if rows - fceil - 1 > fceil - 1 and cols - fceil - 1 > fceil - 1:
        act[fceil - 1:rows - fceil - 1, fceil - 1:cols - fceil - 1] = True
    ii, jj = np.nonzero(act & (wallb == 1))
    if ii.size:
        z = np.zeros(ii.size)
        xs = np.zeros(ii.size)
        ys = np.zeros(ii.size)
        for h, (f1, fb) in enumerate(rotated_filters(fs)):
            ly, lx = np.nonzero(f1 != 0)
            wts = f1[ly, lx]
            cnt = (wallb[ii[:, None] + ly - ffloor, jj[:, None] + lx - ffloor] * wts).sum(axis=1)
            upd = z < cnt
            if not upd.any():
                continue
            u = np.nonzero(upd)[0]
            b1y, b1x = np.nonzero(fb == 1)
            b2y, b2x = np.nonzero(fb == 2)
            s1 = np_sum_f32_rows(a[ii[u, None] + b1y - ffloor, jj[u, None] + b1x - ffloor])
            s2 = np_sum_f32_rows(a[ii[u, None] + b2y - ffloor, jj[u, None] + b2x - ffloor])
            z[u] = cnt[u]
            xs[u] = np.where(s1 > s2, 1, 2)
            ys[u] = 270 - h
        x[ii, jj] = xs
        y[ii, jj] = ys
    y[x == 1] -= 180
    y[y < 0] += 360
    asp_deg = gradient_aspect_deg(a, scale)
    fallback = (wallb == 1) & (y == 0)
    y += fallback * asp_deg
    return (y, fallback) if parts else y
