"""Wall heights and wall aspect of a DSM tile on the GPU.

Drop-in mirrors of ``solweig_gpu/walls_aspect.py``:

* :func:`findwalls` (``walls_aspect.py:27-58``) -> ``sw_find_walls``;
* :func:`filter1Goodwin_as_aspect_v3` (``walls_aspect.py:99-166``) ->
  ``sw_wall_aspect``, with the rotated filter table of :func:`wall_filters`;
* :func:`walls_and_aspect`: both for one tile, the work of
  ``process_file_parallel`` (``walls_aspect.py:168-260``) without the
  GeoTIFF reading and writing.

Inputs may be numpy arrays (the result comes back as numpy float32) or
device tensors (the result stays on the device). Results are float32, the
dtype the reference writes its wall and aspect rasters in
(``walls_aspect.py:214-228``). There is no CPU path.
"""
from __future__ import annotations

import ctypes
import functools

import numpy as np
import torch

from . import _lib as L

F32 = np.float32
#: wall height threshold of the reference (``walls_aspect.py:25``)
WALL_LIMIT = 3.0


def filter_size(scale: float) -> int:
    """Line-filter size for ``scale`` pixels per metre (``walls_aspect.py:114-118``)."""
    fs = int(np.floor((scale + 1e-10) * 9))
    if fs <= 2:
        fs = 3
    elif fs != 9 and fs % 2 == 0:
        fs += 1
    return fs


@functools.lru_cache(maxsize=8)
def wall_filters(fs: int) -> bytes:
    """``sw_wall_filters_t`` for filter size ``fs``, as raw bytes.

    The filters are rotated by ``scipy.ndimage.rotate`` with the reference's
    arguments (line: order 1 then rounded; building halves: order 0) and the
    same special cases at h = 30, 150 and at 135 / 225 degrees
    (``walls_aspect.py:121-142``); each list holds the window offsets of one
    filter value in row-major order.
    """
    from scipy import ndimage
    if fs < 3 or fs % 2 == 0 or fs > 255:
        raise ValueError(f"filter size {fs} not supported (odd, 3..255)")
    n = fs - 1
    half = fs // 2
    filt = np.zeros((fs, fs))
    build = np.zeros((fs, fs))
    filt[:, half] = 1
    build[half, :half] = 1
    build[half, half + 1:] = 2
    t = L.WallFiltersStruct()
    t.size = fs
    for h in range(L.WALL_ROTATIONS):
        f1 = np.round(ndimage.rotate(filt, h, order=1, reshape=False, mode="nearest"))
        fb = np.round(ndimage.rotate(build, h, order=0, reshape=False, mode="nearest"))
        index = 270 - h
        if h in (150, 30):
            fb[:, n] = 0
        if index == 225:
            f1[0, 0] = f1[n, n] = 1
        if index == 135:
            f1[0, n] = f1[n, 0] = 1
        if not np.isin(f1, (0.0, 1.0)).all():
            raise RuntimeError(f"rotated line filter at h={h} is not 0/1")
        for li, sel in enumerate((f1 == 1, fb == 1, fb == 2)):
            ys, xs = np.nonzero(sel)
            if ys.size > L.WALL_LIST_CAP:
                raise ValueError(f"filter size {fs}: {ys.size} cells in one list (cap {L.WALL_LIST_CAP})")
            t.count[h][li] = ys.size
            for k, (yy, xx) in enumerate(zip(ys, xs)):
                t.off[h][li][k][0] = int(yy) - half
                t.off[h][li][k][1] = int(xx) - half
    return bytes(t)


_DEVICE_FILTERS: dict = {}


def device_filters(fs: int, device) -> torch.Tensor:
    """The filter table for ``fs`` resident on ``device`` (uploaded once)."""
    dev = torch.device(device)
    key = (str(dev), fs)
    t = _DEVICE_FILTERS.get(key)
    if t is None:
        raw = np.frombuffer(wall_filters(fs), dtype=np.uint8)
        t = torch.from_numpy(raw.copy()).to(dev)
        _DEVICE_FILTERS[key] = t
    return t


def _to_device(x, name: str):
    """(device float32 contiguous tensor, came-from-host flag)."""
    if isinstance(x, torch.Tensor):
        if x.device.type != "cuda":
            return x.to(torch.float32).contiguous().to("cuda:0"), True
        return x.to(torch.float32).contiguous(), False
    arr = np.ascontiguousarray(np.asarray(x, dtype=F32))
    if arr.ndim != 2:
        raise ValueError(f"{name} must be a 2-D plane")
    return torch.from_numpy(arr).to("cuda:0"), True


def _stream(dev) -> int:
    return torch.cuda.current_stream(dev).cuda_stream


def findwalls(dem_array, walllimit: float = WALL_LIMIT):
    """Wall heights (float32) of a DSM; ``findwalls`` of ``walls_aspect.py:27-58``."""
    L.load()
    a, host = _to_device(dem_array, "dem_array")
    if a.dim() != 2:
        raise ValueError("dem_array must be a 2-D plane")
    rows, cols = a.shape
    out = torch.empty_like(a)
    L.call("sw_find_walls", L.ptr(a), L.ptr(out), rows, cols, float(walllimit), _stream(a.device))
    return out.cpu().numpy() if host else out


def filter1Goodwin_as_aspect_v3(walls, scale, a):
    """Wall aspect in degrees (float32); ``walls_aspect.py:99-166`` on the GPU.

    ``walls`` are wall heights (cells > 0 are walls), ``scale`` pixels per
    metre, ``a`` the DSM. Raises ``ValueError`` for planes smaller than 2 x 2,
    where the reference's ``np.gradient`` raises.
    """
    L.load()
    w, host_w = _to_device(walls, "walls")
    d, host_d = _to_device(a, "a")
    if w.shape != d.shape or w.dim() != 2:
        raise ValueError("walls and a must be 2-D planes of the same shape")
    rows, cols = d.shape
    if rows < 2 or cols < 2:
        raise ValueError("Shape of array too small to calculate a numerical gradient")
    scale = float(scale)
    fs = filter_size(scale)
    filt = device_filters(fs, d.device)
    out = torch.empty_like(d)
    L.call("sw_wall_aspect", L.ptr(w), L.ptr(d), L.ptr(filt), L.ptr(out), rows, cols, scale, _stream(d.device))
    return out.cpu().numpy() if (host_w and host_d) else out


def walls_and_aspect(dsm, scale: float, walllimit: float = WALL_LIMIT):
    """(walls, aspect) of one DSM tile, both float32, computed back to back on the device."""
    L.load()
    d, host = _to_device(dsm, "dsm")
    walls = findwalls(d, walllimit)
    aspect = filter1Goodwin_as_aspect_v3(walls, scale, d)
    if host:
        return walls.cpu().numpy(), aspect.cpu().numpy()
    return walls, aspect


def check_filter_size_abi(scale: float) -> int:
    """``sw_wall_filter_size`` of the library (equal to :func:`filter_size`)."""
    return int(L.load().sw_wall_filter_size(ctypes.c_double(float(scale))))
