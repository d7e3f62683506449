"""SVF cache files, read and written without GDAL.

Interop with the reference pipeline's cache (``solweig_gpu/shadow.py:98-164``,
read back and validated by ``utci_process.py:229-380``):

* ``svfs_<key>.zip``: 15 single-band Float32 GeoTIFFs (``svf.tif``,
  ``svfE.tif`` ... ``svfNaveg.tif``), zip-deflated;
* ``SkyViewFactor_<key>.tif``: SVFtotal;
* ``shadowmats_<key>.npz``: float32 cubes ``shadowmat``, ``vegshadowmat``,
  ``vbshmat`` of shape (rows, cols, 153).

A cache is usable only when all three files exist and SkyViewFactor and
svf.tif have the DSM tile's size and geotransform (within 1e-6); a damaged
file makes the cache unusable instead of raising (as the reference's
``_svf_cache_extent_matches``).

The GeoTIFF codec covers what these caches hold: classic TIFF (either byte
order), one band, strips or tiles, no compression or deflate (no predictor),
8/16/32-bit integer and 32/64-bit float samples; the geotransform comes from
ModelPixelScale + ModelTiepoint or ModelTransformation. Written files are
little-endian, Float32, uncompressed (the reference's GDAL default) or
deflate, with PixelScale/Tiepoint and a GeoKey directory (model type,
PixelIsArea raster, EPSG code when given).

Beside it, :func:`save_packed` / :func:`load_packed` keep the device's own
form: the bit-packed mask cube (20 bytes per pixel instead of 1836) and the
16 field planes.
"""
from __future__ import annotations

import io
import os
import struct
import zipfile
import zlib
from dataclasses import dataclass

import numpy as np

F32 = np.float32

#: raster names inside svfs_<key>.zip, in the reference's order (shadow.py:125-141)
ZIP_FIELDS = ("svf", "svfE", "svfS", "svfW", "svfN", "svfveg", "svfEveg", "svfSveg", "svfWveg", "svfNveg",
              "svfaveg", "svfEaveg", "svfSaveg", "svfWaveg", "svfNaveg")
CUBES = ("shadowmat", "vegshadowmat", "vbshmat")

_TYPES = {1: ("B", 1), 2: ("s", 1), 3: ("H", 2), 4: ("I", 4), 5: ("II", 8), 6: ("b", 1), 7: ("B", 1),
          8: ("h", 2), 9: ("i", 4), 10: ("ii", 8), 11: ("f", 4), 12: ("d", 8), 16: ("Q", 8)}


# ---------------------------------------------------------------------------
# GeoTIFF codec
# ---------------------------------------------------------------------------
def _metadata_xml(band_metadata) -> str:
    """GDAL's per-band metadata document (TIFF tag 42112): one Item per key, ``sample`` = band index."""
    from xml.sax.saxutils import escape, quoteattr
    items = [f"  <Item name={quoteattr(str(k))} sample=\"{b}\">{escape(str(v))}</Item>"
             for b, md in enumerate(band_metadata) for k, v in (md or {}).items()]
    return "<GDALMetadata>\n" + "\n".join(items) + "\n</GDALMetadata>\n" if items else ""


def write_geotiff(path_or_file, array, geotransform, epsg: int | None = None, compress: bool = False,
                  band_metadata=None):
    """Write a 2-D array (one band) or a 3-D array (bands, rows, cols) as a Float32 GeoTIFF (GDAL
    geotransform convention; bands stored one after another, PlanarConfiguration 2).
    ``band_metadata``: one dict per band, written as GDAL band metadata (e.g. ``{"Time": ...}``)."""
    a = np.ascontiguousarray(np.asarray(array, F32))
    if a.ndim == 2:
        a = a[None]
    if a.ndim != 3 or a.shape[0] < 1:
        raise ValueError("a 2-D band or a (bands, rows, cols) stack")
    nb, rows, cols = a.shape
    gt = [float(v) for v in geotransform]
    if len(gt) != 6:
        raise ValueError("geotransform needs 6 values")
    if band_metadata is not None and len(band_metadata) != nb:
        raise ValueError(f"band_metadata has {len(band_metadata)} entries for {nb} bands")
    rps = max(1, min(rows, 65536 // max(1, cols * 4)))
    strips = []
    for b in range(nb):
        for r0 in range(0, rows, rps):
            raw = a[b, r0:r0 + rps].astype("<f4").tobytes()
            strips.append(zlib.compress(raw, 6) if compress else raw)
    tags = []                          # (tag, type, values)
    tags.append((256, 4, [cols]))
    tags.append((257, 4, [rows]))
    tags.append((258, 3, [32] * nb))
    tags.append((259, 3, [8 if compress else 1]))
    tags.append((262, 3, [1]))
    tags.append((273, 4, [0] * len(strips)))           # patched below
    tags.append((277, 3, [nb]))
    tags.append((278, 4, [rps]))
    tags.append((279, 4, [len(s) for s in strips]))
    tags.append((284, 3, [2 if nb > 1 else 1]))
    if nb > 1:
        tags.append((338, 3, [0] * (nb - 1)))           # extra samples: unspecified data
    tags.append((339, 3, [3] * nb))
    xml = _metadata_xml(band_metadata or [])
    if xml:
        tags.append((42112, 2, xml))
    if gt[2] == 0.0 and gt[4] == 0.0:
        tags.append((33550, 12, [gt[1], -gt[5], 0.0]))
        tags.append((33922, 12, [0.0, 0.0, 0.0, gt[0], gt[3], 0.0]))
    else:
        tags.append((34264, 12, [gt[1], gt[2], 0.0, gt[0], gt[4], gt[5], 0.0, gt[3],
                                 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0]))
    keys = [(1024, 0, 1, 1 if epsg else 0), (1025, 0, 1, 1)]
    if epsg:
        keys.append((3072 if epsg not in (4326, 4269, 4258) else 2048, 0, 1, int(epsg)))
        if epsg in (4326, 4269, 4258):
            keys[0] = (1024, 0, 1, 2)
    gk = [1, 1, 0, len(keys)]
    for k in keys:
        gk.extend(k)
    tags.append((34735, 3, gk))
    tags.sort(key=lambda t: t[0])
    # layout: header, IFD, out-of-line values, strips
    n = len(tags)
    ifd_off = 8
    data_off = ifd_off + 2 + 12 * n + 4
    extra = bytearray()
    entries = []
    for tag, typ, vals in tags:
        fmt, size = _TYPES[typ]
        if typ == 2:                    # ASCII, NUL-terminated
            payload = vals.encode("latin-1") + b"\0"
            vals = payload
        else:
            payload = struct.pack("<" + fmt[0] * len(vals), *vals)
        if len(payload) <= 4:
            entries.append([tag, typ, len(vals), payload.ljust(4, b"\0"), None])
        else:
            entries.append([tag, typ, len(vals), None, len(extra)])
            extra += payload
            if len(extra) % 2:
                extra += b"\0"
    strip_off = data_off + len(extra)
    offs = []
    o = strip_off
    for s_ in strips:
        offs.append(o)
        o += len(s_)
    # patch the strip offsets (tag 273)
    for e in entries:
        if e[0] == 273:
            payload = struct.pack("<" + "I" * len(offs), *offs)
            if e[4] is None:
                e[3] = payload.ljust(4, b"\0")
            else:
                extra[e[4]:e[4] + len(payload)] = payload
    out = bytearray(b"II*\0" + struct.pack("<I", ifd_off))
    out += struct.pack("<H", n)
    for tag, typ, cnt, inline, off in entries:
        out += struct.pack("<HHI", tag, typ, cnt)
        out += inline if inline is not None else struct.pack("<I", data_off + off)
    out += struct.pack("<I", 0)
    out += extra
    for s_ in strips:
        out += s_
    if hasattr(path_or_file, "write"):
        path_or_file.write(bytes(out))
    else:
        with open(path_or_file, "wb") as f:
            f.write(bytes(out))


def _parse_ifd(buf: bytes):
    bo = buf[:2]
    if bo == b"II":
        e = "<"
    elif bo == b"MM":
        e = ">"
    else:
        raise ValueError("not a TIFF file")
    magic, ifd = struct.unpack(e + "HI", buf[2:8])
    if magic != 42:
        raise ValueError("not a classic TIFF file (BigTIFF is not supported)")
    n = struct.unpack(e + "H", buf[ifd:ifd + 2])[0]
    tags = {}
    for k in range(n):
        p = ifd + 2 + 12 * k
        tag, typ, cnt = struct.unpack(e + "HHI", buf[p:p + 8])
        if typ not in _TYPES:
            continue
        fmt, size = _TYPES[typ]
        nbytes = size * cnt
        if nbytes <= 4:
            raw = buf[p + 8:p + 8 + nbytes]
        else:
            off = struct.unpack(e + "I", buf[p + 8:p + 12])[0]
            raw = buf[off:off + nbytes]
        if typ == 2:
            tags[tag] = raw.rstrip(b"\0").decode("latin-1")
        elif typ in (5, 10):
            v = struct.unpack(e + fmt[0] * (2 * cnt), raw)
            tags[tag] = [v[2 * i] / v[2 * i + 1] if v[2 * i + 1] else 0.0 for i in range(cnt)]
        else:
            tags[tag] = list(struct.unpack(e + fmt * cnt, raw))
    return e, tags


def read_geotiff(path_or_bytes):
    """(float32 array, geotransform) of a single-band GeoTIFF."""
    r = read_geotiff_full(path_or_bytes)
    if r.data.shape[0] != 1:
        raise ValueError("only single-band rasters")
    return r.data[0], r.geotransform


@dataclass
class GeoRaster:
    """A GeoTIFF's bands as float32 (bands, rows, cols), its geotransform, EPSG code (from the GeoKey
    directory, None when absent) and per-band GDAL metadata."""

    data: np.ndarray
    geotransform: tuple | None
    epsg: int | None
    band_metadata: list

    @property
    def shape(self):
        return self.data.shape[1:]


def _geokeys_epsg(keys):
    """ProjectedCSTypeGeoKey, else GeographicTypeGeoKey, of a GeoKey directory (tag 34735)."""
    if not keys or len(keys) < 4:
        return None
    found = {}
    for k in range(keys[3]):
        kid, loc, cnt, val = keys[4 + 4 * k:8 + 4 * k]
        if loc == 0 and cnt == 1:
            found[kid] = val
    for kid in (3072, 2048):
        v = found.get(kid)
        if v and v != 32767:            # 32767: user-defined
            return int(v)
    return None


def _band_metadata(xml: str | None, nb: int) -> list:
    out = [dict() for _ in range(nb)]
    if not xml:
        return out
    import xml.etree.ElementTree as ET
    try:
        root = ET.fromstring(xml)
    except ET.ParseError:
        return out
    for item in root.iter("Item"):
        b = item.get("sample")
        if b is not None and b.isdigit() and int(b) < nb and item.get("name") is not None:
            out[int(b)][item.get("name")] = item.text or ""
    return out


def read_geotiff_full(path_or_bytes) -> GeoRaster:
    """All bands of a GeoTIFF (strips or tiles, pixel- or band-interleaved) with its georeferencing."""
    if isinstance(path_or_bytes, (bytes, bytearray)):
        buf = bytes(path_or_bytes)
    else:
        with open(path_or_bytes, "rb") as f:
            buf = f.read()
    e, t = _parse_ifd(buf)
    cols, rows = t[256][0], t[257][0]
    spp = t.get(277, [1])[0]
    planar = t.get(284, [1])[0]
    bits = t.get(258, [1])[0]
    fmt = t.get(339, [1])[0]
    comp = t.get(259, [1])[0]
    if comp not in (1, 8, 32946):
        raise ValueError(f"TIFF compression {comp} not supported")
    if t.get(317, [1])[0] != 1:
        raise ValueError("TIFF predictor not supported")
    kind = {(3, 32): "f4", (3, 64): "f8", (1, 8): "u1", (1, 16): "u2", (1, 32): "u4", (2, 8): "i1", (2, 16): "i2",
            (2, 32): "i4"}.get((fmt, bits))
    if kind is None:
        raise ValueError(f"TIFF sample format {fmt}/{bits} not supported")
    dt = np.dtype(kind).newbyteorder(e)

    def chunk(off, cnt):
        raw = buf[off:off + cnt]
        return zlib.decompress(raw) if comp in (8, 32946) else raw
    # one plane per band when band-interleaved; a single plane of spp samples per pixel otherwise
    sep = planar == 2 or spp == 1
    per = 1 if sep else spp
    nplanes = spp if sep else 1
    out = np.zeros((nplanes, rows, cols, per), dt)
    if 322 in t:                       # tiles
        tw, th = t[322][0], t[323][0]
        offs, cnts = t[324], t[325]
        across = -(-cols // tw)
        per_plane = across * -(-rows // th)
        for k, (o, c) in enumerate(zip(offs, cnts)):
            p, kk = divmod(k, per_plane)
            ty, tx = divmod(kk, across)
            tile = np.frombuffer(chunk(o, c), dt, count=tw * th * per).reshape(th, tw, per)
            r0, c0 = ty * th, tx * tw
            h, w = min(th, rows - r0), min(tw, cols - c0)
            out[p, r0:r0 + h, c0:c0 + w] = tile[:h, :w]
    else:
        rps = t.get(278, [rows])[0]
        per_plane = -(-rows // rps)
        for k, (o, c) in enumerate(zip(t[273], t[279])):
            p, kk = divmod(k, per_plane)
            r0 = kk * rps
            h = min(rps, rows - r0)
            out[p, r0:r0 + h] = np.frombuffer(chunk(o, c), dt, count=h * cols * per).reshape(h, cols, per)
    out = out[:, :, :, 0] if sep else np.moveaxis(out[0], -1, 0)
    gt = None
    if 34264 in t:
        m = t[34264]
        gt = (m[3], m[0], m[1], m[7], m[4], m[5])
    elif 33550 in t and 33922 in t:
        sx, sy = t[33550][0], t[33550][1]
        i, j, _, x, y, _ = t[33922][:6]
        gt = (x - i * sx, sx, 0.0, y + j * sy, 0.0, -sy)
    return GeoRaster(np.ascontiguousarray(out.astype(F32)), gt, _geokeys_epsg(t.get(34735)),
                     _band_metadata(t.get(42112), out.shape[0]))


def geotransforms_equal(a, b, tol: float = 1e-6) -> bool:
    """Geotransforms agree within ``tol`` map units (``utci_process.py:250-256``)."""
    return a is not None and b is not None and len(a) == 6 and len(b) == 6 and all(
        abs(x - y) <= tol for x, y in zip(a, b))


# ---------------------------------------------------------------------------
# the reference's cache files
# ---------------------------------------------------------------------------
def cache_paths(svf_dir: str, key: str):
    """(SkyViewFactor_<key>.tif, svfs_<key>.zip, shadowmats_<key>.npz) in ``svf_dir``."""
    return (os.path.join(svf_dir, f"SkyViewFactor_{key}.tif"), os.path.join(svf_dir, f"svfs_{key}.zip"),
            os.path.join(svf_dir, f"shadowmats_{key}.npz"))


def _np(x):
    if hasattr(x, "detach"):
        x = x.detach().cpu().numpy()
    return np.asarray(x, F32)


def save_svf_cache(svf_dir: str, key: str, svf_tuple, geotransform, epsg: int | None = None):
    """Write the 19-tuple of ``svf_calculator`` (reference order, ``shadow.py:625``) as the
    reference's cache files (``shadow.py:98-164``). Returns the three paths."""
    from .api import REF_ORDER
    vals = dict(zip(REF_ORDER, svf_tuple))
    os.makedirs(svf_dir, exist_ok=True)
    tot, zpath, npz = cache_paths(svf_dir, key)
    tmp = zpath + ".tmp"
    with zipfile.ZipFile(tmp, "w", compression=zipfile.ZIP_DEFLATED) as zf:
        for name in ZIP_FIELDS:
            b = io.BytesIO()
            write_geotiff(b, _np(vals[name]), geotransform, epsg)
            zf.writestr(f"{name}.tif", b.getvalue())
    os.replace(tmp, zpath)
    write_geotiff(tot, _np(vals["svftotal"]), geotransform, epsg)
    np.savez_compressed(npz, shadowmat=_np(vals["shmat"]), vegshadowmat=_np(vals["vegshmat"]),
                        vbshmat=_np(vals["vbshvegshmat"]))
    return tot, zpath, npz


def cache_matches(svf_dir: str, key: str, size, geotransform) -> bool:
    """All three files present and SkyViewFactor / svf.tif of ``size`` = (rows, cols) and
    ``geotransform`` (``utci_process.py:259-313``); damaged files count as a mismatch."""
    tot, zpath, npz = cache_paths(svf_dir, key)
    if not (os.path.isfile(tot) and os.path.isfile(zpath) and os.path.isfile(npz)):
        return False
    try:
        a, gt = read_geotiff(tot)
        with zipfile.ZipFile(zpath) as zf:
            b, gt2 = read_geotiff(zf.read("svf.tif"))
        for arr, g in ((a, gt), (b, gt2)):
            if arr.shape != tuple(size) or not geotransforms_equal(g, geotransform):
                return False
    except (ValueError, KeyError, OSError, zipfile.BadZipFile, zlib.error, struct.error, IndexError):
        return False
    return True


def load_svf_cache(svf_dir: str, key: str, size=None, geotransform=None):
    """The 19-tuple (reference order) from the cache files as float32 numpy arrays.

    With ``size`` and ``geotransform`` the cache must match the DSM tile
    (else ``ValueError``, as ``load_cached_svf_outputs``, ``utci_process.py:340-380``).
    The tuple can be passed to ``run_utci_window`` as ``svf_bundle``.
    """
    from .api import REF_ORDER
    if size is not None and not cache_matches(svf_dir, key, size, geotransform):
        raise ValueError(f"Cached SVF outputs for tile {key} do not match the DSM tile's extent; refusing to "
                         "load a stale/foreign SVF cache. Recompute SVF for this tile or restore the matching cache.")
    tot, zpath, npz = cache_paths(svf_dir, key)
    vals = {}
    with zipfile.ZipFile(zpath) as zf:
        for name in ZIP_FIELDS:
            vals[name] = read_geotiff(zf.read(f"{name}.tif"))[0]
    vals["svftotal"] = read_geotiff(tot)[0]
    with np.load(npz, allow_pickle=False) as z:
        vals["shmat"] = np.asarray(z["shadowmat"], F32)
        vals["vegshmat"] = np.asarray(z["vegshadowmat"], F32)
        vals["vbshvegshmat"] = np.asarray(z["vbshmat"], F32)
    return tuple(vals[k] for k in REF_ORDER)


# ---------------------------------------------------------------------------
# packed form (the device's layout)
# ---------------------------------------------------------------------------
def save_packed(path: str, masks, fields, rows: int, cols: int, geotransform=None):
    """Bit-packed masks (uint32 [20][n]) and the 16 field planes (float32 [16][n]) of a scene."""
    m = masks.detach().cpu().numpy() if hasattr(masks, "detach") else np.asarray(masks)
    f = fields.detach().cpu().numpy() if hasattr(fields, "detach") else np.asarray(fields)
    np.savez_compressed(path, masks=m.astype(np.int32).view(np.uint32), fields=f.astype(F32),
                        shape=np.array([rows, cols], np.int64),
                        geotransform=np.array(geotransform if geotransform is not None else [np.nan] * 6))


def load_packed(path: str):
    """(masks uint32 [20][n], fields float32 [16][n], (rows, cols), geotransform or None)."""
    with np.load(path, allow_pickle=False) as z:
        gt = z["geotransform"]
        return (z["masks"].copy(), z["fields"].copy(), tuple(int(v) for v in z["shape"]),
                None if np.isnan(gt).any() else tuple(float(v) for v in gt))
