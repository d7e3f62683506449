"""SVF cache files without GDAL: the GeoTIFF codec against Pillow's TIFF
reader/writer, the reference's cache layout, extent checks and damaged files."""
import io
import os
import zipfile

import numpy as np
import pytest

from solweig_project_amd import svf_cache as SC

GT = (500000.0, 1.0, 0.0, 4200000.0, 0.0, -1.0)


def _arr(rows=37, cols=53, seed=0):
    a = np.random.default_rng(seed).uniform(-5, 50, (rows, cols)).astype(np.float32)
    a[3, 4] = np.nan
    return a


@pytest.mark.parametrize("compress", [False, True])
def test_geotiff_round_trip(tmp_path, compress):
    a = _arr()
    p = tmp_path / "x.tif"
    SC.write_geotiff(p, a, GT, epsg=32633, compress=compress)
    b, gt = SC.read_geotiff(p)
    assert np.array_equal(a, b, equal_nan=True)
    assert SC.geotransforms_equal(gt, GT, 0.0)


def test_rotated_geotransform_round_trip(tmp_path):
    gt0 = (100.0, 0.5, 0.1, 200.0, 0.05, -0.5)
    p = tmp_path / "r.tif"
    SC.write_geotiff(p, _arr(5, 7), gt0)
    assert SC.geotransforms_equal(SC.read_geotiff(p)[1], gt0, 1e-12)


def test_pillow_reads_our_files_and_we_read_pillows():
    PIL = pytest.importorskip("PIL.Image")
    a = _arr(20, 30, seed=1)
    a[3, 4] = 7.0                      # Pillow's float image compare below ignores NaN handling
    for compress in (False, True):
        b = io.BytesIO()
        SC.write_geotiff(b, a, GT, compress=compress)
        im = PIL.open(io.BytesIO(b.getvalue()))
        assert im.mode == "F" and im.size == (30, 20)
        assert np.array_equal(np.asarray(im, np.float32), a)
        tags = im.tag_v2
        assert tuple(tags[33550]) == (1.0, 1.0, 0.0)
        assert tuple(tags[33922])[3:5] == (500000.0, 4200000.0)
    # Pillow-written float TIFF (no geo tags) read back by our codec
    out = io.BytesIO()
    PIL.fromarray(a, mode="F").save(out, format="TIFF")
    got, gt = SC.read_geotiff(out.getvalue())
    assert np.array_equal(got, a) and gt is None


def _tuple(rows=12, cols=9, seed=2):
    from solweig_project_amd.api import REF_ORDER
    rng = np.random.default_rng(seed)
    out = []
    for name in REF_ORDER:
        if name in ("shmat", "vegshmat"):
            out.append(rng.integers(0, 2, (rows, cols, 153)).astype(np.float32))
        elif name == "vbshvegshmat":
            out.append(rng.integers(0, 3, (rows, cols, 153)).astype(np.float32))
        else:
            out.append(rng.uniform(0, 1, (rows, cols)).astype(np.float32))
    return tuple(out)


def test_cache_round_trip_and_layout(tmp_path):
    t = _tuple()
    tot, zpath, npz = SC.save_svf_cache(str(tmp_path), "0_0", t, GT)
    assert os.path.basename(zpath) == "svfs_0_0.zip" and os.path.basename(npz) == "shadowmats_0_0.npz"
    with zipfile.ZipFile(zpath) as zf:
        assert sorted(zf.namelist()) == sorted(f"{n}.tif" for n in SC.ZIP_FIELDS)
    with np.load(npz) as z:
        assert sorted(z.files) == sorted(SC.CUBES)
    assert SC.cache_matches(str(tmp_path), "0_0", (12, 9), GT)
    got = SC.load_svf_cache(str(tmp_path), "0_0", (12, 9), GT)
    for x, y in zip(got, t):
        assert np.array_equal(x, y)


def test_cache_rejects_other_extents_and_damaged_files(tmp_path):
    t = _tuple()
    tot, zpath, npz = SC.save_svf_cache(str(tmp_path), "1_2", t, GT)
    assert not SC.cache_matches(str(tmp_path), "1_2", (12, 10), GT)
    moved = (GT[0] + 1.0,) + GT[1:]
    assert not SC.cache_matches(str(tmp_path), "1_2", (12, 9), moved)
    with pytest.raises(ValueError, match="do not match"):
        SC.load_svf_cache(str(tmp_path), "1_2", (12, 9), moved)
    # a truncated SkyViewFactor tif makes the cache unusable, not an exception
    data = open(tot, "rb").read()
    with open(tot, "wb") as f:
        f.write(data[:40])
    assert not SC.cache_matches(str(tmp_path), "1_2", (12, 9), GT)
    os.remove(npz)
    assert not SC.cache_matches(str(tmp_path), "1_2", (12, 9), GT)


def test_packed_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    masks = rng.integers(0, 2 ** 32, (20, 30), dtype=np.uint64).astype(np.uint32)
    fields = rng.uniform(0, 1, (16, 30)).astype(np.float32)
    SC.save_packed(str(tmp_path / "p.npz"), masks, fields, 5, 6, GT)
    m, f, shape, gt = SC.load_packed(str(tmp_path / "p.npz"))
    assert np.array_equal(m, masks) and np.array_equal(f, fields) and shape == (5, 6) and gt == GT
