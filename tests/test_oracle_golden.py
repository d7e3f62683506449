"""The CPU oracle against the reference's own outputs (golden vectors).

Golden vectors were written by ``tests/golden/make_golden.py`` from the
reference implementation (torch CPU). Bars: shadow/vegetation/wall masks
bit-exact; SVF <= 1e-5; Tmrt <= 0.01 K; UTCI <= 0.01 C.
"""
import numpy as np
import pytest

from oracle import driver as D
from solweig_project_amd import met as M

SCENES = ["a48", "b40x56"]


def _load(golden_dir, name):
    return np.load(golden_dir / f"svf_{name}.npz"), np.load(golden_dir / f"day_{name}.npz")


@pytest.fixture(scope="module", params=SCENES)
def solved(request, golden_dir):
    g, d = _load(golden_dir, request.param)
    sc = D.compose(g["dsm"], g["dem"], g["cdsm"], float(g["scale"]))
    fields, shm, vm, bm = D.svf(sc)
    sun = M.SunTable(d["altitude"][0], d["azimuth"][0], d["zen"][0], d["jday"][0], d["dectime"], d["altmax"][0])
    rec = []
    outs = D.day(sc, fields, shm, vm, bm, g["walls"], g["aspect"], g["landcover"], d["met"], sun,
                 float(d["latitude"]), record=rec)
    return dict(g=g, d=d, sc=sc, fields=fields, cubes=(shm, vm, bm), outs=outs, rec=rec)


def test_amax_and_compose(solved):
    assert np.float32(solved["sc"].amax) == solved["g"]["amax"]


def test_svf_masks_bit_exact(solved):
    g = solved["g"]
    rows, cols = solved["sc"].a.shape
    for mine, key in zip(solved["cubes"], ("shmat", "vegshmat", "vbshvegshmat")):
        got = mine.reshape(153, rows, cols).transpose(1, 2, 0)
        assert np.array_equal(got, g[key].astype(np.float32)), key


def test_svf_fields(solved):
    g = solved["g"]
    for k, name in enumerate(D.FOLD_NAMES):
        assert np.abs(solved["fields"][k] - g[name]).max() <= 1e-5, name


def test_sun_march_planes_bit_exact(solved):
    d = solved["d"]
    names = ["vegsh", "sh", "vbsh", "wallsh", "wallsun", "wallshve", "facesh", "facesun"]
    assert len(solved["rec"]) == d["m_sh"].shape[0]
    for j, r in enumerate(solved["rec"]):
        for q, nm in enumerate(names):
            # bit for bit, signed zeros included
            assert np.array_equal(r["sun"][q].view(np.uint32), d["m_" + nm][j].view(np.uint32)), (j, nm)


def test_day_outputs(solved):
    d = solved["d"]
    o = solved["outs"]
    assert np.array_equal(o["shadow"], d["o_shadow"])
    assert np.abs(o["tmrt"] - d["o_tmrt"]).max() <= 0.01
    ref = d["o_utci"]
    assert np.array_equal(np.isnan(o["utci"]), np.isnan(ref))
    m = ~np.isnan(ref)
    assert np.abs(o["utci"][m] - ref[m]).max() <= 0.01
    for k in ("kdown", "kup", "ldown", "lup"):
        assert np.abs(o[k] - d["o_" + k]).max() <= 0.01, k


def test_utci_grid(golden_dir):
    z = np.load(golden_dir / "utci_grid.npz")
    got = D.utci(z["ta"], z["rh"], z["tmrt"], z["va"])
    ref = z["utci"]
    inv = ref == -999
    assert np.array_equal(got == -999, inv)
    assert np.abs(got[~inv] - ref[~inv]).max() <= 0.01


def test_planted_dz_error_is_caught(golden_dir):
    """A one-ulp change of one dz value of one patch table must flip masks."""
    from solweig_project_amd import skyvault as SV
    g, _ = _load(golden_dir, "a48")
    sc = D.compose(g["dsm"], g["dem"], g["cdsm"], float(g["scale"]))
    rows, cols = sc.a.shape
    vault = SV.build_skyvault()
    offs, dx, dy, dz = SV.svf_tables(vault, sc.scale, np.float32(sc.amax), rows, cols)
    flipped = 0
    for p in (0, 5, 40):
        s, e = offs[p], offs[p + 1]
        tab = SV.StepTable(dx[s:e].copy(), dy[s:e].copy(), dz[s:e].copy(), np.zeros(e - s, np.float32))
        base = D.march_svf_patch(sc, tab)
        tab.dz[:] = np.nextafter(tab.dz, np.float32(-np.inf)) * np.float32(0.5)
        moved = D.march_svf_patch(sc, tab)
        flipped += sum(int((x != y).sum()) for x, y in zip(base, moved))
    assert flipped > 0


@pytest.mark.parametrize("name", SCENES)
def test_recorded_sun_tables_reproduce_the_reference_planes(golden_dir, name):
    """The day-loop tables of ``sun_tabs_day.npz`` (grid-edge tables, amax stop applied by the
    march) give the reference's eight planes of every daytime step through the oracle, bit for bit."""
    g, d = _load(golden_dir, name)
    z = np.load(golden_dir / "sun_tabs_day.npz")
    sc = D.compose(g["dsm"], g["dem"], g["cdsm"], float(g["scale"]))
    rows, cols = sc.a.shape
    n = rows * cols
    walls = np.ascontiguousarray(g["walls"], np.float32)
    asp = ((np.ascontiguousarray(g["aspect"], np.float32) * np.float32(np.pi)) / np.float32(180.0)).astype(np.float32)
    sun = M.SunTable(d["altitude"][0], d["azimuth"][0], d["zen"][0], d["jday"][0], d["dectime"], d["altmax"][0])
    loop = M.DayLoop(d["met"], sun, latitude=float(d["latitude"]), landcover=M.landcover_constants(True),
                     scale=sc.scale, amax=sc.amax, rows=rows, cols=cols)
    tab = z[f"{name}_sun_tab"]
    names = ["vegsh", "sh", "vbsh", "wallsh", "wallsun", "wallshve", "facesh", "facesun"]
    j = 0
    out = np.empty(8 * n, np.float32)
    work = np.empty(25 * n, np.float32)
    for i in np.nonzero(z[f"{name}_is_day"])[0]:
        sp = loop.step(int(i))
        o, c = int(z[f"{name}_sun_off"][i]), int(z[f"{name}_sun_count"][i])
        t = tab[o:o + c]
        dz = np.ascontiguousarray(t[:, 2]).view(np.float32)
        stop = np.nonzero(~(np.float32(sc.amax) >= dz))[0]
        cnt = int(stop[0]) + 1 if stop.size else c
        dx, dy = (np.ascontiguousarray(t[:cnt, k], np.int32) for k in (0, 1))
        dzc = np.ascontiguousarray(dz[:cnt])
        dzp = np.ascontiguousarray(np.ascontiguousarray(t[:cnt, 3]).view(np.float32))
        D.lib().or_march_sun(D._p(sc.a), D._p(sc.vegdsm), D._p(sc.vegdsm2), rows, cols, D._p(dx), D._p(dy),
                             D._p(dzc), D._p(dzp), cnt, D._p(walls), D._p(asp.ravel()), sp.facesh_branch,
                             sp.azilow, sp.azihigh, D._p(out), D._p(work))
        got = out.reshape(8, rows, cols)
        for q, nm in enumerate(names):
            assert np.array_equal(got[q].view(np.uint32), d["m_" + nm][j].view(np.uint32)), (int(i), nm)
        j += 1
    assert j == d["m_sh"].shape[0]
