"""The benchmark configurations at their real sizes on the GPU.

* The seeded 1000 x 1000 x 24 h generator tile (the headline workload) against the oracle: every
  packed sh / vegsh / vbsh mask bit of the 153 patches, the eight sun-march planes of every daytime
  step bit for bit, SVF fields <= 1e-5, shadow exact, Tmrt <= 0.01 K, UTCI <= 0.01 C (same NaN
  pattern), K/L fluxes <= 0.01 W/m2.
* The batched workload at its real shapes: tiles 300^2, 700^2, 1200^2, 2000^2 in one batch, and 64
  edit scenarios of a 500^2 base in another, each item bit-identical to its single-tile solve; the
  300^2 tile also against the oracle.
* The incremental path: after a chain of edits, the edited site against the oracle's full solve of
  the edited inputs.
* The serving loop on a 2000^2 site: every request of a Poisson mix equals the same request served
  alone on the site version it was admitted against.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

OUTS = ("tmrt", "utci", "shadow", "kdown", "kup", "ldown", "lup")


def _say(msg):
    print(f"[gpu-test] {msg}", flush=True)


@pytest.fixture(scope="module")
def solver():
    from solweig_project_amd import _lib
    from solweig_project_amd.solver import HipSolver
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a ROCm device")
    _lib.load()
    return HipSolver("cuda:0")


def _pack_cube(cube, value=1.0):
    """(153, n) float cube -> (5, n) uint32 words: bit p % 32 of word p // 32 where cube == value."""
    n = cube.shape[1]
    bits = np.zeros((n, 160), np.uint8)
    bits[:, :153] = (cube == np.float32(value)).T
    return np.packbits(bits, axis=1, bitorder="little").view("<u4").T.copy()


def _oracle_masks(shm, vm, bm):
    return np.stack([_pack_cube(shm), _pack_cube(vm), _pack_cube(bm, 1.0), _pack_cube(bm, 2.0)])


def _check_outputs(out, ref, label):
    assert np.array_equal(out["shadow"], ref["shadow"]), label
    assert np.abs(out["tmrt"] - ref["tmrt"]).max() <= 0.01, label
    m = ~np.isnan(ref["utci"])
    assert np.array_equal(np.isnan(out["utci"]), ~m), label
    assert np.abs(out["utci"][m] - ref["utci"][m]).max() <= 0.01, label
    for k in ("kdown", "kup", "ldown", "lup"):
        if k in out:
            assert np.abs(out[k] - ref[k]).max() <= 0.01, (label, k)


def _oracle_solve(site_like, met, sun, latitude, record=None):
    from oracle import driver as D
    sc = D.compose(site_like.dsm, site_like.dem, site_like.cdsm, site_like.scale)
    fields, shm, vm, bm = D.svf(sc)
    outs = D.day(sc, fields, shm, vm, bm, site_like.walls, site_like.aspect, site_like.landcover, met, sun,
                 latitude, record=record)
    return fields, (shm, vm, bm), outs


def test_headline_tile_1000x1000x24_matches_oracle(solver):
    from solweig_project_amd import sunpos, synth
    from solweig_project_amd.solver import SUN_PLANES, TileInputs, plan_day
    _say("1000x1000x24 h generator tile (seed 20261203) vs oracle")
    site = synth.make_site(1000, 1000, seed=20261203)
    sun = sunpos.day_sun_table(site.latitude, site.longitude, site.utc_offset, 2021, 180)
    met = synth.make_met(sun.altitude, seed=20261204)
    rec = []
    fields, cubes, ref = _oracle_solve(site, met, sun, site.latitude, record=rec)
    _say("oracle done")
    tile = TileInputs.from_site(site)
    got_f, out, ds = solver.solve(tile, met, sun, site.latitude, outputs=OUTS)
    n = 1000 * 1000
    masks = ds.masks.cpu().numpy().view(np.uint32).reshape(4, 5, n)
    assert np.array_equal(masks, _oracle_masks(*cubes))
    del cubes
    assert np.abs(got_f - fields).max() <= 1e-5
    _check_outputs(out, ref, "1000x1000")
    # the day loop's sun march + wall tail, every daytime step, bit for bit
    plan = plan_day(ds.tile, met, sun, site.latitude)
    dplan = solver.upload_plan(plan)
    days = [i for i in range(plan.n_steps) if plan.is_day[i]]
    assert days == [r["i"] for r in rec]
    for r in rec:
        got = solver.sun_planes(ds, plan, dplan, r["i"]).cpu().numpy()
        for q, nm in enumerate(SUN_PLANES):
            assert np.array_equal(got[q].view(np.uint32), r["sun"][q].view(np.uint32)), (r["i"], nm)


@pytest.fixture(scope="module")
def batch_env(solver):
    from solweig_project_amd import sunpos, synth
    base = synth.make_site(500, 500, seed=20261203 + 20, with_aspect=False)
    sun = sunpos.day_sun_table(base.latitude, base.longitude, base.utc_offset, 2021, 180)
    met = synth.make_met(sun.altitude, seed=20261204)
    return dict(base=base, sun=sun, met=met)


def _with_walls(base, dsm, cdsm, lc):
    from solweig_project_amd import synth
    from solweig_project_amd import walls as WG
    from solweig_project_amd.solver import TileInputs
    w, a = WG.walls_and_aspect(dsm, base.scale)
    return TileInputs.from_site(synth.site_with(base, dsm, cdsm, lc, walls=w, aspect=a))


def _single(solver, tile, env):
    ds = solver.upload(tile)
    _, out = solver.solve_device(ds, env["met"], env["sun"], env["base"].latitude, outputs=OUTS)
    torch.cuda.synchronize()
    return out.clone()


def _bits(t):
    return t.contiguous().view(torch.int32)


def test_batched_tiles_at_real_shapes_match_single_solves_and_oracle(solver, batch_env):
    from solweig_project_amd import synth
    from solweig_project_amd.solver import OUT_ORDER
    tiles = []
    for k, n in enumerate((300, 700, 1200, 2000)):
        s = synth.make_site(n, n, seed=20261203 + 10 + k, with_aspect=False)
        tiles.append((s, _with_walls(s, s.dsm, s.cdsm, s.landcover)))
    _say("batch of tiles 300^2, 700^2, 1200^2, 2000^2")
    refs = [_single(solver, t, batch_env) for _, t in tiles]
    scenes = [solver.upload(t) for _, t in tiles]
    batch = solver.solve_batch(scenes, batch_env["met"], batch_env["sun"], batch_env["base"].latitude, outputs=OUTS)
    torch.cuda.synchronize()
    for (s, _), got, want in zip(tiles, batch.outs, refs):
        assert torch.equal(_bits(got), _bits(want)), s.dsm.shape
    # the smallest tile of the batch against the oracle
    s, t = tiles[0]
    site300 = synth.site_with(s, s.dsm, s.cdsm, s.landcover, walls=t.walls, aspect=t.aspect_deg)
    _, _, ref = _oracle_solve(site300, batch_env["met"], batch_env["sun"], s.latitude)
    names = [nm for nm in OUT_ORDER if nm in OUTS]
    host = batch.outs[0].cpu().numpy()
    _check_outputs({nm: host[:, j] for j, nm in enumerate(names)}, ref, "300^2 in the batch")


def test_64_edit_scenarios_of_500_batched_match_single_solves(solver, batch_env):
    from solweig_project_amd import synth
    base = batch_env["base"]
    edits = [_with_walls(base, d, c, lc) for d, c, lc, _ in synth.edit_scenarios(base, 64, seed=20261203 + 21)]
    _say(f"batch of {len(edits)} edit scenarios of 500^2")
    refs = [_single(solver, t, batch_env) for t in edits]
    scenes = [solver.upload(t) for t in edits]
    batch = solver.solve_batch(scenes, batch_env["met"], batch_env["sun"], base.latitude, outputs=OUTS)
    torch.cuda.synchronize()
    for k, (got, want) in enumerate(zip(batch.outs, refs)):
        assert torch.equal(_bits(got), _bits(want)), k


def test_edited_site_matches_the_oracle(solver):
    """A chain of edits through the incremental path ends on the oracle's full solve of the edited inputs."""
    from solweig_project_amd import incremental as I
    from solweig_project_amd import synth, sunpos
    from solweig_project_amd.solver import OUT_ORDER
    base = synth.make_site(200, 200, seed=83, with_aspect=False)
    base.dsm = (base.dsm * np.float32(0.3)).astype(np.float32)
    sun = sunpos.day_sun_table(base.latitude, base.longitude, base.utc_offset, 2021, 180)
    met = synth.make_met(sun.altitude, seed=84)
    st = I.solve_site(solver, _with_walls(base, base.dsm, base.cdsm, base.landcover), met, sun, base.latitude,
                      outputs=OUTS)
    dsm, cdsm, lc = base.dsm.copy(), base.cdsm.copy(), base.landcover.copy()
    ii, jj = np.mgrid[0:200, 0:200]
    ground = np.argwhere(dsm == 0)
    modes = []
    for k, (ci, cj) in enumerate(ground[[len(ground) // 5, len(ground) // 2, 4 * len(ground) // 5]].tolist()):
        d2 = (ii - ci) ** 2 + (jj - cj) ** 2
        crown = (float(4 + 3 * k) * np.exp(-d2 / 8.0)).astype(np.float32)
        crown[(d2 > 16) | (dsm > 0)] = 0
        cdsm = np.maximum(cdsm, crown)
        modes.append(I.apply_edit(solver, st, dsm, cdsm, lc)["mode"])
    dsm[120:128, 40:52] = np.maximum(dsm[120:128, 40:52], np.float32(5.0))
    lc[120:128, 40:52] = 2
    cdsm[120:128, 40:52] = 0
    modes.append(I.apply_edit(solver, st, dsm, cdsm, lc)["mode"])
    _say(f"edit modes {modes}")
    assert "tiles" in modes
    torch.cuda.synchronize()
    edited = synth.site_with(base, dsm, cdsm, lc, walls=st.tile.walls, aspect=st.tile.aspect_deg)
    fields, _, ref = _oracle_solve(edited, met, sun, base.latitude)
    assert np.abs(st.fields.reshape(16, 200, 200).cpu().numpy() - fields).max() <= 1e-5
    names = [nm for nm in OUT_ORDER if nm in OUTS]
    host = st.out.cpu().numpy()
    _check_outputs({nm: host[:, j] for j, nm in enumerate(names)}, ref, "edited site")


class _InlinePlanner:
    """The planner interface without a child process (planned on submit)."""

    class _Done:
        def __init__(self, v):
            self.v = v

        def ready(self):
            return True

        def get(self, timeout=None):
            return self.v

    def submit(self, tile, job):
        from solweig_project_amd.solver import plan_day
        return self._Done(plan_day(tile, job.met, job.sun, job.latitude))


def test_serving_mix_on_a_2000_site_equals_requests_alone(solver):
    from solweig_project_amd import incremental as I
    from solweig_project_amd import serving as SV
    from solweig_project_amd import synth, sunpos
    base = synth.make_site(2000, 2000, seed=20261203 + 40, with_aspect=False)
    tile = _with_walls(base, base.dsm, base.cdsm, base.landcover)
    sun = sunpos.day_sun_table(base.latitude, base.longitude, base.utc_offset, 2021, 180)
    mets = [synth.make_met(sun.altitude, seed=20261203 + 41 + k) for k in range(3)]
    st = I.solve_site(solver, tile, mets[0], sun, base.latitude)
    reqs = SV.poisson_workload(24, rate=20.0, mix={"edit": 0.6, "met": 0.3, "full": 0.1}, seed=7,
                               site_tile=st.tile, met_pool=mets)
    _say(f"serving {len(reqs)} requests on 2000^2: {[r.kind for r in reqs]}")
    server = SV.Server(solver, st, mets, sun, base.latitude, _InlinePlanner(), slots=3)
    done = server.run(reqs)
    assert all(r.finish > 0 for r in done)
    st2 = I.solve_site(solver, tile, mets[0], sun, base.latitude)
    alone = SV.Server(solver, st2, mets, sun, base.latitude, _InlinePlanner(), slots=1)
    order = sorted(done, key=lambda r: (r.payload["version"], 0 if r.kind == "edit" else 1))
    for r in order:
        twin = SV.Request(r.kind, 0.0, {k: v for k, v in r.payload.items() if k != "version"})
        alone.run([twin])
        assert twin.result == r.result, (r.kind, r.payload)
