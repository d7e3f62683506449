#!/usr/bin/env python3
"""SOLWEIG full-solve benchmark on one MI355X.

One "step" is one full solve of the 1000 x 1000 x 24 h synthetic tile
(scene compose, 153-patch SVF sweep, 24 hourly steps of sun march, ground
view factors, flux assembly, Tmrt and UTCI), timed from host-resident
inputs to host-resident outputs: each solve uploads the six input rasters
from pinned host memory and delivers its 16 SVF fields and every step's
Tmrt, UTCI and shadow planes into pinned host buffers of its own (the
copies stream out while the GPU computes). The host scalar chain of each
solve's day (clearness index, Perez luminance, sun tables, step records) is
computed inside the timed region, in planner processes ahead of the GPU.
The device-resident rate (inputs in HBM, outputs left in HBM) is reported
beside it. Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# Hardware queues per process (read when the HIP runtime starts, so before torch): the runtime puts
# streams on its queues round robin, and with the default 4, two of the host-to-host timed region's
# five streams (null, three lanes, the copy stream) would share a queue, so one lane would wait
# behind the other's work or behind the downloads (11.2-12.7 against 10.3 ms per solve with 8).
os.environ["GPU_MAX_HW_QUEUES"] = os.environ.get("SW_BENCH_HW_QUEUES", "8")

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0      # dense BF16 MFMA
# VALU issue on MI355X (gfx950), measured with single-class FMA/add/compare/select/min-max/bit-field
# loops: FP32 arithmetic, compares, selects, min/max and bit-field ops issue one wave64 instruction
# per CU per clock (562-624 G/s measured of 614.4), 32-bit integer add and logic ops two (1138 G/s
# of 1228.8), FP32 transcendentals one every two clocks. The roofline peak is the hardware maximum,
# two per CU per clock; `frac_full_rate_issue` is the fraction of the one-per-clock rate at which
# all but the integer add/logic ops issue.
N_CU = 256
CLOCK_GHZ = 2.4
VALU_FULL_RATE_GINSTR = N_CU * CLOCK_GHZ
VALU_PEAK_GINSTR = 2 * VALU_FULL_RATE_GINSTR


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rows", type=int, default=1000)
    ap.add_argument("--cols", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=0,
                    help="side of the CPU-oracle tile (0: the benchmark tile itself)")
    ap.add_argument("--workload", choices=("tile", "batch", "edits", "surrogate", "serve"), default="tile",
                    help="tile: one 1000x1000 tile (default); batch: batched tiles 300/700/1200/2000 squared plus "
                         "64 edit scenarios of a 500x500 base; edits: single-tree edits of a 2000x2000 site "
                         "(p50/p99 latency); surrogate: train the SVF surrogate, then gated surrogate vs "
                         "exact SVF on a 2000x2000 city; serve: Poisson mix of edits, met-only and full "
                         "re-solves of a 2000x2000 site at --rate requests/s")
    ap.add_argument("--requests", type=int, default=1000, help="requests served by --workload serve")
    ap.add_argument("--rate", type=float, default=20.0, help="arrival rate of --workload serve (per s)")
    ap.add_argument("--lanes", type=int, default=3,
                    help="tile workload: consecutive solves alternate over this many scene copies on their own streams")
    ap.add_argument("--h2h-lanes", type=int, default=3,
                    help="tile workload: lanes of the host-to-host timed region")
    ap.add_argument("--h2h-copy", choices=("own", "shared", "per-lane", "thread"), default="thread",
                    help="tile workload: where the host-to-host copies run: issued by a copy thread on a stream "
                         "of their own once each chunk is computed (the lanes never wait for a copy), on each "
                         "lane's stream, on one shared copy stream, or on a copy stream per lane")
    ap.add_argument("--h2h-chunk", type=int, default=8,
                    help="tile workload: day steps per chunk of the host-to-host region (each chunk's planes are "
                         "copied out right after it)")
    ap.add_argument("--h2h-fields", choices=("last", "first"), default="last",
                    help="tile workload: copy each solve's SVF fields after its day (default) or before it")
    ap.add_argument("--march", default="",
                    help="tile workload: pin the SVF march launch to CHUNK,SHAPE,SPLIT (e.g. 8,0,2) instead of autotuning")
    ap.add_argument("--slots", type=int, default=3, help="most execution slots of --workload serve")
    ap.add_argument("--serve-hbm-gb", type=float, default=0.0,
                    help="HBM budget of the serving slot pool in GiB (0: free HBM less 10 %% of the device)")
    ap.add_argument("--extras", default="serve,surrogate,batch",
                    help="tile workload: also run these workloads and add their summaries to the line "
                         "(serve: 1000-request mix at --rate; surrogate: shipped weights on a 2000² city; "
                         "batch: the batched tiles and edit scenarios, 3 repetitions); 'none' for the tile alone")
    ap.add_argument("--extra-requests", type=int, default=1000, help="requests of the serve extra")
    ap.add_argument("--chunk-steps", type=int, default=4, help="day steps a serving slot enqueues at a time")
    ap.add_argument("--edits", type=int, default=200, help="edits timed by --workload edits")
    ap.add_argument("--train-seconds", type=float, default=240.0, help="surrogate training budget")
    ap.add_argument("--train-sites", type=int, default=40, help="1024x1024 training cities for the surrogate")
    ap.add_argument("--surrogate-weights",
                    default=str(Path(__file__).resolve().parent / "solweig_project_amd" / "data"
                                / "svf_surrogate_32_64_128_256.npz"),
                    help="trained surrogate weights (.npz; default: the shipped network of the surrogate "
                         "configuration, 1.95 M parameters; its channels are read from the file); '' trains a new one")
    ap.add_argument("--save-weights", default="", help="write the trained surrogate weights here")
    ap.add_argument("--channels", default="32,64,128,256", help="surrogate level channels of a network trained here")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run k_pack_heights over 16 Mi pixels (12 B read + 16 B written per pixel) "
                         "as the known-bytes reference for rocprof FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


PROFILE_KEYS = {"svf_march": "k_svf_march", "svf_fold": "k_svf_fold", "compose": "k_compose",
                "flux_day": "k_flux_step<true> (day)", "flux_night": "k_flux_step<false> (night)",
                "sun_step": "k_sun_step"}
TRAFFIC_FILE = ROOT / "profiles" / "hbm_traffic_full_solve_1000x1000x24.md"
VALU_FILE = ROOT / "profiles" / "valu_issue_full_solve_1000x1000x24.md"


def load_valu(rows, cols, seed):
    """VALU wave-instructions per launch from the committed rocprofv3 SQ_INSTS_VALU summary ({} if the
    summary is absent or was measured on another workload)."""
    if (rows, cols, seed) != (1000, 1000, 20261203) or not VALU_FILE.exists():
        return {}
    text = VALU_FILE.read_text()
    try:
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return {k: int(v) for k, v in json.loads(block)["valu_insts_per_launch"].items()}
    except (IndexError, KeyError, ValueError):
        return {}


def load_traffic(rows, cols, seed):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary ({} if the
    summary is absent or was measured on another workload)."""
    if (rows, cols, seed) != (1000, 1000, 20261203) or not TRAFFIC_FILE.exists():
        return {}
    text = TRAFFIC_FILE.read_text()
    try:
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return {k: int(v) for k, v in json.loads(block)["bytes_per_launch"].items()}
    except (IndexError, KeyError, ValueError):
        return {}


def sun_for_site(lat, lon, utc, year, doy):
    from solweig_project_amd import sunpos
    return sunpos.day_sun_table(lat, lon, utc, year, doy)


def cpu_baseline(site, sun, met):
    """CPU oracle (C restatement of the reference algorithm, OpenMP, all cores) on the same workload."""
    from oracle import driver as D
    rows = site.dsm.shape[0]
    D.lib()
    t0 = time.perf_counter()
    sc = D.compose(site.dsm, site.dem, site.cdsm, site.scale)
    fields, shm, vm, bm = D.svf(sc)
    D.day(sc, fields, shm, vm, bm, site.walls, site.aspect, site.landcover, met, sun, site.latitude)
    dt = time.perf_counter() - t0
    cols = site.dsm.shape[1]
    return dict(value=round(rows * cols * met.shape[0] / dt / 1e6, 5), unit="Mpixel-steps/s", cores=D.threads(),
                kind="port", cpu_model=D.cpu_model(), cpus_visible=len(os.sched_getaffinity(0)),
                sample=f"C oracle (plane-sweep restatement of the reference CPU path, OpenMP over rows) full "
                       f"solve of a {rows}x{cols}x{met.shape[0]}h tile from the same generator, {dt:.1f} s wall")


def c3_inputs(seed, dev):
    """Tiles 300/700/1200/2000 squared (seeds +10..+13) and 64 edit scenarios of a 500 x 500 base
    (seed +20); walls and wall aspect of every DSM computed on the device."""
    from solweig_project_amd import synth
    from solweig_project_amd import walls as WG
    from solweig_project_amd.solver import TileInputs

    def with_walls(base, dsm, cdsm, lc):
        w, a = WG.walls_and_aspect(dsm, base.scale)
        return TileInputs.from_site(synth.site_with(base, dsm, cdsm, lc, walls=w, aspect=a))
    tiles = []
    for k, n in enumerate((300, 700, 1200, 2000)):
        s = synth.make_site(n, n, seed=seed + 10 + k, with_aspect=False)
        tiles.append(with_walls(s, s.dsm, s.cdsm, s.landcover))
    base = synth.make_site(500, 500, seed=seed + 20, with_aspect=False)
    edits = [with_walls(base, d, c, lc) for d, c, lc, _ in synth.edit_scenarios(base, 64, seed=seed + 21)]
    return tiles, edits, base


def run_batch(args, reps=None):
    """Batched exact solves: every group runs its SVF sweeps, then one sw_solve_batch day loop.
    Returns the JSON line (``reps`` timed repetitions, default ``--steps``)."""
    import numpy as np
    import torch
    from solweig_project_amd import _lib as L
    from solweig_project_amd import synth
    from solweig_project_amd.solver import DayBatch, HipSolver

    dev = torch.device("cuda:0")
    log("generating the batch inputs (walls/aspect on the device)")
    tiles, edits, base = c3_inputs(args.seed, dev)
    sun = sun_for_site(base.latitude, base.longitude, base.utc_offset, 2021, 180)
    met = synth.make_met(sun.altitude, seed=args.seed + 1)
    T = met.shape[0]
    solver = HipSolver(dev)
    groups = {"tiles": [solver.upload(t) for t in tiles], "edits": [solver.upload(t) for t in edits]}
    px = {g: sum(ds.n for ds in sc) for g, sc in groups.items()}

    def solve_group(scenes):
        for ds in scenes:
            solver.svf(ds)
        batch = DayBatch(solver, scenes, met, sun, base.latitude)   # plans while the GPU marches
        for ds in scenes:
            solver.tile_static(ds, dict(model=batch.model))
        batch.run()
        return batch

    def sequential(scenes, batch):
        # the same work as one launch per tile and kernel (plans shared with the batch)
        for ds in scenes:
            solver.svf(ds)
            solver.tile_static(ds, dict(model=batch.model))
        for ds, out in zip(scenes, batch.outs):
            plan = batch.plans[ds.tile.shape]
            solver.run_day(ds, plan, batch.dplans[ds.tile.shape], out=out)

    res = {}
    for g, scenes in groups.items():
        log(f"{g}: {len(scenes)} items, {px[g] / 1e6:.2f} Mpx")
        b = solve_group(scenes)
        torch.cuda.synchronize()
        for ds in scenes:
            ds.check_flags()
        times = []
        for _ in range(reps or args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            b = solve_group(scenes)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        ms = float(np.median(times)) * 1e3
        # day loop only: eager batched launches vs captured graph replay vs one launch per tile
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        b.run()
        ev[1].record()
        torch.cuda.synchronize()
        day_ms = ev[0].elapsed_time(ev[1])
        s = torch.cuda.Stream(dev)
        gph = __import__("ctypes").c_void_p()
        L.call("sw_graph_begin", s.cuda_stream)
        rc = L.load().sw_solve_batch(L.ptr(b.items_dev), b.items_host, len(scenes), L.ptr(b.model), b._is_day_c, 0,
                                     b.n_steps, b.mask, b.walk_halo, 1, s.cuda_stream)
        L.call("sw_graph_end", s.cuda_stream, __import__("ctypes").byref(gph))
        if rc:
            raise RuntimeError(L.load().sw_last_error().decode())
        L.call("sw_graph_launch", gph, s.cuda_stream)
        s.synchronize()
        with torch.cuda.stream(s):
            ev[0].record(s)
            L.call("sw_graph_launch", gph, s.cuda_stream)
            ev[1].record(s)
        s.synchronize()
        graph_ms = ev[0].elapsed_time(ev[1])
        L.call("sw_graph_destroy", gph)
        torch.cuda.synchronize()
        # the same day in one persistent launch (sw_solve_batch_persistent): eager and captured, and its
        # outputs against the per-phase launches' (bit for bit)
        ref_outs = [o.clone() for o in b.outs]
        b.run(persistent=True)
        torch.cuda.synchronize()
        b.check()
        ev[0].record()
        b.run(persistent=True)
        ev[1].record()
        torch.cuda.synchronize()
        b.check()
        pers_ms = ev[0].elapsed_time(ev[1])
        pers_equal = all(torch.equal(o.view(torch.int32), r.view(torch.int32)) for o, r in zip(b.outs, ref_outs))
        L.call("sw_graph_begin", s.cuda_stream)
        with torch.cuda.stream(s):
            b.run(stream=s, persistent=True)
        L.call("sw_graph_end", s.cuda_stream, __import__("ctypes").byref(gph))
        L.call("sw_graph_launch", gph, s.cuda_stream)
        s.synchronize()
        with torch.cuda.stream(s):
            ev[0].record(s)
            L.call("sw_graph_launch", gph, s.cuda_stream)
            ev[1].record(s)
        s.synchronize()
        b.check()
        pers_graph_ms = ev[0].elapsed_time(ev[1])
        L.call("sw_graph_destroy", gph)
        pers_equal = pers_equal and all(torch.equal(o.view(torch.int32), r.view(torch.int32))
                                        for o, r in zip(b.outs, ref_outs))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sequential(scenes, b)
        torch.cuda.synchronize()
        seq_ms = (time.perf_counter() - t0) * 1e3
        res[g] = {"items": len(scenes), "mpx": round(px[g] / 1e6, 3), "ms": round(ms, 2),
                  "mpixel_steps_per_s": round(px[g] * T / (ms / 1e3) / 1e6, 1),
                  "day_batched_ms": round(day_ms, 2), "day_graph_replay_ms": round(graph_ms, 2),
                  "day_persistent_ms": round(pers_ms, 2), "day_persistent_graph_replay_ms": round(pers_graph_ms, 2),
                  "persistent_equals_per_phase_launches": bool(pers_equal),
                  "launches_per_day_batched": int(sum(b.is_day)) * 2 + (T - int(sum(b.is_day))),
                  "sequential_one_launch_per_tile_ms": round(seq_ms, 2)}
        log(f"{g}: persistent day {pers_ms:.1f} ms eager, {pers_graph_ms:.1f} ms graph, equal: {pers_equal}")
        log(f"{g}: {ms:.1f} ms per batched solve; day {day_ms:.1f} ms eager, {graph_ms:.1f} ms graph; "
            f"one launch per tile {seq_ms:.1f} ms")
    tot_px = sum(px.values())
    tot_ms = sum(r["ms"] for r in res.values())
    line = {
        "metric": "Mpixel-steps/s (batched exact solves, 4 uneven tiles + 64 edit scenarios of 500², 24 h)",
        "value": round(tot_px * T / (tot_ms / 1e3) / 1e6, 3),
        "unit": "Mpixel-steps/s",
        "n_gpus": 1, "steps": reps or args.steps, "warmup": 1, "ms_per_step": round(tot_ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded city tiles + edit scenarios; inputs resident in HBM, outputs left in HBM)",
        "config": {"workload": "batch: tiles 300², 700², 1200², 2000² and 64 edit scenarios (1-20 tree or "
                               "building edits of 10-400 px) of a 500² base, 24 hourly steps, one plan per tile "
                               "shape, SVF sweep per item, one sun-march and one flux launch per step for all "
                               "items", "time_steps": T},
        "groups": res,
    }
    return line


def run_edits(args):
    """Single-tree edits of a device-resident 2000 x 2000 site: latency from the edit request (new
    window values on the host) to the edited outputs in HBM, beside a met-only day re-run and a full
    re-solve of the site."""
    import numpy as np
    import torch
    from solweig_project_amd import incremental as I
    from solweig_project_amd import synth
    from solweig_project_amd import walls as WG
    from solweig_project_amd.solver import HipSolver, TileInputs

    dev = torch.device("cuda:0")
    n = 2000
    seed = args.seed + 40
    log(f"generating the {n}x{n} site (walls/aspect on the device)")
    base = synth.make_site(n, n, seed=seed, with_aspect=False)
    w, a = WG.walls_and_aspect(base.dsm, base.scale)
    tile = TileInputs.from_site(synth.site_with(base, base.dsm, base.cdsm, base.landcover, walls=w, aspect=a))
    sun = sun_for_site(base.latitude, base.longitude, base.utc_offset, 2021, 180)
    met = synth.make_met(sun.altitude, seed=seed + 1)
    solver = HipSolver(dev)
    st = I.solve_site(solver, tile, met, sun, base.latitude)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = I.solve_site(solver, tile, met, sun, base.latitude)
    full_ms = (time.perf_counter() - t0) * 1e3
    rng = np.random.Generator(np.random.PCG64(seed + 2))
    lat, modes, fracs = [], [], []
    ground = np.flatnonzero((st.tile.dsm == st.tile.dem).ravel())
    for k in range(args.edits + 5):
        ci, cj = divmod(int(ground[rng.integers(0, ground.size)]), n)
        r = int(rng.integers(2, 7))
        r0, r1, c0, c1 = max(0, ci - r), min(n, ci + r + 1), max(0, cj - r), min(n, cj + r + 1)
        cd = st.tile.cdsm[r0:r1, c0:c1].copy()
        if rng.random() < 0.5 or not (cd > 0).any():
            ii, jj = np.mgrid[r0:r1, c0:c1]
            d2 = (ii - ci) ** 2 + (jj - cj) ** 2
            crown = (float(rng.uniform(3.0, 15.0)) * np.exp(-d2 / (2.0 * (r / 2.0) ** 2))).astype(np.float32)
            crown[(d2 > r * r) | (st.tile.dsm[r0:r1, c0:c1] != st.tile.dem[r0:r1, c0:c1])] = 0
            cd = np.maximum(cd, crown)
        else:
            cd[:] = 0
        info = I.apply_edit_window(solver, st, (r0, r1, c0, c1), cdsm=cd)
        if k >= 5 and info["mode"] != "none":          # the first edits warm caches
            lat.append(info["ms"])
            modes.append(info["mode"])
            if info["mode"] == "tiles":
                fracs.append(info["flux_tiles"] / info["grid_tiles"])
    # met-only request: re-plan the day and re-run it over the whole site
    met2 = synth.make_met(sun.altitude, seed=seed + 3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    from solweig_project_amd.solver import plan_day, _owned_plan
    plan2 = plan_day(st.tile, met2, sun, base.latitude)
    dplan2 = _owned_plan(plan2, dev)
    solver.tile_static(st.scene, dplan2)
    solver.run_day(st.scene, plan2, dplan2, st.outputs, out=st.out)
    torch.cuda.synchronize()
    met_ms = (time.perf_counter() - t0) * 1e3
    lat = np.array(lat)
    line = {
        "metric": "p99 single-edit latency (ms, 2000² site x 24 h, edited outputs in HBM)",
        "value": round(float(np.percentile(lat, 99)), 2),
        "unit": "ms",
        "n_gpus": 1, "steps": len(lat), "warmup": 5, "ms_per_step": round(float(lat.mean()), 2),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded 2000² city; random single-tree additions/removals of radius 2-6 px)",
        "config": {"workload": "single-tree edits of a device-resident 2000x2000x24h site: walls/aspect "
                               "window, SVF and day recomputed on the tiles the edit can reach "
                               "(sw_svf_sweep_tiles, sw_solve_day_tiles)", "rows": n, "cols": n},
        "p50_ms": round(float(np.percentile(lat, 50)), 2),
        "p90_ms": round(float(np.percentile(lat, 90)), 2),
        "max_ms": round(float(lat.max()), 2),
        "modes": {m: modes.count(m) for m in sorted(set(modes))},
        "mean_flux_tile_fraction": round(float(np.mean(fracs)), 4) if fracs else None,
        "met_only_day_rerun_ms": round(met_ms, 2),
        "full_resolve_ms": round(full_ms, 2),
    }
    print(json.dumps(line), flush=True)


def run_surrogate(args):
    """Train the SVF surrogate on exact-kernel labels (or load weights), then time the gated
    surrogate against the exact SVF sweep on a 2000 x 2000 city and report their difference."""
    import numpy as np
    import torch
    from solweig_project_amd import surrogate as SG
    from solweig_project_amd import synth
    from solweig_project_amd.solver import HipSolver, TileInputs

    dev = torch.device("cuda:0")
    solver = HipSolver(dev)
    ch = tuple(int(c) for c in args.channels.split(","))
    train_info = None
    if args.surrogate_weights and Path(args.surrogate_weights).exists():
        net = SG.load_weights(args.surrogate_weights, dev)
        ch = net.channels
        log(f"loaded surrogate weights {args.surrogate_weights} (channels {ch}, {net.n_params} parameters)")
    else:
        net = SG.SurrogateNet(ch, dev, seed=args.seed)
        log(f"labelling {args.train_sites} training cities (1024x1024) with the exact SVF sweep")
        sites = SG.exact_sites(solver, args.train_sites, 1024, seed=args.seed + 500)
        # time a few steps to size the run to the budget
        probe = SG.train(net, sites[:2], steps=5, batch=8, seed=args.seed)
        per = probe["seconds"] / 5
        steps = max(50, int(args.train_seconds / per))
        log(f"training {steps} steps ({per * 1e3:.1f} ms per step of 8 windows)")
        train_info = SG.train(net, sites, steps=steps, batch=8, seed=args.seed + 1, log=log)
        del sites
        if args.save_weights:
            Path(args.save_weights).parent.mkdir(parents=True, exist_ok=True)
            np.savez_compressed(args.save_weights, **net.state_dict())
    # evaluation city (seed +30), exact sweep as the reference and as the timing baseline
    n = 2000
    site = synth.make_site(n, n, seed=args.seed + 30, with_aspect=False)
    ds = solver.upload(TileInputs.from_site(site))
    solver.svf(ds)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ex_ms = []
    for _ in range(3):
        ev[0].record()
        solver.svf(ds)
        ev[1].record()
        torch.cuda.synchronize()
        ex_ms.append(ev[0].elapsed_time(ev[1]))
    exact = ds.fields[:SG.N_FIELDS].clone()
    exact_all = ds.fields.clone()

    def gated(model, reps=5, criterion="mean", samples=4):
        """Median wall time of ``reps`` gated passes after a warm one (each starts from the exact fields)."""
        gate = SG.SvfSurrogate(model, tol=0.02, samples=samples, seed=args.seed, criterion=criterion)
        gate.svf(solver, ds)                     # warm
        walls = []
        for _ in range(reps):
            ds.fields.copy_(exact_all)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rep = gate.svf(solver, ds)
            torch.cuda.synchronize()
            walls.append((time.perf_counter() - t0) * 1e3)
        err = (ds.fields[:SG.N_FIELDS] - exact).abs()
        return rep, float(np.median(walls)), err
    # derived Tmrt: the day on the exact fields against the day on the gated fields (network fields 0..9 and
    # SVFtotal; the per-patch masks and the aveg fields stay those of the exact sweep)
    from solweig_project_amd import sunpos
    from solweig_project_amd.solver import plan_day
    sun = sunpos.day_sun_table(site.latitude, site.longitude, site.utc_offset, 2021, 180)
    met = synth.make_met(sun.altitude, seed=args.seed + 31)
    plan = plan_day(ds.tile, met, sun, site.latitude)
    dplan = solver.upload_plan(plan)

    def day_tmrt():
        solver.tile_static(ds, dplan)
        out = solver.run_day(ds, plan, dplan, ("tmrt",))
        torch.cuda.synchronize()
        return out[:, 0]
    tmrt_exact = day_tmrt()
    tmrt_exact_keep = tmrt_exact.clone()
    rep, wall, err = gated(net)
    tmrt_gated = day_tmrt()
    d3 = (tmrt_gated - tmrt_exact).abs()
    per_px = d3.mean(dim=0)
    core_means = [float(per_px[r0:r0 + SG.CORE, c0:c0 + SG.CORE].mean())
                  for r0, c0 in (SG.core_origins(n, n) + SG.HALO).tolist()]
    dt = d3.flatten()
    dq = torch.quantile(dt[torch.randperm(dt.numel(), device=dt.device)[:8_000_000]], 0.99).item()
    derived = {"core_mean_max_K": round(max(core_means), 4), "max_K": round(float(dt.max()), 4),
               "mean_K": round(float(dt.mean()), 5), "p99_K": round(dq, 4),
               "frac_over_0.5K": round(float((dt > 0.5).float().mean()), 6), "pixel_steps": int(dt.numel()),
               "policy": "network fields 0-9 + SVFtotal from the gate; per-patch masks and aveg fields of the "
                         "exact sweep (the anisotropic sky reads per-patch occlusion the network does not predict)"}
    del tmrt_exact, tmrt_gated, dt, d3, per_px
    solver.svf(ds)
    torch.cuda.synchronize()
    e = err.flatten()
    q = torch.quantile(e[torch.randperm(e.numel(), device=e.device)[:4_000_000]], 0.99).item()
    raw = SG.SvfSurrogate(net)
    surr = torch.empty_like(ds.fields)
    raw.predict(ds, surr)
    raw_err = (surr[:SG.N_FIELDS] - exact).abs()
    untrained = SG.SurrogateNet(ch, dev, seed=args.seed + 7)
    rep_u, wall_u, err_u = gated(untrained)
    exact_ms = float(np.median(ex_ms))
    # the bar: max |dSVF| <= 0.02 over every pixel and field, max |dTmrt| <= 0.5 K over every
    # pixel-step of the day derived from the gated fields; gates of increasing strictness
    gates = []
    for crit, samp in (("mean", 4), ("max", 4), ("max", 16), ("max", 64)):
        g_rep, g_wall, g_err = gated(net, criterion=crit, samples=samp)
        g_tmrt = day_tmrt()
        g_dt = (g_tmrt - tmrt_exact_keep).abs()
        gmax, tmax = float(g_err.max()), float(g_dt.max())
        gates.append({"criterion": crit, "samples_per_core": samp, "accepted": g_rep.accepted,
                      "windows": g_rep.windows, "gated_ms": round(g_wall, 3), "speedup": round(exact_ms / g_wall, 3),
                      "svf_error_max": round(gmax, 4), "svf_error_mean": round(float(g_err.mean()), 5),
                      "tmrt_error_max_K": round(tmax, 4), "tmrt_frac_over_0.5K": round(float((g_dt > 0.5).float().mean()), 6),
                      "meets_bar": bool(gmax <= 0.02 and tmax <= 0.5)})
        log(f"gate {crit}/{samp}: {g_rep.accepted}/{g_rep.windows} accepted, {g_wall:.2f} ms "
            f"({exact_ms / g_wall:.2f}x), max dSVF {gmax:.4f}, max dTmrt {tmax:.3f} K")
        del g_tmrt, g_dt, g_err
        solver.svf(ds)
        torch.cuda.synchronize()
    ok = [g for g in gates if g["meets_bar"]]
    at_bar = max(ok, key=lambda g: g["speedup"]) if ok else None
    del tmrt_exact_keep
    # MFMA roofline of the network: 2 x MACs per 384^2 window as inference runs it (model channels; the
    # last two layers over the core) x windows / network time
    fpw = 2 * SG.macs_per_window(ch, core=SG.CORE)
    flops = float(fpw) * rep.windows
    ach = flops / (rep.ms_network * 1e-3) / 1e12
    roof = {"kernel": "sw_cnn_fwd (window pack + 15 conv layers + scatter)", "bound": "mfma",
            "achieved": round(ach, 1), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(ach / BF16_PEAK_TFLOPS, 4), "traffic": None,
            "flop_per_window": fpw, "windows": rep.windows,
            "network_ms": round(rep.ms_network, 3)}
    line = {
        "metric": "surrogate vs exact SVF (speedup of the gated surrogate path, 2000² city, 10 SVF fields)",
        # the speed-up at the accuracy bar (max |dSVF| <= 0.02 and derived max |dTmrt| <= 0.5 K) of the
        # fastest gate that meets it; the mean gate's speed-up, which misses the bar, stays beside it
        "value": at_bar["speedup"] if at_bar else None,
        "value_mean_gate": round(exact_ms / wall, 3),
        "unit": "x (exact ms / gated surrogate ms)",
        "n_gpus": 1, "steps": 5, "warmup": 1, "ms_per_step": round(wall, 3),
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "bf16 (MFMA, f32 accumulate)",
        "data": "synthetic (seeded cities; labels from the exact SVF sweep of whole 1024² cities)",
        "config": {"workload": "4-level conv encoder-decoder on 384² windows (256² core + 64 px halo), gate: "
                               "exact kernel on 4 sampled 16x16 tiles per core, accepted when their mean absolute "
                               "error over the 10 fields is <= 0.02, exact re-solve of rejected cores", "channels": list(ch), "params": net.n_params, "rows": n,
                   "cols": n},
        "roofline": roof,
        "exact_svf_ms": round(exact_ms, 3),
        "gated_ms": round(wall, 3), "network_ms": round(rep.ms_network, 3),
        "windows": rep.windows, "accepted": rep.accepted, "sampled_tiles": rep.sampled_tiles,
        "exact_tiles_after_gate": rep.exact_tiles,
        "gated_error": {"max": round(float(err.max()), 4), "mean": round(float(err.mean()), 5), "p99": round(q, 4)},
        "network_error_ungated": {"max": round(float(raw_err.max()), 4), "mean": round(float(raw_err.mean()), 5)},
        "derived_tmrt_error": derived,
        "gates": gates,
        "at_bar": at_bar,
        "untrained_same_gate": {"accepted": rep_u.accepted, "gated_ms": round(wall_u, 3),
                                "speedup": round(exact_ms / wall_u, 3), "max_error": round(float(err_u.max()), 4)},
        "training": None if train_info is None else {
            "seconds": round(train_info["seconds"], 1), "steps": train_info["steps"], "batch": train_info["batch"],
            "windows": train_info["windows"], "first_loss": train_info["losses"][0][1],
            "last_loss": train_info["losses"][-1][1]},
    }
    return line


def run_serve(args, pipeline, requests=None):
    """Serving mix: Poisson arrivals (60 % single-tree edits, 30 % met-only, 10 % full re-solves) against a
    device-resident 2000 x 2000 x 24 h site; latency = arrival to results in pinned host memory, queueing
    included; nearest-rank percentiles. Slots are admitted by HBM footprint against ``--serve-hbm-gb``
    (default: the free HBM less 10 % of the device). Returns the JSON line."""
    import torch
    from solweig_project_amd import incremental as I
    from solweig_project_amd import serving as SV
    from solweig_project_amd import synth
    from solweig_project_amd import walls as WG
    from solweig_project_amd.solver import HipSolver, TileInputs

    dev = torch.device("cuda:0")
    n = 2000
    seed = args.seed + 40
    base = synth.make_site(n, n, seed=seed, with_aspect=False)
    w, a = WG.walls_and_aspect(base.dsm, base.scale)
    tile = TileInputs.from_site(synth.site_with(base, base.dsm, base.cdsm, base.landcover, walls=w, aspect=a))
    sun = sun_for_site(base.latitude, base.longitude, base.utc_offset, 2021, 180)
    mets = [synth.make_met(sun.altitude, seed=seed + 1 + k) for k in range(4)]
    solver = HipSolver(dev)
    st = I.solve_site(solver, tile, mets[0], sun, base.latitude)
    from solweig_project_amd.pipeline import Job
    pipeline.warm(st.tile, Job(mets[0], sun, base.latitude))
    budget = int(args.serve_hbm_gb * 2 ** 30) if args.serve_hbm_gb > 0 else None
    server = SV.Server(solver, st, mets, sun, base.latitude, pipeline, slots=args.slots,
                       chunk_steps=args.chunk_steps, hbm_budget=budget)
    # warm every path once (and create the slots the mix will use)
    warm = SV.poisson_workload(3 * args.slots, rate=1000.0, mix={"edit": 0.34, "met": 0.33, "full": 0.33},
                               seed=seed + 7, site_tile=st.tile, met_pool=mets)
    server.run(warm, checksums=False)
    server.deferred = 0
    nreq = requests or args.requests
    reqs = SV.poisson_workload(nreq, rate=args.rate, mix={"edit": 0.6, "met": 0.3, "full": 0.1},
                               seed=seed + 8, site_tile=st.tile, met_pool=mets)
    log(f"serving {len(reqs)} requests at {args.rate}/s over {reqs[-1].arrival:.1f} s; HBM budget "
        f"{server.budget / 2 ** 30:.1f} GiB, slot {server.slot_bytes['site'] / 2 ** 30:.2f} GiB, "
        f"{len(server.slots)} slots warm")
    t0 = time.perf_counter()
    done = server.run(reqs, checksums=False)
    wall = time.perf_counter() - t0
    allp = SV.percentiles(done)
    T = st.plan.n_steps
    px_steps = sum(n * n * T for r in done if r.kind in ("met", "full"))
    line = {
        "metric": "p99 request latency (ms), serving mix on a 2000² x 24 h site, results on the host",
        "value": allp["p99_ms"], "unit": "ms",
        "n_gpus": 1, "steps": len(done), "warmup": len(warm), "ms_per_step": allp["mean_ms"],
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded 2000² city; Poisson arrivals, seeded request mix)",
        "config": {"workload": f"Poisson arrivals at {args.rate}/s: 60% single-tree edits, 30% met-only day "
                               f"re-runs, 10% full re-solves; continuous batching over up to {args.slots} slots "
                               f"admitted by HBM footprint (own stream and scene copy each, days fed "
                               f"{args.chunk_steps} steps at a time, each chunk copied to pinned host memory on the "
                               "DMA engines), edits first and in order on the site, each edit's recomputed "
                               "output window copied to the host; latency = arrival to results on the host, "
                               "nearest-rank percentiles",
                   "rows": n, "cols": n, "requests": len(done), "rate_per_s": args.rate},
        "all": allp,
        "edit": SV.percentiles(done, ("edit",)),
        "met": SV.percentiles(done, ("met",)),
        "full": SV.percentiles(done, ("full",)),
        "admission": {"hbm_budget_bytes": server.budget, "slot_bytes": server.slot_bytes["site"],
                      "edit_reserve_bytes": server.edit_reserve, "slots": len(server.slots),
                      "peak_busy_slots": server.peak_busy, "deferred_requests": server.deferred,
                      "peak_pool_bytes": server.peak_pool_bytes},
        "wall_s": round(wall, 2),
        "completed_per_s": round(len(done) / wall, 2),
        "met_full_mpixel_steps_per_s": round(px_steps / wall / 1e6, 1),
        "edit_modes": {m: sum(1 for r in done if r.kind == "edit" and r.payload.get("mode") == m)
                       for m in ("tiles", "full", "none")},
        "edit_result_mb_mean": round(sum(r.payload.get("result_bytes", 0) for r in done if r.kind == "edit")
                                     / max(1, sum(1 for r in done if r.kind == "edit")) / 1e6, 3),
    }
    del server, st
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return line


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU path is "
              "design-only (tiles with read-only halos, no collectives)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        print(json.dumps(run_batch(args)), flush=True)
        return
    if args.workload == "edits":
        run_edits(args)
        return
    if args.workload == "surrogate":
        print(json.dumps(run_surrogate(args)), flush=True)
        return
    from solweig_project_amd.pipeline import Job, PlanPipeline, solve_many
    extras = [x for x in args.extras.split(",") if x and x != "none"]
    serve_pipeline = None
    if args.workload == "serve" or "serve" in extras:
        serve_pipeline = PlanPipeline(workers=2)   # CPU-only planner processes, started before the GPU is used
    if args.workload == "serve":
        try:
            print(json.dumps(run_serve(args, serve_pipeline)), flush=True)
        finally:
            serve_pipeline.close()
        return
    pipeline = PlanPipeline(workers=3)      # CPU-only planner processes, started before the GPU is used
    import numpy as np
    import torch
    from solweig_project_amd import skyvault as SV
    from solweig_project_amd import synth
    from solweig_project_amd.solver import (DeviceScene, HipSolver, KernelTimer, TileInputs, _NoTimer,
                                            upload_inputs)

    rows, cols = args.rows, args.cols
    log(f"generating {rows}x{cols} tile (seed {args.seed})")
    site = synth.make_site(rows, cols, seed=args.seed)
    sun = sun_for_site(site.latitude, site.longitude, site.utc_offset, 2021, 180)
    met = synth.make_met(sun.altitude, seed=args.seed + 1)
    T = met.shape[0]
    tile = TileInputs.from_site(site)
    dev = torch.device("cuda:0")
    solver = HipSolver(dev, march_config=tuple(int(x) for x in args.march.split(",")) if args.march else None)
    inputs = upload_inputs(tile, dev)
    ds = DeviceScene(tile, dev, inputs)
    out = torch.empty((T, 3, rows, cols), dtype=torch.float32, device=dev)
    job = Job(met, sun, site.latitude)
    # lanes: consecutive solves alternate over scene copies on their own streams, so one solve's
    # kernel tails overlap the next solve's kernels (every solve is still a complete SVF + day)
    lanes = None
    if args.lanes > 1:
        lanes = [(ds, torch.cuda.Stream(dev), out)]
        for _ in range(args.lanes - 1):
            lanes.append((DeviceScene(tile, dev, upload_inputs(tile, dev)), torch.cuda.Stream(dev),
                          torch.empty_like(out)))
    log("warmup")
    pipeline.warm(ds.tile, job)
    n_warm = max(args.warmup, 1, len(lanes) if lanes else 1)
    solve_many(solver, ds, [job] * n_warm, pipeline, out=out, lanes=lanes)
    torch.cuda.synchronize()
    ds.check_flags()
    mt = solver.march_tuner
    mkey = (rows, cols, float(tile.scale))
    march_info = {"autotuned": mt.enabled}
    if mt.fixed is not None:
        march_info.update(chunk=mt.fixed[0], shape=["8x8", "4x16", "16x4"][mt.fixed[1]], split=mt.fixed[2])
    elif mkey in mt.best:
        ch, sh, spl = mt.best[mkey]
        tm = mt.timings[mkey]
        march_info.update(chunk=ch, shape=["8x8", "4x16", "16x4"][sh], split=spl, tuned_ms=tm[mt.best[mkey]],
                          default_ms=tm[mt.DEFAULT], candidates=len(tm))
        log(f"march autotuner: chunk {ch}, wave {march_info['shape']}, split {spl}: "
            f"{tm[mt.best[mkey]]:.3f} ms (default {tm[mt.DEFAULT]:.3f} ms, {len(tm)} candidates)")
    timer = KernelTimer(reserve=args.steps * (6 + 3 * T))
    solver.timer = timer
    log(f"timing {args.steps} steps over {max(args.lanes, 1)} lane(s)")
    solver.plan_ms.clear()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    # every solve plans its own day (in the planner processes, two solves ahead)
    solve_many(solver, ds, [job] * args.steps, pipeline, out=out, lanes=lanes)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    solver.timer = _NoTimer()
    marks = np.array(getattr(solver, "pipeline_marks", [[0, 0, 0, 0]])) * 1e3
    log("host ms per solve (mean): svf enqueue %.2f, wait for plan %.2f, upload %.2f, day enqueue %.2f"
        % tuple(marks.mean(axis=0)))
    plan_ms = sum(solver.plan_ms) / max(len(solver.plan_ms), 1)   # per solve, in the planner process
    dev_ms = (t1 - t0) / args.steps * 1e3
    dev_value = rows * cols * T / (dev_ms / 1e3) / 1e6
    ksum = timer.summary()
    log("kernels: " + ", ".join(f"{k}: {n} x {t / n:.3f} ms" for k, (n, t) in ksum.items()))
    log(f"device-resident: {dev_ms:.3f} ms per solve, {dev_value:.1f} Mpixel-steps/s")

    # ---- the timed region of `value`: host-resident inputs -> host-resident outputs ----
    from solweig_project_amd.pipeline import HostBuffers, pinned_inputs, solve_many_host
    if lanes is None:
        lanes = [(ds, torch.cuda.Stream(dev), out)]
    host_in = pinned_inputs(tile)
    # host-to-host lanes: each copies its outputs in order on its own stream (lanes + the null stream
    # stay within the 4 hardware queues); scene copies beyond the device-resident lanes are added
    h2h_lanes = [(sc, st, None) for sc, st, _ in lanes]
    while len(h2h_lanes) < args.h2h_lanes:
        h2h_lanes.append((DeviceScene(tile, dev, upload_inputs(tile, dev)), torch.cuda.Stream(dev), None))
    h2h_lanes = h2h_lanes[:args.h2h_lanes]
    if args.h2h_copy == "shared":
        cs = torch.cuda.Stream(dev)
        h2h_lanes = [(sc, st, cs) for sc, st, _ in h2h_lanes]
    elif args.h2h_copy == "per-lane":
        h2h_lanes = [(sc, st, torch.cuda.Stream(dev)) for sc, st, _ in h2h_lanes]
    n_spare = len(h2h_lanes) + 1              # host buffers the warm-up solves write in turn
    n_h2h = args.steps + n_spare
    gb = n_h2h * (16 + 3 * T) * rows * cols * 4 / 1e9
    log(f"pinned host outputs for {n_h2h} solves ({gb:.1f} GB)")
    hbufs = [HostBuffers(torch.empty((16, rows, cols), dtype=torch.float32, pin_memory=True),
                         torch.empty((T, 3, rows, cols), dtype=torch.float32, pin_memory=True))
             for _ in range(n_h2h)]
    threaded = args.h2h_copy == "thread"
    n_ring = 6 if threaded else 3
    dev_outs = [lanes[0][2]] + [torch.empty_like(lanes[0][2]) for _ in range(n_ring - 1)]   # ring of day outputs
    if threaded:
        from solweig_project_amd.pipeline import solve_many_host_threaded
        fields_ring = [torch.empty_like(ds.fields) for _ in range(n_ring)]
        # the lanes' streams and the copy stream created one after another: the runtime puts
        # consecutive streams on different hardware queues (a copy stream sharing a lane's queue
        # would serialise that lane behind the downloads)
        h2h_streams = [torch.cuda.Stream(dev) for _ in range(len(h2h_lanes) + 1)]
        for st in h2h_streams:               # first use in creation order
            with torch.cuda.stream(st):
                torch.zeros(1, device=dev)
        torch.cuda.synchronize()
        tl = [(sc, st) for (sc, _, _), st in zip(h2h_lanes, h2h_streams)]
        copy_stream = h2h_streams[-1]

        def run_h2h(jobs_, outs_):
            return solve_many_host_threaded(solver, tl, jobs_, pipeline, host_in, outs_, dev_outs, fields_ring,
                                            chunk_steps=args.h2h_chunk, copy_stream=copy_stream)
    else:
        def run_h2h(jobs_, outs_):
            return solve_many_host(solver, h2h_lanes, jobs_, pipeline, host_in, outs_, dev_outs,
                                   chunk_steps=args.h2h_chunk, fields_last=args.h2h_fields == "last")
    # warm-up: every lane and every plan staging buffer once (first-use allocations stay out of the timing)
    n_warm_h2h = max(n_warm, len(h2h_lanes) * solver.STAGING_DEPTH, n_ring)
    spare = hbufs[args.steps:]
    run_h2h([job] * n_warm_h2h, [spare[i % n_spare] for i in range(n_warm_h2h)])
    torch.cuda.synchronize()
    log(f"timing {args.steps} host-to-host solves over {len(h2h_lanes)} lane(s)")
    h0 = time.perf_counter()
    done = run_h2h([job] * args.steps, hbufs[:args.steps])
    torch.cuda.synchronize()
    h1 = time.perf_counter()
    hm = np.array(solver.pipeline_marks) * 1e3
    log("host ms per host-to-host solve (mean): upload + svf + fields copy enqueue %.2f, wait for plan %.2f, "
        "plan upload %.2f, day + copies enqueue %.2f" % tuple(hm.mean(axis=0)))
    ms = (h1 - h0) / args.steps * 1e3
    value = rows * cols * T / (ms / 1e3) / 1e6
    gaps = [done[i - 1].elapsed_time(done[i]) for i in range(1, len(done))]
    gap_ms = float(np.median(gaps)) if gaps else ms
    # every timed solve delivered the same (deterministic) outputs to its own host buffers
    f_ref = ds.fields.reshape(16, rows, cols).cpu().view(torch.int32)
    o_ref = dev_outs[(args.steps - 1) % len(dev_outs)].cpu().view(torch.int32)   # bits: UTCI NaN off ground
    delivered = all(torch.equal(hb.fields.view(torch.int32), f_ref) and torch.equal(hb.out.view(torch.int32), o_ref)
                    for hb in hbufs[:args.steps])
    if not delivered:
        raise RuntimeError("a host-to-host solve delivered outputs different from the device-resident solve")
    log(f"host-to-host: {ms:.3f} ms per solve ({value:.1f} Mpixel-steps/s), median gap between solve "
        f"completions {gap_ms:.3f} ms; outputs of all {args.steps} solves checked on the host")
    # this box's device-to-host rate on the DMA engines (the copy path of the timed region), alone
    from solweig_project_amd import _lib as LP
    cs = torch.cuda.Stream(dev)
    src, dst = dev_outs[0], hbufs[0].out
    nb = src.numel() * src.element_size()
    pe = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    LP.call("sw_copy_to_host", dst.data_ptr(), src.data_ptr(), nb, cs.cuda_stream)
    pe[0].record(cs)
    for _ in range(3):
        LP.call("sw_copy_to_host", dst.data_ptr(), src.data_ptr(), nb, cs.cuda_stream)
    pe[1].record(cs)
    cs.synchronize()
    d2h_gbs = 3 * nb / (pe[0].elapsed_time(pe[1]) / 1e3) / 1e9
    out_bytes = (16 + 3 * T) * rows * cols * 4
    bound_ms = max(dev_ms, out_bytes / (d2h_gbs * 1e9) * 1e3)
    log(f"device-to-host {d2h_gbs:.1f} GB/s alone: {out_bytes / 1e6:.0f} MB of outputs per solve take "
        f"{out_bytes / d2h_gbs / 1e6:.2f} ms; host-to-host bound max(device, copy) = {bound_ms:.2f} ms")
    del hbufs

    # rooflines: kernels timed alone (one lane, HIP events on the launch stream, 3 solves), each against
    # VALU issue (their limiter) with the measured HBM fraction beside it
    alone = KernelTimer(reserve=3 * (6 + 3 * T))
    solver.timer = alone
    solve_many(solver, ds, [job] * 3, pipeline, out=out)
    pipeline.close()
    torch.cuda.synchronize()
    solver.timer = _NoTimer()
    ksum_alone = alone.summary()
    log("kernels alone: " + ", ".join(f"{k}: {n} x {t / n:.3f} ms" for k, (n, t) in ksum_alone.items()))
    amax, hmax = ds.amax_hmax()
    offs, dx, dy, dz = SV.svf_tables(solver.vault, tile.scale, np.float32(amax), rows, cols)
    n = rows * cols
    pixel_steps = int(offs[-1]) * n
    alg = {
        # 3 float32 heights per executed march step (the reference's per-step shifted-slice read)
        # + heights in (16 B/px) + packed masks out (80 B/px)
        "svf_march": 12.0 * pixel_steps + (16 + 80) * n,
        # 18 x 22 walk gathers of 6 float planes + ~220 B/px compulsory per step
        "flux_day": (24.0 * 18 * 22 + 220) * n,
        "flux_night": 220.0 * n,
        # packed masks in, 16 float fields out
        "svf_fold": (80 + 64) * n,
        "compose": (24 + 16 + 4 + 4 + 4 + 1) * n,
    }
    traffic = load_traffic(rows, cols, args.seed)
    valu = load_valu(rows, cols, args.seed)
    rooflines = {}
    for name, (launches, tot) in ksum_alone.items():
        key = PROFILE_KEYS.get(name, "")
        if name not in alg or key not in valu:
            continue
        avg_ms = tot / launches
        ach = valu[key] / (avg_ms / 1e3) / 1e9
        tb = traffic.get(key)
        rooflines[name] = {"bound": "valu", "achieved": round(ach, 1), "peak": VALU_PEAK_GINSTR,
                           "unit": "G wave64 VALU instr/s", "frac": round(ach / VALU_PEAK_GINSTR, 4),
                           "frac_full_rate_issue": round(ach / VALU_FULL_RATE_GINSTR, 4), "traffic": tb,
                           "hbm_frac": round(tb / (avg_ms / 1e3) / (HBM_PEAK_GBS * 1e9), 4) if tb else None,
                           "valu_insts_per_launch": valu[key], "avg_launch_ms": round(avg_ms, 4),
                           "launches_per_solve": launches // 3,
                           "algorithmic_bytes_per_launch": int(alg[name])}
    # UTCI on the vector ALUs: the standalone kernel (the function the flux kernel fuses) over this
    # workload's pixel-steps, 450 FLOP per pixel-step (210 polynomial terms as FMAs + the saturation
    # vapour pressure chain, log and exp counted once each)
    from solweig_project_amd import _lib as LU
    m = rows * cols * met.shape[0]
    g = torch.Generator(device=dev).manual_seed(args.seed)
    u_in = [torch.rand(m, device=dev, generator=g) * (hi - lo) + lo
            for lo, hi in ((22.0, 36.0), (30.0, 70.0), (20.0, 70.0), (0.5, 4.0))]
    u_out = torch.empty(m, device=dev)
    us = torch.cuda.current_stream().cuda_stream
    for _ in range(2):
        LU.call("sw_utci", *[LU.ptr(x) for x in u_in], LU.ptr(u_out), m, us)
    ue = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ue[0].record()
    for _ in range(5):
        LU.call("sw_utci", *[LU.ptr(x) for x in u_in], LU.ptr(u_out), m, us)
    ue[1].record()
    torch.cuda.synchronize()
    u_ms = ue[0].elapsed_time(ue[1]) / 5
    u_tf = 450.0 * m / (u_ms / 1e3) / 1e12
    rooflines["utci"] = {"bound": "valu", "achieved": round(u_tf, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(u_tf / FP32_PEAK_TFLOPS, 4), "traffic": None, "avg_launch_ms": round(u_ms, 4),
                         "pixel_steps": m, "flop_per_pixel_step": 450,
                         "note": "standalone k_utci over the day's pixel-steps; the flux kernel fuses the same function"}
    del u_in, u_out
    # dominant kernel = largest total time per solve
    dom = max(ksum_alone, key=lambda k: ksum_alone[k][1])
    roof = dict(kernel=dom, **rooflines[dom]) if dom in rooflines else None
    if args.calibrate_copy:
        from solweig_project_amd import _lib as L
        m = 16 * 1024 * 1024
        planes = [torch.ones(m, dtype=torch.float32, device=dev) for _ in range(3)]
        hp = torch.empty((m, 4), dtype=torch.float32, device=dev)
        L.call("sw_pack_heights", L.ptr(planes[0]), L.ptr(planes[1]), L.ptr(planes[2]), L.ptr(hp), 4096, 4096,
               torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        log(f"copy calibration: k_pack_heights read {12 * m} B, wrote {16 * m} B")
    # tile inputs on the device: wall heights + wall aspect of the same DSM (not part of the metric)
    from solweig_project_amd import walls as WG
    dsm_dev = torch.from_numpy(site.dsm).to(dev)
    WG.walls_and_aspect(dsm_dev, tile.scale)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    reps = 5
    wa_ms = [0.0, 0.0]
    for _ in range(reps):
        ev[0].record()
        w_dev = WG.findwalls(dsm_dev)
        ev[1].record()
        WG.filter1Goodwin_as_aspect_v3(w_dev, tile.scale, dsm_dev)
        ev[2].record()
        torch.cuda.synchronize()
        wa_ms[0] += ev[0].elapsed_time(ev[1]) / reps
        wa_ms[1] += ev[1].elapsed_time(ev[2]) / reps
    walls_info = {"find_walls_ms": round(wa_ms[0], 4), "wall_aspect_ms": round(wa_ms[1], 4),
                  "wall_cells": int((w_dev > 0).sum().item()), "filter_size": WG.filter_size(tile.scale)}
    log(f"walls/aspect on the device: {wa_ms[0]:.3f} + {wa_ms[1]:.3f} ms")
    cpu = None
    if not args.no_cpu_baseline:
        if args.cpu_rows:
            csite = synth.make_site(args.cpu_rows, args.cpu_rows, seed=args.seed)
            csun = sun_for_site(csite.latitude, csite.longitude, csite.utc_offset, 2021, 180)
            cmet = synth.make_met(csun.altitude, seed=args.seed + 1)
        else:
            csite, csun, cmet = site, sun, met
        log(f"CPU oracle baseline on {csite.dsm.shape}")
        cpu = cpu_baseline(csite, csun, cmet)
        # walls/aspect CPU leg: the numpy oracle (vectorised over wall cells) on the same DSM
        from oracle import walls_aspect as OW
        c0 = time.perf_counter()
        OW.wall_aspect(OW.findwalls(site.dsm, 3.0), tile.scale, site.dsm)
        walls_info["cpu_oracle_ms"] = round((time.perf_counter() - c0) * 1e3, 1)
    line = {
        "metric": "Mpixel-steps/s (Tmrt+UTCI, 1000²×24h)",
        "value": round(value, 3),
        "unit": "Mpixel-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": n_warm,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded city tile generator)",
        "config": {"workload": f"full SOLWEIG solve {rows}x{cols}x{T}h (SVF sweep + sun march + GVF + "
                               f"flux -> Tmrt -> UTCI), 30N DOY 180", "rows": rows, "cols": cols,
                   "time_steps": T, "pixel_size_m": site.pixel_size, "lanes": len(h2h_lanes),
                   "lanes_device_resident": len(lanes), "h2h_chunk_steps": args.h2h_chunk,
                   "h2h_copies": args.h2h_copy,
                   "h2h_fields_copy": "after the day" if args.h2h_copy == "thread" else args.h2h_fields,
                   "timed_region": "host-resident inputs (6 rasters, pinned) -> host-resident outputs (16 SVF "
                                   "fields + 24 x {tmrt, utci, shadow}, pinned, one buffer set per solve): "
                                   "upload, SVF sweep, day loop, download; day plans in planner processes",
                   "march": march_info},
        "roofline": roof,
        "rooflines": rooflines,
        "cpu_baseline": cpu,
        "value_device_resident": round(dev_value, 3),
        "ms_per_step_device_resident": round(dev_ms, 3),
        "median_solve_gap_ms": round(gap_ms, 3),
        "pcie_d2h_gbs": round(d2h_gbs, 1),
        "h2h_bound_ms": round(bound_ms, 3),
        "kernel_ms": {k: round(t / nn, 4) for k, (nn, t) in ksum_alone.items()},
        "kernel_ms_lanes": {k: round(t / nn, 4) for k, (nn, t) in ksum.items()},
        "host_plan_ms": round(plan_ms, 3),
        "walls_aspect": walls_info,
    }
    line.update(run_extras(args, extras, serve_pipeline))
    print(json.dumps(line), flush=True)


def run_extras(args, extras, serve_pipeline):
    """Summaries of the other configurations, measured after the tile workload in the same run: the
    serving mix (edit p99 to the host), the gated SVF surrogate with the shipped weights, and the batched
    tiles and edit scenarios. Each full record is logged to stderr."""
    import torch
    out = {}
    for name in extras:
        t0 = time.perf_counter()
        if name == "serve":
            try:
                r = run_serve(args, serve_pipeline, requests=args.extra_requests)
            finally:
                serve_pipeline.close()
            out["serve"] = {"edit_p99_ms": r["edit"]["p99_ms"], "edit_p50_ms": r["edit"]["p50_ms"],
                            "edits": r["edit"]["n"], "all_p99_ms": r["all"]["p99_ms"],
                            "met_p99_ms": (r["met"] or {}).get("p99_ms"), "full_p99_ms": (r["full"] or {}).get("p99_ms"),
                            "requests": r["config"]["requests"], "rate_per_s": r["config"]["rate_per_s"],
                            "completed_per_s": r["completed_per_s"], "admission": r["admission"],
                            "latency": "arrival to results in pinned host memory; nearest-rank percentiles",
                            "site": "2000x2000x24h"}
        elif name == "surrogate":
            import copy
            fast = copy.copy(args)
            fast.surrogate_weights = str(ROOT / "solweig_project_amd" / "data" / "svf_surrogate_16_32_64_128.npz")
            rf = run_surrogate(fast)
            log(f"surrogate (fast network): {json.dumps(rf)}")
            r = run_surrogate(args)
            bar = {"svf_max_abs": 0.02, "tmrt_max_abs_K": 0.5}
            ge, te = r["gated_error"]["max"], r["derived_tmrt_error"]["max_K"]
            out["surrogate"] = {"speedup_at_bar": r["value"], "speedup_mean_gate": r["value_mean_gate"],
                                "exact_svf_ms": r["exact_svf_ms"], "gated_ms": r["gated_ms"],
                                "gated_svf_error_max": ge, "derived_tmrt_error_max_K": te, "bar": bar,
                                "meets_bar": bool(ge <= bar["svf_max_abs"] and te <= bar["tmrt_max_abs_K"]),
                                "accepted_windows": f"{r['accepted']}/{r['windows']}", "gate": "mean, 4 tiles per core",
                                "at_bar": r["at_bar"],
                                "gates": [{k: g[k] for k in ("criterion", "samples_per_core", "speedup", "svf_error_max",
                                                             "tmrt_error_max_K", "meets_bar")} for g in r["gates"]],
                                "params": r["config"]["params"], "site": "2000x2000",
                                "fast_network": {"params": rf["config"]["params"], "speedup_at_bar": rf["value"],
                                                 "speedup_mean_gate": rf["value_mean_gate"],
                                                 "gated_svf_error_max": rf["gated_error"]["max"],
                                                 "derived_tmrt_error_max_K": rf["derived_tmrt_error"]["max_K"],
                                                 "at_bar": rf["at_bar"]}}
        elif name == "batch":
            r = run_batch(args, reps=3)
            out["batch"] = {"value": r["value"], "unit": r["unit"], "ms": r["ms_per_step"],
                            "groups": {g: {k: v[k] for k in ("items", "mpx", "ms", "mpixel_steps_per_s")}
                                       for g, v in r["groups"].items()}}
        else:
            raise SystemExit(f"unknown extra {name!r}")
        log(f"{name}: {json.dumps(r)}")
        out[name]["seconds"] = round(time.perf_counter() - t0, 1)
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    main()
