"""The C ABI: header declarations, exported symbols, ctypes signatures, struct layout.

Runs on the CPU: loads libsolweig_hip.so and calls entry points only with arguments they refuse
before any device work.
"""
import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from solweig_project_amd import _lib as L

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "solweig_hip.h"


def _declarations():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decls = {}
    for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sw_\w+)\s*\(([^;]*?)\)\s*;", text, flags=re.M | re.S):
        args = m.group(2).strip()
        n = 0 if args in ("void", "") else len([a for a in args.split(",") if a.strip()])
        decls[m.group(1)] = n
    return decls


def test_header_declares_entry_points():
    decls = _declarations()
    for name in ("sw_abi_version", "sw_last_error", "sw_svf_march", "sw_svf_fold", "sw_compose",
                 "sw_sun_step", "sw_flux_step", "sw_utci", "sw_run_march_svf_shadow",
                 "sw_run_march_wallheight23", "sw_unpack_masks", "sw_tile_static"):
        assert name in decls, name


def test_library_exports_every_declared_symbol():
    if not L.LIB_PATH.exists():
        pytest.fail(f"{L.LIB_PATH} not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", str(L.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [d for d in _declarations() if d not in exported]
    assert not missing, missing


def test_ctypes_signatures_match_header():
    decls = _declarations()
    assert set(L.SIGNATURES) == set(decls), set(L.SIGNATURES) ^ set(decls)
    for name, (_, args) in L.SIGNATURES.items():
        assert len(args) == decls[name], (name, len(args), decls[name])


def test_library_loads_and_reports_version():
    lib = L.load()
    assert lib.sw_abi_version() == 1
    assert b"gfx950" in lib.sw_build_flags()


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "solweig_hip.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(sw_step_t), sizeof(sw_model_t),'
                   ' offsetof(sw_step_t, pk), offsetof(sw_step_t, pl), offsetof(sw_model_t, walk_hi),'
                   ' sizeof(sw_tile_t), offsetof(sw_tile_t, state), sizeof(sw_wall_filters_t),'
                   ' offsetof(sw_wall_filters_t, off));'
                   'printf("%zu %zu %zu\\n", sizeof(sw_batch_item_t), offsetof(sw_batch_item_t, out),'
                   ' offsetof(sw_batch_item_t, flux_block0));'
                   'printf("%zu %zu %zu %zu\\n", sizeof(sw_march_config_t), offsetof(sw_march_config_t, split),'
                   ' sizeof(sw_cnn_op_t), offsetof(sw_cnn_op_t, site_cols));'
                   'return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(exe), str(src)], check=True)
    vals = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()))
    assert vals == [ctypes.sizeof(L.StepStruct), ctypes.sizeof(L.ModelStruct), L.StepStruct.pk.offset,
                    L.StepStruct.pl.offset, L.ModelStruct.walk_hi.offset, ctypes.sizeof(L.TileStruct),
                    L.TileStruct.state.offset, ctypes.sizeof(L.WallFiltersStruct), L.WallFiltersStruct.off.offset,
                    ctypes.sizeof(L.BatchItemStruct), L.BatchItemStruct.out.offset,
                    L.BatchItemStruct.flux_block0.offset, ctypes.sizeof(L.MarchConfig), L.MarchConfig.split.offset,
                    ctypes.sizeof(L.CnnOp), L.CnnOp.site_cols.offset]


def test_loader_fails_loudly_without_library(tmp_path):
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        L.load(tmp_path / "missing.so")


def test_argument_refusals_return_before_any_launch():
    """Entry points refuse bad arguments with a message and an error code before touching the
    device (pointers here are never dereferenced): oversized tiles for the padded march buffer, an
    empty output mask, a malformed convolution op, bad gate sizes."""
    L.load()
    fake = ctypes.c_void_p(256)
    with pytest.raises(L.HipError, match="too large"):
        L.call("sw_svf_march", fake, fake, fake, fake, fake, fake, fake, 14000, 14000, None)
    tile = L.TileStruct()
    tile.rows, tile.cols = 8, 8
    day = (ctypes.c_int32 * 2)(0, 1)
    with pytest.raises(L.HipError, match="selects nothing"):
        L.call("sw_solve_day", ctypes.byref(tile), fake, fake, fake, day, 0, 2, fake, 0, 0, 0, None)
    with pytest.raises(L.HipError, match="step range"):
        L.call("sw_solve_day", ctypes.byref(tile), fake, fake, fake, day, 2, 1, fake, 1, 0, 0, None)
    op = L.CnnOp(kind=L.CNN_CONV, n=1, h=8, w=8, c1=8, c2=0, o1=16, cout=16, ks=3)
    op.in0, op.out0, op.weight = fake, fake, fake
    arr = (L.CnnOp * 1)(op)
    with pytest.raises(L.HipError, match="multiple of 16"):
        L.call("sw_cnn_fwd", arr, 1, None)
    with pytest.raises(L.HipError, match="bad sizes"):
        L.call("sw_gate_write", fake, fake, fake, fake, fake, 1, 64, 16, 32, 16, 4, 8, 8, None)
    for bad in ((5, 0, 1), (8, 3, 1), (8, 0, 3), (12, -1, 2)):
        cfg = L.MarchConfig(chunk=bad[0], shape=bad[1], split=bad[2])
        with pytest.raises(L.HipError, match="march configuration"):
            L.call("sw_svf_march_cfg", *[fake] * 7, 8, 8, ctypes.byref(cfg), None)
        with pytest.raises(L.HipError, match="march configuration"):
            L.call("sw_svf_sweep_cfg", *[fake] * 14, 8, 8, ctypes.byref(cfg), None)


def test_march_tuner_pinned_configuration():
    """A pinned march configuration is used for every tile without timing; one outside the
    candidates is refused."""
    pytest.importorskip("torch")
    from solweig_project_amd.solver import MarchTuner
    t = MarchTuner(fixed=(8, 0, 2))
    assert not t.enabled
    c = t.config(None, None, None, None)
    assert (c.chunk, c.shape, c.split) == (8, 0, 2) and t.best == {}
    with pytest.raises(ValueError):
        MarchTuner(fixed=(16, 0, 1))


def test_build_flags_keep_one_rounding_per_operation():
    """The device library and the oracle are built without FMA contraction, fast math or
    flush-to-zero (the reference gates the same build mutations), and without packed-FP32
    instruction selection."""
    from solweig_project_amd import build as B
    flags = B.HIPCC_FLAGS
    assert "-ffp-contract=off" in flags and "-fno-fast-math" in flags
    assert "-packed-fp32-ops" in flags and flags[flags.index("-packed-fp32-ops") - 1] == "-Xclang"
    for bad in ("-ffast-math", "-ffp-contract=fast", "-ffp-contract=on", "-fgpu-flush-denormals-to-zero",
                "-funsafe-math-optimizations", "-ffinite-math-only", "-O3 -ffast-math"):
        assert bad not in flags, bad
    make = (Path(__file__).resolve().parents[1] / "oracle" / "Makefile").read_text()
    assert "-ffp-contract=off" in make and "-ffast-math" not in make and "-Ofast" not in make


def test_persistent_day_refusals_return_before_any_launch():
    """sw_solve_batch_persistent refuses a missing barrier workspace, a time-out outside (0, 60] s,
    an empty output mask and items not laid out by sw_batch_layout before touching the device."""
    L.load()
    fake = ctypes.c_void_p(256)
    items = (L.BatchItemStruct * 1)()
    items[0].tile.rows, items[0].tile.cols = 16, 16
    for name in ("hpack", "walls", "aspect_rad", "buildings", "cls", "stats", "masks", "svfbuveg", "tan_asvf",
                 "tan_svfalfa", "walkface", "src", "state"):
        setattr(items[0].tile, name, 256)
    items[0].steps, items[0].sun_tab, items[0].out = 256, 256, 256
    day = (ctypes.c_int32 * 2)(0, 1)
    args = lambda sync, mask, timeout: (fake, items, 1, fake, day, 0, 2, mask, 22, 0, sync, timeout, None)
    with pytest.raises(L.HipError, match="NULL pointer"):
        L.call("sw_solve_batch_persistent", *args(None, 1, 5.0))
    for bad in (0.0, -1.0, 61.0, float("nan")):
        with pytest.raises(L.HipError, match="timeout"):
            L.call("sw_solve_batch_persistent", *args(fake, 1, bad))
    with pytest.raises(L.HipError, match="selects nothing"):
        L.call("sw_solve_batch_persistent", *args(fake, 0, 5.0))
    items[0].flux_block0 = 3
    with pytest.raises(L.HipError, match="sw_solve_batch_persistent: items not laid out"):
        L.call("sw_solve_batch_persistent", *args(fake, 1, 5.0))


def test_device_upload_refusals():
    """sw_copy_to_device_cu refuses unaligned pointers and a source that is not pinned host memory before
    any launch (its kernel dereferences the source on the device)."""
    L.load()
    buf = np.zeros(64, np.uint8)
    with pytest.raises(L.HipError, match="16-byte aligned"):
        L.call("sw_copy_to_device_cu", ctypes.c_void_p(256 + 4), ctypes.c_void_p(512), 64, None)
    src = buf.ctypes.data + (-buf.ctypes.data) % 16
    with pytest.raises(L.HipError, match="not pinned host memory"):
        L.call("sw_copy_to_device_cu", ctypes.c_void_p(4096), ctypes.c_void_p(src), 16, None)
