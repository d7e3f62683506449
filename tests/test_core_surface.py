"""Core request surface: request validation, explicit dispatch, buffer views, adapter refusals (CPU)."""
import numpy as np
import pytest

from solweig_project_amd.solweig_core import (AbiValidationError, ArrayView, DispatchBlockedError, DispatchStatus,
                                              PhysicalWindow, RefusalReason, RequestValidationError,
                                              SolveRequestView, StaleBufferError, TimeCoverage, resolve)

DOM = "site/test/64x48"


def _req(**over):
    kw = dict(logical_domain_id=DOM, rows=64, cols=48, origin_x_m=0.0, origin_y_m=0.0, pixel_size_m=1.0,
              read_window=PhysicalWindow(0, 64, 0, 48), write_window=PhysicalWindow(8, 40, 4, 44),
              time=TimeCoverage(0, 24, 24), profile_id="mi355x_v1", device="cuda")
    kw.update(over)
    return SolveRequestView(**kw)


@pytest.mark.parametrize("over,match", [
    (dict(write_window=PhysicalWindow(0, 70, 0, 48)), "exceeds the logical domain"),
    (dict(read_window=PhysicalWindow(10, 64, 0, 48)), "not contained in the read window"),
    (dict(time=TimeCoverage(5, 30, 24)), "exceeds the forcing series"),
    (dict(time=TimeCoverage(5, 5, 24)), "replay range"),
    (dict(profile_id=""), "profile_id"),
    (dict(device="rocm"), "vocabulary"),
    (dict(pixel_size_m=0.0), "pixel_size_m"),
    (dict(requested_variables=()), "requested_variables"),
])
def test_request_refusals(over, match):
    with pytest.raises(RequestValidationError, match=match) as e:
        _req(**over).validate()
    assert e.value.refusal.reason is RefusalReason.REQUEST_INVALID


def test_dispatch_is_explicit():
    r = resolve(_req())
    assert r.status is DispatchStatus.READY and r.plan.backend == "hip-gfx950"
    blocked = resolve(_req(device="cpu"))
    assert blocked.status is DispatchStatus.BLOCKED
    assert blocked.refusal.reason is RefusalReason.UNSUPPORTED_DEVICE
    with pytest.raises(DispatchBlockedError):
        blocked.raise_if_blocked()
    unk = resolve(_req(), prefer_backend="numba-full-solve")
    assert unk.refusal.reason is RefusalReason.UNKNOWN_BACKEND


def test_views_freeze_and_detect_mutation():
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    v = ArrayView.from_numpy(a, logical_domain_id=DOM, freeze=True)
    v.assert_not_mutated()
    a[1, 1] = -1
    with pytest.raises(StaleBufferError) as e:
        v.assert_not_mutated()
    assert e.value.refusal.reason is RefusalReason.STALE_BUFFER
    nc = ArrayView.from_numpy(a[:, ::2], logical_domain_id=DOM)
    assert not nc.is_contiguous()
    with pytest.raises(AbiValidationError):
        ArrayView.from_numpy(a, logical_domain_id="")


def _planes(shape=(64, 48), origin=(0, 0), dom=DOM, dtype=np.float32, freeze=True, drop=None):
    names = ("building_dsm", "dem", "veg_canopy", "walls", "wall_aspect")
    out = {n: ArrayView.from_numpy(np.zeros(shape, dtype), global_origin=origin, logical_domain_id=dom,
                                   freeze=freeze) for n in names if n != drop}
    return out


@pytest.mark.parametrize("kw,match", [
    (dict(dom="other"), "belongs to domain"),
    (dict(shape=(60, 48)), "read window is"),
    (dict(origin=(1, 0)), "global origin"),
    (dict(dtype=np.float64), "dtype"),
    (dict(freeze=False), "frozen view"),
    (dict(drop="walls"), "missing input planes"),
])
def test_adapter_refuses_bad_planes(kw, match):
    from solweig_project_amd.adapter import solve_with_core_adapter
    with pytest.raises(AbiValidationError, match=match) as e:
        solve_with_core_adapter(_req(), _planes(**kw), met=np.zeros((24, 25)), sun=None, latitude=30.0)
    assert e.value.refusal.reason is RefusalReason.ABI_INVALID


def test_adapter_blocks_cpu_requests():
    from solweig_project_amd.adapter import solve_with_core_adapter
    with pytest.raises(DispatchBlockedError) as e:
        solve_with_core_adapter(_req(device="cpu"), _planes(), met=np.zeros((24, 25)), sun=None, latitude=30.0)
    assert e.value.refusal.reason is RefusalReason.UNSUPPORTED_DEVICE


def test_core_package_does_not_import_torch():
    import subprocess
    import sys
    code = ("import sys; import solweig_project_amd.solweig_core as c; "
            "print('torch' in sys.modules)")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True,
                         cwd=str(__import__('pathlib').Path(__file__).resolve().parents[1]))
    assert out.stdout.strip() == "False"
