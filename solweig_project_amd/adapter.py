"""Solve through the core request surface (mirror of ``solve_with_core_adapter``,
``solweig_gpu/incremental/solver.py:2695-2829``).

A :class:`~solweig_project_amd.solweig_core.SolveRequestView` declares the
logical domain, the read window (inputs incl. halo), the write window
(outputs), the GLOBAL time coverage, the requested variables, dtype plans
and the device. :func:`solve_with_core_adapter`:

1. validates the request and resolves its declared device through
   ``solweig_core.dispatch`` (``'cuda'`` -> ``'hip-gfx950'``; anything else
   is a typed BLOCKED refusal, never a silent re-target);
2. checks the input planes as frozen ABI views: each covers the read
   window, carries the request's logical domain id and the read window's
   global origin, and has the planned dtype;
3. runs :func:`~solweig_project_amd.api.run_utci_window` on the read
   window with ``out_window`` = the write window and
   ``[time_start, time_stop)`` = the request's global time coverage;
4. returns frozen read-only output views whose ``global_origin`` is
   (first global timestep, write-window row, write-window column).

Unsupported inputs raise :class:`UnsupportedInputError` with the
``unsupported_input`` refusal reason.
"""
from __future__ import annotations

from typing import Any, Mapping, MutableMapping

import numpy as np

from . import api
from .solver import UnsupportedInput
from .solweig_core import dispatch as core_dispatch
from .solweig_core.abi import ArrayView
from .solweig_core.request import SolveRequestView
from .solweig_core.status import (AbiValidationError, DispatchBlockedError, RefusalReason, TypedRefusal,
                                  UnsupportedInputError)

#: request plane name -> run_utci_window argument
PLANES = {"building_dsm": "a", "dem": "temp2", "veg_canopy": "temp1", "walls": "walls", "wall_aspect": "dirwalls",
          "landcover": "landcover_grid"}
REQUIRED = ("building_dsm", "dem", "veg_canopy", "walls", "wall_aspect")


def _abi_refusal(detail: str, **ctx) -> AbiValidationError:
    return AbiValidationError(TypedRefusal(RefusalReason.ABI_INVALID, detail, ctx or None))


def _check_plane(name: str, view: ArrayView, request: SolveRequestView) -> None:
    rw = request.read_window
    if not isinstance(view, ArrayView):
        raise _abi_refusal(f"plane {name!r} is not an ArrayView", plane=name)
    view.validate()
    if not view.frozen:
        raise _abi_refusal(f"plane {name!r} must cross the boundary as a frozen view (freeze=True)", plane=name)
    if view.logical_domain_id != request.logical_domain_id:
        raise _abi_refusal(f"plane {name!r} belongs to domain {view.logical_domain_id!r}, the request to "
                           f"{request.logical_domain_id!r}", plane=name)
    if tuple(view.shape) != (rw.height, rw.width):
        raise _abi_refusal(f"plane {name!r} has shape {tuple(view.shape)}, the read window is "
                           f"{(rw.height, rw.width)}", plane=name)
    if tuple(view.global_origin) != tuple(rw.origin):
        raise _abi_refusal(f"plane {name!r} has global origin {tuple(view.global_origin)}, the read window "
                           f"starts at {tuple(rw.origin)}", plane=name)
    want = request.dtype_plan.get(name)
    if want is not None and np.dtype(view.dtype) != np.dtype(want):
        raise _abi_refusal(f"plane {name!r} has dtype {np.dtype(view.dtype).str}, the plan says {want}",
                           plane=name)
    view.assert_not_mutated()


def _array(view: ArrayView):
    if view.residency == "hip":
        import torch
        return torch.as_tensor(view.owner, device="cuda")
    return np.asarray(view.to_numpy())


def solve_with_core_adapter(request: SolveRequestView, planes: Mapping[str, ArrayView], *, met: np.ndarray,
                            sun, latitude: float, svf_bundle=None, initial_state: Mapping[str, Any] | None = None,
                            lc_class=None, output_views: MutableMapping[str, ArrayView] | None = None,
                            prefer_backend: str | None = None) -> dict[str, np.ndarray]:
    """Solve the request's write window over its time coverage on the HIP backend.

    planes: ArrayViews over the read window keyed by the request's plane
    names (``building_dsm``, ``dem``, ``veg_canopy``, ``walls``,
    ``wall_aspect``, optional ``landcover``). met: the full forcing table
    (rows = global timesteps); sun: a SunTable (or any object with
    altitude/azimuth/zen/jday/dectime/altmax arrays) over the same rows.
    svf_bundle: SVF of the read window (computed here when None).
    """
    plan = core_dispatch.resolve(request, prefer_backend=prefer_backend).raise_if_blocked()
    if plan.backend != core_dispatch.HIP_BACKEND:
        raise DispatchBlockedError(TypedRefusal(RefusalReason.UNKNOWN_BACKEND,
                                                f"backend {plan.backend!r} is not served by this adapter"))
    missing = [p for p in REQUIRED if p not in planes]
    if missing:
        raise _abi_refusal(f"missing input planes {missing}", missing=missing)
    unknown = sorted(set(planes) - set(PLANES))
    if unknown:
        raise _abi_refusal(f"unknown input planes {unknown}", unknown=unknown)
    for name, view in planes.items():
        _check_plane(name, view, request)
    met = np.asarray(met, dtype=np.float64)
    t = request.time
    if met.shape[0] < t.time_stop:
        raise _abi_refusal(f"forcing has {met.shape[0]} rows, the request needs {t.time_stop}")
    rw, ww = request.read_window, request.write_window
    out_window = (ww.row_start - rw.row_start, ww.row_stop - rw.row_start,
                  ww.col_start - rw.col_start, ww.col_stop - rw.col_start)
    args = {PLANES[k]: _array(v) for k, v in planes.items()}
    scale = 1.0 / float(request.pixel_size_m)
    try:
        if svf_bundle is None:
            # SVF inputs composed as the reference's driver does (utci_process.py:667-676)
            host = {k: np.asarray(v.cpu() if hasattr(v, "cpu") else v, np.float32)
                    for k, v in args.items() if k in ("a", "temp1", "temp2")}
            a, t2 = host["a"], host["temp2"]
            t1 = np.where(host["temp1"] < 0, np.float32(0), host["temp1"])
            v1 = t1 + a
            v1[v1 == a] = 0
            v2 = t1 * np.float32(0.25) + a
            v2[v2 == a] = 0
            vegdem = t1 + t2
            vegdem2 = t1 * np.float32(0.25) + t2
            bush = np.logical_not(vegdem2 * vegdem) * vegdem
            amax = max(float(a.max()), float(vegdem.max()))
            svf_bundle = api.svf_calculator(2, amaxvalue=amax, a=a, vegdem=v1, vegdem2=v2, bush=bush, scale=scale)
        out = api.run_utci_window(
            svf_bundle=svf_bundle, met_file=met, altitude=np.reshape(sun.altitude, (1, -1)),
            azimuth=np.reshape(sun.azimuth, (1, -1)), zen=np.reshape(sun.zen, (1, -1)),
            jday=np.reshape(sun.jday, (1, -1)), dectime=np.asarray(sun.dectime),
            altmax=np.reshape(sun.altmax, (1, -1)), location={"latitude": float(latitude)}, scale=scale,
            lc_class=lc_class, time_start=t.time_start, time_stop=t.time_stop,
            requested_variables=tuple(request.requested_variables), out_window=out_window,
            initial_state=initial_state, **args)
    except UnsupportedInput as exc:
        raise UnsupportedInputError(TypedRefusal(RefusalReason.UNSUPPORTED_INPUT, str(exc))) from exc
    result = {}
    for name, arr in out.items():
        if name not in request.requested_variables and name != "utci":
            continue
        want = request.output_dtype_plan.get(name, "<f4")
        arr = np.ascontiguousarray(arr, dtype=np.dtype(want))
        arr.setflags(write=False)
        result[name] = arr
        if output_views is not None:
            output_views[name] = ArrayView.from_numpy(arr, global_origin=(t.time_start, ww.row_start, ww.col_start),
                                                      logical_domain_id=request.logical_domain_id, freeze=True)
    for view in planes.values():
        view.assert_not_mutated()
    return result
