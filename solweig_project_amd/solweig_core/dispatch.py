"""Explicit device/backend resolution (mirror of ``solweig_core/dispatch.py:51-191``).

The request DECLARES its device; :func:`resolve` answers READY with a plan
or BLOCKED with a typed refusal. No availability probing, no silent
re-targeting. Backends per declared device: ``'cuda'`` (a ROCm/HIP device
under PyTorch-ROCm's naming) resolves to the MI355X backend; this package
carries no CPU solver, so ``'cpu'`` is refused with a typed reason (a
caller keeps its own CPU backends in its own table).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any

from .request import SolveRequestView
from .status import DispatchStatus, FallbackReason, RefusalReason, TypedRefusal

HIP_BACKEND = "hip-gfx950"

_BACKENDS: dict[str, tuple[str, ...]] = {
    "cpu": (),
    "cuda": (HIP_BACKEND,),
}


@dataclass(frozen=True)
class DispatchPlan:
    device: str
    backend: str
    profile_id: str
    logical_domain_id: str
    fallback_reason: FallbackReason | None = None
    status: DispatchStatus = DispatchStatus.READY

    def describe(self) -> dict[str, Any]:
        return {"device": self.device, "backend": self.backend, "profile_id": self.profile_id,
                "logical_domain_id": self.logical_domain_id,
                "fallback_reason": self.fallback_reason.value if self.fallback_reason else None,
                "status": self.status.value}


@dataclass(frozen=True)
class DispatchResult:
    status: DispatchStatus
    plan: DispatchPlan | None = None
    refusal: TypedRefusal | None = None

    @property
    def ready(self) -> bool:
        return self.status is DispatchStatus.READY

    def raise_if_blocked(self) -> DispatchPlan:
        if self.status is not DispatchStatus.READY or self.plan is None:
            from .status import DispatchBlockedError
            raise DispatchBlockedError(self.refusal or TypedRefusal(RefusalReason.UNKNOWN_BACKEND,
                                                                    "dispatch blocked"))
        return self.plan

    def describe(self) -> dict[str, Any]:
        return {"status": self.status.value, "plan": self.plan.describe() if self.plan else None,
                "refusal": self.refusal.to_dict() if self.refusal else None}


def supported_devices() -> tuple[str, ...]:
    return tuple(sorted(_BACKENDS))


def resolve(request: SolveRequestView, prefer_backend: str | None = None) -> DispatchResult:
    """Resolve the DECLARED device of a request (validated here again)."""
    request.validate()
    if request.device not in _BACKENDS:
        return DispatchResult(DispatchStatus.BLOCKED, refusal=TypedRefusal(
            RefusalReason.UNSUPPORTED_DEVICE,
            f"device {request.device!r} is outside the dispatch vocabulary {sorted(_BACKENDS)}",
            context={"device": request.device}))
    backends = _BACKENDS[request.device]
    if not backends:
        return DispatchResult(DispatchStatus.BLOCKED, refusal=TypedRefusal(
            RefusalReason.UNSUPPORTED_DEVICE,
            f"device {request.device!r}: this package provides only the MI355X backend "
            f"({HIP_BACKEND!r} under device 'cuda'); a CPU request is never run on the GPU",
            context={"device": request.device, "logical_domain_id": request.logical_domain_id}))
    backend = backends[0]
    if prefer_backend is not None:
        if prefer_backend not in backends:
            return DispatchResult(DispatchStatus.BLOCKED, refusal=TypedRefusal(
                RefusalReason.UNKNOWN_BACKEND,
                f"preferred backend {prefer_backend!r} is not registered for device {request.device!r} "
                f"(registered: {list(backends)})",
                context={"device": request.device, "prefer_backend": prefer_backend}))
        backend = prefer_backend
    return DispatchResult(DispatchStatus.READY, plan=DispatchPlan(
        device=request.device, backend=backend, profile_id=request.profile_id,
        logical_domain_id=request.logical_domain_id, fallback_reason=None))
