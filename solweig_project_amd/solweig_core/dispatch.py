"""Explicit device/backend resolution (the interface of ``solweig_core/dispatch.py:51-191``).

A request declares its device; :func:`resolve` answers READY with a plan or BLOCKED with a typed
refusal, never probing availability and never re-targeting. Kept from the reference: the record
types ``DispatchPlan`` / ``DispatchResult``, their ``describe()`` keys (run-record format) and the
refusal reasons. The registry is ours: the declared device ``'cuda'`` (PyTorch-ROCm's name for a HIP
device) maps to the MI355X backend ``'hip-gfx950'``. This package carries no CPU solver, so
``'cpu'`` has no backend here and is refused (a caller keeps its own CPU backends in its own table).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any

from .request import SolveRequestView
from .status import DispatchBlockedError, DispatchStatus, FallbackReason, RefusalReason, TypedRefusal

HIP_BACKEND = "hip-gfx950"

#: declared device -> backends registered for it, first = default
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
        record = {key: getattr(self, key) for key in ("device", "backend", "profile_id", "logical_domain_id")}
        record["fallback_reason"] = None if self.fallback_reason is None else self.fallback_reason.value
        record["status"] = self.status.value
        return record


@dataclass(frozen=True)
class DispatchResult:
    status: DispatchStatus
    plan: DispatchPlan | None = None
    refusal: TypedRefusal | None = None

    @property
    def ready(self) -> bool:
        return self.status is DispatchStatus.READY and self.plan is not None

    def raise_if_blocked(self) -> DispatchPlan:
        """The plan of a READY result; a BLOCKED one raises its refusal."""
        if self.ready:
            return self.plan
        why = self.refusal if self.refusal is not None else TypedRefusal(
            RefusalReason.UNKNOWN_BACKEND, "the dispatch result carries no plan")
        raise DispatchBlockedError(why)

    def describe(self) -> dict[str, Any]:
        return {"status": self.status.value,
                "plan": None if self.plan is None else self.plan.describe(),
                "refusal": None if self.refusal is None else self.refusal.to_dict()}


def supported_devices() -> tuple[str, ...]:
    return tuple(sorted(_BACKENDS))


def _blocked(reason: RefusalReason, detail: str, **context) -> DispatchResult:
    return DispatchResult(DispatchStatus.BLOCKED, refusal=TypedRefusal(reason, detail, context=context))


def resolve(request: SolveRequestView, prefer_backend: str | None = None) -> DispatchResult:
    """Resolve the declared device of a request (validating it first)."""
    request.validate()
    device = request.device
    registered = _BACKENDS.get(device)
    if registered is None:
        return _blocked(RefusalReason.UNSUPPORTED_DEVICE,
                        f"device {device!r} is outside the dispatch vocabulary {supported_devices()}", device=device)
    if not registered:
        return _blocked(RefusalReason.UNSUPPORTED_DEVICE,
                        f"no backend of this package serves device {device!r}: it provides {HIP_BACKEND!r} under "
                        "'cuda' only, and a CPU request is never moved to the GPU",
                        device=device, logical_domain_id=request.logical_domain_id)
    chosen = registered[0] if prefer_backend is None else prefer_backend
    if chosen not in registered:
        return _blocked(RefusalReason.UNKNOWN_BACKEND,
                        f"backend {chosen!r} is not registered for device {device!r} (registered: {list(registered)})",
                        device=device, prefer_backend=chosen)
    return DispatchResult(DispatchStatus.READY, plan=DispatchPlan(
        device=device, backend=chosen, profile_id=request.profile_id,
        logical_domain_id=request.logical_domain_id))
