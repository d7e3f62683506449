"""Typed refusal / fallback vocabulary (mirror of ``solweig_core/status.py:21-172``).

String values are the interface: run records quote them verbatim, so they
are identical to the reference's.
"""
from __future__ import annotations

from dataclasses import dataclass
from enum import Enum
from typing import Any


class RefusalReason(str, Enum):
    CUDA_UNAVAILABLE = "cuda_unavailable"
    UNSUPPORTED_DEVICE = "unsupported_device"
    ABI_INVALID = "abi_invalid"
    REQUEST_INVALID = "request_invalid"
    STALE_BUFFER = "stale_buffer"
    UNKNOWN_BACKEND = "unknown_backend"
    NON_CONTIGUOUS = "non_contiguous"
    #: inputs outside the domain the HIP backend reproduces exactly (bush
    #: layer, non-finite heights, non-default model switches)
    UNSUPPORTED_INPUT = "unsupported_input"


class FallbackReason(str, Enum):
    LEGACY_ROUNDTRIP = "legacy_roundtrip"
    FULL_TILE_UNSAFE_LOCAL = "full_tile_unsafe_local"
    KERNEL_REFUSED = "kernel_refused"


class DispatchStatus(str, Enum):
    READY = "ready"
    BLOCKED = "blocked"


@dataclass(frozen=True)
class TypedRefusal:
    reason: RefusalReason
    detail: str = ""
    context: dict[str, Any] | None = None

    def to_dict(self) -> dict[str, Any]:
        return {"reason": self.reason.value, "detail": self.detail,
                "context": dict(self.context) if self.context else None}

    def __str__(self) -> str:
        ctx = f" context={self.context}" if self.context else ""
        return f"TypedRefusal({self.reason.value}: {self.detail}{ctx})"


class SolweigCoreError(RuntimeError):
    """Base class of every typed error raised by the core surface."""


class _WithRefusal:
    def __init__(self, refusal: TypedRefusal) -> None:
        super().__init__(str(refusal))
        self.refusal = refusal


class AbiError(SolweigCoreError):
    """A buffer/view violated the buffer ABI."""


class AbiValidationError(_WithRefusal, AbiError):
    """Structural ABI validation failed (owner/shape/strides/residency)."""


class StaleBufferError(_WithRefusal, AbiError):
    """A frozen view detected mutation of its owner after view creation."""


class NonContiguousError(_WithRefusal, AbiError):
    """A strided buffer reached a consumer that cannot honour strides."""


class RequestValidationError(_WithRefusal, SolweigCoreError):
    """A solve request failed structural validation."""


class DispatchBlockedError(_WithRefusal, SolweigCoreError):
    """A BLOCKED dispatch result was consumed as if READY."""


class UnsupportedInputError(_WithRefusal, SolweigCoreError):
    """The inputs lie outside the supported (bit-faithful) domain."""
