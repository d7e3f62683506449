"""Request / dispatch / buffer-ABI surface of the SOLWEIG core, with the HIP backend.

Mirrors the reference's ``solweig_core`` package (``solweig_core/status.py``,
``request.py``, ``abi.py``, ``dispatch.py``): typed refusals, validated solve
requests with logical domain + physical windows + global time coverage, a
strided buffer ABI with freeze/stale detection, and explicit device
dispatch. The difference is the backend table: the declared device
``'cuda'`` (PyTorch-ROCm's name for a HIP device) resolves to
``'hip-gfx950'``. This package imports numpy only (no torch).
"""
from .status import (AbiError, AbiValidationError, DispatchBlockedError, DispatchStatus, FallbackReason,
                     NonContiguousError, RefusalReason, RequestValidationError, SolweigCoreError,
                     StaleBufferError, TypedRefusal)
from .request import DEFAULT_DTYPE_PLAN, DEFAULT_OUTPUT_DTYPE_PLAN, PhysicalWindow, SolveRequestView, TimeCoverage
from .abi import KNOWN_RESIDENCIES, RESIDENCY_CPU, RESIDENCY_HIP, ArrayView
from .dispatch import DispatchPlan, DispatchResult, resolve, supported_devices

__all__ = [
    "AbiError", "AbiValidationError", "DispatchBlockedError", "DispatchStatus", "FallbackReason",
    "NonContiguousError", "RefusalReason", "RequestValidationError", "SolweigCoreError",
    "StaleBufferError", "TypedRefusal", "DEFAULT_DTYPE_PLAN", "DEFAULT_OUTPUT_DTYPE_PLAN",
    "PhysicalWindow", "SolveRequestView", "TimeCoverage", "KNOWN_RESIDENCIES", "RESIDENCY_CPU",
    "RESIDENCY_HIP", "ArrayView", "DispatchPlan", "DispatchResult", "resolve", "supported_devices",
]
