"""Typed solve requests (mirror of ``solweig_core/request.py:66-259``).

The logical domain (whole grid) and the physical windows (cells read /
written) are separate objects; time coverage is in GLOBAL timestep indices.
Validation raises :class:`RequestValidationError` with a typed refusal.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Mapping

import numpy as np

from .status import RefusalReason, RequestValidationError, TypedRefusal

DEFAULT_DTYPE_PLAN: dict[str, str] = {
    "building_dsm": "<f4", "dem": "<f4", "veg_canopy": "<f4", "walls": "<f4",
    "wall_aspect": "<f4", "landcover": "|u1",
}
DEFAULT_OUTPUT_DTYPE_PLAN: dict[str, str] = {"utci": "<f4", "tmrt": "<f4", "shadow": "<f4"}

#: request vocabulary of devices (dispatch resolves, never guesses)
DEVICES = ("cpu", "cuda")


def _refusal(detail: str, context: dict[str, Any] | None = None):
    return RequestValidationError(TypedRefusal(RefusalReason.REQUEST_INVALID, detail, context))


@dataclass(frozen=True)
class PhysicalWindow:
    """Half-open cell window [row_start, row_stop) x [col_start, col_stop)."""

    row_start: int
    row_stop: int
    col_start: int
    col_stop: int

    def validate(self, *, rows: int, cols: int, label: str = "window") -> None:
        if self.row_start < 0 or self.col_start < 0:
            raise _refusal(f"{label} origin ({self.row_start}, {self.col_start}) is negative",
                           {"window": self.list()})
        if self.row_stop <= self.row_start or self.col_stop <= self.col_start:
            raise _refusal(f"{label} is empty or inverted: {self}", {"window": self.list()})
        if self.row_stop > rows or self.col_stop > cols:
            raise _refusal(f"{label} {self} exceeds the logical domain grid ({rows} rows x {cols} cols)",
                           {"window": self.list(), "grid": [rows, cols]})

    @property
    def height(self) -> int:
        return self.row_stop - self.row_start

    @property
    def width(self) -> int:
        return self.col_stop - self.col_start

    @property
    def origin(self) -> tuple[int, int]:
        return (self.row_start, self.col_start)

    def list(self) -> list[int]:
        return [self.row_start, self.row_stop, self.col_start, self.col_stop]

    def __iter__(self):
        return iter((self.row_start, self.row_stop, self.col_start, self.col_stop))

    def contains(self, other: "PhysicalWindow") -> bool:
        return (self.row_start <= other.row_start and self.row_stop >= other.row_stop
                and self.col_start <= other.col_start and self.col_stop >= other.col_stop)


@dataclass(frozen=True)
class TimeCoverage:
    """GLOBAL time indices [time_start, time_stop) of the forcing series."""

    time_start: int
    time_stop: int | None
    total_steps: int

    def __post_init__(self) -> None:
        if self.time_stop is None:
            object.__setattr__(self, "time_stop", int(self.total_steps))

    def validate(self) -> None:
        if int(self.total_steps) <= 0:
            raise _refusal(f"total_steps must be positive, got {self.total_steps}")
        if int(self.time_start) < 0 or int(self.time_start) >= int(self.time_stop):
            raise _refusal(f"time_start {self.time_start} outside the replay range [0, {self.time_stop})",
                           {"time": [self.time_start, self.time_stop]})
        if int(self.time_stop) > int(self.total_steps):
            raise _refusal(f"time_stop {self.time_stop} exceeds the forcing series ({self.total_steps} global "
                           "timesteps)", {"time": [self.time_start, self.time_stop]})

    @property
    def global_indices(self) -> tuple[int, ...]:
        return tuple(range(int(self.time_start), int(self.time_stop)))


@dataclass(frozen=True)
class SolveRequestView:
    """Everything the core needs to route one window solve; the device is DECLARED."""

    logical_domain_id: str
    rows: int
    cols: int
    origin_x_m: float
    origin_y_m: float
    pixel_size_m: float
    read_window: PhysicalWindow
    write_window: PhysicalWindow
    time: TimeCoverage
    requested_variables: tuple[str, ...] = ("utci", "tmrt", "shadow")
    dtype_plan: Mapping[str, str] = field(default_factory=lambda: dict(DEFAULT_DTYPE_PLAN))
    output_dtype_plan: Mapping[str, str] = field(default_factory=lambda: dict(DEFAULT_OUTPUT_DTYPE_PLAN))
    profile_id: str = ""
    device: str = "cpu"

    def validate(self) -> None:
        if not isinstance(self.logical_domain_id, str) or not self.logical_domain_id:
            raise _refusal("logical_domain_id must be a non-empty string")
        if self.rows <= 0 or self.cols <= 0:
            raise _refusal(f"logical domain shape ({self.rows}, {self.cols}) is empty")
        if not (self.pixel_size_m > 0.0):
            raise _refusal(f"pixel_size_m must be positive, got {self.pixel_size_m!r}")
        self.read_window.validate(rows=self.rows, cols=self.cols, label="read window")
        self.write_window.validate(rows=self.rows, cols=self.cols, label="write window")
        if not self.read_window.contains(self.write_window):
            raise _refusal(f"write window {self.write_window} is not contained in the read window "
                           f"{self.read_window}", {"read": self.read_window.list(),
                                                   "write": self.write_window.list()})
        self.time.validate()
        if not self.requested_variables:
            raise _refusal("requested_variables must be non-empty")
        for plan_name, plan in (("dtype_plan", self.dtype_plan), ("output_dtype_plan", self.output_dtype_plan)):
            for plane, dtype_str in dict(plan).items():
                try:
                    np.dtype(dtype_str)
                except TypeError as error:
                    raise _refusal(f"{plan_name}[{plane!r}] = {dtype_str!r} is not a numpy dtype") from error
        if not self.profile_id:
            raise _refusal("profile_id is required (profile identity of the run)")
        if self.device not in DEVICES:
            raise _refusal(f"device {self.device!r} is outside the request vocabulary ('cpu' or 'cuda'); "
                           "dispatch never guesses")

    def describe(self) -> dict[str, Any]:
        return {
            "logical_domain_id": self.logical_domain_id, "rows": self.rows, "cols": self.cols,
            "origin_x_m": self.origin_x_m, "origin_y_m": self.origin_y_m, "pixel_size_m": self.pixel_size_m,
            "read_window": self.read_window.list(), "write_window": self.write_window.list(),
            "time_start": self.time.time_start, "time_stop": self.time.time_stop,
            "total_steps": self.time.total_steps, "requested_variables": list(self.requested_variables),
            "dtype_plan": dict(self.dtype_plan), "output_dtype_plan": dict(self.output_dtype_plan),
            "profile_id": self.profile_id, "device": self.device,
        }
