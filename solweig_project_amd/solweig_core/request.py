"""Typed solve requests of the core surface (the interface of ``solweig_core/request.py:66-259``).

A request names a logical domain (the whole grid), two physical cell windows (read and write) and a
coverage of GLOBAL forcing timesteps. What is kept from the reference is the interface: the field
names and defaults, the refusal reason (``request_invalid``), the ``describe()`` keys that run records
quote, and the phrases its tests match ("exceeds the logical domain", "empty or inverted",
"replay range", "exceeds the forcing series", "not contained in the read", "numpy dtype",
"vocabulary", ...). The checks themselves are rule tables: each object lists the problems it can
have as (predicate, message) pairs, and :func:`_first_problem` raises the first one that applies.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Callable, Iterable, Mapping

import numpy as np

from .status import RefusalReason, RequestValidationError, TypedRefusal

DEFAULT_DTYPE_PLAN: dict[str, str] = {
    "building_dsm": "<f4", "dem": "<f4", "veg_canopy": "<f4", "walls": "<f4",
    "wall_aspect": "<f4", "landcover": "|u1",
}
DEFAULT_OUTPUT_DTYPE_PLAN: dict[str, str] = {"utci": "<f4", "tmrt": "<f4", "shadow": "<f4"}

#: devices a request may declare; dispatch maps them to backends and never substitutes one
DEVICES = ("cpu", "cuda")

# A rule: (applies() -> bool, message() -> str, context() -> dict | None). Messages are built only
# for the rule that fires.
Rule = tuple[Callable[[], bool], Callable[[], str], Callable[[], "dict[str, Any] | None"]]


def _first_problem(rules: Iterable[Rule]) -> None:
    for applies, message, context in rules:
        if applies():
            raise RequestValidationError(TypedRefusal(RefusalReason.REQUEST_INVALID, message(), context()))


def _no_context():
    return None


def _is_dtype(text: str) -> bool:
    try:
        np.dtype(text)
    except TypeError:
        return False
    return True


@dataclass(frozen=True)
class PhysicalWindow:
    """Cells [row_start, row_stop) x [col_start, col_stop) of the logical grid."""

    row_start: int
    row_stop: int
    col_start: int
    col_stop: int

    @property
    def height(self) -> int:
        return self.row_stop - self.row_start

    @property
    def width(self) -> int:
        return self.col_stop - self.col_start

    @property
    def origin(self) -> tuple[int, int]:
        return self.row_start, self.col_start

    def list(self) -> list[int]:
        return [self.row_start, self.row_stop, self.col_start, self.col_stop]

    def __iter__(self):
        return iter(self.list())

    def _span(self) -> str:
        return f"rows {self.row_start}:{self.row_stop}, cols {self.col_start}:{self.col_stop}"

    def contains(self, other: "PhysicalWindow") -> bool:
        rows_in = self.row_start <= other.row_start and other.row_stop <= self.row_stop
        cols_in = self.col_start <= other.col_start and other.col_stop <= self.col_stop
        return rows_in and cols_in

    def validate(self, *, rows: int, cols: int, label: str = "window") -> None:
        ctx = lambda: {"window": self.list(), "grid": [rows, cols]}  # noqa: E731
        _first_problem((
            (lambda: min(self.row_start, self.col_start) < 0,
             lambda: f"{label} ({self._span()}) starts at a negative cell", ctx),
            (lambda: self.height <= 0 or self.width <= 0,
             lambda: f"{label} ({self._span()}) is empty or inverted", ctx),
            (lambda: self.row_stop > rows or self.col_stop > cols,
             lambda: f"{label} ({self._span()}) exceeds the logical domain of {rows} x {cols} cells", ctx),
        ))


@dataclass(frozen=True)
class TimeCoverage:
    """GLOBAL forcing timesteps [time_start, time_stop) of a series of ``total_steps`` (None = to the end)."""

    time_start: int
    time_stop: int | None
    total_steps: int

    def __post_init__(self) -> None:
        if self.time_stop is None:
            object.__setattr__(self, "time_stop", int(self.total_steps))

    @property
    def global_indices(self) -> tuple[int, ...]:
        return tuple(range(int(self.time_start), int(self.time_stop)))

    def validate(self) -> None:
        start, stop, total = int(self.time_start), int(self.time_stop), int(self.total_steps)
        ctx = lambda: {"time": [start, stop], "total_steps": total}  # noqa: E731
        _first_problem((
            (lambda: total < 1, lambda: f"total_steps is {total}; a forcing series has at least one step", ctx),
            (lambda: not 0 <= start < stop,
             lambda: f"steps {start}..{stop} leave no step of the replay range [0, {stop})", ctx),
            (lambda: stop > total,
             lambda: f"time_stop {stop} exceeds the forcing series of {total} global timesteps", ctx),
        ))


@dataclass(frozen=True)
class SolveRequestView:
    """What the core needs to route one window solve. The device is declared by the caller."""

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

    def _bad_dtypes(self) -> list[str]:
        plans = {"dtype_plan": self.dtype_plan, "output_dtype_plan": self.output_dtype_plan}
        return [f"{name}[{plane!r}] = {text!r}" for name, plan in plans.items()
                for plane, text in dict(plan).items() if not _is_dtype(text)]

    def validate(self) -> None:
        # identity of the run first, then the grid, the windows, time and the outputs
        _first_problem((
            (lambda: not (isinstance(self.logical_domain_id, str) and self.logical_domain_id),
             lambda: "logical_domain_id must name the domain (a non-empty string)", _no_context),
            (lambda: not self.profile_id,
             lambda: "profile_id is required: every solve runs under a named profile", _no_context),
            (lambda: self.device not in DEVICES,
             lambda: f"device {self.device!r} is not in the request vocabulary {list(DEVICES)}; "
                     "dispatch never guesses a device",
             lambda: {"device": self.device}),
            (lambda: self.rows < 1 or self.cols < 1,
             lambda: f"the logical domain of {self.rows} x {self.cols} cells is empty", _no_context),
            (lambda: not self.pixel_size_m > 0.0,
             lambda: f"pixel_size_m is {self.pixel_size_m!r}; cells need a positive size", _no_context),
        ))
        self.read_window.validate(rows=self.rows, cols=self.cols, label="read window")
        self.write_window.validate(rows=self.rows, cols=self.cols, label="write window")
        _first_problem((
            (lambda: not self.read_window.contains(self.write_window),
             lambda: f"write window {self.write_window.list()} is not contained in the read window "
                     f"{self.read_window.list()}",
             lambda: {"read": self.read_window.list(), "write": self.write_window.list()}),
        ))
        self.time.validate()
        _first_problem((
            (lambda: len(self.requested_variables) == 0,
             lambda: "requested_variables lists no output", _no_context),
            (lambda: bool(self._bad_dtypes()),
             lambda: "not a numpy dtype: " + ", ".join(self._bad_dtypes()), _no_context),
        ))

    def describe(self) -> dict[str, Any]:
        """The run-record fields of the request (the reference's record keys)."""
        record: dict[str, Any] = {name: getattr(self, name) for name in (
            "logical_domain_id", "rows", "cols", "origin_x_m", "origin_y_m", "pixel_size_m")}
        record["read_window"] = self.read_window.list()
        record["write_window"] = self.write_window.list()
        record.update(time_start=self.time.time_start, time_stop=self.time.time_stop,
                      total_steps=self.time.total_steps)
        record["requested_variables"] = list(self.requested_variables)
        record["dtype_plan"] = dict(self.dtype_plan)
        record["output_dtype_plan"] = dict(self.output_dtype_plan)
        record.update(profile_id=self.profile_id, device=self.device)
        return record
