"""Strided buffer views of the core surface with device residency.

Interface kept from ``solweig_core/abi.py:86-365``: an :class:`ArrayView` records owner, dtype, shape,
byte strides, global origin, logical domain id, residency and borrow mode; inconsistencies are
refused with ``abi_invalid``; frozen views detect later mutation (``stale_buffer``, with the
``frozen_sha256`` / ``current_sha256`` context keys); non-contiguous buffers are refused or copied
(``non_contiguous``). The phrases the reference's tests match are kept ("declared shape", "strides",
"dtype", "global_origin rank", "negative global origin", "zero byte stride", "read-only owner",
"residency", "never frozen").

Our own design: the owner's real layout is read once into a :class:`_Layout` by a reader per
residency ('cpu': numpy arrays; 'hip': any device buffer exposing ``__cuda_array_interface__``, as
PyTorch-ROCm tensors do), and validation is a rule table comparing the declared layout with it. The
module imports numpy only.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field
from typing import Any, NamedTuple

import numpy as np

from .status import AbiError, AbiValidationError, NonContiguousError, RefusalReason, StaleBufferError, TypedRefusal

RESIDENCY_CPU = "cpu"
RESIDENCY_HIP = "hip"
KNOWN_RESIDENCIES = (RESIDENCY_CPU, RESIDENCY_HIP)


class _Layout(NamedTuple):
    dtype: np.dtype
    shape: tuple[int, ...]
    strides: tuple[int, ...]
    read_only: bool


def _c_strides(shape, itemsize: int) -> tuple[int, ...]:
    """Byte strides of a C-ordered array of ``shape``."""
    if len(shape) == 0:
        return ()
    sizes = [int(s) for s in shape][1:] + [1]
    return tuple(int(v) * itemsize for v in np.cumprod(sizes[::-1])[::-1])


def _numpy_layout(owner) -> _Layout | None:
    if not isinstance(owner, np.ndarray):
        return None
    return _Layout(owner.dtype, tuple(map(int, owner.shape)), tuple(map(int, owner.strides)),
                   not owner.flags.writeable)


def _device_layout(owner) -> _Layout | None:
    iface = getattr(owner, "__cuda_array_interface__", None)
    if iface is None:
        return None
    dtype = np.dtype(iface["typestr"])
    shape = tuple(map(int, iface["shape"]))
    strides = iface.get("strides")
    strides = _c_strides(shape, dtype.itemsize) if strides is None else tuple(map(int, strides))
    return _Layout(dtype, shape, strides, bool(iface["data"][1]))


_READERS = {RESIDENCY_CPU: (_numpy_layout, "a numpy ndarray"),
            RESIDENCY_HIP: (_device_layout, "a buffer with a device array interface")}


def _host_copy(owner) -> np.ndarray:
    if isinstance(owner, np.ndarray):
        return owner
    if hasattr(owner, "cpu"):
        return owner.detach().cpu().numpy()
    raise AbiError(f"a {type(owner).__name__} owner cannot be read on the host")


def _content_sha256(owner) -> str:
    """sha256 of the logical values in C order (independent of the owner's strides)."""
    values = np.ascontiguousarray(_host_copy(owner))
    digest = hashlib.sha256()
    flat = values.reshape(-1).view(np.uint8) if values.size else values.reshape(-1)
    step = 1 << 24                                   # bytes per update
    for lo in range(0, flat.size, step):
        digest.update(memoryview(flat[lo:lo + step]))
    return digest.hexdigest()


def _abi_refusal(detail: str, domain: str | None = None) -> AbiValidationError:
    return AbiValidationError(TypedRefusal(RefusalReason.ABI_INVALID, detail,
                                           context={"plane_domain": domain} if domain is not None else None))


def _require_domain(logical_domain_id) -> None:
    if not (isinstance(logical_domain_id, str) and logical_domain_id):
        raise _abi_refusal("logical_domain_id must name the plane's domain (a non-empty string)")


@dataclass
class ArrayView:
    owner: Any
    dtype: np.dtype
    shape: tuple[int, ...]
    strides: tuple[int, ...]
    global_origin: tuple[int, ...]
    logical_domain_id: str
    residency: str = RESIDENCY_CPU
    read_only: bool = False
    _frozen_digest: str | None = field(default=None, repr=False, compare=False)

    def __post_init__(self) -> None:
        self.validate()

    # -- construction ------------------------------------------------------------------------
    @classmethod
    def _borrow(cls, owner, layout: _Layout, residency: str, global_origin, logical_domain_id: str,
                read_only: bool, freeze: bool) -> "ArrayView":
        view = cls(owner=owner, dtype=layout.dtype, shape=layout.shape, strides=layout.strides,
                   global_origin=tuple(int(v) for v in global_origin), logical_domain_id=logical_domain_id,
                   residency=residency, read_only=read_only)
        return view.freeze() if freeze else view

    @classmethod
    def from_numpy(cls, array: np.ndarray, *, global_origin=(0, 0), logical_domain_id: str,
                   read_only: bool | None = None, freeze: bool = False) -> "ArrayView":
        """Borrow a host array as it is (strides verbatim, no copy)."""
        layout = _numpy_layout(array)
        if layout is None:
            raise _abi_refusal(f"a view borrows a numpy.ndarray, not a {type(array).__name__}")
        _require_domain(logical_domain_id)
        ro = layout.read_only if read_only is None else bool(read_only)
        return cls._borrow(array, layout, RESIDENCY_CPU, global_origin, logical_domain_id, ro, freeze)

    @classmethod
    def from_device(cls, buf, *, global_origin=(0, 0), logical_domain_id: str, read_only: bool = False,
                    freeze: bool = False) -> "ArrayView":
        """Borrow a HIP device buffer through its ``__cuda_array_interface__``."""
        layout = _device_layout(buf)
        if layout is None:
            raise _abi_refusal(f"a {type(buf).__name__} exposes no device array interface")
        _require_domain(logical_domain_id)
        return cls._borrow(buf, layout, RESIDENCY_HIP, global_origin, logical_domain_id,
                           bool(read_only or layout.read_only), freeze)

    # -- validation --------------------------------------------------------------------------
    def _owner_layout(self) -> _Layout:
        reader, kind = _READERS[self.residency]
        layout = reader(self.owner)
        if layout is None:
            raise _abi_refusal(f"the owner of a {self.residency!r} view must be {kind}", self.logical_domain_id)
        return layout

    def validate(self) -> None:
        if self.residency not in _READERS:
            raise _abi_refusal(f"residency {self.residency!r} cannot host a view (hostable: "
                               f"{list(KNOWN_RESIDENCIES)})", self.logical_domain_id)
        own = self._owner_layout()
        rank = len(own.shape)
        zero_axes = [k for k, (n, s) in enumerate(zip(self.shape, self.strides)) if n > 1 and s == 0]
        rules = (
            (len(self.shape) != rank or len(self.strides) != rank,
             lambda: f"declared shape/strides have rank {len(self.shape)}/{len(self.strides)}, the owner {rank}"),
            (tuple(self.shape) != own.shape, lambda: f"declared shape {tuple(self.shape)} differs from the "
                                                     f"owner's {own.shape}"),
            (tuple(self.strides) != own.strides,
             lambda: f"declared byte strides {tuple(self.strides)} differ from the owner's {own.strides}: "
                     "strides are read from the owner, never assumed"),
            (np.dtype(self.dtype) != own.dtype, lambda: f"declared dtype {np.dtype(self.dtype)} differs from the "
                                                        f"owner's {own.dtype}"),
            (len(self.global_origin) != rank,
             lambda: f"global_origin rank {len(self.global_origin)} differs from the owner rank {rank}"),
            (any(int(v) < 0 for v in self.global_origin),
             lambda: f"negative global origin {tuple(self.global_origin)}"),
            (bool(zero_axes), lambda: f"zero byte stride on axis {zero_axes[0]}: a view of overlapping memory "
                                      "is not a valid borrow"),
            (own.read_only and not self.read_only,
             lambda: "a writable view of a read-only owner: declare the view read_only or copy at the boundary"),
        )
        for broken, detail in rules:
            if broken:
                raise _abi_refusal(detail(), self.logical_domain_id)

    # -- contiguity ----------------------------------------------------------------------------
    def is_contiguous(self) -> bool:
        """C-contiguous up to axes of extent 1 (whose strides never matter)."""
        want = _c_strides(self.shape, np.dtype(self.dtype).itemsize)
        return all(n <= 1 or s == w for n, s, w in zip(self.shape, self.strides, want))

    def _non_contiguous(self, detail: str) -> NonContiguousError:
        return NonContiguousError(TypedRefusal(RefusalReason.NON_CONTIGUOUS, detail,
                                               context={"plane_domain": self.logical_domain_id}))

    def require_contiguous(self) -> "ArrayView":
        if self.is_contiguous():
            return self
        raise self._non_contiguous(f"byte strides {tuple(self.strides)} of shape {tuple(self.shape)} are not "
                                   "C order: honour the strides, copy at the boundary (to_contiguous) or refuse")

    def to_contiguous(self) -> "ArrayView":
        if self.is_contiguous():
            return self
        if self.residency != RESIDENCY_CPU:
            raise self._non_contiguous("a device view is copied by its owner's runtime, not by the view")
        packed = np.ascontiguousarray(self.owner)
        return ArrayView._borrow(packed, _numpy_layout(packed), RESIDENCY_CPU, self.global_origin,
                                 self.logical_domain_id, self.read_only, False)

    # -- access ----------------------------------------------------------------------------------
    def to_numpy(self) -> np.ndarray:
        return self.owner if self.residency == RESIDENCY_CPU else _host_copy(self.owner)

    @property
    def nbytes(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64)) * np.dtype(self.dtype).itemsize

    # -- freezing --------------------------------------------------------------------------------
    def freeze(self) -> "ArrayView":
        self._frozen_digest = _content_sha256(self.owner)
        return self

    @property
    def frozen(self) -> bool:
        return self._frozen_digest is not None

    def assert_not_mutated(self) -> None:
        if self._frozen_digest is None:
            raise AbiError("assert_not_mutated needs a view that was frozen; this one was never frozen")
        now = _content_sha256(self.owner)
        if now != self._frozen_digest:
            raise StaleBufferError(TypedRefusal(
                RefusalReason.STALE_BUFFER, "the owner changed after the view was frozen",
                context={"plane_domain": self.logical_domain_id, "frozen_sha256": self._frozen_digest,
                         "current_sha256": now}))

    def same_logical_values(self, other: "ArrayView") -> bool:
        if not isinstance(other, ArrayView):
            return NotImplemented
        if np.dtype(self.dtype) != np.dtype(other.dtype) or tuple(self.shape) != tuple(other.shape):
            return False
        return bool(np.array_equal(self.to_numpy(), other.to_numpy()))

    def describe(self) -> dict[str, Any]:
        return {"dtype": np.dtype(self.dtype).str, "shape": list(self.shape), "strides": list(self.strides),
                "global_origin": list(self.global_origin), "logical_domain_id": self.logical_domain_id,
                "residency": self.residency, "read_only": self.read_only, "contiguous": self.is_contiguous(),
                "frozen": self.frozen}
