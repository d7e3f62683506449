"""Strided buffer ABI (mirror of ``solweig_core/abi.py:86-365``) with device residency.

An :class:`ArrayView` carries owner, dtype, shape, byte strides, global
origin, logical domain id, residency and borrow mode as data, and refuses
any inconsistency with a typed refusal. Residency ``'cpu'`` owners are numpy
arrays; residency ``'hip'`` owners are device buffers exposing
``__cuda_array_interface__`` (PyTorch-ROCm tensors do), so this module stays
numpy-only. Frozen views pin a digest of the logical bytes and detect
mutation after the solve (host views; device views digest through a host
copy).
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from .status import AbiError, AbiValidationError, NonContiguousError, RefusalReason, StaleBufferError, TypedRefusal

RESIDENCY_CPU = "cpu"
RESIDENCY_HIP = "hip"
KNOWN_RESIDENCIES = (RESIDENCY_CPU, RESIDENCY_HIP)
_DIGEST_CHUNK = 1 << 20


def _host_array(owner) -> np.ndarray:
    if isinstance(owner, np.ndarray):
        return owner
    if hasattr(owner, "cpu"):
        return owner.detach().cpu().numpy()
    raise AbiError(f"cannot read owner of type {type(owner).__name__} on the host")


def _logical_digest(array) -> str:
    a = _host_array(array)
    h = hashlib.sha256()
    for start in range(0, int(a.size), _DIGEST_CHUNK):
        stop = min(start + _DIGEST_CHUNK, int(a.size))
        h.update(np.ascontiguousarray(a.flat[start:stop]).tobytes())
    return h.hexdigest()


def _device_meta(owner):
    cai = getattr(owner, "__cuda_array_interface__", None)
    if cai is None:
        return None
    dtype = np.dtype(cai["typestr"])
    shape = tuple(int(s) for s in cai["shape"])
    strides = cai.get("strides")
    if strides is None:
        strides, acc = [], dtype.itemsize
        for s in reversed(shape):
            strides.insert(0, acc)
            acc *= s
    return dtype, shape, tuple(int(s) for s in strides), bool(cai["data"][1])


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

    @classmethod
    def from_numpy(cls, array: np.ndarray, *, global_origin=(0, 0), logical_domain_id: str,
                   read_only: bool | None = None, freeze: bool = False) -> "ArrayView":
        if not isinstance(array, np.ndarray):
            raise AbiValidationError(TypedRefusal(RefusalReason.ABI_INVALID,
                                                  f"owner must be numpy.ndarray, got {type(array).__name__}"))
        if not isinstance(logical_domain_id, str) or not logical_domain_id:
            raise AbiValidationError(TypedRefusal(RefusalReason.ABI_INVALID,
                                                  "logical_domain_id must be a non-empty string"))
        view = cls(owner=array, dtype=array.dtype, shape=tuple(int(s) for s in array.shape),
                   strides=tuple(int(s) for s in array.strides),
                   global_origin=tuple(int(o) for o in global_origin), logical_domain_id=logical_domain_id,
                   read_only=(not array.flags.writeable) if read_only is None else bool(read_only))
        if freeze:
            view._frozen_digest = _logical_digest(array)
        return view

    @classmethod
    def from_device(cls, buf, *, global_origin=(0, 0), logical_domain_id: str, read_only: bool = False,
                    freeze: bool = False) -> "ArrayView":
        """View of a HIP device buffer exposing ``__cuda_array_interface__``."""
        meta = _device_meta(buf)
        if meta is None:
            raise AbiValidationError(TypedRefusal(RefusalReason.ABI_INVALID,
                                                  f"{type(buf).__name__} exposes no device array interface"))
        if not isinstance(logical_domain_id, str) or not logical_domain_id:
            raise AbiValidationError(TypedRefusal(RefusalReason.ABI_INVALID,
                                                  "logical_domain_id must be a non-empty string"))
        dtype, shape, strides, ro = meta
        view = cls(owner=buf, dtype=dtype, shape=shape, strides=strides,
                   global_origin=tuple(int(o) for o in global_origin), logical_domain_id=logical_domain_id,
                   residency=RESIDENCY_HIP, read_only=bool(read_only or ro))
        if freeze:
            view._frozen_digest = _logical_digest(buf)
        return view

    def _owner_meta(self):
        if self.residency == RESIDENCY_CPU:
            if not isinstance(self.owner, np.ndarray):
                raise self._refusal("owner of a 'cpu' view is not a numpy ndarray")
            o = self.owner
            return o.dtype, tuple(int(s) for s in o.shape), tuple(int(s) for s in o.strides), not o.flags.writeable
        meta = _device_meta(self.owner)
        if meta is None:
            raise self._refusal("owner of a 'hip' view exposes no device array interface")
        return meta

    def validate(self) -> None:
        if self.residency not in KNOWN_RESIDENCIES:
            raise self._refusal(f"residency {self.residency!r} is not hostable (known: {list(KNOWN_RESIDENCIES)})")
        dtype, shape, strides, owner_ro = self._owner_meta()
        ndim = len(shape)
        if len(self.shape) != ndim or len(self.strides) != ndim:
            raise self._refusal(f"shape/strides rank ({len(self.shape)}/{len(self.strides)}) does not match "
                                f"owner rank {ndim}")
        if tuple(self.shape) != shape:
            raise self._refusal(f"declared shape {self.shape} != owner shape {shape}")
        if tuple(self.strides) != strides:
            raise self._refusal(f"declared strides {self.strides} != owner strides {strides} (bytes); strides "
                                "are honoured, not assumed")
        if self.dtype != dtype:
            raise self._refusal(f"declared dtype {self.dtype} != owner dtype {dtype}")
        if len(self.global_origin) != ndim:
            raise self._refusal(f"global_origin rank {len(self.global_origin)} != owner rank {ndim}")
        if any(int(o) < 0 for o in self.global_origin):
            raise self._refusal(f"negative global origin {self.global_origin}")
        for axis, (size, stride) in enumerate(zip(self.shape, self.strides)):
            if size > 1 and stride == 0:
                raise self._refusal(f"zero byte stride on axis {axis} of size {size}: overlapping memory is not "
                                    "a valid borrow")
        if not self.read_only and owner_ro:
            raise self._refusal("writable borrow of a read-only owner; declare the view read_only or copy at "
                                "the boundary")

    def _refusal(self, detail: str) -> AbiValidationError:
        return AbiValidationError(TypedRefusal(RefusalReason.ABI_INVALID, detail,
                                               context={"plane_domain": self.logical_domain_id}))

    def is_contiguous(self) -> bool:
        expected = self.dtype.itemsize
        for size, stride in zip(reversed(self.shape), reversed(self.strides)):
            if size > 1 and stride != expected:
                return False
            if size > 1:
                expected *= size
        return True

    def require_contiguous(self) -> "ArrayView":
        if not self.is_contiguous():
            raise NonContiguousError(TypedRefusal(
                RefusalReason.NON_CONTIGUOUS,
                f"strides {self.strides} are not C-contiguous for shape {self.shape}; handle strides, copy at "
                "the boundary (to_contiguous), or refuse", context={"plane_domain": self.logical_domain_id}))
        return self

    def to_contiguous(self) -> "ArrayView":
        if self.is_contiguous():
            return self
        if self.residency != RESIDENCY_CPU:
            raise NonContiguousError(TypedRefusal(RefusalReason.NON_CONTIGUOUS,
                                                  "device views are copied by their owner, not here"))
        copy = np.ascontiguousarray(self.owner)
        return ArrayView(owner=copy, dtype=copy.dtype, shape=tuple(copy.shape), strides=tuple(copy.strides),
                         global_origin=self.global_origin, logical_domain_id=self.logical_domain_id,
                         residency=self.residency, read_only=self.read_only)

    def to_numpy(self) -> np.ndarray:
        if self.residency == RESIDENCY_CPU:
            return self.owner
        return _host_array(self.owner)

    @property
    def nbytes(self) -> int:
        return int(np.prod(self.shape)) * self.dtype.itemsize

    def freeze(self) -> "ArrayView":
        self._frozen_digest = _logical_digest(self.owner)
        return self

    @property
    def frozen(self) -> bool:
        return self._frozen_digest is not None

    def assert_not_mutated(self) -> None:
        if self._frozen_digest is None:
            raise AbiError("assert_not_mutated on a view that was never frozen")
        current = _logical_digest(self.owner)
        if current != self._frozen_digest:
            raise StaleBufferError(TypedRefusal(
                RefusalReason.STALE_BUFFER, "owner mutated after view creation (frozen digest mismatch)",
                context={"plane_domain": self.logical_domain_id, "frozen_sha256": self._frozen_digest,
                         "current_sha256": current}))

    def same_logical_values(self, other: "ArrayView") -> bool:
        if not isinstance(other, ArrayView):
            return NotImplemented
        if self.dtype != other.dtype or self.shape != other.shape:
            return False
        return np.array_equal(np.ascontiguousarray(self.to_numpy()), np.ascontiguousarray(other.to_numpy()))

    def describe(self) -> dict[str, Any]:
        return {"dtype": str(self.dtype.str), "shape": list(self.shape), "strides": list(self.strides),
                "global_origin": list(self.global_origin), "logical_domain_id": self.logical_domain_id,
                "residency": self.residency, "read_only": self.read_only, "contiguous": self.is_contiguous(),
                "frozen": self.frozen}
