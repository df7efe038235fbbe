"""Columnar cone metadata (reference python/pdhcg/cones.py): ConeType codes
(the C enum values) and ConeSpec, one entry per cone block."""
from __future__ import annotations

from enum import IntEnum
from numbers import Integral
from typing import Any, Mapping, Optional

import numpy as np


class ConeType(IntEnum):
    RSOC = 0
    SOC = 1
    EXP = 2
    POWER = 3
    PSD = 4
    ROTATED_SOC = 0
    STANDARD_SOC = 1
    EXPONENTIAL = 2


_NAMES = {"rsoc": ConeType.RSOC, "soc": ConeType.SOC, "exp": ConeType.EXP, "power": ConeType.POWER,
          "psd": ConeType.PSD}
_I32 = np.iinfo(np.int32)


def _type_code(v: Any) -> int:
    if isinstance(v, str):
        if v.lower() not in _NAMES:
            raise ValueError("cone type must be 'soc', 'rsoc', 'exp', 'power', or 'psd'")
        return int(_NAMES[v.lower()])
    if not isinstance(v, Integral):
        raise TypeError("cone type must be a ConeType, integer code, or string")
    if not int(ConeType.RSOC) <= int(v) <= int(ConeType.PSD):
        raise ValueError(f"invalid cone type code {int(v)}")
    return int(v)


def _ints(v: Any, count: int, name: str, default: Optional[int] = None) -> np.ndarray:
    if v is None:
        if default is None:
            raise ValueError(f"{name} is required")
        return np.full(count, default, dtype=np.int32)
    a = np.asarray(v)
    if a.dtype.kind not in "iu":
        raise TypeError(f"{name} must contain integers")
    if a.ndim == 0:
        a = np.full(count, int(a))
    elif a.ndim != 1 or a.size != count:
        raise ValueError(f"{name} must be a scalar or a 1D array of length {count}")
    if a.size and (a.min() < _I32.min or a.max() > _I32.max):
        raise OverflowError(f"{name} contains a value that does not fit int32")
    return np.ascontiguousarray(a, dtype=np.int32)


def _floats(v: Any, count: int, name: str, default: float) -> np.ndarray:
    if v is None:
        return np.full(count, default, dtype=np.float64)
    a = np.asarray(v, dtype=np.float64)
    if a.ndim == 0:
        return np.full(count, float(a))
    if a.ndim != 1 or a.size != count:
        raise ValueError(f"{name} must be a scalar or a 1D array of length {count}")
    return np.ascontiguousarray(a)


def block_lengths(types: np.ndarray, v_dims: np.ndarray) -> np.ndarray:
    """Slots of each block: SOC / RSOC v + 2, EXP / POWER 3, PSD v (v + 1) / 2."""
    v = v_dims.astype(np.int64)
    out = v + 2
    out = np.where((types == ConeType.EXP) | (types == ConeType.POWER), 3, out)
    return np.where(types == ConeType.PSD, v * (v + 1) // 2, out)


class ConeSpec:
    """Cone blocks as columns: ``types`` (ConeType, code or name), ``starts``
    (first variable, or first row of F for affine cones), ``v_dims`` (vector
    slots; the matrix order for PSD), ``power_alphas`` and an optional
    ambient ``fixed_mask``.  Scalars broadcast over the blocks."""

    __slots__ = ("types", "starts", "v_dims", "power_alphas", "fixed_mask")

    def __init__(self, types: Any, starts: Any, v_dims: Any = 1, power_alphas: Any = 0.0,
                 fixed_mask: Optional[Any] = None) -> None:
        s = np.asarray(starts)
        if s.ndim != 1:
            raise ValueError("starts must be a 1D array")
        count = int(s.size)
        self.starts = _ints(s, count, "starts")
        if isinstance(types, (str, ConeType, int, np.integer)):
            self.types = np.full(count, _type_code(types), dtype=np.int32)
        else:
            t = np.asarray(types)
            if t.ndim != 1 or t.size != count:
                raise ValueError(f"types must be a scalar or a 1D array of length {count}")
            self.types = np.fromiter((_type_code(x if not isinstance(x, np.integer) else int(x)) for x in t),
                                     dtype=np.int32, count=count)
        self.v_dims = _ints(v_dims, count, "v_dims", default=1)
        self.power_alphas = _floats(power_alphas, count, "power_alphas", 0.0)
        if np.any(self.starts < 0):
            raise ValueError("starts must be nonnegative")
        if np.any(self.v_dims <= 0):
            raise ValueError("v_dims must be positive")
        three = (self.types == ConeType.EXP) | (self.types == ConeType.POWER)
        if np.any(self.v_dims[three] != 1):
            raise ValueError("EXP and POWER cones require v_dim == 1")
        a = self.power_alphas[self.types == ConeType.POWER]
        if np.any(~np.isfinite(a) | (a <= 0.0) | (a >= 1.0)):
            raise ValueError("POWER cone alphas must lie in (0, 1)")
        if fixed_mask is None:
            self.fixed_mask = None
        else:
            m = np.asarray(fixed_mask)
            if m.ndim != 1:
                raise ValueError("fixed_mask must be a 1D ambient-coordinate mask")
            self.fixed_mask = np.ascontiguousarray(m, dtype=np.uint8)

    def __len__(self) -> int:
        return int(self.starts.size)

    def validate_ambient(self, ambient_dimension: int, *, allow_fixed: bool, require_cover: bool = False) -> None:
        n = int(ambient_dimension)
        if n < 0:
            raise ValueError("ambient_dimension must be nonnegative")
        lengths = block_lengths(self.types, self.v_dims)
        if np.any(self.starts.astype(np.int64) + lengths > n):
            raise ValueError("a cone block extends beyond the ambient dimension")
        if require_cover and int(lengths.sum()) != n:
            raise ValueError("affine cones must cover every row of F")
        if self.fixed_mask is not None:
            if not allow_fixed:
                raise ValueError("affine cones do not support fixed slots")
            if self.fixed_mask.size != n:
                raise ValueError(f"fixed_mask length {self.fixed_mask.size} != ambient dimension {n}")
            for st, ln, t in zip(self.starts, lengths, self.types):
                if t == ConeType.PSD and np.any(self.fixed_mask[int(st):int(st + ln)]):
                    raise ValueError("PSD cones do not support fixed slots")

    @classmethod
    def from_columnar(cls, payload: Mapping[str, Any]) -> "ConeSpec":
        return cls(payload["types"], payload["starts"], payload["v_dims"], payload["power_alphas"],
                   payload.get("fixed_mask"))

    def as_tuples(self):
        """(type, start, v_dim, alpha) per block: the form QPData.cones takes."""
        return [(int(t), int(s), int(v), float(a)) for t, s, v, a in
                zip(self.types, self.starts, self.v_dims, self.power_alphas)]
