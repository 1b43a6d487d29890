"""Full-state health gate of the dycore state (reference
gpuwm/core/health.py: rules :177-330, collection :376-503, the CPU mirror
:506-585, StateHealthValidator :611-770; kernel kernels/health.cu:224-297).

Every checked field has a :class:`FieldRule` (a status class, bounds, a
strict lower bound for denominators); theta' is checked as total theta
(plus the base profile) and mu' as the coupled column mass (plus mub2d).
The device gate (``StateHealthValidator``) makes one pass over every field
on the GPU (``arw_health_gate``) and copies back a two-word record: the OR
of the failing classes' bits and the first failing (field, flat index).
Nothing is clipped or repaired; ``require_healthy`` raises
:class:`HealthCheckError`.  ``validate_fields_cpu`` is the host mirror of
the same rules (corruption tests, host-built tables).

The state this path owns has no integer fields, so the reference's int32
storage policy (``gpu_integer_policy``, for physics integer maps) has no
counterpart here: every field is of the state's float dtype.
"""
from __future__ import annotations

import ctypes
import math
from collections.abc import Mapping, Sequence
from dataclasses import dataclass
from typing import Any

import numpy as np
import torch

from . import _lib

#: Stable status classes (the reference's bit positions, so records read
#: the same).  The dycore state uses the first eight and "generic".
STATUS_BITS = {
    "wind": 1 << 0,
    "geopotential": 1 << 1,
    "coupled_mass": 1 << 2,
    "pressure": 1 << 3,
    "specific_volume": 1 << 4,
    "theta": 1 << 5,
    "moisture": 1 << 6,
    "moment": 1 << 7,
    "held_tendency": 1 << 9,
    "generic": 1 << 15,
    "particle_volume": 1 << 16,
}

MAX_HEALTH_FIELDS = 1024
_INDEX_MASK = (1 << 48) - 1
_NONE = (1 << 64) - 1
#: chunking of the device launch: about this many blocks over all fields,
#: each at least this many elements
_CHUNK_TARGET_BLOCKS = 8192
_MIN_CHUNK_ELEMENTS = 4096


@dataclass(frozen=True)
class FieldRule:
    """Bounds and status class of one health field."""
    status_class: str
    lower: float | None = None
    upper: float | None = None
    strict_lower: bool = False

    @property
    def status_bit(self) -> int:
        return STATUS_BITS[self.status_class]


@dataclass(frozen=True)
class HealthField:
    """One array, optionally checked as ``values + auxiliary``:
    ``aux_mode='direct'`` elementwise, ``'level'`` one auxiliary value per
    horizontal plane of ``plane_size`` elements (the 1-D base theta)."""
    name: str
    values: Any
    rule: FieldRule
    auxiliary: Any | None = None
    aux_mode: str | None = None
    plane_size: int = 0


@dataclass(frozen=True)
class ValidationReport:
    """Compact result of one full-state validation pass."""
    ok: bool
    status_bits: int
    first_bad_field: str | None = None
    first_bad_index: tuple | None = None
    first_bad_flat_index: int | None = None
    first_bad_value: float | None = None
    reason: str | None = None
    phase: str | None = None

    @property
    def failing_classes(self) -> tuple:
        return tuple(n for n, b in STATUS_BITS.items()
                     if self.status_bits & b)


class HealthCheckError(FloatingPointError):
    """A synchronized model-state health gate failed (terminal)."""

    def __init__(self, report: ValidationReport):
        self.report = report
        where = ("unknown" if report.first_bad_index is None
                 else str(report.first_bad_index))
        phase = "" if report.phase is None else f" during {report.phase}"
        super().__init__(
            f"full-state health gate failed{phase}: "
            f"{report.first_bad_field}{where}: {report.reason}; "
            f"classes={report.failing_classes}")


_FINITE = FieldRule("generic")


def rule_for_field(name: str) -> FieldRule:
    """The rule of a named field (reference health.py:247-326, the
    dycore-state fields)."""
    leaf = name.lower().split(".")[-1]
    if name.lower().startswith("held."):
        return FieldRule("held_tendency")
    if leaf in ("u", "v"):
        return FieldRule("wind", -500.0, 500.0)
    if leaf == "w":
        return FieldRule("wind", -200.0, 200.0)
    if leaf == "thp":
        # checked as total theta (the base profile is the auxiliary)
        return FieldRule("theta", 100.0, 600.0)
    if leaf == "php":
        return FieldRule("geopotential", -1.0e8, 1.0e8)
    if leaf == "mup":
        # checked as the coupled column mass mub2d + mu' (a denominator)
        return FieldRule("coupled_mass", 0.0, 2.0e6, strict_lower=True)
    if leaf == "p":
        return FieldRule("pressure", 0.0, 2.0e6, strict_lower=True)
    if leaf == "al":
        return FieldRule("specific_volume", -1.0e4, 1.0e4)
    if leaf == "alt":
        return FieldRule("specific_volume", 0.0, 1.0e4, strict_lower=True)
    if leaf in ("qv", "qc", "qr", "qi", "qs", "qg", "qh"):
        return FieldRule("moisture", 0.0, 1.0)
    if leaf in ("nc", "nr", "ni", "ns", "ng", "qndrop", "qnr", "qni", "qns",
                "qng", "qnh", "qnn", "nwfa", "nifa"):
        # number moments per kilogram: a blow-up ceiling, not the
        # schemes' own bounds
        return FieldRule("moment", 0.0, 1.0e15)
    if leaf in ("qvolg", "qvolh"):
        return FieldRule("particle_volume", 0.0, 1.0)
    return _FINITE


def field_from_array(name: str, values: Any, *, auxiliary: Any = None,
                     aux_mode: str | None = None,
                     plane_size: int = 0) -> HealthField:
    return HealthField(name, values, rule_for_field(name), auxiliary,
                       aux_mode, int(plane_size))


#: the cross-step fields of the state (the restart manifest's serialized
#: set), in gate order
STATE_FIELDS = ("u", "v", "w", "thp", "php", "mup", "p", "al", "alt",
                "qv", "qc", "qr", "qi", "qs", "qg", "nc", "nr", "ni", "ns",
                "ng", "nwfa", "nifa", "qh", "qndrop", "qnr", "qni", "qns",
                "qng", "qnh", "qnn", "qvolg", "qvolh", "h_diabatic")


def collect_state_fields(state: Any) -> tuple:
    """The state's health descriptors: theta' with the base theta, mu'
    with the base column mass, the rest as they are (absent ones
    skipped)."""
    out = []
    for name in STATE_FIELDS:
        value = getattr(state, name, None)
        if value is None:
            continue
        if name == "mup" and getattr(state, "mub2d", None) is not None:
            out.append(field_from_array(name, value, auxiliary=state.mub2d,
                                        aux_mode="direct"))
        elif name == "thp" and getattr(state, "thb", None) is not None:
            thb = state.thb
            if thb.ndim == 1:
                plane = int(np.prod(tuple(value.shape[1:])))
                out.append(field_from_array(name, value, auxiliary=thb,
                                            aux_mode="level",
                                            plane_size=plane))
            else:
                out.append(field_from_array(name, value, auxiliary=thb,
                                            aux_mode="direct"))
        else:
            out.append(field_from_array(name, value))
    if len(out) > MAX_HEALTH_FIELDS:
        raise ValueError(f"{len(out)} health fields exceed "
                         f"MAX_HEALTH_FIELDS={MAX_HEALTH_FIELDS}")
    return tuple(out)


def _host(x) -> np.ndarray:
    if torch.is_tensor(x):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def _cpu_values(f: HealthField) -> np.ndarray:
    values = _host(f.values)
    if f.auxiliary is None:
        return values
    aux = _host(f.auxiliary)
    if f.aux_mode == "direct":
        if aux.shape != values.shape:
            raise ValueError(f"{f.name}: auxiliary shape {aux.shape} != "
                             f"{values.shape}")
        return values + aux
    if f.aux_mode == "level":
        if aux.ndim != 1 or values.shape[0] != aux.size:
            raise ValueError(f"{f.name}: level auxiliary {aux.shape} does "
                             f"not fit {values.shape}")
        return values + aux.reshape((-1,) + (1,) * (values.ndim - 1))
    raise ValueError(f"unknown auxiliary mode {f.aux_mode!r}")


def _bad_mask(values: np.ndarray, rule: FieldRule) -> np.ndarray:
    bad = ~np.isfinite(values)
    if rule.lower is not None:
        bad |= values <= rule.lower if rule.strict_lower \
            else values < rule.lower
    if rule.upper is not None:
        bad |= values > rule.upper
    return bad


def _reason(value: float, rule: FieldRule) -> str:
    if not math.isfinite(value):
        return f"non-finite value {value!r}"
    if rule.lower is not None and (
            value <= rule.lower if rule.strict_lower else value < rule.lower):
        op = ">" if rule.strict_lower else ">="
        return f"value {value!r} violates lower bound {op} {rule.lower!r}"
    if rule.upper is not None and value > rule.upper:
        return f"value {value!r} exceeds upper bound {rule.upper!r}"
    return "unknown invariant failure"


def validate_fields_cpu(fields: Mapping | Sequence, *,
                        phase: str | None = None) -> ValidationReport:
    """The host mirror of the gate: the same rules, the same attribution
    (the first failing field in order, its first flat index)."""
    if isinstance(fields, Mapping):
        fields = tuple(field_from_array(n, v) for n, v in fields.items())
    status, first = 0, None
    for number, f in enumerate(fields):
        values = _cpu_values(f)
        bad = _bad_mask(values, f.rule)
        if bool(np.any(bad)):
            status |= f.rule.status_bit
            if first is None:
                flat = int(np.flatnonzero(np.ravel(bad))[0])
                first = (flat, values, f)
    if first is None:
        return ValidationReport(True, status, phase=phase)
    flat, values, f = first
    value = float(np.ravel(values)[flat])
    return ValidationReport(
        False, status, f.name,
        tuple(int(i) for i in np.unravel_index(flat, values.shape)), flat,
        value, _reason(value, f.rule), phase)


def validate_state_cpu(state: Any, *, phase: str | None = None,
                       raise_on_error: bool = False) -> ValidationReport:
    report = validate_fields_cpu(collect_state_fields(state), phase=phase)
    if raise_on_error and not report.ok:
        raise HealthCheckError(report)
    return report


class StateHealthValidator:
    """The device gate of one DomainState: ``validate()`` is one pass over
    every field (two launches: the record's initialisation and the gate)
    and one 16-byte copy back."""

    def __init__(self, state):
        self.state = state
        self.fields: tuple = ()
        self._desc = None
        self._result = None
        self._chunk = _MIN_CHUNK_ELEMENTS
        self._nchunks = 1

    def _refresh(self) -> None:
        st = self.state
        fields = collect_state_fields(st)
        n = len(fields)
        arr = (_lib.ArwHealthField * max(n, 1))()
        sizes = []
        for i, f in enumerate(fields):
            v = f.values
            if not v.is_cuda or not v.is_contiguous() or v.dtype != st.dtype:
                raise TypeError(f"health field {f.name!r} must be a "
                                f"contiguous {st.dtype} device tensor")
            d = arr[i]
            d.values = v.data_ptr()
            d.size = v.numel()
            sizes.append(v.numel())
            r = f.rule
            flags = 0
            if r.lower is not None:
                flags |= _lib.HEALTH_LOWER
                d.lower = float(r.lower)
            if r.upper is not None:
                flags |= _lib.HEALTH_UPPER
                d.upper = float(r.upper)
            if r.strict_lower:
                flags |= _lib.HEALTH_STRICT_LOWER
            if f.auxiliary is not None:
                a = f.auxiliary
                if not a.is_contiguous() or a.dtype != st.dtype:
                    raise TypeError(f"health auxiliary of {f.name!r} must "
                                    f"be a contiguous {st.dtype} tensor")
                d.aux = a.data_ptr()
                flags |= _lib.HEALTH_ADD_AUX
                if f.aux_mode == "level":
                    flags |= _lib.HEALTH_AUX_LEVEL
                    d.plane = f.plane_size
                elif f.aux_mode != "direct":
                    raise ValueError(f"unknown auxiliary mode {f.aux_mode!r}")
            d.flags = flags
            d.status_bit = r.status_bit
        raw = np.frombuffer(bytes(arr), dtype=np.uint8)
        self._desc = st.scratch((raw.size,), "health_fields",
                                dtype=torch.uint8)
        self._desc.copy_(torch.from_numpy(raw.copy()))
        self._result = st.scratch((2,), "health_result", dtype=torch.int64)
        total = int(sum(sizes))
        self._chunk = max(_MIN_CHUNK_ELEMENTS, total // _CHUNK_TARGET_BLOCKS)
        largest = max(sizes, default=0)
        self._nchunks = max(1, (largest + self._chunk - 1) // self._chunk)
        self.fields = fields

    def validate(self, *, phase: str | None = None) -> ValidationReport:
        """One gate launch; the compact record comes back to the host."""
        from .dycore import _stream
        self._refresh()
        lib = _lib.load()
        fn = (lib.arw64_health_gate if self.state.dtype == torch.float64
              else lib.arw_health_gate)
        _lib.check(fn(ctypes.c_void_p(self._desc.data_ptr()),
                      len(self.fields), self._chunk, self._nchunks,
                      ctypes.c_void_p(self._result.data_ptr()), _stream()),
                   "arw_health_gate")
        words = self._result.cpu().numpy().view(np.uint64)
        status, packed = int(words[0]), int(words[1])
        if packed == _NONE:
            return ValidationReport(True, status, phase=phase)
        number, flat = packed >> 48, packed & _INDEX_MASK
        if number >= len(self.fields):
            raise RuntimeError(f"health gate returned field {number}")
        f = self.fields[number]
        value = float(f.values.reshape(-1)[flat])
        if f.auxiliary is not None:
            k = flat if f.aux_mode == "direct" else flat // f.plane_size
            value = float((f.values.reshape(-1)[flat]
                           + f.auxiliary.reshape(-1)[k]))
        return ValidationReport(
            False, status, f.name,
            tuple(int(i) for i in np.unravel_index(flat,
                                                   tuple(f.values.shape))),
            flat, value, _reason(value, f.rule), phase)

    def require_healthy(self, *, phase: str | None = None) -> ValidationReport:
        report = self.validate(phase=phase)
        if not report.ok:
            raise HealthCheckError(report)
        return report

    def phase_observer(self, phase: str) -> ValidationReport:
        """Validate at a named phase (raises on failure)."""
        return self.require_healthy(phase=phase)


__all__ = [
    "FieldRule", "HealthCheckError", "HealthField", "MAX_HEALTH_FIELDS",
    "STATUS_BITS", "StateHealthValidator", "ValidationReport",
    "collect_state_fields", "field_from_array", "rule_for_field",
    "validate_fields_cpu", "validate_state_cpu",
]
