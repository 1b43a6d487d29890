"""Bit-exact restart files of the dycore state (reference
gpuwm/io/restart.py:1-60 for the contract, :395-470 and
gpuwm/state_serialization_contract.py:16-66 for the state's part of the
manifest, :2061-2166 write_restart, :3401-3415 restore_restart).

A restart file is a NumPy NPZ archive (uncompressed zip of raw ``.npy``
members): each member's payload is the array's memory bytes, so float32
values -- negative zeros, denormals and NaN payloads included -- round-trip
bit for bit.  ``write_restart`` then ``restore_restart`` into a freshly
prepared state continues the trajectory bit-identically: N steps, a
restart, N more steps equal 2N uninterrupted steps on every field.

Completeness is enforced.  Every attribute of the ``DomainState`` and every
scratch slot is classified here as

* SERIALIZED -- read across step boundaries and not reconstructable: the
  prognostics, the end-of-step EOS diagnostics (so a restored state equals
  the live one for any pre-step reader), the moist scalars and the clock;
* REBUILT -- overwritten before every read: the RK stage buffers (the
  ``*0`` arrays), the slow tendencies, the acoustic perturbations and every
  per-call work buffer of the library;
* SETUP -- deterministic from the configuration and the host setup (base
  state, coordinate, map factors): rebuilt by the normal preparation and
  validated against a SHA-256 fingerprint in the header, so a restart into
  a different setup fails loudly instead of drifting;
* INFRASTRUCTURE -- the configuration (compared field by field), device,
  dtype, host copies and caches.

An attribute or slot in none of these classes makes ``write_restart`` raise
:class:`RestartManifestError`, so new state cannot silently drop out of
restarts.
"""
from __future__ import annotations

import dataclasses
import hashlib
import json
import os
from pathlib import Path

import numpy as np
import torch

from .config import EXTRA_SPECIES, RunConfig
from .state import DomainState

#: Bump on any change to the member layout or the classification tables;
#: readers refuse other versions.
RESTART_FORMAT_VERSION = 2

_HEADER_KEY = "__header__"

#: Cross-step state written to the file (present ones only: the moist
#: scalars exist on moist states).
#: every mixed-phase scheme's transported species, in registry order
_EXTRA = tuple(dict.fromkeys(q for mp in sorted(EXTRA_SPECIES)
                             for q in EXTRA_SPECIES[mp]))

STATE_SERIALIZED_ATTRS = (
    "u", "v", "w", "thp", "php", "mup",
    "p", "al", "alt",
    "qv", "qc", "qr", "h_diabatic",
) + _EXTRA

#: Overwritten before every read.  The step rotates its stage buffers
#: through the *0 arrays, writes every tendency and perturbation of a stage
#: before reading it, and rebuilds p'' and Omega'' every substep.
STATE_REBUILT_ATTRS = frozenset({
    "u0", "v0", "w0", "thp0", "php0", "mup0", "qv0", "qc0", "qr0",
    "ru_t", "rv_t", "rw_t", "rth_t", "rph_t", "rmu_t",
    "u_pp", "v_pp", "w_pp", "th_pp", "ph_pp", "mu_pp",
    "p_pp", "p_pp_old", "ww_pp", "al_pp",
}) | frozenset(q + "0" for q in _EXTRA)

#: Deterministic setup arrays and scalars covered by the fingerprint.
STATE_SETUP_ARRAYS = (
    "thb", "pb", "alb", "phb", "mub2d", "ht",
    "c1h", "c2h", "c1f", "c2f", "c3h", "c4h", "c3f", "c4f",
    "msft", "msfu", "msfv", "f", "e",
    "dnw", "rdnw", "dn", "rdn", "fnp", "fnm", "znu", "znw",
)
STATE_SETUP_SCALARS = (
    "mub", "p_top", "cf1", "cf2", "cf3", "cfn", "cfn1",
    "has_msf", "rotational",
)

#: Handled by dedicated header entries (configuration, clock) or rebuilt
#: by the preparation (host copies of the setup, caches, scratch pool).
STATE_INFRA_ATTRS = frozenset({
    "cfg", "device", "dtype", "elapsed_seconds", "_scratch", "_phb_host",
    "_coord", "_base", "_diff_rdz", "_dom_memo",
})

#: Scratch slots outside the library's own work areas (prefix ``arw_``,
#: all per-call buffers written before they are read): the stage mass
#: fluxes of ``dycore.stage_fluxes`` and the health gate's descriptor
#: table and result.
REBUILT_SCRATCH_SLOTS = frozenset({"rk_ru", "rk_rv", "health_fields",
                                   "health_result"})
_REBUILT_SCRATCH_PREFIX = "arw_"

#: Configuration fields that may change at a restart boundary without
#: changing the model (run length and output cadence).
CONFIG_RUN_LENGTH_FIELDS = frozenset({"run_seconds", "output_interval_s"})


class RestartManifestError(RuntimeError):
    """State the restart manifest does not classify."""


class RestartMismatchError(ValueError):
    """A restart file that cannot continue this state and configuration."""


@dataclasses.dataclass(frozen=True)
class RestartInfo:
    """What ``restore_restart`` restored."""
    path: Path
    elapsed_seconds: float
    members: tuple


def classify_state_attr(name: str) -> str:
    """The manifest class of one DomainState attribute."""
    if name in STATE_SERIALIZED_ATTRS:
        return "serialized"
    if name in STATE_REBUILT_ATTRS:
        return "rebuilt"
    if name in STATE_SETUP_ARRAYS or name in STATE_SETUP_SCALARS:
        return "setup"
    if name in STATE_INFRA_ATTRS:
        return "infrastructure"
    raise RestartManifestError(
        f"DomainState attribute {name!r} is not classified by the restart "
        "manifest (serialized, rebuilt, setup or infrastructure); classify "
        "it before writing restarts")


def classify_scratch_slot(slot: str) -> str:
    if slot.startswith(_REBUILT_SCRATCH_PREFIX) or slot in REBUILT_SCRATCH_SLOTS:
        return "rebuilt"
    raise RestartManifestError(
        f"scratch slot {slot!r} is not classified by the restart manifest")


def state_manifest(state: DomainState) -> dict:
    """{"state/<name>": array} of the serialized attributes; raises for any
    unclassified attribute or scratch slot."""
    out = {}
    for name, value in sorted(vars(state).items()):
        cls = classify_state_attr(name)
        if cls == "serialized" and value is not None:
            out["state/" + name] = value
    for slot in sorted(getattr(state, "_scratch", {})):
        classify_scratch_slot(slot)
    return out


def _host(value) -> np.ndarray:
    if torch.is_tensor(value):
        return value.detach().cpu().numpy()
    return np.asarray(value)


def setup_fingerprint(state: DomainState) -> str:
    """SHA-256 over the setup arrays (name, shape, dtype, bytes) and
    scalars of the state."""
    digest = hashlib.sha256()
    for name in STATE_SETUP_ARRAYS:
        host = np.ascontiguousarray(_host(getattr(state, name)))
        digest.update(name.encode())
        digest.update(str(host.shape).encode())
        digest.update(str(host.dtype).encode())
        digest.update(host.tobytes())
    for name in STATE_SETUP_SCALARS:
        value = getattr(state, name)
        if value is not None and not isinstance(value, bool):
            value = float(value)
        digest.update(f"{name}={value!r};".encode())
    return digest.hexdigest()


def write_restart(path, state: DomainState, cfg: RunConfig) -> Path:
    """Serialize the cross-step state to ``path`` (published atomically:
    written to ``<path>.tmp``, flushed to disk, then renamed)."""
    path = Path(path)
    elapsed = float(state.elapsed_seconds)
    if not np.isfinite(elapsed) or elapsed < 0.0:
        raise ValueError(f"elapsed_seconds {elapsed!r} cannot be restarted")
    arrays = {k: _host(v) for k, v in state_manifest(state).items()}
    header = {
        "format_version": RESTART_FORMAT_VERSION,
        "case": cfg.case,
        "shape": list(state.p.shape),
        "dtype": str(state.np_dtype),
        "elapsed_seconds": elapsed,
        "config": dataclasses.asdict(cfg),
        "setup_fingerprint": setup_fingerprint(state),
        "array_manifest": {k: {"shape": list(a.shape), "dtype": str(a.dtype)}
                           for k, a in sorted(arrays.items())},
    }
    payload = {_HEADER_KEY: np.frombuffer(
        json.dumps(header, allow_nan=False).encode("utf-8"), dtype=np.uint8)}
    payload.update(arrays)
    path.parent.mkdir(parents=True, exist_ok=True)
    temp = path.with_name(path.name + ".tmp")
    try:
        with temp.open("wb") as stream:
            np.savez(stream, **payload)
            stream.flush()
            os.fsync(stream.fileno())
        os.replace(temp, path)
    except BaseException:
        temp.unlink(missing_ok=True)
        raise
    return path


def _load(path: Path, with_arrays: bool):
    import zipfile
    try:
        with np.load(path, allow_pickle=False) as data:
            if _HEADER_KEY not in data.files:
                raise RestartMismatchError(
                    f"{path} is not a restart file of this build (no header)")
            header = json.loads(bytes(bytearray(data[_HEADER_KEY]))
                                .decode("utf-8"))
            stored = ({k: data[k] for k in data.files if k != _HEADER_KEY}
                      if with_arrays else None)
    except RestartMismatchError:
        raise
    except FileNotFoundError:
        raise
    except (zipfile.BadZipFile, OSError, EOFError, ValueError) as exc:
        raise RestartMismatchError(
            f"restart file {path} is unreadable (truncated or corrupt "
            "archive)") from exc
    return header, stored


def read_restart_header(path) -> dict:
    """The JSON header: format version, configuration echo, clock, setup
    fingerprint and array manifest."""
    return _load(Path(path), False)[0]


def _require_config_match(stored: dict, cfg: RunConfig, path) -> None:
    live = dataclasses.asdict(cfg)
    missing = object()
    diffs = []
    for key in sorted(set(stored) | set(live)):
        if key in CONFIG_RUN_LENGTH_FIELDS:
            continue
        a, b = stored.get(key, missing), live.get(key, missing)
        if a is not b and a != b:
            diffs.append(f"{key}: restart={a!r} run={b!r}")
    if diffs:
        raise RestartMismatchError(
            f"restart file {path} was written under a different "
            "configuration; refusing to continue a different model:\n  "
            + "\n  ".join(diffs))


def restore_restart(path, state: DomainState, cfg: RunConfig) -> RestartInfo:
    """Restore ``path`` into a freshly prepared state, in place.

    The caller has completed the normal setup (the base state and
    coordinate loaded).  The configuration echo, the state's shape and
    dtype and the setup fingerprint are checked first; then every
    serialized array is copied into the state's existing tensors, and the
    clock is set last."""
    path = Path(path)
    header, stored = _load(path, True)
    if header.get("format_version") != RESTART_FORMAT_VERSION:
        raise RestartMismatchError(
            f"{path}: restart format {header.get('format_version')!r}, this "
            f"build reads {RESTART_FORMAT_VERSION}")
    _require_config_match(header.get("config", {}), cfg, path)
    if list(state.p.shape) != header["shape"] \
            or str(state.np_dtype) != header["dtype"]:
        raise RestartMismatchError(
            f"{path}: restart state {header['shape']} {header['dtype']} does "
            f"not match {list(state.p.shape)} {state.np_dtype}")
    if setup_fingerprint(state) != header["setup_fingerprint"]:
        raise RestartMismatchError(
            f"{path}: the setup (base state, coordinate, map factors) "
            "differs from the one the restart was written under")
    expected = set(state_manifest(state))
    if set(stored) != expected:
        raise RestartMismatchError(
            f"{path}: members {sorted(set(stored) ^ expected)} do not match "
            "the state's serialized inventory")
    for key in sorted(stored):
        target = getattr(state, key[len("state/"):])
        host = stored[key]
        if tuple(host.shape) != tuple(target.shape) \
                or host.dtype != state.np_dtype:
            raise RestartMismatchError(
                f"{key}: restart {host.shape} {host.dtype} does not match "
                f"state {tuple(target.shape)} {state.np_dtype}")
        target.copy_(torch.from_numpy(np.ascontiguousarray(host)))
    state.elapsed_seconds = float(header["elapsed_seconds"])
    return RestartInfo(path, state.elapsed_seconds, tuple(sorted(stored)))


__all__ = [
    "RESTART_FORMAT_VERSION", "RestartInfo", "RestartManifestError",
    "RestartMismatchError", "STATE_INFRA_ATTRS", "STATE_REBUILT_ATTRS",
    "STATE_SERIALIZED_ATTRS", "STATE_SETUP_ARRAYS", "STATE_SETUP_SCALARS",
    "classify_scratch_slot", "classify_state_attr", "read_restart_header",
    "restore_restart", "setup_fingerprint", "state_manifest",
    "write_restart",
]
