"""The dycore state's restart contract on the host (no device): every
DomainState attribute and scratch slot is classified, the NPZ round trip is
bit-exact (negative zeros, denormals and NaN payloads included), and a
restart refuses a different configuration, setup, shape or a damaged file
(reference gpuwm/io/restart.py:1-60; its CPU manifest tests are the
model)."""
import dataclasses

import numpy as np
import pytest
import torch

from arwen_amd import cases
from arwen_amd import restart as R
from arwen_amd.state import DomainState


def _state(moist=False, seed=3, mp=0):
    if mp:
        case = cases.mixed_phase(mp, nx=12, ny=10, nz=8)
    elif moist:
        case = cases.moist_tracer(nx=12, ny=10, nz=8)
    else:
        case = cases.ensemble_member(12, 10, 8, seed)
    return case, DomainState.from_host(case, device="cpu")


def test_every_attribute_is_classified():
    for moist, mp in ((False, 0), (True, 0), (True, 10), (True, 18),
                      (True, 28)):
        _, st = _state(moist, mp=mp)
        for name in vars(st):
            assert R.classify_state_attr(name) in (
                "serialized", "rebuilt", "setup", "infrastructure")


def test_unclassified_attribute_and_slot_refuse_the_write(tmp_path):
    case, st = _state()
    st.new_field = torch.zeros(3)
    with pytest.raises(R.RestartManifestError, match="new_field"):
        R.write_restart(tmp_path / "a.npz", st, case.cfg)
    del st.new_field
    st.scratch((4,), "mystery")
    with pytest.raises(R.RestartManifestError, match="mystery"):
        R.write_restart(tmp_path / "a.npz", st, case.cfg)


@pytest.mark.parametrize("mp", [0, 18])
def test_round_trip_is_bit_exact(tmp_path, mp):
    case, a = _state(moist=True, mp=mp)
    special = np.array([-0.0, 1e-45, np.float32(np.nan), -np.inf],
                       dtype=np.float32)
    # a NaN with a payload: a round trip through arithmetic would lose it
    special[2] = np.frombuffer(np.uint32(0x7fc01234).tobytes(),
                               dtype=np.float32)[0]
    a.thp.view(-1)[:4] = torch.from_numpy(special)
    a.qv.view(-1)[5] = 3.5e-39                      # denormal
    a.elapsed_seconds = 1234.5
    path = R.write_restart(tmp_path / "rst.npz", a, case.cfg)
    assert not (tmp_path / "rst.npz.tmp").exists()
    _, b = _state(moist=True, mp=mp, seed=4)
    info = R.restore_restart(path, b, case.cfg)
    assert info.elapsed_seconds == 1234.5 == b.elapsed_seconds
    present = [n for n in R.STATE_SERIALIZED_ATTRS
               if getattr(a, n, None) is not None]
    assert set(info.members) == {"state/" + n for n in present}
    assert set(a.extra_species) <= set(present)
    for n in present:
        x, y = getattr(a, n).numpy(), getattr(b, n).numpy()
        assert x.tobytes() == y.tobytes(), n
    hdr = R.read_restart_header(path)
    assert hdr["format_version"] == R.RESTART_FORMAT_VERSION
    assert hdr["setup_fingerprint"] == R.setup_fingerprint(a)


def test_dry_state_serializes_no_scalars(tmp_path):
    case, a = _state()
    path = R.write_restart(tmp_path / "d.npz", a, case.cfg)
    members = set(np.load(path, allow_pickle=False).files)
    assert "state/qv" not in members and "state/u" in members


def test_run_length_may_change_the_model_may_not(tmp_path):
    case, a = _state()
    path = R.write_restart(tmp_path / "c.npz", a, case.cfg)
    _, b = _state()
    R.restore_restart(path, b, case.cfg.with_(run_seconds=7200.0))
    with pytest.raises(R.RestartMismatchError, match="epssm"):
        R.restore_restart(path, b, case.cfg.with_(epssm=0.2))


def test_different_setup_is_refused(tmp_path):
    case, a = _state()
    path = R.write_restart(tmp_path / "s.npz", a, case.cfg)
    _, b = _state()
    b.phb[3] += 1.0
    with pytest.raises(R.RestartMismatchError, match="setup"):
        R.restore_restart(path, b, case.cfg)


def test_different_shape_or_dtype_is_refused(tmp_path):
    case, a = _state()
    path = R.write_restart(tmp_path / "x.npz", a, case.cfg)
    big = cases.ensemble_member(14, 10, 8, 3)
    with pytest.raises(R.RestartMismatchError):
        R.restore_restart(path, DomainState.from_host(big, device="cpu"),
                          big.cfg.with_(nx=12))
    b64 = DomainState.from_host(case, device="cpu", dtype=torch.float64)
    with pytest.raises(R.RestartMismatchError):
        R.restore_restart(path, b64, case.cfg)


def test_truncated_file_is_refused(tmp_path):
    case, a = _state()
    path = R.write_restart(tmp_path / "t.npz", a, case.cfg)
    raw = path.read_bytes()
    bad = tmp_path / "bad.npz"
    bad.write_bytes(raw[: len(raw) // 2])
    with pytest.raises(R.RestartMismatchError, match="unreadable"):
        R.restore_restart(bad, a, case.cfg)


def test_fingerprint_covers_every_setup_array():
    case, a = _state()
    f0 = R.setup_fingerprint(a)
    for name in R.STATE_SETUP_ARRAYS:
        t = getattr(a, name)
        old = t.view(-1)[0].clone()
        t.view(-1)[0] = old + 1.0
        assert R.setup_fingerprint(a) != f0, name
        t.view(-1)[0] = old
    assert R.setup_fingerprint(a) == f0
    assert dataclasses.is_dataclass(case.cfg)
