"""The health gate's host mirror and rules (reference tests/test_health.py
:42-110 are the model): every corruption class of the dycore state is
attributed to its field and first flat index, classes OR while the first
failing field stays deterministic, theta' and mu' are checked with their
base auxiliaries, and a failed gate raises a terminal error that carries
the report."""
import numpy as np
import pytest
import torch

from arwen_amd import cases
from arwen_amd.health import (HealthCheckError, HealthField, STATUS_BITS,
                              collect_state_fields, rule_for_field,
                              validate_fields_cpu, validate_state_cpu)
from arwen_amd.state import DomainState

CORRUPTIONS = (
    ("u", -501.0), ("v", 501.0), ("w", 201.0), ("php", np.nan),
    ("mup", 0.0), ("p", 0.0), ("al", np.inf), ("alt", 0.0),
    ("qv", -1.0e-6), ("qc", -1.0e-6), ("qr", 1.5),
    ("held.diff6.u", np.nan), ("anything_else", -np.inf),
)


@pytest.mark.parametrize(("field", "corruption"), CORRUPTIONS)
def test_cpu_mirror_attributes_every_corruption(field, corruption):
    values = np.full((2, 3), 1.0, dtype=np.float32)
    values.flat[4] = corruption
    report = validate_fields_cpu({field: values}, phase="injected")
    assert not report.ok
    assert report.first_bad_field == field
    assert report.first_bad_index == (1, 1)
    assert report.first_bad_flat_index == 4
    assert report.status_bits == rule_for_field(field).status_bit
    assert report.reason and report.phase == "injected"


def test_classes_or_but_the_first_field_is_deterministic():
    report = validate_fields_cpu({
        "v": np.array([0.0, 900.0], dtype=np.float32),
        "p": np.array([1.0e5, np.nan], dtype=np.float32),
        "qv": np.array([2.0], dtype=np.float32),
    })
    assert not report.ok and report.first_bad_field == "v"
    assert set(report.failing_classes) == {"wind", "pressure", "moisture"}


def test_bounds_are_inclusive_except_the_strict_denominators():
    ok = validate_fields_cpu({"u": np.array([500.0, -500.0], np.float32),
                              "al": np.array([-1.0e4], np.float32),
                              "qv": np.array([0.0, 1.0], np.float32)})
    assert ok.ok
    for name in ("mup", "p", "alt"):
        assert not validate_fields_cpu({name: np.array([0.0], np.float32)}).ok
        assert validate_fields_cpu({name: np.array([1e-30], np.float32)}).ok


def test_theta_and_column_mass_checked_with_their_bases():
    thp = np.array([[[0.0]], [[-250.0]]], dtype=np.float32)
    thb = np.array([300.0, 300.0], dtype=np.float32)
    mup = np.array([[-101.0]], dtype=np.float32)
    mub = np.array([[100.0]], dtype=np.float32)
    fields = (HealthField("thp", thp, rule_for_field("thp"), thb, "level", 1),
              HealthField("mup", mup, rule_for_field("mup"), mub, "direct", 0))
    report = validate_fields_cpu(fields)
    assert report.first_bad_field == "thp"
    assert report.first_bad_index == (1, 0, 0)
    assert report.first_bad_value == 50.0
    assert set(report.failing_classes) == {"theta", "coupled_mass"}


def test_theta_bound_admits_a_legal_upper_state():
    assert validate_fields_cpu({"thp": np.array([505.0], np.float32)}).ok
    assert not validate_fields_cpu({"thp": np.array([601.0], np.float32)}).ok


def test_error_is_terminal_and_carries_the_report():
    report = validate_fields_cpu({"p": np.array([0.0], np.float32)},
                                 phase="post-step")
    err = HealthCheckError(report)
    assert err.report is report
    assert isinstance(err, FloatingPointError)
    assert "post-step" in str(err) and "p" in str(err)


def test_state_inventory_follows_the_restart_manifest():
    from arwen_amd import restart
    dry = DomainState.from_host(cases.ensemble_member(8, 6, 5, 1),
                                device="cpu")
    moist = DomainState.from_host(cases.moist_tracer(nx=8, ny=6, nz=5),
                                  device="cpu")
    names = [f.name for f in collect_state_fields(dry)]
    assert names == ["u", "v", "w", "thp", "php", "mup", "p", "al", "alt"]
    mnames = [f.name for f in collect_state_fields(moist)]
    assert set(mnames) == {n for n in restart.STATE_SERIALIZED_ATTRS
                           if getattr(moist, n, None) is not None}
    assert set(mnames) == {"u", "v", "w", "thp", "php", "mup", "p", "al",
                           "alt", "qv", "qc", "qr", "h_diabatic"}
    # an NSSL state: every transported species is gated, with the
    # reference's moisture / moment / particle-volume classes
    nssl = DomainState.from_host(cases.mixed_phase(18, nx=8, ny=6, nz=5),
                                 device="cpu")
    nf = {f.name: f for f in collect_state_fields(nssl)}
    assert set(nf) == {n for n in restart.STATE_SERIALIZED_ATTRS
                       if getattr(nssl, n, None) is not None}
    assert set(nssl.extra_species) <= set(nf)
    assert nf["qh"].rule.status_class == "moisture"
    assert nf["qnn"].rule.status_class == "moment"
    assert nf["qnn"].rule.upper == 1.0e15
    assert nf["qvolg"].rule.status_class == "particle_volume"
    by = {f.name: f for f in collect_state_fields(dry)}
    assert by["thp"].aux_mode == "level" and by["thp"].plane_size == 48
    assert by["mup"].aux_mode == "direct"
    for st in (dry, moist):            # the EOS runs on the device only
        st.p.fill_(9.0e4)
        st.alt.fill_(1.0)
    assert validate_state_cpu(dry).ok and validate_state_cpu(moist).ok
    moist.qv.view(-1)[7] = -1.0
    with pytest.raises(HealthCheckError, match="qv"):
        validate_state_cpu(moist, raise_on_error=True)


def test_status_bits_are_distinct():
    bits = list(STATUS_BITS.values())
    assert len(set(bits)) == len(bits) and all(b & (b - 1) == 0 for b in bits)
