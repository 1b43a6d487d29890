"""The device health gate (arw_health_gate through StateHealthValidator)
against its host mirror (reference tests/test_health.py:309-440, the GPU
twins): on a stepped state it passes; every injected corruption is
reported with the host mirror's field, index, value and status classes;
a fault far beyond the first chunk is found; with several faults the
first field in gate order wins, and within it the smallest index; FP32
and FP64 states."""
import numpy as np
import pytest
import torch

from arwen_amd import cases
from arwen_amd.health import (HealthCheckError, StateHealthValidator,
                              validate_state_cpu)
from arwen_amd.state import DomainState

pytestmark = pytest.mark.gpu


def _state(dtype=torch.float32, moist=True):
    from arwen_amd import dycore
    case = (cases.moist_tracer(nx=96, ny=80, nz=40) if moist
            else cases.ensemble_member(96, 80, 40, 13))
    st = DomainState.from_host(case, dtype=dtype)
    dycore.run_steps(st, case.cfg, 2)
    torch.cuda.synchronize()
    return st


class _Host:
    """A host copy of the state's checked fields for the mirror."""

    def __init__(self, st):
        for n in ("u", "v", "w", "thp", "php", "mup", "p", "al", "alt",
                  "qv", "qc", "qr", "h_diabatic", "thb", "mub2d"):
            t = getattr(st, n, None)
            setattr(self, n, None if t is None else t.cpu())


def _same(rep, ref):
    assert rep.ok == ref.ok
    assert rep.status_bits == ref.status_bits
    assert rep.first_bad_field == ref.first_bad_field
    assert rep.first_bad_flat_index == ref.first_bad_flat_index
    assert rep.first_bad_index == ref.first_bad_index
    if ref.first_bad_value is not None:
        a, b = rep.first_bad_value, ref.first_bad_value
        assert (np.isnan(a) and np.isnan(b)) or a == b


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_stepped_state_passes(dtype):
    st = _state(dtype)
    print(f"health gate on a stepped state {dtype}", flush=True)
    rep = StateHealthValidator(st).require_healthy(phase="post-step")
    assert rep.ok and rep.status_bits == 0


CORRUPT = [("u", 1234, 600.0), ("v", 77, np.nan), ("w", 200000, -250.0),
           ("thp", 5, 400.0), ("php", 99999, np.inf), ("mup", 3, -1.0e6),
           ("p", 150000, 0.0), ("al", 10, 2.0e4), ("alt", 250000, -1.0),
           ("qv", 300000, -1.0e-3), ("qr", 42, 2.0)]


@pytest.mark.parametrize("name,index,value", CORRUPT)
def test_injected_corruption_matches_the_host_mirror(name, index, value):
    st = _state()
    print(f"health gate corruption {name}[{index}] = {value}", flush=True)
    t = getattr(st, name)
    t.view(-1)[index] = value
    torch.cuda.synchronize()
    rep = StateHealthValidator(st).validate(phase="injected")
    ref = validate_state_cpu(_Host(st), phase="injected")
    assert not rep.ok and rep.first_bad_field == name
    assert rep.first_bad_flat_index == index
    _same(rep, ref)
    with pytest.raises(HealthCheckError):
        StateHealthValidator(st).require_healthy()


def test_first_field_then_smallest_index_wins():
    st = _state(torch.float64)
    print("health gate attribution", flush=True)
    st.v.view(-1)[5] = np.nan            # later field, small index
    st.u.view(-1)[250000] = 900.0        # earlier field, far chunk
    st.u.view(-1)[200001] = -900.0       # same field, smaller index
    st.qc.view(-1)[1] = -1.0
    torch.cuda.synchronize()
    rep = StateHealthValidator(st).validate()
    ref = validate_state_cpu(_Host(st))
    assert rep.first_bad_field == "u" and rep.first_bad_flat_index == 200001
    assert set(rep.failing_classes) == {"wind", "moisture"}
    _same(rep, ref)
