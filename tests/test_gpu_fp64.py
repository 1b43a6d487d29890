"""The FP64 build of the step (arw64_*) against the float64 oracle: the
same discretisation in double precision.  Divisions and square roots are
correctly rounded on both sides; pow (EOS) and sin (damper) come from
different libm implementations and may differ in the last bit, so the
bar is a stated relative tolerance instead of bit identity."""
import numpy as np
import pytest
import torch

from oracle import dycore_np as O
from tests import pins
from tests.helpers import PROG

pytestmark = pytest.mark.gpu

#: relative (normwise, per field) FP64 agreement after the steps below
FP64_TOL = 1e-9


def _fp64_pair(case, set_w=False):
    from arwen_amd.state import DomainState
    o = O.from_case(case, np.float64)
    if set_w:
        O.set_w_surface(o, case.cfg.dx, case.cfg.dy)
    g = DomainState.from_host(case, dtype=torch.float64)
    for name in PROG:
        getattr(g, name).copy_(torch.as_tensor(getattr(o, name)))
    return o, g


def _rel(o, g, names):
    out = {}
    for n in names:
        a = getattr(g, n).detach().cpu().numpy()
        b = getattr(o, n)
        out[n] = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
    return out


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
def test_fp64_step_matches_fp64_oracle(name):
    from arwen_amd import dycore
    case = pins.CASES[name]()
    o, g = _fp64_pair(case, set_w=(name == "terrain"))
    print(f"fp64 parity {name}", flush=True)
    for _ in range(3):
        O.step(o, case.cfg)
        dycore.step(g, case.cfg)
    rel = _rel(o, g, PROG + ("p", "alt"))
    assert max(rel.values()) < FP64_TOL, rel


def test_fp64_moist_step_matches_fp64_oracle():
    from arwen_amd import dycore
    from tests.test_gpu_moist import _case
    case = _case()
    o, g = _fp64_pair(case)
    print("fp64 moist parity", flush=True)
    for _ in range(2):
        O.step(o, case.cfg)
        dycore.step(g, case.cfg)
    rel = _rel(o, g, PROG + ("qv", "qc", "qr", "p"))
    assert max(rel.values()) < FP64_TOL, rel
