"""The precision form of the float32 step (compensated column sums, FP64
vertical tridiagonal) against the oracle's restatement of the same form:
bit-identical, like the plain step."""
import numpy as np
import pytest

from oracle import dycore_np as O
from tests import pins
from tests.helpers import DIAG, PROG, compare, oracle_and_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
def test_precision_form_matches_oracle_bitwise(name):
    from arwen_amd import dycore
    case = pins.CASES[name]()
    cfg = case.cfg.with_(precision_form=True)
    o, g = oracle_and_gpu(case, set_w=(name == "terrain"))
    print(f"precision form parity {name}", flush=True)
    for _ in range(3):
        O.step(o, cfg)
        dycore.step(g, cfg)
    res = compare(o, g, PROG + DIAG)
    bad = {k: v for k, v in res.items() if v[1] < 1.0}
    assert not bad, bad


def test_precision_form_moist_matches_oracle_bitwise():
    from arwen_amd import dycore
    from tests.test_gpu_moist import MOIST, _case
    case = _case()
    cfg = case.cfg.with_(precision_form=True)
    o, g = oracle_and_gpu(case)
    print("precision form moist parity", flush=True)
    for _ in range(2):
        O.step(o, cfg)
        dycore.step(g, cfg)
    res = compare(o, g, MOIST)
    bad = {k: v for k, v in res.items() if v[1] < 1.0}
    assert not bad, bad
