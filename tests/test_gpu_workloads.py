"""The benchmark workloads at their full sizes on the device:

- C2 (3-D stratified flow over a 1 km hill, 250x200x49, dx = 12 km,
  dt = 60 s): the first step bit-identical to the float32 oracle, then an
  hour of steps with bounded winds and dry mass conserved;
- C3 (1000x1000x50 warm bubbles): the first step bit-identical to the
  float32 oracle at full size, in the plain form and in the precision
  form;
- a long run of the small dry pin case: 50 steps still bit-identical to
  the oracle (the drift between the device step and the restatement of
  the reference step stays exactly zero).
"""
import numpy as np
import pytest
import torch

from arwen_amd import cases
from oracle import dycore_np as O
from tests import pins
from tests.helpers import PROG, DIAG, compare, oracle_and_gpu

pytestmark = pytest.mark.gpu


def _bitwise(o, g, names=PROG + DIAG):
    res = compare(o, g, names)
    return {k: v for k, v in res.items() if v[1] < 1.0}


def test_c2_first_step_bitwise_then_one_hour_stable():
    from arwen_amd import dycore
    case = cases.stratified_hill()
    print("C2 250x200x49 full size", flush=True)
    o, g = oracle_and_gpu(case)
    O.step(o, case.cfg)
    dycore.step(g, case.cfg)
    torch.cuda.synchronize()
    assert not _bitwise(o, g)
    del o
    m0 = dycore.domain_mass_measure(g)
    dycore.run_steps(g, case.cfg, 59)          # one hour at dt = 60 s
    rep = dycore.stability_report(g, case.cfg)
    m1 = dycore.domain_mass_measure(g)
    print(rep, (m1 - m0) / m0, flush=True)
    assert not rep["nan"]
    assert rep["u_max"] < 30.0 and rep["w_max"] < 5.0
    assert rep["cfl"] < 1.0
    assert abs(m1 - m0) <= 1e-6 * abs(m0)


def test_c3_first_step_bitwise_full_size():
    from arwen_amd import dycore
    case = cases.warm_bubbles()
    print("C3 1000x1000x50 full size", flush=True)
    o, g = oracle_and_gpu(case)
    O.step(o, case.cfg)
    dycore.step(g, case.cfg)
    torch.cuda.synchronize()
    assert not _bitwise(o, g)


def test_fifty_steps_stay_bitwise():
    from arwen_amd import dycore
    case = pins.CASES["dry_flat"]()
    print("dry_flat 50 steps vs oracle", flush=True)
    o, g = oracle_and_gpu(case)
    for n in range(50):
        O.step(o, case.cfg)
        dycore.step(g, case.cfg)
    torch.cuda.synchronize()
    assert not _bitwise(o, g)


def test_c3_precision_form_first_step_bitwise_full_size():
    """The precision form (compensated column sums, FP64 vertical
    tridiagonal, perturbation-form EOS) at the C3 size: the first step
    bit-identical to the oracle's restatement of the same form."""
    from arwen_amd import dycore
    case = cases.warm_bubbles()
    cfg = case.cfg.with_(precision_form=True)
    print("C3 1000x1000x50 precision form full size", flush=True)
    o, g = oracle_and_gpu(case)
    O.step(o, cfg)
    dycore.step(g, cfg)
    torch.cuda.synchronize()
    assert not _bitwise(o, g)
