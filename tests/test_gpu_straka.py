"""Straka et al. (1993) density current on the device: the reference's
pass gates at t = 900 s (gpuwm/verify/cases/straka.py:30-41 -- theta'_min,
surface -1 K front position, x-symmetry, bounded w; tests/
test_case_straka.py), dry-mass conservation to FP32 round-off over the
run, and bit-identity with the float32 oracle over the first steps of
the same x-z (ny = 1) case with constant-K diffusion."""
import numpy as np
import pytest
import torch

from arwen_amd import cases
from oracle import dycore_np as O
from tests.helpers import compare, oracle_and_gpu

pytestmark = pytest.mark.gpu

#: the reference's gates (lo, hi), open bounds as None
GATES = {"theta_min": (-11.0, -8.0), "front_km": (14.0, 17.0),
         "symmetry_err": (None, 0.05), "w_max": (None, 40.0)}
#: relative change of the FP64 dry-mass sum over the 900 s run
MASS_TOL = 2e-6


def front_km(th_sfc, x):
    """Position (km) of the outermost surface crossing of theta' = -1 K on
    the x > 0 side, interpolated linearly between cell centres."""
    cold = th_sfc <= -1.0
    edges = np.flatnonzero(cold[:-1] & ~cold[1:])
    if edges.size == 0:
        return float("nan")
    i = edges[-1]
    a, b = th_sfc[i], th_sfc[i + 1]
    return float(x[i] + (-1.0 - a) / (b - a) * (x[i + 1] - x[i])) / 1e3


def _gate(metrics):
    bad = {}
    for k, (lo, hi) in GATES.items():
        v = metrics[k]
        if not np.isfinite(v) or (lo is not None and not v > lo) \
                or (hi is not None and not v < hi):
            bad[k] = v
    return bad


@pytest.mark.parametrize("dt,ns", [(1.0, 6), (0.5, 6)])
def test_straka_gates_and_mass(dt, ns):
    from arwen_amd import dycore
    from arwen_amd.state import DomainState
    case = cases.straka(dt=dt, time_step_sound=ns)
    cfg = case.cfg
    g = DomainState.from_host(case)
    print(f"straka dt={dt} ns={ns}", flush=True)
    m0 = dycore.domain_mass_measure(g)
    n = int(round(900.0 / dt))
    for k in range(0, n, 100):
        dycore.run_steps(g, cfg, min(100, n - k))
    rep = dycore.stability_report(g, cfg)
    m1 = dycore.domain_mass_measure(g)
    thp = g.thp.double().cpu().numpy()
    x = cases.cell_x(cfg)
    metrics = {"theta_min": float(thp.min()),
               "front_km": front_km(thp[0, 0], x),
               "symmetry_err": float(np.abs(thp - thp[..., ::-1]).max()),
               "w_max": float(rep["w_max"])}
    print(metrics, "mass drift", (m1 - m0) / m0, flush=True)
    assert not rep["nan"]
    assert not _gate(metrics), metrics
    assert abs(m1 - m0) <= MASS_TOL * abs(m0), (m0, m1)


def test_straka_first_steps_match_oracle_bitwise():
    from arwen_amd import dycore
    case = cases.straka()
    o, g = oracle_and_gpu(case)
    print("straka oracle parity", flush=True)
    for _ in range(5):
        O.step(o, case.cfg)
        dycore.step(g, case.cfg)
    res = compare(o, g)
    bad = {k: v for k, v in res.items() if v[1] < 1.0}
    assert not bad, bad
