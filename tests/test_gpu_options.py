"""GPU parity of two configuration options against the float32 oracle,
bit for bit over whole steps: h_sca_adv_order = 5 (rhs_ph's horizontal
geopotential advection, reference dycore.cu:668-748) and
hypsometric_opt = 2 (the EOS's log-pressure specific volume, reference
diagnostics.cu:75-80), dry on flat and terrain bases and moist; the FP64
build against the float64 oracle; the standalone EOS."""
import numpy as np
import pytest
import torch

from oracle import dycore_np as O
from tests.helpers import compare, oracle_and_gpu, random_state_case

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "thp", "php", "mup", "p", "al", "alt")


def _case(hill=0.0, **opts):
    case, _ = random_state_case(seed=21, nz=12, ny=14, nx=18, hill=hill)
    case.cfg = case.cfg.with_(**opts)
    return case


@pytest.mark.parametrize("opts,hill", [
    ({"h_sca_adv_order": 5}, 0.0),
    ({"h_sca_adv_order": 5}, 300.0),
    ({"hypsometric_opt": 2}, 0.0),
    ({"hypsometric_opt": 2}, 300.0),
    ({"hypsometric_opt": 2, "h_sca_adv_order": 5}, 0.0),
])
def test_option_step_matches_oracle_bitwise(opts, hill):
    from arwen_amd import dycore
    case = _case(hill, **opts)
    o, g = oracle_and_gpu(case, set_w=hill > 0)
    print(f"option step parity {opts} hill={hill}", flush=True)
    for _ in range(3):
        O.step(o, case.cfg)
        dycore.step(g, case.cfg)
    res = compare(o, g, FIELDS)
    bad = {k: v for k, v in res.items() if v[1] < 1.0}
    assert not bad, bad


def test_moist_hypsometric_step_matches_oracle_bitwise():
    from arwen_amd import cases, dycore
    case = cases.mixed_phase(6, nx=20, ny=14, nz=16)
    case.cfg = case.cfg.with_(hypsometric_opt=2, h_sca_adv_order=5)
    o, g = oracle_and_gpu(case)
    print("moist option step parity", flush=True)
    for _ in range(2):
        O.step(o, case.cfg)
        dycore.step(g, case.cfg)
    res = compare(o, g, FIELDS + ("qv", "qi"))
    bad = {k: v for k, v in res.items() if v[1] < 1.0}
    assert not bad, bad


def test_standalone_eos_hypsometric_two():
    from arwen_amd import dycore
    case = _case(300.0, hypsometric_opt=2)
    o, g = oracle_and_gpu(case, set_w=True)
    O.calc_p_alpha(o, False, 2)
    dycore.update_diagnostics(g, 2)
    res = compare(o, g, ("p", "al", "alt"))
    assert all(v[1] == 1.0 for v in res.values()), res
    o1 = O.from_case(case, np.float32)
    O.calc_p_alpha(o1, False, 1)
    assert not np.array_equal(o1.alt, o.alt)


def test_fp64_options_match_fp64_oracle():
    from arwen_amd import dycore
    from arwen_amd.state import DomainState
    case = _case(300.0, hypsometric_opt=2, h_sca_adv_order=5)
    o = O.from_case(case, np.float64)
    O.set_w_surface(o, case.cfg.dx, case.cfg.dy)
    g = DomainState.from_host(case, dtype=torch.float64)
    g.w.copy_(torch.as_tensor(o.w))
    print("fp64 option parity", flush=True)
    for _ in range(2):
        O.step(o, case.cfg)
        dycore.step(g, case.cfg)
    for n in FIELDS:
        a = getattr(g, n).cpu().numpy()
        b = getattr(o, n)
        rel = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
        assert rel < 1e-9, (n, rel)


@pytest.mark.parametrize("mp", [0, 1, 10, 18])
def test_moist_cq_step_matches_oracle_bitwise(mp):
    """moist_cq (WRF calc_cq, reference acoustic.cu:84-158): cqu/cqv scale
    the horizontal acoustic pressure gradient, cqw the implicit w rows and
    the w pressure forcing; qv only for mp_physics 0, the loading masses
    otherwise (NSSL's hail mass included)."""
    from arwen_amd import cases, dycore
    if mp == 0:
        case = cases.moist_tracer(nx=20, ny=14, nz=16)
    elif mp == 1:
        case = cases.add_species(cases.moist_tracer(nx=20, ny=14, nz=16), 1)
    else:
        case = cases.mixed_phase(mp, nx=20, ny=14, nz=16)
    case.cfg = case.cfg.with_(moist_cq=True)
    o, g = oracle_and_gpu(case)
    plain = cases.moist_tracer(nx=20, ny=14, nz=16) if mp == 0 else None
    print(f"moist_cq step parity mp={mp}", flush=True)
    for _ in range(2):
        O.step(o, case.cfg)
        dycore.step(g, case.cfg)
    res = compare(o, g, FIELDS + ("qv",))
    bad = {k: v for k, v in res.items() if v[1] < 1.0}
    assert not bad, bad
    if plain is not None:                # the option is live on the device
        o0, g0 = oracle_and_gpu(plain)
        for _ in range(2):
            dycore.step(g0, plain.cfg)
        assert not torch.equal(g0.w, g.w)
