"""The held 6th-order filter (diff_6th_opt 1 and 2; the slope taper on a
terrain base) in the fused device step, against the float32 oracle step
bit for bit and the float64 oracle to rounding, on states with grid-scale
noise so that the filter acts.  The oracle's filter is pinned by its
closed forms (tests/test_oracle_diff6.py)."""
import numpy as np
import pytest
import torch

import pins
from helpers import DIAG, PROG, oracle_and_gpu
from oracle import dycore_np as O

pytestmark = pytest.mark.gpu


def _noisy(name, dtype, seed):
    case = pins.CASES[name]()
    o, g = oracle_and_gpu(case, dtype=dtype)
    rng = np.random.default_rng(seed)
    for n, amp in (("u", 0.5), ("v", 0.5), ("thp", 0.3)):
        x = getattr(o, n)
        x[...] = (x + rng.standard_normal(x.shape) * amp).astype(dtype)
    x = o.w
    x[1:-1] = (x[1:-1] + rng.standard_normal(x[1:-1].shape) * 0.1).astype(dtype)
    O.calc_p_alpha(o)
    return case, o, g


@pytest.mark.parametrize("name,opt,slope", [("dry_flat", 1, 0),
                                            ("dry_flat", 2, 0),
                                            ("terrain", 2, 0),
                                            ("terrain", 2, 1)])
def test_step_with_diff6_bit_identical_to_oracle(name, opt, slope):
    from arwen_amd import dycore
    case, o, g = _noisy(name, np.float32, 97)
    cfg = case.cfg.with_(diff_6th_opt=opt, diff_6th_factor=0.25,
                         diff_6th_slopeopt=slope, diff_6th_thresh=0.02)
    print(f"diff6 step parity {name} opt={opt} slope={slope}", flush=True)
    for n in PROG:
        getattr(g, n).copy_(torch.as_tensor(getattr(o, n)))
    dycore.update_diagnostics(g)
    plain = {n: getattr(o, n).copy() for n in PROG}
    for _ in range(2):
        O.step(o, cfg)
        dycore.step(g, cfg)
    for n in PROG + DIAG:
        got = getattr(g, n).cpu().numpy()
        np.testing.assert_array_equal(got, getattr(o, n), err_msg=n)
    # the filter acted: the same two steps without it differ
    o2 = O.from_case(case, np.float32)
    for n in PROG:
        getattr(o2, n)[...] = plain[n]
    O.calc_p_alpha(o2)
    for _ in range(2):
        O.step(o2, case.cfg)
    assert not np.array_equal(o2.thp, o.thp)


def test_step_with_diff6_fp64_matches_oracle():
    from arwen_amd import dycore
    from arwen_amd.state import DomainState
    case, o, _ = _noisy("terrain", np.float64, 101)
    cfg = case.cfg.with_(diff_6th_opt=2, diff_6th_factor=0.25,
                         diff_6th_slopeopt=1, diff_6th_thresh=0.02)
    print("diff6 step FP64", flush=True)
    g = DomainState.from_host(case, dtype=torch.float64)
    for n in PROG:
        getattr(g, n).copy_(torch.as_tensor(getattr(o, n)))
    dycore.update_diagnostics(g)
    O.step(o, cfg)
    dycore.step(g, cfg)
    for n in PROG:
        a = getattr(g, n).cpu().numpy()
        b = getattr(o, n)
        scale = np.max(np.abs(b)) or 1.0
        assert np.max(np.abs(a - b)) <= 1e-12 * scale, n


def test_diff6_rejections():
    from arwen_amd import dycore
    from arwen_amd._lib import ArwenError
    from arwen_amd.state import DomainState
    case = pins.CASES["dry_flat"]()
    g = DomainState.from_host(case)
    with pytest.raises(ValueError, match="diff_6th_opt"):
        dycore.step(g, case.cfg.with_(diff_6th_opt=3))
    with pytest.raises(NotImplementedError, match="plain form"):
        dycore.step(g, case.cfg.with_(diff_6th_opt=2, precision_form=True))
    from arwen_amd import adjoint
    with pytest.raises(ArwenError, match="held mixing"):
        adjoint.step_adjoint(g, case.cfg.with_(diff_6th_opt=2),
                             {"u": torch.ones_like(g.u)})
