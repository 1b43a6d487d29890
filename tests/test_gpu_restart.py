"""The restart contract on the device: N steps, a restart file, a freshly
prepared state restored from it and N more steps equal 2N uninterrupted
steps bit for bit on every serialized field and the clock (reference
gpuwm/io/restart.py:1-10, its 6 h + restart + 6 h == 12 h gate), dry on a
flat and a terrain base with the upper damper, moist with the PD
limiter, and an NSSL state with every species and retained heating."""
import numpy as np
import pytest
import torch

from arwen_amd import cases
from arwen_amd import restart as R
from arwen_amd.state import DomainState

pytestmark = pytest.mark.gpu


def _cases():
    return {
        "dry_flat": cases.ensemble_member(40, 32, 20, 7),
        "terrain_damped": cases.stratified_hill(nx=36, ny=30, nz=24,
                                                seed=9, radius=40000.0),
        "moist_pd": cases.moist_tracer(nx=32, ny=24, nz=20),
        "nssl_heated": cases.mixed_phase(18, nx=32, ny=24, nz=20),
    }


@pytest.mark.parametrize("name", ["dry_flat", "terrain_damped", "moist_pd",
                                  "nssl_heated"])
def test_restart_continues_bitwise(name, tmp_path):
    from arwen_amd import dycore
    case = _cases()[name]
    cfg = case.cfg
    print(f"restart continuation {name}", flush=True)
    whole = DomainState.from_host(case)
    dycore.run_steps(whole, cfg, 6)
    first = DomainState.from_host(case)
    dycore.run_steps(first, cfg, 3)
    path = R.write_restart(tmp_path / f"{name}.npz", first, cfg)
    del first
    resumed = DomainState.from_host(case)
    R.restore_restart(path, resumed, cfg)
    dycore.run_steps(resumed, cfg, 3)
    torch.cuda.synchronize()
    assert resumed.elapsed_seconds == whole.elapsed_seconds
    for n in R.STATE_SERIALIZED_ATTRS:
        a, b = getattr(whole, n, None), getattr(resumed, n, None)
        if a is None:
            assert b is None
            continue
        assert torch.equal(a, b), n
