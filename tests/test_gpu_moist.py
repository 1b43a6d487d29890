"""GPU parity of the moist step (qv/qc/qr with mp_physics = 0) against
the float32 oracle: bit-identical fields after whole steps with the PD
limiter on and off, on flat and terrain bases, plus the reference's
uniform-scalar invariance on the device."""
import numpy as np
import pytest

from arwen_amd import cases
from arwen_amd.config import RunConfig
from arwen_amd.grid import bell_hill, make_base_state, make_vertical_coord
from oracle import dycore_np as O
from tests.helpers import compare, oracle_and_gpu

pytestmark = pytest.mark.gpu

MOIST = ("u", "v", "w", "thp", "php", "mup", "p", "al", "alt", "qv", "qc",
         "qr")


def _case(moist_adv_opt=1, qv_const=None, seed=5, hill=0.0, u0=8.0):
    nx, ny, nz = 16, 12, 16
    cfg = RunConfig(nx=nx, ny=ny, nz=nz, dx=500.0, dy=500.0, ztop=8000.0,
                    dt=2.0, moist=True, moist_adv_opt=moist_adv_opt,
                    terrain_opt=1 if hill > 0 else 0)
    vc = make_vertical_coord(nz)
    th = lambda z: 300.0 * np.exp(1e-4 * np.asarray(z, float) / 9.81)
    b = make_base_state(vc, th, p_surf=cfg.p_surf, ztop=cfg.ztop)
    z = cases.height_half(b)
    x = cases.cell_x(cfg)
    L = np.sqrt((x[None, None, :] / 2000.0) ** 2
                + ((z[:, None, None] - 2000.0) / 1500.0) ** 2)
    thp = np.where(L < 1.0, 5.0 * np.cos(np.pi * L / 2) ** 2, 0.0) \
        * np.ones((nz, ny, nx))
    if qv_const is not None:
        qv = np.full((nz, ny, nx), qv_const)
    else:
        rng = np.random.default_rng(seed)
        qv = rng.uniform(0.0, 0.016, (nz, ny, nx))
        qv[rng.uniform(size=qv.shape) < 0.3] = 0.0
        qv = qv * np.exp(-z / 2500.0)[:, None, None]
    case = cases.init_moist_balanced(cfg, vc, b, qv, thp)
    case.fields["u"][...] = u0
    case.fields["qc"][...] = 0.3 * qv        # nonzero cloud and rain too
    case.fields["qr"][...] = 0.1 * qv
    if hill > 0:
        # same fields over a bell hill: rebuild the base with terrain
        tz = bell_hill(nx, ny, 500.0, hill, 1500.0)
        bt = make_base_state(vc, th, p_surf=cfg.p_surf, ztop=cfg.ztop,
                             terrain_z=tz)
        case = cases.HostCase(cfg, vc, bt, case.fields)
    return case


@pytest.mark.parametrize("opt", [1, 0])
@pytest.mark.parametrize("nsteps", [1, 3])
def test_moist_step_matches_oracle_bitwise(opt, nsteps):
    from arwen_amd import dycore
    case = _case(moist_adv_opt=opt)
    o, g = oracle_and_gpu(case)
    print(f"moist step parity opt={opt} steps={nsteps}", flush=True)
    for _ in range(nsteps):
        O.step(o, case.cfg)
        dycore.step(g, case.cfg)
    res = compare(o, g, MOIST)
    bad = {k: v for k, v in res.items() if v[1] < 1.0}
    assert not bad, bad
    assert float(g.qv.min()) >= 0.0


def test_moist_terrain_step_matches_oracle_bitwise():
    from arwen_amd import dycore
    case = _case(hill=300.0)
    o, g = oracle_and_gpu(case, set_w=True)
    print("moist terrain step parity", flush=True)
    for _ in range(2):
        O.step(o, case.cfg)
        dycore.step(g, case.cfg)
    res = compare(o, g, MOIST)
    bad = {k: v for k, v in res.items() if v[1] < 1.0}
    assert not bad, bad


def test_uniform_qv_stays_uniform_on_device():
    """Reference tests/test_moist.py:517-577 on the GPU path."""
    from arwen_amd import dycore
    from arwen_amd.state import DomainState
    case = _case(qv_const=0.005, u0=0.0)
    case.fields["qc"][...] = 0.0
    case.fields["qr"][...] = 0.0
    g = DomainState.from_host(case)
    print("uniform qv on device", flush=True)
    for _ in range(10):
        dycore.step(g, case.cfg)
    assert float(g.w.abs().max()) > 0.5
    dev = float((g.qv.double() - 0.005).abs().max()) / 0.005
    assert dev < 2e-5, dev
