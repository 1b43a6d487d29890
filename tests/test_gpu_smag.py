"""GPU parity of km_opt = 4 (WRF 2-D Smagorinsky, reference
dycore.py:1223-1343 and kernels/smag2d.cu) against the float32 oracle:
the held tendencies themselves and whole steps, bit for bit, on flat and
terrain bases, for each surface policy, with the dycore's vertical mixing
and without it (a PBL scheme's), alone and under the 6th-order filter;
the FP64 build against the float64 oracle; a state at rest steps exactly
as with km_opt = 1."""
import numpy as np
import pytest
import torch

from oracle import dycore_np as O
from oracle import smag_np as S
from tests.helpers import compare, oracle_and_gpu, random_state_case

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "thp", "php", "mup", "p", "al", "alt")
HELD = {"ru_t": "arw_d6_u", "rv_t": "arw_d6_v", "rw_t": "arw_d6_w",
        "rth_t": "arw_d6_t"}


def _case(hill=0.0, seed=31, **opts):
    case, _ = random_state_case(seed=seed, nz=10, ny=14, nx=18, hill=hill)
    case.fields["u"] = 40.0 * case.fields["u"]      # ~2 m/s of shear
    case.fields["v"] = 40.0 * case.fields["v"]
    case.cfg = case.cfg.with_(km_opt=4, **opts)
    return case


OPTS = [
    ({}, 0.0),
    ({}, 300.0),
    ({"isfflx": 0, "tke_drag_coefficient": 0.01, "tke_heat_flux": 0.05},
     300.0),
    ({"isfflx": 2, "tke_heat_flux": -0.03}, 0.0),
    ({"bl_pbl_physics": 1}, 300.0),
    ({"diff_6th_opt": 2, "diff_6th_slopeopt": 1}, 300.0),
]


@pytest.mark.parametrize("opts,hill", OPTS)
def test_held_tendencies_match_oracle_bitwise(opts, hill):
    from arwen_amd import dycore
    case = _case(hill, **opts)
    o, g = oracle_and_gpu(case, set_w=hill > 0)
    print(f"smag held tendencies {opts} hill={hill}", flush=True)
    for n in O.PROGNOSTICS:
        getattr(o, n + "0")[...] = getattr(o, n)
    O.calc_p_alpha(o)
    want = S.smag_held(o, case.cfg)
    if case.cfg.diff_6th_opt:
        d6 = O.diff6_held(o, case.cfg)
        want = {n: want[n] + d6[n] for n in d6}
    dycore.step(g, case.cfg)
    for name, slot in HELD.items():
        got = g._scratch[slot].cpu().numpy()
        assert np.array_equal(got, want[name]), (
            name, float(np.abs(got - want[name]).max()))
        assert np.abs(want[name]).max() > 0.0, name


@pytest.mark.parametrize("opts,hill", OPTS)
def test_step_matches_oracle_bitwise(opts, hill):
    from arwen_amd import dycore
    case = _case(hill, **opts)
    o, g = oracle_and_gpu(case, set_w=hill > 0)
    print(f"smag step parity {opts} hill={hill}", flush=True)
    for _ in range(3):
        O.step(o, case.cfg)
        dycore.step(g, case.cfg)
    res = compare(o, g, FIELDS)
    bad = {k: v for k, v in res.items() if v[1] < 1.0}
    assert not bad, bad


def test_fp64_step_matches_fp64_oracle():
    from arwen_amd import dycore
    from arwen_amd.state import DomainState
    case = _case(300.0, isfflx=0, tke_drag_coefficient=0.01,
                 tke_heat_flux=0.05)
    o = O.from_case(case, np.float64)
    O.set_w_surface(o, case.cfg.dx, case.cfg.dy)
    g = DomainState.from_host(case, dtype=torch.float64)
    g.w.copy_(torch.as_tensor(o.w))
    print("fp64 smag parity", flush=True)
    for _ in range(2):
        O.step(o, case.cfg)
        dycore.step(g, case.cfg)
    for n in FIELDS:
        a = getattr(g, n).cpu().numpy()
        b = getattr(o, n)
        rel = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
        assert rel < 1e-9, (n, rel)


def test_rest_state_steps_as_constant_k():
    """K is formed from the time-t winds: a state at rest has K = 0 and
    its first step is bitwise the km_opt = 1 step (reference
    tests/test_smag2d.py:663-723)."""
    from arwen_amd import cases, dycore
    from arwen_amd.config import RunConfig
    from arwen_amd.grid import make_base_state, make_vertical_coord
    from arwen_amd.state import DomainState
    cfg = RunConfig(nx=16, ny=16, nz=8, dx=1000.0, dy=1000.0, ztop=8000.0,
                    dt=6.0)
    vc = make_vertical_coord(cfg.nz)
    b = make_base_state(vc, lambda z: 300.0 + 0.003 * np.asarray(z, float),
                        cfg.p_surf, cfg.ztop)

    def bubble(x, z):
        r = np.sqrt((x[None, None, :] / 2000.0) ** 2
                    + ((z[:, None, None] - 2000.0) / 1000.0) ** 2)
        return 2.0 * np.maximum(0.0, 1.0 - r) * np.ones((1, cfg.ny, 1))

    case = cases.host_theta_perturbation(cfg, vc, b, bubble)
    g1 = DomainState.from_host(case)
    g4 = DomainState.from_host(case)
    c4 = cfg.with_(km_opt=4)
    print("smag at rest vs constant K", flush=True)
    dycore.step(g1, cfg)
    dycore.step(g4, c4)
    for n in ("u", "v", "w", "thp", "php", "mup"):
        assert torch.equal(getattr(g1, n), getattr(g4, n)), n
    for _ in range(3):
        dycore.step(g1, cfg)
        dycore.step(g4, c4)
    assert not torch.equal(g1.u, g4.u)


def test_refusals():
    from arwen_amd import _lib, dycore
    case = _case()
    case.cfg = case.cfg.with_(km_opt=3)
    o, g = oracle_and_gpu(_case())
    with pytest.raises(NotImplementedError):
        dycore.step(g, case.cfg)
    with pytest.raises(NotImplementedError):
        dycore.step(g, _case().cfg.with_(precision_form=True))


# --- moist states: vapour-loaded density, held species tendencies ---------

MOIST = [
    (0, {"km_opt": 4}),
    (10, {"km_opt": 4}),
    (10, {"km_opt": 4, "moist_adv_opt": 0}),
    (6, {"diff_6th_opt": 2}),
    (18, {"km_opt": 4, "diff_6th_opt": 1, "moist_mix6_off": True}),
]


@pytest.mark.parametrize("mp,opts", MOIST)
def test_moist_step_matches_oracle_bitwise(mp, opts):
    from arwen_amd import dycore
    from arwen_amd.config import EXTRA_SPECIES
    from tests.test_oracle_smag import _sheared_moist
    case = _sheared_moist(mp, **opts)
    o, g = oracle_and_gpu(case)
    print(f"moist held mixing step parity mp={mp} {opts}", flush=True)
    for _ in range(2):
        O.step(o, case.cfg)
        dycore.step(g, case.cfg)
    names = FIELDS + ("qv", "qc", "qr") + EXTRA_SPECIES[mp]
    res = compare(o, g, names)
    bad = {k: v for k, v in res.items() if v[1] < 1.0}
    assert not bad, bad


def test_moist_held_species_tendencies_bitwise():
    """The unlimited scheme leaves the held species tendencies in place:
    Smagorinsky plus the filter of every species, against the oracle."""
    from arwen_amd import dycore
    from tests.test_oracle_smag import _sheared_moist
    case = _sheared_moist(10, km_opt=4, diff_6th_opt=2, moist_adv_opt=0)
    o, g = oracle_and_gpu(case)
    for n in O.PROGNOSTICS:
        getattr(o, n + "0")[...] = getattr(o, n)
    names = ("qv", "qc", "qr") + o.extra
    for q in names:
        getattr(o, q + "0")[...] = getattr(o, q)
    O.calc_p_alpha(o)
    want = S.smag_held(o, case.cfg)["q"]
    d6 = O.diff6_held(o, case.cfg)["q"]
    print("moist held species tendencies", flush=True)
    dycore.step(g, case.cfg)
    got = g._scratch["arw_held_q"].cpu().numpy()
    for e, n in enumerate(names):
        assert np.array_equal(got[e], want[n] + d6[n]), n
        assert np.abs(want[n]).max() > 0.0, n


@pytest.mark.parametrize("opt", [1, 0])
def test_column_march_scalars_with_held_mixing(opt):
    """512 x 512 columns: the column-march scalar updates add the held
    tendencies, and the column limiter advects the folded time-t scalar."""
    from arwen_amd import cases, dycore
    from arwen_amd.config import EXTRA_SPECIES
    case = cases.mixed_phase(6, nx=512, ny=512, nz=6, dx=3000.0,
                             moist_adv_opt=opt)
    y = (np.arange(512) + 0.5) / 512
    case.fields["u"] = case.fields["u"] + 3.0 * np.sin(2 * np.pi * y)[
        None, :, None]
    case.cfg = case.cfg.with_(km_opt=4, diff_6th_opt=2)
    o, g = oracle_and_gpu(case)
    print(f"column-march held mixing 512x512x6 moist_adv_opt={opt}",
          flush=True)
    O.step(o, case.cfg)
    dycore.step(g, case.cfg)
    res = compare(o, g, FIELDS + ("qv", "qc", "qr") + EXTRA_SPECIES[6])
    bad = {k: v for k, v in res.items() if v[1] < 1.0}
    assert not bad, bad
