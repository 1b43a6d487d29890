"""The precision form's perturbation-from-base arithmetic in the oracle
(calc_p_alpha_pert, and the prec branches of slow_geopotential,
small_step_init and small_step_finish), pinned on the CPU:

- each branch is an exact rearrangement of the plain expression given the
  base state's hydrostatic relation and its own EOS, so in float64 the two
  agree to rounding, on a flat and a terrain base, dry and moist;
- at rest the perturbation EOS returns al = 0, p' = 0, p = pb exactly;
- in float32 the perturbation EOS's p', al and alt are closer to the
  float64 EOS than the plain float32 EOS is (the reason for the form);
- an alb off the base state's hydrostatic relation breaks the float64
  agreement (so the agreement tests see the relation the form rests on).
"""
import numpy as np
import pytest

from oracle import dycore_np as D
from tests import pins


def _state(name, dtype, seed=3, moist=False):
    case = pins.CASES[name]()
    s = D.from_case(case, dtype)
    rng = np.random.default_rng(seed)
    # a disturbed state: perturbations of the mass and the geopotential
    s.mup[...] += (rng.standard_normal(s.mup.shape) * 30.0).astype(dtype)
    s.php[1:] += (rng.standard_normal(s.php[1:].shape) * 40.0).astype(dtype)
    s.thp[...] += (rng.standard_normal(s.thp.shape) * 0.5).astype(dtype)
    if moist:
        s.qv = (rng.uniform(0.0, 0.015, s.thp.shape)).astype(dtype)
        s.qc = np.zeros_like(s.qv)
        s.qr = np.zeros_like(s.qv)
    return case, s


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
@pytest.mark.parametrize("moist", [False, True])
def test_perturbation_eos_equals_plain_eos_in_fp64(name, moist):
    _, s = _state(name, np.float64, moist=moist)
    D.calc_p_alpha(s)
    p, al, alt = s.p.copy(), s.al.copy(), s.alt.copy()
    D.calc_p_alpha(s, prec=True)
    assert _rel(s.alt, alt) < 1e-13
    # al and p' are differences of large numbers in the plain form: compare
    # on the scale of the base values they perturb
    assert np.abs(s.al - al).max() < 1e-11 * np.abs(D._b3(s.alb)).max()
    assert np.abs(s.p - p).max() < 1e-10 * np.abs(D._b3(s.pb)).max()
    assert np.abs(s.pp - (p - D._b3(s.pb))).max() \
        < 1e-10 * np.abs(D._b3(s.pb)).max()


def test_perturbation_eos_at_rest_is_the_base_state():
    case = pins.CASES["dry_flat"]()
    s = D.from_case(case, np.float32)
    s.thp[...] = 0.0
    s.php[...] = 0.0
    s.mup[...] = 0.0
    D.calc_p_alpha(s, prec=True)
    assert np.all(s.al == 0.0)
    assert np.all(s.pp == 0.0)
    assert np.array_equal(s.p, np.broadcast_to(D._b3(s.pb), s.p.shape))
    assert np.array_equal(s.alt, np.broadcast_to(D._b3(s.alb), s.alt.shape))


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
def test_fp32_perturbation_eos_is_closer_to_fp64(name):
    _, s64 = _state(name, np.float64)
    _, s32 = _state(name, np.float32)
    # the same float32-representable inputs on both sides
    for f in ("thp", "php", "mup"):
        getattr(s64, f)[...] = getattr(s32, f).astype(np.float64)
    D.calc_p_alpha(s64)
    pp64 = s64.p - D._b3(s64.pb)
    D.calc_p_alpha(s32)
    e_plain = np.abs((s32.p.astype(np.float64) - D._b3(s32.pb)) - pp64).max()
    a_plain = np.abs(s32.al - s64.al).max()
    D.calc_p_alpha(s32, prec=True)
    e_pert = np.abs(s32.pp - pp64).max()
    a_pert = np.abs(s32.al - s64.al).max()
    assert e_pert < 0.1 * e_plain, (e_pert, e_plain)
    assert a_pert < 0.1 * a_plain, (a_pert, a_plain)


def test_inconsistent_base_breaks_agreement():
    """The perturbation EOS rests on alb being the base state's own
    hydrostatic alpha; an alb off that relation by 1e-6 (a plausible input
    mistake, e.g. a base rounded apart from its geopotential) shows up as a
    p difference far above the float64 agreement bar."""
    _, s = _state("terrain", np.float64)
    D.calc_p_alpha(s)
    p = s.p.copy()
    s.alb = s.alb * (1.0 + 1e-6)
    D.calc_p_alpha(s, prec=True)
    assert np.abs(s.p - p).max() > 1e-8 * np.abs(D._b3(s.pb)).max()


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
def test_prec_rhs_ph_equals_plain_in_fp64(name):
    case, s = _state(name, np.float64)
    nz = s.thp.shape[0]
    rng = np.random.default_rng(5)
    ww = rng.standard_normal((nz + 1,) + s.mup.shape) * 1e-2
    ww[0] = 0.0
    ww[nz] = 0.0
    s.rph_t[...] = 0.0
    D.slow_geopotential(s, ww, case.cfg.dx, case.cfg.dy)
    plain = s.rph_t.copy()
    s.rph_t[...] = 0.0
    D.slow_geopotential(s, ww, case.cfg.dx, case.cfg.dy, prec=True)
    assert np.abs(s.rph_t - plain).max() < 1e-9 * np.abs(plain).max()


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
def test_prec_init_and_finish_equal_plain_in_fp64(name):
    _, s = _state(name, np.float64)
    rng = np.random.default_rng(7)
    for f in D.PROGNOSTICS:
        x = getattr(s, f)
        getattr(s, f + "0")[...] = x + rng.standard_normal(x.shape) * (
            0.01 * (np.abs(x).max() + 1.0))
    D.calc_p_alpha(s)
    D.small_step_init(s)
    plain = {f: getattr(s, f).copy() for f in ("th_pp", "p_pp")}
    D.small_step_init(s, prec=True)
    for f, v in plain.items():
        assert np.abs(getattr(s, f) - v).max() < 1e-9 * np.abs(v).max(), f
    s.mu_pp[...] = rng.standard_normal(s.mu_pp.shape) * 5.0
    s2 = D.OState(**{k: (v.copy() if isinstance(v, np.ndarray) else v)
                     for k, v in vars(s).items()})
    D.small_step_finish(s)
    D.small_step_finish(s2, prec=True)
    assert np.abs(s2.thp - s.thp).max() < 1e-12 * np.abs(D._b3(s.thb)).max()
    for f in ("u", "v", "w", "php", "mup"):
        assert np.array_equal(getattr(s2, f), getattr(s, f)), f
