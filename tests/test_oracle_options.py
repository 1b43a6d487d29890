"""Pins of two of the oracle's configuration options (CPU).

* h_sca_adv_order = 5 (reference dycore.cu:668-748): rhs_ph's horizontal
  geopotential advection with the centred (45 d1 - 9 d2 + d3) / 60
  difference approximates u dphi/dx to sixth order, against second order
  for the default; the bound rejects a wrong weight (-8 instead of -9).
* hypsometric_opt = 2 (reference diagnostics.cu:75-80): on an isothermal
  column whose geopotential follows the hypsometric equation exactly,
  phi = phi_s + Rd T ln(p_s / p) over the dry hybrid pressures, the EOS's
  specific volume is the ideal-gas Rd T / p at the layer's middle
  pressure to round-off; the hypsometric_opt = 1 form is only a
  finite-difference approximation of it.
"""
import numpy as np

from oracle import dycore_np as D
from tests.helpers import random_state_case


def _geopotential_mode_error(nx, order):
    """max |rhs_ph horizontal term - exact| for phi' = A sin(kx), uniform
    u, v = 0, w = 0, Omega = 0, on a flat base (float64)."""
    case, _ = random_state_case(seed=1, nz=6, ny=4, nx=nx)
    case.fields["u"][...] = 12.0
    case.fields["v"][...] = 0.0
    case.fields["w"][...] = 0.0
    case.fields["mup"][...] = 0.0
    s = D.from_case(case, np.float64)
    dx = case.cfg.dx
    L = nx * dx
    x = np.arange(nx) * dx
    amp = 50.0
    s.php[...] = amp * np.sin(2.0 * np.pi * x / L)[None, None, :]
    s.rph_t[...] = 0.0
    ww = np.zeros_like(s.w)
    D.slow_geopotential(s, ww, dx, case.cfg.dy, False, order)
    # with w = 0 and Omega = 0 only the horizontal term is left
    mut = s.mub2d + s.mup
    nz = s.thp.shape[0]
    k = 3                                # an interior w level
    fc = (s.c1f[k] * mut + s.c2f[k]) * (2.0 * 12.0)     # coupled u at faces
    exact = -0.25 / dx * 2.0 * fc * dx * (
        amp * 2.0 * np.pi / L * np.cos(2.0 * np.pi * x / L))[None, :]
    got = s.rph_t[k]
    assert nz > k
    return float(np.abs(got - exact).max() / np.abs(exact).max())


def test_order_five_is_sixth_order_accurate():
    e1 = _geopotential_mode_error(16, 5)
    e2 = _geopotential_mode_error(32, 5)
    assert e1 < 1e-4
    assert e1 / e2 > 50.0                # ~2^6
    f1 = _geopotential_mode_error(16, 2)
    f2 = _geopotential_mode_error(32, 2)
    assert 3.0 < f1 / f2 < 5.0           # the default: second order
    assert e1 < 0.01 * f1


def test_a_wrong_weight_would_break_the_bound():
    """The bound above is sensitive to the weights: the same centred
    difference with -8 in place of -9 is inconsistent (its weights sum to
    64/60) and misses it by orders of magnitude."""
    nx = 32
    x = np.arange(nx)
    p = np.sin(2.0 * np.pi * x / nx)
    d = [np.roll(p, -o) - np.roll(p, o) for o in (1, 2, 3)]
    exact = 2.0 * np.pi / nx * np.cos(2.0 * np.pi * x / nx)
    good = ((45.0 * d[0] - 9.0 * d[1]) + d[2]) / 60.0
    bad = ((45.0 * d[0] - 8.0 * d[1]) + d[2]) / 60.0
    assert np.abs(good - exact).max() < 1e-6
    assert np.abs(bad - exact).max() > 1e-2
    assert _geopotential_mode_error(nx, 5) < 1e-6


def _hypsometric_state(T0=250.0):
    case, _ = random_state_case(seed=2, nz=12, ny=5, nx=7)
    case.fields["mup"] = 0.0 * case.fields["mup"] + 150.0
    s = D.from_case(case, np.float64)
    mu = s.mub2d + s.mup
    pf = s.c3f[:, None, None] * mu[None] + s.c4f[:, None, None] + s.p_top
    phi = D.RD * T0 * np.log(pf[0][None] / pf)          # phi_s = 0
    s.php[...] = phi - D._b3(s.phb)
    return s, mu, T0


def test_hypsometric_two_is_exact_on_an_isothermal_column():
    s, mu, T0 = _hypsometric_state()
    D.calc_p_alpha(s, False, 2)
    phm = s.c3h[:, None, None] * mu[None] + s.c4h[:, None, None] + s.p_top
    np.testing.assert_allclose(s.alt, D.RD * T0 / phm, rtol=1e-13)
    np.testing.assert_allclose(s.al, s.alt - D._b3(s.alb), rtol=0,
                               atol=1e-12 * np.abs(s.alt).max())
    # the default form is only an approximation of it
    s1, _, _ = _hypsometric_state()
    D.calc_p_alpha(s1, False, 1)
    assert np.abs(s1.alt / s.alt - 1.0).max() > 1e-6


def test_planted_pressure_shift_is_caught():
    s, mu, T0 = _hypsometric_state()
    c3f, c4f = s.c3f.copy(), s.c4f.copy()
    # evaluating the layer pressures one level off breaks the identity
    s.c3f = np.concatenate([c3f[1:], c3f[-1:]])
    s.c4f = np.concatenate([c4f[1:], c4f[-1:]])
    with np.errstate(divide="ignore"):
        D.calc_p_alpha(s, False, 2)
    phm = s.c3h[:, None, None] * mu[None] + s.c4h[:, None, None] + s.p_top
    assert np.abs(s.alt[:-1] / (D.RD * T0 / phm[:-1]) - 1.0).max() > 1e-3
