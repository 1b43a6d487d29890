"""Pins of the oracle's moist scalar transport (CPU): the reference's
uniform-scalar invariance, exact coupled-mass conservation, the PD
limiter's positivity and its reduction to the unlimited flux form, the
upwind flux, the discretely balanced moist rest state, and planted
mistakes that each of these catches."""
import numpy as np
import pytest

from arwen_amd import cases
from arwen_amd.config import RunConfig
from arwen_amd.grid import make_base_state, make_vertical_coord
from oracle import dycore_np as D
from oracle import moist_np as M


def _moist_case(qv, thp_amp=5.0, nx=16, ny=12, nz=16, u0=0.0):
    cfg = RunConfig(nx=nx, ny=ny, nz=nz, dx=500.0, dy=500.0, ztop=8000.0,
                    dt=2.0, moist=True)
    vc = make_vertical_coord(nz)
    th = lambda z: 300.0 * np.exp(1e-4 * np.asarray(z, float) / 9.81)
    b = make_base_state(vc, th, p_surf=cfg.p_surf, ztop=cfg.ztop)
    z = cases.height_half(b)
    x = cases.cell_x(cfg)
    L = np.sqrt((x[None, None, :] / 2000.0) ** 2
                + ((z[:, None, None] - 2000.0) / 1500.0) ** 2)
    thp = np.where(L < 1.0, thp_amp * np.cos(np.pi * L / 2) ** 2, 0.0) \
        * np.ones((nz, ny, nx))
    qv3 = qv(z, cfg) if callable(qv) else np.full((nz, ny, nx), qv)
    case = cases.init_moist_balanced(cfg, vc, b, qv3, thp)
    case.fields["u"][...] = u0
    return case


def _coupled_mass(s):
    chm = s.c1h[:, None, None] * (s.mub2d + s.mup)[None] \
        + s.c2h[:, None, None]
    return float(np.sum((-s.dnw[:, None, None] * chm * s.qv)
                        .astype(np.float64)))


@pytest.mark.parametrize("dtype,tol", [(np.float32, 2e-5),
                                       (np.float64, 1e-10)])
def test_uniform_qv_stays_uniform(dtype, tol):
    """Reference tests/test_moist.py:517-577: uniform qv = 5 g/kg under a
    5 K bubble's divergent flow stays uniform (10 steps)."""
    case = _moist_case(0.005)
    s = D.from_case(case, dtype)
    D.run_steps(s, case.cfg, 10)
    assert float(np.abs(s.w).max()) > 0.5
    dev = float(np.abs(s.qv.astype(np.float64) - 0.005).max()) / 0.005
    assert dev < tol, dev


def test_planted_stage_flux_mistake_breaks_uniformity(monkeypatch):
    """Advecting with the stage fluxes instead of the substep-averaged
    ones (the reference's pre-fix defect) must break the invariance."""
    monkeypatch.setattr(D, "sumflux_finish",
                        lambda ru_m, rv_m, ww_m, ru, rv, ww, nsub:
                        (ru, rv, ww))
    case = _moist_case(0.005)
    s = D.from_case(case, np.float64)
    D.run_steps(s, case.cfg, 3)
    dev = float(np.abs(s.qv - 0.005).max()) / 0.005
    assert dev > 1e-6, dev


def _rough(z, cfg, seed=7):
    rng = np.random.default_rng(seed)
    q = rng.uniform(0.0, 0.016, (cfg.nz, cfg.ny, cfg.nx))
    q[rng.uniform(size=q.shape) < 0.3] = 0.0
    return q * np.exp(-z / 2500.0)[:, None, None]


def test_coupled_scalar_mass_conserved_fp64():
    case = _moist_case(_rough, u0=8.0)
    s = D.from_case(case, np.float64)
    m0 = _coupled_mass(s)
    D.run_steps(s, case.cfg, 3)
    assert abs(_coupled_mass(s) - m0) / m0 < 1e-12


def _pd_inputs(seed=3, wind=30.0):
    """Rough scalar, strong winds: the limiter is active."""
    rng = np.random.default_rng(seed)
    case = _moist_case(lambda z, cfg: _rough(z, cfg, seed))
    s = D.from_case(case, np.float64)
    nz, ny, nx = s.qv.shape
    ru = wind * (1.0 + rng.uniform(-0.5, 0.5, (nz, ny, nx + 1))) * 3e4
    ru[..., -1] = ru[..., 0]
    rv = wind * rng.uniform(-1.0, 1.0, (nz, ny + 1, nx)) * 3e4
    rv[:, -1] = rv[:, 0]
    ww = np.zeros((nz + 1, ny, nx))
    ww[1:nz] = rng.uniform(-1.0, 1.0, (nz - 1, ny, nx)) * 30.0
    mu = s.mub2d + s.mup
    return s, ru, rv, ww, mu


def _pd_update(s, ru, rv, ww, mu, dt=6.0, apply=M.pd_apply):
    fl = M.pd_fluxes(s.qv, s.qv, ru, rv, ww, mu, s.c1h, s.c2h, s.rdnw,
                     s.fnm, s.fnp, 500.0, 500.0, dt)
    sc = M.pd_scale(s.qv, mu, s.c1h, s.c2h, s.rdnw, fl, 1 / 500.0,
                    1 / 500.0, dt)
    tend = apply(sc, fl, s.rdnw, 1 / 500.0, 1 / 500.0)
    return M.coupled_update(s.qv, tend, s.c1h, s.c2h, mu, mu, dt,
                            clamp=False), sc


def test_pd_limiter_keeps_scalar_positive():
    s, ru, rv, ww, mu = _pd_inputs()
    q, sc = _pd_update(s, ru, rv, ww, mu)
    assert float(sc.min()) < 0.5                  # limiter is active
    assert float(q.min()) > -1e-12 * float(q.max())
    unl = M.coupled_update(
        s.qv, D.flux_div_scalar(s.qv, ru, rv, ww, s.rdnw, s.fnm, s.fnp,
                                1 / 500.0, 1 / 500.0),
        s.c1h, s.c2h, mu, mu, 6.0, clamp=False)
    assert float(unl.min()) < -1e-6 * float(unl.max())   # it had work


def _apply_receiver_scaled(sc, fl, rdnw, dx_inv, dy_inv):
    """pd_apply with one planted mistake: each x-face correction takes
    the scale of the cell it flows INTO instead of its donor."""
    fxl, fxc, fyl, fyc, fzl, fzc = fl
    zero, one = 0.0, 1.0
    fxc_l, fxc_r = fxc[..., :-1], fxc[..., 1:]
    sx_l = np.where(fxc_l > zero, sc,
                    np.where(fxc_l < zero, M._pm(sc, -1, 2), one))
    sx_r = np.where(fxc_r > zero, M._pm(sc, 1, 2),
                    np.where(fxc_r < zero, sc, one))
    rest = M.pd_apply(sc, (np.zeros_like(fxl), np.zeros_like(fxc), fyl,
                           fyc, fzl, fzc), rdnw, dx_inv, dy_inv)
    gx = (sx_r * fxc_r + fxl[..., 1:]) - (sx_l * fxc_l + fxl[..., :-1])
    return rest + (-dx_inv) * gx


def test_planted_wrong_donor_breaks_positivity():
    s, ru, rv, ww, mu = _pd_inputs()
    q, _ = _pd_update(s, ru, rv, ww, mu, apply=_apply_receiver_scaled)
    assert float(q.min()) < -1e-9 * float(q.max())


def test_pd_reduces_to_unlimited_form_without_limiting():
    """Smooth, strictly positive scalar: every scale is 1 and the PD
    divergence equals the unlimited flux form (pinned by the reference's
    advection fixture)."""
    s, ru, rv, ww, mu = _pd_inputs(wind=5.0)
    nz, ny, nx = s.qv.shape
    kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx),
                             indexing="ij")
    s.qv[...] = 0.01 * (1.0 + 0.2 * np.sin(2 * np.pi * ii / nx)
                        * np.cos(2 * np.pi * jj / ny)) * np.exp(-kk / 20.0)
    fl = M.pd_fluxes(s.qv, s.qv, ru, rv, ww, mu, s.c1h, s.c2h, s.rdnw,
                     s.fnm, s.fnp, 500.0, 500.0, 6.0)
    sc = M.pd_scale(s.qv, mu, s.c1h, s.c2h, s.rdnw, fl, 1 / 500.0,
                    1 / 500.0, 6.0)
    assert float(sc.min()) == 1.0
    pd = M.pd_apply(sc, fl, s.rdnw, 1 / 500.0, 1 / 500.0)
    un = D.flux_div_scalar(s.qv, ru, rv, ww, s.rdnw, s.fnm, s.fnp,
                           1 / 500.0, 1 / 500.0)
    assert np.allclose(pd, un, rtol=1e-10, atol=1e-12 * np.abs(un).max())


def test_upwind_flux_closed_form():
    """Below Courant 0.5 the PD low-order flux is the donor-cell flux
    vel * q_upwind."""
    rng = np.random.default_rng(1)
    qm1, q0 = rng.uniform(0, 1, 100), rng.uniform(0, 1, 100)
    cr = rng.uniform(-0.49, 0.49, 100)
    fu = M.flux_upwind(qm1, q0, cr)
    assert np.allclose(fu, np.where(cr > 0, cr * qm1, cr * q0), rtol=1e-14)


def test_moist_balanced_rest_state_has_no_w_forcing():
    """The balanced moist init inverts pg_buoy_w row by row, so the slow
    w forcing (vertical PGF + moist buoyancy) vanishes at rest at every
    interior w level.  At the lid the reference's builder uses the top
    layer's qv alone while pg_buoy_w averages the top two layers
    (moist.py:754-758 vs dycore.cu:258-268), so the lid balances exactly
    only where qv is uniform over those two layers."""
    case = _moist_case(lambda z, cfg: _rough(z, cfg), thp_amp=0.0)
    s = D.from_case(case, np.float64)
    D.calc_p_alpha(s)
    s.rw_t[...] = 0.0
    D.slow_buoyancy(s)
    scale = float(np.abs(s.mub2d).max()) * 9.81
    nz = s.qv.shape[0]
    assert float(np.abs(s.rw_t[1:nz]).max()) < 1e-9 * scale
    # column-uniform qv: the lid balances too
    case = _moist_case(lambda z, cfg: np.broadcast_to(
        _rough(z, cfg)[:1], (cfg.nz, cfg.ny, cfg.nx)), thp_amp=0.0)
    s = D.from_case(case, np.float64)
    D.calc_p_alpha(s)
    s.rw_t[...] = 0.0
    D.slow_buoyancy(s)
    assert float(np.abs(s.rw_t[1:]).max()) < 1e-9 * scale


def test_planted_missing_loading_breaks_moist_balance(monkeypatch):
    orig = D._q_total
    monkeypatch.setattr(D, "_q_total", lambda s: orig(s) * 0.0)
    case = _moist_case(lambda z, cfg: _rough(z, cfg), thp_amp=0.0)
    s = D.from_case(case, np.float64)
    D.calc_p_alpha(s)
    s.rw_t[...] = 0.0
    D.slow_buoyancy(s)
    scale = float(np.abs(s.mub2d).max()) * 9.81
    nz = s.qv.shape[0]
    assert float(np.abs(s.rw_t[1:nz]).max()) > 1e-6 * scale
