"""Closed forms and invariants that pin the oracle where the reference's
fixtures cannot (CPU only)."""
import numpy as np
import pytest

from conftest import rel_err
from oracle import dycore_np as O


def _tridiag(n, rng, dtype):
    a = np.zeros(n + 1, dtype)
    b = np.zeros(n + 1, dtype)
    c = np.zeros(n + 1, dtype)
    a[2:] = rng.uniform(-0.4, -0.1, n - 1)
    b[1:] = rng.uniform(1.5, 2.5, n)
    c[1:n] = rng.uniform(-0.4, -0.1, n - 1)
    return a, b, c


@pytest.mark.parametrize("n", [3, 16, 50])
def test_thomas_matches_dense_solve(n):
    rng = np.random.default_rng(n)
    a, b, c = _tridiag(n, rng, np.float64)
    rhs = rng.standard_normal(n + 1)
    rhs[0] = 0.37                     # boundary value coupled through a[1]
    a[1] = rng.uniform(-0.4, -0.1)
    m = np.zeros((n, n))
    for r in range(1, n + 1):
        m[r - 1, r - 1] = b[r]
        if r > 1:
            m[r - 1, r - 2] = a[r]
        if r < n:
            m[r - 1, r] = c[r]
    want = np.linalg.solve(m, rhs[1:] - np.r_[a[1] * rhs[0], np.zeros(n - 1)])
    alpha, gam = O.lu_factors(a, b, c)
    w = rhs.copy()
    O.thomas_solve(w, a, alpha, gam)
    np.testing.assert_allclose(w[1:], want, rtol=1e-12, atol=1e-12)
    assert w[0] == 0.37


def test_thomas_planted_index_mistake_is_caught():
    rng = np.random.default_rng(1)
    a, b, c = _tridiag(20, rng, np.float64)
    m = np.diag(b[1:]) + np.diag(a[2:], -1) + np.diag(c[1:20], 1)
    rhs = rng.standard_normal(21)
    rhs[0] = 0.0
    want = np.linalg.solve(m, rhs[1:])
    alpha = np.zeros_like(a)
    gam = np.zeros_like(a)
    for k in range(1, 20):                      # gam[k] instead of gam[k-1]
        alpha[k] = 1.0 / (b[k] - a[k] * gam[k])
        gam[k] = c[k] * alpha[k]
    alpha[20] = 1.0 / (b[20] - a[20] * gam[19])
    w = rhs.copy()
    O.thomas_solve(w, a, alpha, gam)
    assert not np.allclose(w[1:], want, rtol=1e-6)


def _rest_case(nx=64, nz=32, n_bv=0.01):
    """The reference's rest setup (tests/test_dycore_full.py:74-82)."""
    from arwen_amd.cases import host_at_rest, stratified_sounding
    from arwen_amd.config import RunConfig
    from arwen_amd.grid import make_base_state, make_vertical_coord
    cfg = RunConfig(nx=nx, ny=1, nz=nz, dx=1000.0, dy=1000.0, ztop=10000.0,
                    dt=6.0)
    vc = make_vertical_coord(nz)
    b = make_base_state(vc, stratified_sounding(n_bv), cfg.p_surf, cfg.ztop)
    return host_at_rest(cfg, vc, b)


def test_rest_state_stays_at_rest():
    """Discrete hydrostatic balance: a resting stratified atmosphere stays
    at rest to FP32 round-off (reference test_dycore_full.py:86-98 gates:
    w_max < 1e-3, |theta'| < 1e-2)."""
    case = _rest_case()
    s = O.from_case(case, np.float32)
    O.run_steps(s, case.cfg, 60)
    rep = O.stability_report(s, case.cfg)
    assert not rep["nan"]
    assert rep["w_max"] < 1e-3
    assert np.abs(s.thp).max() < 1e-2


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dry_mass_conserved(dtype):
    """Periodic domain: the column-mass update is a flux divergence whose
    domain sum telescopes, so sum(mu) is conserved to round-off."""
    import pins
    case = pins.build_dry_flat()
    s = O.from_case(case, dtype)
    m0 = O.domain_mass(s)
    O.run_steps(s, case.cfg, 5)
    m1 = O.domain_mass(s)
    npts = s.mup.size
    ulp_mu = np.finfo(dtype).eps * float(np.max(s.mub2d))
    # every column rounds at most a few ulps per substep
    assert abs(m1 - m0) <= 64 * 35 * npts * ulp_mu
    if dtype == np.float64:
        assert abs(m1 - m0) / m0 < 1e-12


def _isothermal_pulse(nx=400, nz=16, dx=100.0, T=300.0, amp=1.0, x0=20000.0,
                      hw=500.0, dtype=np.float64):
    """Isothermal atmosphere at rest with a Gaussian p'' pulse planted in
    (mu theta)'' (reference verify/npref.py:2971-3016)."""
    from arwen_amd.cases import host_at_rest
    from arwen_amd.config import RunConfig
    from arwen_amd.grid import make_base_state, make_vertical_coord
    from arwen_amd import constants as c
    cfg = RunConfig(nx=nx, ny=1, nz=nz, dx=dx, dy=dx, ztop=8000.0, dt=1.0)
    vc = make_vertical_coord(nz)
    b = make_base_state(vc, lambda z: T * np.exp(c.G * np.asarray(z, float)
                                                 / (c.CP * T)),
                        cfg.p_surf, cfg.ztop)
    s = O.from_case(host_at_rest(cfg, vc, b), dtype)
    O.calc_p_alpha(s)
    x = (np.arange(nx) + 0.5) * dx
    g = amp * np.exp(-((x - x0) / hw) ** 2)
    th_ref = O._b3(s.thb) + s.thp
    mut = s.mub2d + s.mup
    s.th_pp[...] = g[None, None, :] * mut[None] * th_ref / (c.GAMMA * s.p)
    coefs = O.calc_coefs(s, 0.05, cfg.epssm)
    O.diagnose_p(s, coefs["c2a"])
    s.p_pp_old[...] = s.p_pp
    return s, cfg


def test_sound_speed():
    """Acoustic-only stepping propagates a pulse at sqrt(gamma R T) within
    3 % (reference tests/test_acoustic.py:345-362)."""
    s, cfg = _isothermal_pulse()
    t = 40.0
    O.run_acoustic_only(s, cfg, 0.05, int(round(t / 0.05)))
    p = s.p_pp[s.p_pp.shape[0] // 2, 0, :]
    x = (np.arange(400) + 0.5) * 100.0
    xr = x[220 + int(np.argmax(p[220:]))]
    c_meas = (xr - 20000.0) / t
    c_theory = np.sqrt(1.4 * 287.0 * 300.0)
    assert abs(c_meas - c_theory) / c_theory < 0.03


def test_fp32_oracle_tracks_fp64():
    """Same discretisation, FP32 vs FP64: after one step the FP32 error is
    at the FP32-noise level of the p' = p - pb cancellation."""
    import pins
    case = pins.build_dry_flat()
    s32 = O.from_case(case, np.float32)
    s64 = O.from_case(case, np.float64)
    O.step(s32, case.cfg)
    O.step(s64, case.cfg)
    assert rel_err(s32.thp, s64.thp) < 1e-4
    assert rel_err(s32.u, s64.u) < 1e-4
    assert rel_err(s32.php, s64.php) < 1e-3
