"""Pins of the oracle's moist_cq (WRF calc_cq and its three uses), after
the reference's own closed forms (tests/test_acoustic_npref.py:218-420):

* calc_cq sums exactly the active Registry ``moist`` masses: qv for
  mp_physics 0, qv/qc/qr/qi/qs/qg for Morrison (never its number
  moments), plus qh for NSSL (never its hail number);
* the horizontal acoustic pressure gradient, not the forcing, is scaled:
  u'' = u''_old + dtau (ru_t - cqu * dpxy);
* the implicit w rows and the w pressure forcing take cqw: the oracle's
  substep equals a dense solve of the cq-scaled tridiagonal system;
* with no moisture cq is exactly 1 and the step is bitwise the one
  without moist_cq.
"""
import numpy as np
import pytest

from arwen_amd import cases
from oracle import dycore_np as D
from tests.test_oracle_acoustic_pins import DTAU, EPSSM, G, _column


def _moist_ns(names_values, extra=(), n_load=0):
    s = D.OState(dtype=np.dtype(np.float64))
    for n, v in names_values.items():
        setattr(s, n, np.asarray(v, dtype=np.float64))
    s.thp = np.zeros_like(next(iter(names_values.values())))
    s.extra, s.n_load = tuple(extra), n_load
    return s


def _expected(qtot):
    nz = qtot.shape[0]
    eu = 1.0 / (1.0 + 0.5 * (qtot + np.roll(qtot, 1, axis=2)))
    ev = 1.0 / (1.0 + 0.5 * (qtot + np.roll(qtot, 1, axis=1)))
    ew = np.ones((nz + 1,) + qtot.shape[1:])
    ew[1:nz] = 1.0 / (1.0 + 0.5 * (qtot[1:] + qtot[:-1]))
    return (np.concatenate([eu, eu[:, :, :1]], axis=2),
            np.concatenate([ev, ev[:, :1, :]], axis=1), ew)


def _check(got, want):
    for g, w in zip(got, want):
        np.testing.assert_allclose(g, w, rtol=2e-16 * 8, atol=0)


def test_calc_cq_sums_the_active_moist_masses():
    base = np.arange(3 * 2 * 3, dtype=np.float64).reshape(3, 2, 3)
    names = ("qv", "qc", "qr", "qi", "qs", "qg")
    q = {n: (i + 1.0) * 1e-4 + base * (i + 1.0) * 1e-6
         for i, n in enumerate(names)}
    q["nr"] = np.full_like(base, 1e8)            # a number moment
    s = _moist_ns(q, extra=("qi", "qs", "qg", "nr"), n_load=3)
    _check(D.calc_cq(s, 10), _expected(sum(q[n] for n in names)))


def test_calc_cq_nssl_includes_hail_mass_not_hail_number():
    shape = (3, 2, 3)
    qv = np.linspace(0.001, 0.018, np.prod(shape)).reshape(shape)
    names = ("qv", "qc", "qr", "qi", "qs", "qg", "qh")
    q = {n: sc * qv for n, sc in zip(names, (1.0, 0.2, 0.1, 0.05, 0.03,
                                             0.02, 0.01))}
    q["qnh"] = np.full(shape, 1e8)
    s = _moist_ns(q, extra=("qi", "qs", "qg", "qh", "qnh"), n_load=4)
    _check(D.calc_cq(s, 18), _expected(sum(q[n] for n in names)))


def test_calc_cq_mp0_uses_vapour_only():
    qv = np.array([[[0.001, 0.002], [0.003, 0.004]],
                   [[0.005, 0.006], [0.007, 0.008]]])
    s = _moist_ns({"qv": qv, "qc": np.full_like(qv, 0.2),
                   "qr": np.full_like(qv, 0.1)})
    _check(D.calc_cq(s, 0), _expected(qv))
    assert not np.allclose(D.calc_cq(s, 1)[0], _expected(qv)[0])


def test_cq_scales_the_horizontal_acoustic_pgf_not_the_forcing():
    from tests.helpers import random_state_case
    case, rng = random_state_case(seed=4, nz=6, ny=5, nx=7)
    s = D.from_case(case, np.float64)
    D.calc_p_alpha(s)
    s.p_pp[...] = 30.0 * rng.standard_normal(s.p_pp.shape)
    s.p_pp_old[...] = s.p_pp
    s.ph_pp[...] = 5.0 * rng.standard_normal(s.ph_pp.shape)
    s.mu_pp[...] = 3.0 * rng.standard_normal(s.mu_pp.shape)
    s.ru_t[...] = rng.standard_normal(s.ru_t.shape)
    s.rv_t[...] = rng.standard_normal(s.rv_t.shape)
    u0, v0 = s.u_pp.copy(), s.v_pp.copy()
    qv = rng.uniform(0.0, 0.02, s.thp.shape)
    cq = D.calc_cq(_moist_ns({"qv": qv}), 0)
    s.cq = None
    D.advance_uv(s, DTAU, 0.1, case.cfg.dx, case.cfg.dy)
    u_off, v_off = s.u_pp.copy(), s.v_pp.copy()
    s.u_pp[...], s.v_pp[...] = u0, v0
    s.cq = cq
    D.advance_uv(s, DTAU, 0.1, case.cfg.dx, case.cfg.dy)
    du = s.ru_t - (u_off - u0) / DTAU
    dv = s.rv_t - (v_off - v0) / DTAU
    np.testing.assert_allclose(s.u_pp, u0 + DTAU * (s.ru_t - cq[0] * du),
                               rtol=0, atol=1e-12)
    np.testing.assert_allclose(s.v_pp, v0 + DTAU * (s.rv_t - cq[1] * dv),
                               rtol=0, atol=1e-12)
    assert np.abs(s.u_pp - u_off).max() > 1e-6


def test_cqw_scales_the_implicit_rows_and_the_w_pressure_forcing():
    """The reference's tests/test_acoustic_npref.py:354-420 column."""
    s = _column([90000.0, 70000.0, 50000.0, 30000.0])
    s.ph_pp[1:, 0, 0] = np.array([0.2, -0.1, 0.15, 0.0])
    qv = np.array([0.002, 0.008, 0.014, 0.020])[:, None, None]
    q = _moist_ns({"qv": qv, "qc": 0.25 * qv, "qr": 0.10 * qv})
    s.cq = D.calc_cq(q, 1)
    coefs = D.calc_coefs(s, DTAU, EPSSM, True)
    D.advance_w_phi(s, coefs, DTAU, EPSSM, 1000.0, 1000.0, True, 0.0, 5000.0)
    nz = 4
    mut = 90000.0
    c2a = 1.4 * s.p[:, 0, 0] / s.alt[:, 0, 0]
    cq = s.cq[2][:, 0, 0]
    cof = (0.5 * DTAU * G * (1.0 + EPSSM)) ** 2
    rdn, rdnw = s.rdn, s.rdnw
    rhs = np.array([0.0, 0.2, -0.1, 0.15, 0.0])
    m = np.zeros((nz, nz))
    forcing = np.zeros(nz)
    for k in range(1, nz):
        r = k - 1
        m[r, r] = 1.0 + cq[k] * cof * rdn[k] * (
            rdnw[k] * c2a[k] / mut ** 2 + rdnw[k - 1] * c2a[k - 1] / mut ** 2)
        if k > 1:
            m[r, r - 1] = -cq[k] * cof * rdn[k] * rdnw[k - 1] * c2a[k - 1] \
                / mut ** 2
        m[r, r + 1] = -cq[k] * cof * rdn[k] * rdnw[k] * c2a[k] / mut ** 2
        dph_up = 2.0 * (rhs[k + 1] - rhs[k])
        dph_dn = 2.0 * (rhs[k] - rhs[k - 1])
        forcing[r] = cq[k] * 0.5 * DTAU * G * rdn[k] * (
            c2a[k] * rdnw[k] / mut * dph_up
            - c2a[k - 1] * rdnw[k - 1] / mut * dph_dn)
    m[-1, -1] = 1.0 + 2.0 * cof * rdnw[-1] ** 2 * c2a[-1] / mut ** 2
    want = np.linalg.solve(m, forcing)
    want[-1] = 0.0
    np.testing.assert_allclose(s.w_pp[1:, 0, 0], want, rtol=2e-13,
                               atol=2e-13)
    np.testing.assert_allclose(
        s.ph_pp[1:, 0, 0],
        rhs[1:] + 0.5 * DTAU * G * (1.0 + EPSSM) * want / mut,
        rtol=2e-13, atol=2e-13)
    assert np.abs(cq[1:nz] - 1.0).min() > 1e-3


def test_no_moisture_makes_cq_exactly_one():
    case = cases.mixed_phase(6, nx=12, ny=10, nz=8)
    for name in ("qv", "qc", "qr", "qi", "qs", "qg"):
        case.fields[name][...] = 0.0
    off = D.from_case(case, np.float32)
    on = D.from_case(case, np.float32)
    D.step(off, case.cfg)
    D.step(on, case.cfg.with_(moist_cq=True))
    for n in ("u", "v", "w", "thp", "php", "mup", "p"):
        np.testing.assert_array_equal(getattr(on, n), getattr(off, n),
                                      err_msg=n)
    wet = cases.mixed_phase(6, nx=12, ny=10, nz=8)
    a = D.from_case(wet, np.float32)
    b = D.from_case(wet, np.float32)
    D.step(a, wet.cfg)
    D.step(b, wet.cfg.with_(moist_cq=True))
    assert np.abs(a.w - b.w).max() > 0.0
