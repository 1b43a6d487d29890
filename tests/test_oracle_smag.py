"""Pins of the oracle's km_opt = 4 (2-D Smagorinsky) package (CPU, float64
unless stated), from the reference's own gates (tests/test_smag2d.py,
tests/test_smag2d_wrf_host.py) and what the mathematics fixes:

* the slope-corrected deformation: horizontally uniform u rising one unit
  per level under zx = 0.5 gives D11 = -1 and K = 0.03125 (the
  reference's host gate, test_smag2d_wrf_host.py:14-34), and the same
  construction pins D22, D12's two slope terms and D13;
* the analytic pure-shear K on a periodic sinusoid and the zero K of a
  solid-body rotation (test_smag2d.py:231-300), the 10 l cap and
  K_h = K_m / Pr;
* the metric terms of the u stress divergence and of the scalar flux
  cancel the coordinate-surface differences of fields that depend on
  height only, to round-off;
* flux form: on a flat periodic grid every horizontal divergence sums to
  zero over the domain and every vertical one over a column
  (test_smag2d_wrf_host.py:162-186);
* the flat cross-component u stress reduces to the divergence of
  -K dv/dx (test_smag2d_wrf_host.py:37-61);
* t_2 = theta - 300 carries the base-theta gradient through the slope
  correction (test_smag2d_wrf_host.py:108-138);
* the isfflx surface policy with a prescribed drag and heat flux
  (test_smag2d.py:585-647, the constant-coefficient arms);
* the step: from rest K = 0 and the step equals the km_opt = 1 one bit
  for bit, with winds they differ (test_smag2d.py:663-723).
"""
import numpy as np
import pytest

from oracle import dycore_np as D
from oracle import smag_np as S

G = D.G


def _state(nz, ny, nx, ph, dtype=np.float64):
    """A synthetic oracle state on the geopotential ``ph`` (nz+1, ny, nx):
    uniform eta levels (dnw = -1/nz, fnm = fnp = 1/2) and the uniform-
    spacing surface extrapolation (15, -10, 3)/8."""
    s = D.OState(dtype=np.dtype(dtype))
    s.phb = np.asarray(ph, dtype=dtype)
    s.dnw = np.full(nz, -1.0 / nz, dtype=dtype)
    s.dn = np.r_[0.0, np.full(nz - 1, -1.0 / nz)].astype(dtype)
    s.fnm = np.r_[0.0, np.full(nz - 1, 0.5)].astype(dtype)
    s.fnp = s.fnm.copy()
    s.cf1, s.cf2, s.cf3 = (np.dtype(dtype).type(c)
                           for c in (15 / 8, -10 / 8, 3 / 8))
    return s


def _sloped(nz, ny, nx, dx=1.0, dy=1.0, bx=0.5, by=0.0):
    """phi = G (k + bx x + by y): rdzw = 1, zx = bx, zy = by (away from the
    periodic seam of the linear slope)."""
    k = np.arange(nz + 1)[:, None, None]
    x = (np.arange(nx) * dx)[None, None, :]
    y = (np.arange(ny) * dy)[None, :, None]
    return G * (k + bx * x + by * y) * np.ones((1, ny, nx))


def _grid(s, dx=1.0, dy=1.0, alt=None):
    nz, ny, nx = s.phb.shape[0] - 1, s.phb.shape[1], s.phb.shape[2]
    alt = np.ones((nz, ny, nx)) if alt is None else alt
    return S.Grid(s, dx, dy, np.zeros_like(s.phb), alt)


def _levels_u(nz, ny, nx):
    return np.broadcast_to(np.arange(nz, dtype=float)[:, None, None],
                           (nz, ny, nx + 1)).copy()


INT = (slice(1, -2), slice(2, -3), slice(2, -3))   # interior mass points


def _eq(a, want):
    """Equal to round-off of the metric arithmetic."""
    np.testing.assert_allclose(a, want, rtol=0, atol=1e-13)


def test_d11_slope_correction_and_k():
    nz, ny, nx = 6, 8, 10
    s = _state(nz, ny, nx, _sloped(nz, ny, nx))
    gr = _grid(s)
    u = _levels_u(nz, ny, nx)
    v = np.zeros((nz, ny + 1, nx))
    d11, d22, d12 = S.deformation(gr, u, v)
    _eq(d11[INT], -1.0)
    _eq(d22[INT], 0.0)
    _eq(d12[INT], 0.0)
    km, kh = S.smag_km(gr, d11, d22, d12, 0.25)
    _eq(km[INT], 0.03125)
    np.testing.assert_allclose(kh[INT], 3.0 * 0.03125, rtol=1e-15)
    # the same-level (flat) reduction sees no horizontal difference
    flat = _grid(_state(nz, ny, nx, _sloped(nz, ny, nx, bx=0.0)))
    f11, _, _ = S.deformation(flat, u, v)
    _eq(f11[INT], 0.0)


def test_d22_and_both_d12_slope_terms():
    nz, ny, nx = 6, 10, 8
    s = _state(nz, ny, nx, _sloped(nz, ny, nx, bx=0.0, by=0.5))
    gr = _grid(s)
    v = np.broadcast_to(np.arange(nz, dtype=float)[:, None, None],
                        (nz, ny + 1, nx)).copy()
    u = np.zeros((nz, ny, nx + 1))
    d11, d22, d12 = S.deformation(gr, u, v)
    _eq(d22[INT], -1.0)
    _eq(d11[INT], 0.0)
    _eq(d12[INT], 0.0)      # zx = 0: no v slope
    # u rising with the level under zy = 0.5: the u slope of D12
    _, _, d12 = S.deformation(gr, _levels_u(nz, ny, nx),
                              np.zeros((nz, ny + 1, nx)))
    _eq(d12[INT], -0.5)
    # v rising with the level under zx = 0.5: the v slope of D12
    gx = _grid(_state(nz, ny, nx, _sloped(nz, ny, nx)))
    _, _, d12 = S.deformation(gx, np.zeros((nz, ny, nx + 1)), v)
    _eq(d12[INT], -0.5)


def test_d13_slope_term_through_tau13():
    """w rising one unit per w level under zx = 0.5 (rdz = 1): D13 = -0.5,
    so with K = 2 and rho = 1, tau13 = -rho K D13 = 1 at interior levels."""
    nz, ny, nx = 6, 8, 10
    gr = _grid(_state(nz, ny, nx, _sloped(nz, ny, nx)))
    w = np.broadcast_to(np.arange(nz + 1, dtype=float)[:, None, None],
                        (nz + 1, ny, nx)).copy()
    u = np.zeros((nz, ny, nx + 1))
    v = np.zeros((nz, ny + 1, nx))
    z = np.zeros((nz, ny, nx))
    km = np.full((nz, ny, nx), 2.0)
    tau = S.stresses(gr, km, z, z, z, u, v, w)
    t13, t23 = tau[3], tau[4]
    _eq(t13[1:-1, 2:-3, 2:-3], 1.0)
    np.testing.assert_array_equal(t13[0], 0.0)
    np.testing.assert_array_equal(t13[-1], 0.0)
    _eq(t23[1:-1, 2:-3, 2:-3], 0.0)   # zy = 0
    # tau33 = -rho K 2 dw rdzw = -4
    _eq(tau[5], -4.0)


def test_pure_shear_analytic_k():
    nz, ny, nx = 4, 16, 16
    dx = dy = 1000.0
    s = _state(nz, ny, nx, _sloped(nz, ny, nx, dx, dy, bx=0.0))
    gr = _grid(s, dx, dy)
    yu = (np.arange(ny) + 0.5) * dy
    prof = 8.0 * np.sin(2.0 * np.pi * yu / (ny * dy))
    u = np.broadcast_to(prof[None, :, None], (nz, ny, nx + 1)).copy()
    v = np.zeros((nz, ny + 1, nx))
    km, kh = S.smag_km(gr, *S.deformation(gr, u, v), 0.25)
    corner = (prof - np.roll(prof, 1)) / dy
    shear = 0.5 * (corner + np.roll(corner, -1))
    want = np.broadcast_to((0.0625 * dx * dy * np.abs(shear))[None, :, None],
                           km.shape)
    np.testing.assert_allclose(km, want, rtol=1e-12, atol=1e-12 * want.max())
    np.testing.assert_allclose(kh, 3.0 * want, rtol=1e-12,
                               atol=3e-12 * want.max())


def test_solid_body_rotation_has_zero_k():
    nz, ny, nx = 4, 16, 16
    dx = dy = 1000.0
    gr = _grid(_state(nz, ny, nx, _sloped(nz, ny, nx, dx, dy, bx=0.0)),
               dx, dy)
    om = 2.0 ** -13
    yu = (np.arange(ny) + 0.5) * dy
    xv = (np.arange(nx) + 0.5) * dx
    u = np.broadcast_to((-om * yu)[None, :, None], (nz, ny, nx + 1)).copy()
    v = np.broadcast_to((om * xv)[None, None, :], (nz, ny + 1, nx)).copy()
    km, _ = S.smag_km(gr, *S.deformation(gr, u, v), 0.25)
    assert np.abs(km[:, 2:-2, 2:-2]).max() < 1e-8 * 0.0625 * dx * dy * om


def test_cap_and_prandtl_on_random_winds():
    nz, ny, nx = 6, 8, 12
    dx, dy = 100.0, 200.0
    gr = _grid(_state(nz, ny, nx, _sloped(nz, ny, nx, dx, dy, bx=0.0)),
               dx, dy)
    rng = np.random.default_rng(1)
    u = 500.0 * rng.standard_normal((nz, ny, nx + 1))
    v = 500.0 * rng.standard_normal((nz, ny + 1, nx))
    km, kh = S.smag_km(gr, *S.deformation(gr, u, v), 0.25)
    cap = 10.0 * np.sqrt(dx * dy)
    assert (km == cap).any() and (km < cap).any() and km.max() <= cap
    np.testing.assert_array_equal(kh, km / (1.0 / 3.0))


def test_u_stress_metric_term_cancels_height_only_stress():
    """tau11 = z (height only) on a sloping grid: its divergence along
    constant height vanishes at the interior levels; the coordinate-surface
    difference alone does not."""
    nz, ny, nx = 7, 6, 12
    dx = dy = 1.0
    a, b = 1.0, 0.3
    ph = G * (a * np.arange(nz + 1)[:, None, None]
              + b * (np.arange(nx) * dx)[None, None, :]) * np.ones((1, ny, 1))
    s = _state(nz, ny, nx, ph)
    gr = _grid(s)
    zm = 0.5 * (ph[1:] + ph[:-1]) / G
    z0 = np.zeros((nz, ny, nx))
    tau = (zm, z0, z0, np.zeros((nz + 1, ny, nx)),
           np.zeros((nz + 1, ny, nx)), z0)
    ru, rv, rw = S.momentum_tendencies(gr, tau, False, None, None, None)
    inner = ru[1:nz - 1, :, 2:nx - 2]
    assert np.abs(inner).max() < 1e-12
    # the same without the metric correction is the raw slope
    flat = _grid(_state(nz, ny, nx, _sloped(nz, ny, nx, bx=0.0)))
    ru0, _, _ = S.momentum_tendencies(flat, tau, False, None, None, None)
    assert np.abs(ru0[1:nz - 1, :, 2:nx - 2]).min() > 1e-3


def test_scalar_flux_vanishes_for_height_only_fields():
    nz, ny, nx = 7, 10, 10
    a, bx, by = 1.0, 0.3, -0.2
    ph = G * (a * np.arange(nz + 1)[:, None, None]
              + bx * np.arange(nx)[None, None, :]
              + by * np.arange(ny)[None, :, None])
    gr = _grid(_state(nz, ny, nx, ph))
    zm = 0.5 * (ph[1:] + ph[:-1]) / G
    kh = np.full((nz, ny, nx), 5.0)
    fx, fy = S.scalar_fluxes(gr, kh, 0.01 * zm)
    assert np.abs(fx[1:nz - 1, 2:-2, 2:-2]).max() < 1e-13
    assert np.abs(fy[1:nz - 1, 2:-2, 2:-2]).max() < 1e-13
    # a wrong sign on the slope term leaves twice the raw gradient
    raw = 0.01 * bx * 5.0
    fz, _ = S.scalar_fluxes(_grid(_state(nz, ny, nx, G * np.arange(
        nz + 1)[:, None, None] * np.ones((1, ny, nx)))), kh, 0.01 * zm)
    np.testing.assert_allclose(fz[1:nz - 1, 2:-2, 2:-2], -raw, rtol=1e-12)


def _flat_random(seed, nz=5, ny=6, nx=8):
    rng = np.random.default_rng(seed)
    ph = G * 500.0 * np.arange(nz + 1)[:, None, None] * np.ones((1, ny, nx))
    s = _state(nz, ny, nx, ph)
    gr = S.Grid(s, 900.0, 1100.0, np.zeros_like(ph),
                rng.uniform(0.8, 1.25, (nz, ny, nx)))
    u = rng.standard_normal((nz, ny, nx + 1))
    v = rng.standard_normal((nz, ny + 1, nx))
    w = rng.standard_normal((nz + 1, ny, nx))
    u[..., -1] = u[..., 0]
    v[:, -1] = v[:, 0]
    return gr, u, v, w, rng


def test_flat_periodic_divergences_telescope():
    gr, u, v, w, rng = _flat_random(812)
    d = S.deformation(gr, u, v)
    km, kh = S.smag_km(gr, *d, 0.25)
    tau = S.stresses(gr, km, *d, u, v, w)
    ru, rv, rw = S.momentum_tendencies(gr, tau, False, None, u, v)
    dnw = gr.dnw[:, None, None]
    for r in (ru[..., :-1], rv[:, :-1]):
        scale = np.abs(r).sum()
        assert abs((dnw * r).sum()) < 1e-13 * scale
    assert abs(rw.sum()) < 1e-13 * np.abs(rw).sum()
    # vertical divergences telescope per column (tau13 = 0 at both ends)
    ru2, rv2, rw2 = S.momentum_tendencies(gr, tau, True, None, u, v)
    col = ((ru2 - ru) * dnw).sum(axis=0)
    assert np.abs(col).max() < 1e-13 * np.abs(ru2 - ru).max()
    # w: the interior levels leave the end stresses tau33(nz-1) - tau33(0)
    dn = gr.dn[1:, None, None]
    colw = ((rw2 - rw)[1:-1] * dn).sum(axis=0) / G
    t33 = tau[5]
    np.testing.assert_allclose(colw, t33[-1] - t33[0], rtol=1e-12,
                               atol=1e-12 * np.abs(t33).max())
    thb = np.full(5, 300.0)
    thp = 0.02 * rng.standard_normal((5, 6, 8))
    rth = S.theta_tendency(gr, kh, thp, thb, None)
    assert abs((dnw * rth).sum()) < 1e-13 * np.abs(dnw * rth).sum()


def test_flat_cross_component_u_stress():
    """u = 0, v = sin x sin y, non-uniform K, rho = 1, flat: the u
    tendency is g dz/dnw d/dy(-K_corner dv/dx) with the corner K the
    four-cell mean (WRF horizontal_diffusion_u_2's cross term)."""
    nz, ny, nx = 3, 32, 32
    dx, dy = 2.0 * np.pi / nx, 2.0 * np.pi / ny
    ph = G * np.arange(nz + 1)[:, None, None] * np.ones((1, ny, nx))
    gr = _grid(_state(nz, ny, nx, ph), dx, dy)
    x = (np.arange(nx) + 0.5) * dx
    yf = np.arange(ny + 1) * dy
    v = np.broadcast_to(np.sin(yf)[:, None] * np.sin(x)[None, :],
                        (nz, ny + 1, nx)).copy()
    u = np.zeros((nz, ny, nx + 1))
    w = np.zeros((nz + 1, ny, nx))
    km = np.broadcast_to(1.0 + 0.1 * np.cos(2.0 * np.pi * np.arange(nx)
                                            / nx), (nz, ny, nx)).copy()
    d = S.deformation(gr, u, v)
    tau = S.stresses(gr, km, *d, u, v, w)
    ru, _, _ = S.momentum_tendencies(gr, tau, False, None, u, v)
    want = np.zeros((nz, ny, nx))
    for j in range(ny):
        for i in range(nx):
            def t12(jj):
                jj %= ny
                kc = 0.25 * (km[0, jj, i % nx] + km[0, jj, (i - 1) % nx]
                             + km[0, (jj - 1) % ny, (i - 1) % nx]
                             + km[0, (jj - 1) % ny, i % nx])
                dvdx = (v[0, jj, i % nx] - v[0, jj, (i - 1) % nx]) / dx
                return -kc * dvdx
            want[:, j, i] = G * 1.0 / gr.dnw[0] * (t12(j + 1) - t12(j)) / dy
    np.testing.assert_allclose(ru[..., :-1], want, rtol=1e-12,
                               atol=1e-12 * np.abs(want).max())
    assert np.abs(want).max() > 1e-3


def test_full_theta_carries_the_base_gradient_on_sloping_levels():
    nz, ny, nx = 4, 6, 8
    x, y = np.arange(nx), np.arange(ny)
    terrain = 80.0 * np.sin(2.0 * np.pi * x / nx)[None, :] \
        + 50.0 * np.cos(2.0 * np.pi * y / ny)[:, None]
    sig = 1.0 - np.arange(nz + 1) / nz
    ph = G * (1000.0 * np.arange(nz + 1)[:, None, None]
              + sig[:, None, None] * terrain[None])
    gr = _grid(_state(nz, ny, nx, ph), 1000.0, 1200.0)
    kh = np.full((nz, ny, nx), 30.0)
    thp = np.zeros((nz, ny, nx))
    full = S.theta_tendency(gr, kh, thp,
                            np.array([300.0, 301.0, 304.0, 309.0]), None)
    pert = S.theta_tendency(gr, kh, thp, np.full(nz, 300.0), None)
    np.testing.assert_array_equal(pert, 0.0)
    assert np.abs(full).max() > 0.0


@pytest.mark.parametrize("isfflx", [0, 1, 2])
def test_surface_policy(isfflx):
    nz, ny, nx = 4, 5, 6
    ph = G * 700.0 * np.arange(nz + 1)[:, None, None] * np.ones((1, ny, nx))
    s = _state(nz, ny, nx, ph)
    alt = np.full((nz, ny, nx), 0.9)
    gr = S.Grid(s, 1000.0, 1000.0, np.zeros_like(ph), alt)
    u = np.full((nz, ny, nx + 1), 4.0)
    v = np.full((nz, ny + 1, nx), 3.0)
    w = np.zeros((nz + 1, ny, nx))
    d = S.deformation(gr, u, v)
    km, kh = S.smag_km(gr, *d, 0.25)
    np.testing.assert_array_equal(km, 0.0)
    cd0, heat = 0.01, 0.2
    tau = S.stresses(gr, km, *d, u, v, w)
    ru, rv, rw = S.momentum_tendencies(gr, tau, True,
                                       cd0 if isfflx == 0 else None, u, v)
    rth = S.theta_tendency(gr, kh, np.zeros((nz, ny, nx)), np.full(nz, 300.),
                           heat if isfflx in (0, 2) else None)
    rho, dnw0 = 1.0 / 0.9, s.dnw[0]
    if isfflx == 0:
        np.testing.assert_allclose(ru[0], G * 0.01 * 5.0 * 4.0 * rho / dnw0,
                                   rtol=1e-14)
        np.testing.assert_allclose(rv[0], G * 0.01 * 5.0 * 3.0 * rho / dnw0,
                                   rtol=1e-14)
    else:
        np.testing.assert_array_equal(ru[0], 0.0)
        np.testing.assert_array_equal(rv[0], 0.0)
    if isfflx in (0, 2):
        np.testing.assert_allclose(rth[0], -G * 0.2 * rho / dnw0, rtol=1e-14)
    else:
        np.testing.assert_array_equal(rth[0], 0.0)
    for r in (ru[1:], rv[1:], rw, rth[1:]):
        np.testing.assert_array_equal(r, 0.0)


def test_step_from_rest_equals_constant_k_then_diverges():
    from arwen_amd import cases
    from arwen_amd.config import RunConfig
    from arwen_amd.grid import make_base_state, make_vertical_coord
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
    s1 = D.from_case(case, np.float32)
    s4 = D.from_case(case, np.float32)
    c4 = cfg.with_(km_opt=4)
    D.step(s1, cfg)
    D.step(s4, c4)
    for n in ("u", "v", "w", "thp", "php", "mup"):
        np.testing.assert_array_equal(getattr(s1, n), getattr(s4, n))
    for _ in range(3):
        D.step(s1, cfg)
        D.step(s4, c4)
    assert np.abs(s1.u - s4.u).max() > 0.0
    assert np.abs(s1.thp - s4.thp).max() > 0.0
    assert np.isfinite(s4.thp).all()


@pytest.mark.parametrize("rise,want", [(1.0, 0.0625 * 3.0 / 3.0),
                                       (10.0, 0.0625 * 30.0 / 9.0)])
def test_slope_factor_alpha_and_its_square_beyond_the_limit(rise, want):
    """zx = 3 on a unit grid: alpha = sqrt(sx^2 + sy^2) = 3.  u rising
    ``rise`` per level gives D11 = -6 rise and |S| = 3 rise: K is divided
    by alpha below max(10 / l, 1e-3) = 10 and by alpha^2 above it."""
    nz, ny, nx = 6, 8, 10
    gr = _grid(_state(nz, ny, nx, _sloped(nz, ny, nx, bx=3.0)))
    u = rise * _levels_u(nz, ny, nx)
    km, _ = S.smag_km(gr, *S.deformation(gr, u, np.zeros((nz, ny + 1, nx))),
                      0.25)
    np.testing.assert_allclose(km[INT], want, rtol=1e-13)


# --- moist states: the held species tendencies ----------------------------

def _sheared_moist(mp=10, nx=20, ny=14, nz=12, **opts):
    from arwen_amd import cases
    case = cases.mixed_phase(mp, nx=nx, ny=ny, nz=nz) if mp \
        else cases.moist_tracer(nx=nx, ny=ny, nz=nz)
    y = (np.arange(ny) + 0.5) / ny
    x = (np.arange(nx) + 0.5) / nx
    case.fields["u"] = case.fields["u"] + 3.0 * np.sin(2 * np.pi * y)[
        None, :, None]
    case.fields["v"] = case.fields["v"] + 2.0 * np.cos(2 * np.pi * x)[
        None, None, :]
    case.cfg = case.cfg.with_(**opts)
    return case


def test_vapour_loads_the_smagorinsky_density():
    """rho = (1 + qv) / alt: a uniform qv = c scales every momentum stress
    by exactly (1 + c) against the dry density (smag2d.cu:394-404)."""
    nz, ny, nx = 5, 6, 8
    gr_d, u, v, w, rng = _flat_random(5, nz, ny, nx)
    c = 0.0125
    s = _state(nz, ny, nx, gr_d.ph)
    alt = 1.0 / gr_d.rho
    gr_m = S.Grid(s, 900.0, 1100.0, np.zeros_like(gr_d.ph), alt,
                  np.full((nz, ny, nx), c))
    out = []
    for gr in (gr_d, gr_m):
        d = S.deformation(gr, u, v)
        km, _ = S.smag_km(gr, *d, 0.25)
        tau = S.stresses(gr, km, *d, u, v, w)
        out.append(S.momentum_tendencies(gr, tau, True, None, u, v))
    for a, b in zip(*out):
        np.testing.assert_allclose(b, (1.0 + c) * a, rtol=1e-12,
                                   atol=1e-12 * np.abs(a).max())


def test_species_rows_conserve_and_mix():
    """Every species takes K_h's flux divergence of its own mixing ratio:
    on the flat sheared moist case each row's coupled sum is zero and the
    rows are live."""
    from oracle import dycore_np as D
    case = _sheared_moist(10, km_opt=4)
    s = D.from_case(case, np.float64)
    s.php[...] = 0.0
    for n in D.PROGNOSTICS:
        getattr(s, n + "0")[...] = getattr(s, n)
    for q in ("qv", "qc", "qr") + s.extra:
        getattr(s, q + "0")[...] = getattr(s, q)
    D.calc_p_alpha(s)
    held = S.smag_held(s, case.cfg)
    dnw = s.dnw[:, None, None]
    assert set(held["q"]) == {"qv", "qc", "qr"} | set(s.extra)
    for n, h in held["q"].items():
        rdzw = D.G / (s.phb[1:] - s.phb[:-1])
        w = dnw / rdzw[:, None, None] if rdzw.ndim == 1 else dnw / rdzw
        assert abs((h * w).sum()) < 1e-12 * np.abs(h * w).sum(), n
        assert np.abs(h).max() > 0.0, n


@pytest.mark.parametrize("pd", [True, False])
def test_held_species_tendency_at_rest(pd):
    """With no advecting flux a stage moves q by dt fixed / C(mu): folded
    into q_t ahead of the limiter, or added to the unlimited update
    (reference moist.py:293-346, :384-386); leaving the fold out fails."""
    from oracle import dycore_np as D
    from oracle import moist_np as M
    case = _sheared_moist(6)
    s = D.from_case(case, np.float64)
    for q in ("qv", "qc", "qr") + s.extra:
        getattr(s, q + "0")[...] = getattr(s, q)
    s.mup0[...] = s.mup
    z = (np.zeros_like(s.u), np.zeros_like(s.v), np.zeros_like(s.w))
    rng = np.random.default_rng(3)
    fixed = {n: 1e-3 * rng.standard_normal(s.qv.shape)
             for n in ("qv", "qc", "qr") + s.extra}
    dt = 18.0
    M.advance_scalars_stage(s, *z, dt, True, case.cfg.dx, case.cfg.dy,
                            1 if pd else 0, fixed)
    chm = s.c1h[:, None, None] * (s.mub2d + s.mup)[None] \
        + s.c2h[:, None, None]
    for n, fx in fixed.items():
        want = np.maximum(getattr(s, n + "0") + dt * fx / chm, 0.0)
        np.testing.assert_allclose(getattr(s, n), want, rtol=1e-13,
                                   atol=1e-18)
        assert np.abs(getattr(s, n) - getattr(s, n + "0")).max() > 1e-9


def test_moist_mix6_off_exempts_the_moist_array_only():
    from oracle import dycore_np as D
    case = _sheared_moist(10, diff_6th_opt=1, moist_mix6_off=True)
    s = D.from_case(case, np.float32)
    for n in D.PROGNOSTICS:
        getattr(s, n + "0")[...] = getattr(s, n)
    for q in ("qv", "qc", "qr") + s.extra:
        getattr(s, q + "0")[...] = getattr(s, q)
    held = D.diff6_held(s, case.cfg)
    kept = set(held["q"])
    assert not kept & {"qv", "qc", "qr", "qi", "qs", "qg"}
    assert "nr" in kept and kept <= set(s.extra)
    all_rows = D.diff6_held(s, case.cfg.with_(moist_mix6_off=False))
    assert set(all_rows["q"]) == {"qv", "qc", "qr"} | set(s.extra)
