"""Pins of the oracle's constant-K diffusion ``add_diff2`` (CPU only).

The reference checks its kernel against the decay of a sine mode
(tests/test_diffusion.py:22-35: amplitude exp(-K k^2 t) within 2 %).  The
oracle is pinned tighter, by what the discretisation fixes exactly:

* horizontally the operator is the periodic 3-point Laplacian, whose symbol
  on exp(i a x/dx) is -4 K/dx^2 sin^2(a/2);
* vertically, on a uniform grid, the mass-level operator (zero flux through
  the bottom and top) has the Neumann eigenvectors cos(pi m (k + 1/2)/nz)
  with eigenvalues -4 K/dz^2 sin^2(pi m/(2 nz)); the w-level operator
  (rows 0 and nz held) has the Dirichlet ones sin(pi m k/nz);
* on a stretched grid the flux form conserves the column integral
  sum_k dz_k f_k (the interface fluxes telescope), and a profile linear in
  height has the same gradient at every interface, hence no interior
  tendency.

A vertical spacing taken one level off fails the last two (planted below).
"""
import math

import numpy as np
import pytest

from oracle import dycore_np as O

K = 75.0


def _zf_uniform(nz, dz):
    return np.arange(nz + 1) * dz


def _zf_stretched(nz, ztop=8000.0, ratio=1.08):
    d = ratio ** np.arange(nz)
    return np.r_[0.0, np.cumsum(d)] * (ztop / d.sum())


@pytest.mark.parametrize("stagger", ["", "x", "y"])
def test_horizontal_symbol(stagger):
    nz, ny, nx, dx, dy = 3, 8, 32, 100.0, 250.0
    zf = _zf_uniform(nz, 500.0)
    a, b = 2 * math.pi * 3 / nx, 2 * math.pi * 2 / ny
    i = np.arange(nx + (stagger == "x"))
    j = np.arange(ny + (stagger == "y"))
    f = np.cos(a * i)[None, None, :] * np.cos(b * j)[None, :, None] \
        * np.ones((nz, 1, 1))
    out = O.add_diff2(f, K, 0.0, dx, dy, zf, stagger, np.float64)
    lam = -4.0 * K * (math.sin(a / 2) ** 2 / dx ** 2
                      + math.sin(b / 2) ** 2 / dy ** 2)
    np.testing.assert_allclose(out, lam * f, rtol=0, atol=1e-12 * abs(lam))


@pytest.mark.parametrize("m", [1, 3, 7])
def test_vertical_neumann_modes_mass_levels(m):
    nz, dz = 16, 400.0
    zf = _zf_uniform(nz, dz)
    col = np.cos(math.pi * m * (np.arange(nz) + 0.5) / nz)
    f = np.broadcast_to(col[:, None, None], (nz, 2, 4)).copy()
    out = O.add_diff2(f, 0.0, K, 1000.0, 1000.0, zf, "", np.float64)
    lam = -4.0 * K / dz ** 2 * math.sin(math.pi * m / (2 * nz)) ** 2
    np.testing.assert_allclose(out, lam * f, rtol=0, atol=1e-12 * abs(lam))


@pytest.mark.parametrize("m", [1, 4, 9])
def test_vertical_dirichlet_modes_w_levels(m):
    nz, dz = 16, 400.0
    zf = _zf_uniform(nz, dz)
    col = np.sin(math.pi * m * np.arange(nz + 1) / nz)
    f = np.broadcast_to(col[:, None, None], (nz + 1, 2, 4)).copy()
    out = O.add_diff2(f, 0.0, K, 1000.0, 1000.0, zf, "z", np.float64)
    lam = -4.0 * K / dz ** 2 * math.sin(math.pi * m / (2 * nz)) ** 2
    want = lam * f
    want[0] = want[-1] = 0.0
    np.testing.assert_allclose(out, want, rtol=0, atol=1e-12 * abs(lam))


def _stretched_checks(out_mass, out_w, zf, fm, fw):
    """(conservation residual, linear-profile residual) of the mass and w
    operators, each relative to the operator's own scale."""
    zh = 0.5 * (zf[1:] + zf[:-1])
    dzm = np.diff(zf)                       # mass layer thickness
    dzw = np.diff(zh)                       # w layer thickness, 1..nz-1
    scale_m = K * np.abs(fm).max() / dzm.min() ** 2
    scale_w = K * np.abs(fw).max() / dzw.min() ** 2
    cons_m = abs(np.sum(dzm * out_mass[0])) / (scale_m * zf[-1])
    # w: rows 0 and nz are held, so the interior sum is the two boundary
    # interface fluxes
    bflux = K * ((fw[-1] - fw[-2]) / dzm[-1] - (fw[1] - fw[0]) / dzm[0])
    cons_w = abs(np.sum(dzw * out_w[1:-1]) - bflux) / (scale_w * zf[-1])
    return cons_m, cons_w


def _random_columns(zf, seed=0):
    rng = np.random.default_rng(seed)
    nz = zf.size - 1
    return rng.standard_normal(nz), rng.standard_normal(nz + 1)


def _apply(zf, fm, fw):
    nz = zf.size - 1
    m3 = np.broadcast_to(fm[:, None, None], (nz, 1, 2)).copy()
    w3 = np.broadcast_to(fw[:, None, None], (nz + 1, 1, 2)).copy()
    om = O.add_diff2(m3, 0.0, K, 1e3, 1e3, zf, "", np.float64)[:, 0, 0]
    ow = O.add_diff2(w3, 0.0, K, 1e3, 1e3, zf, "z", np.float64)[:, 0, 0]
    return om, ow


def test_stretched_flux_form_conserves_the_column_integral():
    zf = _zf_stretched(24)
    fm, fw = _random_columns(zf)
    om, ow = _apply(zf, fm, fw)
    cm, cw = _stretched_checks(om[None], ow, zf, fm, fw)
    assert cm < 1e-13 and cw < 1e-13, (cm, cw)


def test_stretched_linear_profile_has_no_interior_tendency():
    zf = _zf_stretched(24)
    zh = 0.5 * (zf[1:] + zf[:-1])
    om, ow = _apply(zf, 3e-3 * zh, 3e-3 * zf)
    scale = K * 3e-3 / np.diff(zf).min()
    assert np.abs(om[1:-1]).max() < 1e-12 * scale
    assert np.abs(ow[1:-1]).max() < 1e-12 * scale


def _shift_rdzc(zf, stagger):
    rdzf, rdzc = _ORIG(zf, stagger)
    return rdzf, np.roll(rdzc, 1)


def _shift_rdzf(zf, stagger):
    rdzf, rdzc = _ORIG(zf, stagger)
    return np.roll(rdzf, 1), rdzc


_ORIG = O._diff2_spacings


@pytest.mark.parametrize("mistake", ["cell spacing one level off",
                                     "interface spacing one level off"])
def test_planted_spacing_mistakes_are_caught(mistake, monkeypatch):
    monkeypatch.setattr(O, "_diff2_spacings",
                        _shift_rdzc if mistake.startswith("cell")
                        else _shift_rdzf)
    zf = _zf_stretched(24)
    zh = 0.5 * (zf[1:] + zf[:-1])
    fm, fw = _random_columns(zf)
    om, ow = _apply(zf, fm, fw)
    cm, cw = _stretched_checks(om[None], ow, zf, fm, fw)
    lm, lw = _apply(zf, 3e-3 * zh, 3e-3 * zf)
    scale = K * 3e-3 / np.diff(zf).min()
    lin = max(np.abs(lm[1:-1]).max(), np.abs(lw[1:-1]).max()) / scale
    assert max(cm, cw) > 1e-6 or lin > 1e-6, (cm, cw, lin)
