"""Pins of the oracle's transport adjoint: the dot-product identity for
the flux-form operator and for a whole RK3 transport step (float64), the
linear forward agreeing with flux_div_scalar, and a planted transpose
mistake failing the identity."""
import numpy as np
import pytest

from oracle import adjoint_np as A
from oracle import dycore_np as D


def _fields(seed=0, nz=12, ny=9, nx=11):
    rng = np.random.default_rng(seed)
    ru = rng.standard_normal((nz, ny, nx + 1)) * 3e4
    ru[..., -1] = ru[..., 0]
    rv = rng.standard_normal((nz, ny + 1, nx)) * 3e4
    rv[:, -1] = rv[:, 0]
    rw = np.zeros((nz + 1, ny, nx))
    rw[1:nz] = rng.standard_normal((nz - 1, ny, nx)) * 30.0
    rdnw = -1.0 / rng.uniform(0.02, 0.12, nz)
    fnm = np.r_[0.0, rng.uniform(0.4, 0.6, nz - 1)]
    fnp = np.r_[0.0, 1.0 - fnm[1:]]
    return rng, ru, rv, rw, rdnw, fnm, fnp


def _ckpt(rng, nz, ny, nx, ru, rv, rw):
    stages = []
    for dt in (6.0, 9.0, 18.0):
        stages.append((ru * rng.uniform(0.9, 1.1), rv * rng.uniform(0.9, 1.1),
                       rw * rng.uniform(0.9, 1.1),
                       rng.uniform(-200, 200, (ny, nx)), dt))
    return {"mup_t": rng.uniform(-200, 200, (ny, nx)), "stages": stages}


def test_flux_div_adjoint_dot_product():
    rng, ru, rv, rw, rdnw, fnm, fnp = _fields()
    x = rng.standard_normal(ru[..., :-1].shape)
    y = rng.standard_normal(x.shape)
    dx = D.flux_div_scalar(x, ru, rv, rw, rdnw, fnm, fnp, 1e-3, 1e-3)
    dty = A.flux_div_scalar_adjoint(y, ru, rv, rw, rdnw, fnm, fnp, 1e-3,
                                    1e-3)
    lhs, rhs = np.vdot(dx, y), np.vdot(x, dty)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0), (lhs, rhs)


def test_planted_transpose_mistake_fails_dot_product(monkeypatch):
    """Using the flux5 weights of the mirrored face (i.e. applying D
    instead of D^T's gather order) must break the identity."""
    rng, ru, rv, rw, rdnw, fnm, fnp = _fields(1)
    orig = A.flux5_weights
    monkeypatch.setattr(A, "flux5_weights",
                        lambda vel: tuple(reversed(orig(vel))))
    x = rng.standard_normal(ru[..., :-1].shape)
    y = rng.standard_normal(x.shape)
    dx = D.flux_div_scalar(x, ru, rv, rw, rdnw, fnm, fnp, 1e-3, 1e-3)
    dty = A.flux_div_scalar_adjoint(y, ru, rv, rw, rdnw, fnm, fnp, 1e-3,
                                    1e-3)
    assert abs(np.vdot(dx, y) - np.vdot(x, dty)) > 1e-6 * abs(np.vdot(dx, y))


def test_flux5_weights_reproduce_flux5():
    rng = np.random.default_rng(3)
    q = rng.standard_normal((6, 50))
    vel = rng.standard_normal(50)
    w = A.flux5_weights(vel)
    assert np.allclose(sum(wi * qi for wi, qi in zip(w, q)),
                       D.flux5(*q, vel), rtol=1e-13, atol=1e-13)
    w3 = A.flux3_weights(vel)
    assert np.allclose(sum(wi * qi for wi, qi in zip(w3, q[:4])),
                       D.flux3(*q[:4], vel), rtol=1e-13, atol=1e-13)


def test_tracer_step_adjoint_dot_product():
    rng, ru, rv, rw, rdnw, fnm, fnp = _fields(2)
    nz, ny, nx1 = ru.shape
    nx = nx1 - 1
    ck = _ckpt(rng, nz, ny, nx, ru, rv, rw)
    c1h = rng.uniform(0.5, 1.0, nz)
    c2h = rng.uniform(0.0, 2000.0, nz)
    mub = np.full((ny, nx), 9e4)
    args = (ck, c1h, c2h, mub, rdnw, fnm, fnp, 1e-3, 1e-3)
    x = rng.standard_normal((nz, ny, nx))
    y = rng.standard_normal((nz, ny, nx))
    fx = A.tracer_step(x, *args)
    fty = A.tracer_step_adjoint(y, *args)
    lhs, rhs = np.vdot(fx, y), np.vdot(x, fty)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)
    # finite differences of the forward (linear map: exact up to rounding)
    eps = 1e-3
    fd = (A.tracer_step(x + eps * y, *args)
          - A.tracer_step(x - eps * y, *args)) / (2 * eps)
    assert np.allclose(fd, A.tracer_step(y, *args), rtol=1e-9, atol=1e-12)
