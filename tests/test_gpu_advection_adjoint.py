"""GPU checks of the advection-kernel adjoints (arw_flux_div_*_adjoint,
arw64_flux_div_*_adjoint) and the autograd wrapper ``FluxDivergence``:

- FP64: every gradient (advected field, ru, rv, rw) equals the oracle's
  hand adjoint (oracle/adjoint_np.flux_div_adjoint, itself equal to the
  transpose of the complex-step Jacobian), ragged and one-cell-wide grids
  included, with independent duplicate faces;
- at a 50 x 64 x 80 grid, two size-independent identities: the operator is
  linear in the field (<D dq, y> = <dq, D^T y> for a random dq) and
  positively homogeneous of degree 1 in (ru, rv, rw) (Euler:
  sum <r, dL/dr> = <D, lam>), both to 1e-12;
- the FP32 build agrees with FP64 to 1e-5;
- torch.autograd.gradcheck of ``FluxDivergence`` (finite differences of
  the device forward against the hand-written backward).
"""
import types

import numpy as np
import pytest
import torch

from oracle import adjoint_np as A

pytestmark = pytest.mark.gpu

OPS = ("scalar", "u", "v", "w")
ORACLE_TOL = 1e-12
IDENTITY_TOL = 1e-12
FP32_TOL = 1e-5


def _inputs(op, nz, ny, nx, seed, graded=False):
    """Field, mass fluxes (random signs, independent duplicate faces) and
    vertical coefficients.  graded: per-level signs with magnitudes
    growing 3x per level, so that every face velocity (including the
    level-mixing ones of the w operator) stays away from zero, as finite
    differences of |vel| need."""
    rng = np.random.default_rng(seed)
    from arwen_amd.advection import field_shape
    q = rng.standard_normal(field_shape(op, nz, ny, nx))

    def flux(shape, levels):
        mag = rng.uniform(0.5, 2.0, shape)
        if not graded:
            return np.where(rng.uniform(size=shape) < 0.5, -1.0, 1.0) * mag
        sign = np.where(rng.uniform(size=levels) < 0.5, -1.0, 1.0)
        scale = sign * 3.0 ** np.arange(levels)
        return scale[:, None, None] * rng.uniform(0.9, 1.1, shape)

    ru = flux((nz, ny, nx + 1), nz)
    rv = flux((nz, ny + 1, nx), nz)
    rw = flux((nz + 1, ny, nx), nz + 1)
    rw[0] = 0.0
    rw[nz] = 0.0
    fnm = np.r_[0.0, rng.uniform(0.4, 0.6, nz)]
    fnp = np.r_[0.0, 1.0 - fnm[1:]]
    coord = types.SimpleNamespace(
        rdnw=-1.0 / rng.uniform(0.01, 0.05, nz),
        rdn=np.r_[0.0, -1.0 / rng.uniform(0.01, 0.05, nz)],
        fnm=fnm, fnp=fnp)
    return q, ru, rv, rw, coord


def _dev(x, dtype=torch.float64):
    return torch.tensor(x, dtype=dtype, device="cuda")


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64)


def _oracle(op, lam, q, ru, rv, rw, coord, dx, dy):
    rdz = coord.rdn if op == "w" else coord.rdnw
    return A.flux_div_adjoint(op, lam, q, ru, rv, rw, rdz, coord.fnm,
                              coord.fnp, 1.0 / dx, 1.0 / dy)


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("shape", [(10, 12, 16), (3, 1, 7), (5, 2, 3),
                                   (4, 6, 1)])
def test_adjoint_matches_oracle(op, shape):
    from arwen_amd.advection import field_shape, flux_div_adjoint
    nz, ny, nx = shape
    print(f"advection adjoint vs oracle op={op} shape={shape}", flush=True)
    q, ru, rv, rw, coord = _inputs(op, nz, ny, nx, seed=sum(shape))
    lam = np.random.default_rng(9).standard_normal(field_shape(op, *shape))
    dx, dy = 1500.0, 1200.0
    got = flux_div_adjoint(op, _dev(lam), _dev(q), _dev(ru), _dev(rv),
                           _dev(rw), coord, dx, dy)
    torch.cuda.synchronize()
    ref = _oracle(op, lam, q, ru, rv, rw, coord, dx, dy)
    for name in ("q", "ru", "rv", "rw"):
        r = ref[name]
        err = np.abs(_np(got[name]) - r).max()
        assert err <= ORACLE_TOL * max(np.abs(r).max(), 1e-300), (name, err)


@pytest.mark.parametrize("op", OPS)
def test_identities_at_size(op):
    """Field linearity and flux homogeneity at 50 x 64 x 80."""
    from arwen_amd.advection import field_shape, flux_div, flux_div_adjoint
    nz, ny, nx = 50, 64, 80
    print(f"advection adjoint identities op={op}", flush=True)
    q, ru, rv, rw, coord = _inputs(op, nz, ny, nx, seed=21)
    rng = np.random.default_rng(22)
    lam = _dev(rng.standard_normal(field_shape(op, nz, ny, nx)))
    dq = _dev(rng.standard_normal(q.shape))
    Q, RU, RV, RW = _dev(q), _dev(ru), _dev(rv), _dev(rw)
    dx, dy = 3000.0, 3000.0
    g = flux_div_adjoint(op, lam, Q, RU, RV, RW, coord, dx, dy)
    d_dq = flux_div(op, dq, RU, RV, RW, coord, dx, dy)
    lhs, rhs = torch.vdot(d_dq.flatten(), lam.flatten()), \
        torch.vdot(dq.flatten(), g["q"].flatten())
    assert abs(float(lhs - rhs)) <= IDENTITY_TOL * abs(float(lhs)), \
        (float(lhs), float(rhs))
    d = flux_div(op, Q, RU, RV, RW, coord, dx, dy)
    dl = float(torch.vdot(d.flatten(), lam.flatten()))
    euler = sum(float(torch.vdot(x.flatten(), g[n].flatten()))
                for n, x in (("ru", RU), ("rv", RV), ("rw", RW)))
    lin = float(torch.vdot(Q.flatten(), g["q"].flatten()))
    assert abs(euler - dl) <= IDENTITY_TOL * abs(dl), (euler, dl)
    assert abs(lin - dl) <= IDENTITY_TOL * abs(dl), (lin, dl)


@pytest.mark.parametrize("op", OPS)
def test_fp32_close_to_fp64(op):
    from arwen_amd.advection import field_shape, flux_div_adjoint
    nz, ny, nx = 20, 24, 32
    print(f"advection adjoint fp32 vs fp64 op={op}", flush=True)
    q, ru, rv, rw, coord = _inputs(op, nz, ny, nx, seed=31)
    lam = np.random.default_rng(32).standard_normal(
        field_shape(op, nz, ny, nx))
    args = (q, ru, rv, rw)
    g64 = flux_div_adjoint(op, _dev(lam), *[_dev(x) for x in args], coord,
                           2000.0, 2000.0)
    g32 = flux_div_adjoint(op, _dev(lam, torch.float32),
                           *[_dev(x, torch.float32) for x in args], coord,
                           2000.0, 2000.0)
    for name in ("q", "ru", "rv", "rw"):
        r = _np(g64[name])
        err = np.abs(_np(g32[name]) - r).max()
        assert err <= FP32_TOL * np.abs(r).max(), (name, err)


@pytest.mark.parametrize("op", OPS)
def test_gradcheck(op):
    from arwen_amd.advection import FluxDivergence
    print(f"advection gradcheck op={op}", flush=True)
    q, ru, rv, rw, coord = _inputs(op, 5, 3, 4, seed=41, graded=True)
    xs = [_dev(x).requires_grad_() for x in (q, ru, rv, rw)]
    # grid spacings of order one keep the horizontal terms as large as the
    # vertical ones; tendencies reach ~1e4, so central differences with
    # eps = 1e-6 carry ~1e4 * 2e-16 / 1e-6 = 2e-6 of rounding: atol 1e-5
    fn = lambda a, b, c, d: FluxDivergence.apply(op, a, b, c, d, coord,
                                                 0.5, 0.4)
    assert torch.autograd.gradcheck(fn, xs, eps=1e-6, atol=1e-5, rtol=1e-6)


def test_launcher_adds_into_tend():
    """launch_flux_div_* ADD the raw kernel's tendency into tend (FP32,
    bitwise), and the unsupported boundary options raise."""
    from arwen_amd import _lib
    from arwen_amd.advection import launch_flux_div_u
    nz, ny, nx = 6, 7, 9
    q, ru, rv, rw, coord = _inputs("u", nz, ny, nx, seed=51)
    f = lambda x: _dev(x, torch.float32)
    U, RU, RV, RW = f(q), f(ru), f(rv), f(rw)
    base = torch.randn(U.shape, dtype=torch.float32, device="cuda")
    tend = base.clone()
    launch_flux_div_u(U, RU, RV, RW, tend, coord, 1000.0, 1000.0)
    raw = torch.zeros_like(U)
    lib = _lib.load()
    c = [f(coord.rdnw), f(coord.fnm), f(coord.fnp)]
    _lib.check(lib.arw_flux_div_u(U.data_ptr(), RU.data_ptr(), RV.data_ptr(),
                                  RW.data_ptr(), raw.data_ptr(),
                                  *[t.data_ptr() for t in c], 1.0 / 1000.0,
                                  1.0 / 1000.0, nz, ny, nx,
                                  torch.cuda.current_stream().cuda_stream),
               "arw_flux_div_u")
    assert torch.equal(tend, base + raw)
    with pytest.raises(NotImplementedError):
        launch_flux_div_u(U, RU, RV, RW, tend, coord, 1000.0, 1000.0,
                          open_x=True)


def test_adjoint_argument_errors():
    from arwen_amd import _lib
    lib = _lib.load()
    rc = lib.arw_flux_div_u_adjoint(None, None, None, None, None, None, None,
                                    None, None, None, None, None, 1.0, 1.0, 4,
                                    4, 4, None)
    assert rc != 0 and b"null" in lib.arw_last_error()
    x = torch.zeros(64, device="cuda")
    p = x.data_ptr()
    rc = lib.arw_flux_div_w_adjoint(*([p] * 12), 1.0, 1.0, 0, 4, 4, None)
    assert rc != 0 and b"extents" in lib.arw_last_error()
