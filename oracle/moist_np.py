"""Moist scalar transport for the oracle -- TEST INFRASTRUCTURE ONLY.

NumPy restatement (float32 or float64) of the reference's qv/qc/qr stage
update: the coupled update with the acoustic time-averaged mass fluxes
(gpuwm/core/moist.py:434-698, fused update kernel :353-402) and WRF's
positive-definite limiter on the final RK stage
(gpuwm/core/kernels/pd_advection.cu, periodic path, msf = 1).

Pinning: uniform-qv invariance (a constant mixing ratio stays constant to
round-off, the reference's tests/test_moist.py:517-577 property), exact
periodic conservation of sum(C(mu) q), positivity with the limiter on,
the limiter reducing to the unlimited flux form (itself pinned by the
reference's advection fixture) wherever no cell needs limiting, and
planted mistakes in the limiter (tests/test_oracle_moist.py).
"""
from __future__ import annotations

import numpy as np

from oracle import dycore_np as D

SPECIES = ("qv", "qc", "qr")

# The species each mixed-phase scheme transports beyond qv/qc/qr, in the
# reference's registry order (moist.py:107-240 extra_moist_species; the
# allocation state.py:428-470), and how many of them (the leading condensate
# masses) enter the buoyancy's total water (dycore.cu:213-230 q_total and
# dycore.py:398-414: qi, qs, qg for moist_mode 2, plus qh for NSSL).
_ICE = ("qi", "qs", "qg")
EXTRA_SPECIES = {
    0: (), 1: (),
    6: _ICE,
    8: _ICE + ("nr", "ni"),
    10: _ICE + ("nr", "ni", "ns", "ng"),
    18: ("qi", "qs", "qg", "qh", "qndrop", "qnr", "qni", "qns", "qng",
         "qnh", "qnn", "qvolg", "qvolh"),
    28: _ICE + ("nr", "ni", "nc", "nwfa", "nifa"),
}
LOADING_EXTRA = {0: 0, 1: 0, 6: 3, 8: 3, 10: 3, 18: 4, 28: 3}


def _pm(a, shift, axis):
    """a[(i + shift) mod n] along axis (periodic neighbour)."""
    return np.roll(a, -shift, axis=axis)


# min, max, |x| and ordering tests that also carry complex-step
# perturbations: on real arrays they are np.fmin, np.fmax, np.abs and the
# plain comparisons; on complex arrays the branch follows the real part
# and the derivative rides in the imaginary part.
def _re(x):
    return np.real(x) if np.iscomplexobj(x) else x


def _abs(x):
    return np.sign(np.real(x)) * x if np.iscomplexobj(x) else np.abs(x)


def _fmin(a, b):
    if np.iscomplexobj(a) or np.iscomplexobj(b):
        return np.where(_re(a) <= _re(b), a, b)
    return np.fmin(a, b)


def _fmax(a, b):
    if np.iscomplexobj(a) or np.iscomplexobj(b):
        return np.where(_re(a) >= _re(b), a, b)
    return np.fmax(a, b)


def coupled_update(q0, tend, c1h, c2h, mu0, mu, dt_eff, clamp):
    """WRF rk_update_scalar (reference moist.py:385-398): q = (C(mu_t) q_t
    + dt tend) / C(mu_new); the final stage clamps at zero (NaN kept)."""
    f = q0.dtype.type
    c1, c2 = c1h[:, None, None], c2h[:, None, None]
    chm0 = c1 * mu0[None] + c2
    chm = c1 * mu[None] + c2
    v = chm0 * q0
    t = f(dt_eff) * tend
    v = v + t
    v = v / chm
    if clamp:
        v = np.where(np.isnan(v), v, np.maximum(v, f(0.0)))
    return v


def flux_upwind(qm1, q0v, cr):
    """WRF flux_upwind (reference pd_advection.cu:101-108)."""
    f = cr.dtype.type
    half, one = f(0.5), f(1.0)
    return ((half * _fmin(one, cr + _abs(cr))) * qm1
            + (half * _fmax(-one, cr - _abs(cr))) * q0v)


def pd_fluxes(q, q0, ru, rv, rw, mut, c1h, c2h, rdnw, fnm, fnp, dx, dy,
              dt):
    """Upwind fluxes of q0 and high-order corrections of q at every face
    (reference pd_advection.cu pd_fluxes, periodic branches).  Returns
    (fxl, fxc, fyl, fyc, fzl, fzc)."""
    f = q.dtype.type
    half, two = f(0.5), f(2.0)
    nz, ny, nx = q.shape
    dxf, dyf, dtf = f(dx), f(dy), f(dt)
    c1, c2 = c1h[:, None, None], c2h[:, None, None]
    # x faces 0..nx: cells i-1 and i (periodic), duplicate face nx == 0
    mx = (c1 * half) * (mut + _pm(mut, -1, 1))[None] + c2
    mx = np.concatenate([mx, mx[..., :1]], axis=-1)
    cr = ((ru * dtf) / dxf) / mx
    qm1 = np.concatenate([_pm(q0, -1, 2), _pm(q0, -1, 2)[..., :1]], axis=-1)
    qc0 = np.concatenate([q0, q0[..., :1]], axis=-1)
    fxl = (mx * (dxf / dtf)) * flux_upwind(qm1, qc0, cr)
    fxc = D._xfaces(q, ru) - fxl
    # y faces 0..ny
    my = (c1 * half) * (mut + _pm(mut, -1, 0))[None] + c2
    my = np.concatenate([my, my[:, :1]], axis=1)
    cr = ((rv * dtf) / dyf) / my
    qm1 = np.concatenate([_pm(q0, -1, 1), _pm(q0, -1, 1)[:, :1]], axis=1)
    qc0 = np.concatenate([q0, q0[:, :1]], axis=1)
    fyl = (my * (dyf / dtf)) * flux_upwind(qm1, qc0, cr)
    fyc = D._yfaces(q, rv) - fyl
    # eta faces 0..nz; boundary faces carry nothing
    fzl = np.zeros((nz + 1, ny, nx), dtype=q.dtype)
    fzc = np.zeros_like(fzl)
    fh = D._zfaces_half(q, rw, fnm, fnp)
    for k in range(1, nz):
        dz = two / (rdnw[k] + rdnw[k - 1])
        muf = c1h[k] * mut + c2h[k]
        crz = (rw[k] * dtf) / (dz * muf)
        fzl[k] = (muf * (dz / dtf)) * flux_upwind(q0[k - 1], q0[k], crz)
        fzc[k] = fh[k] - fzl[k]
    return fxl, fxc, fyl, fyc, fzl, fzc


def pd_budget(q0, mu_old, c1h, c2h, rdnw, fl, dx_inv, dy_inv, dt):
    """(ph_low, fo) of every cell (reference pd_advection.cu pd_scale,
    periodic no-msf branch): the coupled scalar after the upwind update,
    and the outgoing correction flux over the stage."""
    fxl, fxc, fyl, fyc, fzl, fzc = fl
    f = q0.dtype.type
    zero = f(0.0)
    dxi, dyi, dtf = f(dx_inv), f(dy_inv), f(dt)
    rz = rdnw[:, None, None]
    chm0 = c1h[:, None, None] * mu_old[None] + c2h[:, None, None]
    dl = (dxi * (fxl[..., 1:] - fxl[..., :-1])
          + dyi * (fyl[:, 1:] - fyl[:, :-1])) + rz * (fzl[1:] - fzl[:-1])
    ph_low = chm0 * q0 - dtf * dl
    # fminf/fmaxf semantics (a NaN operand yields the other one)
    fo = dtf * ((dxi * (_fmax(zero, fxc[..., 1:])
                        - _fmin(zero, fxc[..., :-1]))
                 + dyi * (_fmax(zero, fyc[:, 1:])
                          - _fmin(zero, fyc[:, :-1])))
                + rz * (_fmin(zero, fzc[1:]) - _fmax(zero, fzc[:-1])))
    return ph_low, fo


def pd_scale(q0, mu_old, c1h, c2h, rdnw, fl, dx_inv, dy_inv, dt):
    """Per-cell renormalisation factor (reference pd_advection.cu
    pd_scale, periodic no-msf branch): where the outgoing correction flux
    exceeds the upwind-updated coupled scalar, scale it down to fit."""
    f = q0.dtype.type
    zero, one = f(0.0), f(1.0)
    ph_low, fo = pd_budget(q0, mu_old, c1h, c2h, rdnw, fl, dx_inv, dy_inv,
                           dt)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        r = _fmin(one, _fmax(zero, ph_low / (fo + f(1e-20))))
    return np.where(_re(fo) > _re(ph_low), r, one)


def pd_donor_scales(sc, fl):
    """The donor cell's factor for each correction flux as each cell sees
    it (reference pd_advection.cu pd_renorm_apply): (sx_l, sx_r, sy_l,
    sy_r, sz_b, sz_t), 1 where the correction vanishes."""
    _, fxc, _, fyc, _, fzc = fl
    f = sc.dtype.type
    zero, one = f(0.0), f(1.0)
    nz = sc.shape[0]
    fxc_l, fxc_r = _re(fxc[..., :-1]), _re(fxc[..., 1:])
    sx_l = np.where(fxc_l > zero, _pm(sc, -1, 2),
                    np.where(fxc_l < zero, sc, one))
    sx_r = np.where(fxc_r > zero, sc,
                    np.where(fxc_r < zero, _pm(sc, 1, 2), one))
    fyc_l, fyc_r = _re(fyc[:, :-1]), _re(fyc[:, 1:])
    sy_l = np.where(fyc_l > zero, _pm(sc, -1, 1),
                    np.where(fyc_l < zero, sc, one))
    sy_r = np.where(fyc_r > zero, sc,
                    np.where(fyc_r < zero, _pm(sc, 1, 1), one))
    below = np.concatenate([sc[:1], sc[:-1]], axis=0)      # scale(k-1)
    above = np.concatenate([sc[1:], sc[-1:]], axis=0)      # scale(k+1)
    fzc_b, fzc_t = _re(fzc[:nz]), _re(fzc[1:])
    sz_b = np.where(fzc_b > zero, sc, np.where(fzc_b < zero, below, one))
    sz_t = np.where(fzc_t < zero, sc, np.where(fzc_t > zero, above, one))
    return sx_l, sx_r, sy_l, sy_r, sz_b, sz_t


def pd_apply(sc, fl, rdnw, dx_inv, dy_inv):
    """Renormalised PD flux divergence (reference pd_advection.cu
    pd_renorm_apply, periodic no-msf branch): each correction flux is
    scaled by its donor cell's factor."""
    fxl, fxc, fyl, fyc, fzl, fzc = fl
    f = sc.dtype.type
    dxi, dyi = f(dx_inv), f(dy_inv)
    nz = sc.shape[0]
    sx_l, sx_r, sy_l, sy_r, sz_b, sz_t = pd_donor_scales(sc, fl)
    gx = (sx_r * fxc[..., 1:] + fxl[..., 1:]) - (sx_l * fxc[..., :-1]
                                                 + fxl[..., :-1])
    gy = (sy_r * fyc[:, 1:] + fyl[:, 1:]) - (sy_l * fyc[:, :-1]
                                             + fyl[:, :-1])
    gz = (sz_t * fzc[1:] + fzl[1:]) - (sz_b * fzc[:nz] + fzl[:nz])
    return ((-dxi) * gx - dyi * gy) - rdnw[:, None, None] * gz


def advance_scalars_stage(s, ru_m, rv_m, ww_m, dt_stage, final, dx, dy,
                          moist_adv_opt=1, fixed=None):
    """qv/qc/qr one RK stage from their time-t copies (reference moist.py:
    434-698): unlimited flux form, or the PD limiter on the final stage
    when moist_adv_opt = 1.  ``dt_stage`` is the host double nsub*dtau.
    ``fixed`` ({name: tendency}) holds held mixing tendencies: the
    unlimited update adds them to the advective tendency (moist.py:
    384-386), the limited one folds them into the time-t scalar first,
    q0 + dt fixed / C(mu_t) (``_pd_fold_sources`` moist.py:293-346)."""
    f = s.dtype.type
    mu0 = s.mub2d + s.mup0
    mu = s.mub2d + s.mup
    dx_inv = float(np.real(f(1.0 / dx)))
    dy_inv = float(np.real(f(1.0 / dy)))
    pd = final and moist_adv_opt == 1
    for name in SPECIES + tuple(getattr(s, "extra", ())):
        q, q0 = getattr(s, name), getattr(s, name + "0")
        fx = fixed.get(name) if fixed else None
        if pd:
            if fx is not None:
                chm0 = s.c1h[:, None, None] * mu0[None] \
                    + s.c2h[:, None, None]
                q0 = q0 + (f(dt_stage) * fx) / chm0
            fl = pd_fluxes(q, q0, ru_m, rv_m, ww_m, mu, s.c1h, s.c2h,
                           s.rdnw, s.fnm, s.fnp, dx, dy, dt_stage)
            sc = pd_scale(q0, mu0, s.c1h, s.c2h, s.rdnw, fl, dx_inv,
                          dy_inv, dt_stage)
            tend = f(0.0) + pd_apply(sc, fl, s.rdnw, dx_inv, dy_inv)
            q[...] = coupled_update(q0, tend, s.c1h, s.c2h, mu0, mu,
                                    dt_stage, clamp=True)
        else:
            tend = f(0.0) + D.flux_div_scalar(q, ru_m, rv_m, ww_m, s.rdnw,
                                              s.fnm, s.fnp, dx_inv, dy_inv)
            if fx is not None:
                tend = tend + fx
            q[...] = coupled_update(q0, tend, s.c1h, s.c2h, mu0, mu,
                                    dt_stage, clamp=final)
