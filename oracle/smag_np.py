"""WRF diff_opt = 2, km_opt = 4 (2-D Smagorinsky) held mixing tendencies,
restated in NumPy for a periodic domain with unit map factors.

TEST INFRASTRUCTURE: this module is part of the oracle.  Only tests/,
the smoke check and bench.py's CPU baseline may import it; the device path
never does.

The reference computes these once per step from the time-t fields
(gpuwm/core/dycore.py:1223-1343 ``_compute_wrf_smag_tendencies``,
:1345-1420 ``prepare_fixed_tendencies``) with the kernels of
gpuwm/core/kernels/smag2d.cu:272-1340:

* deformation D11, D22 (mass points) and D12 (the u-v corner) with the
  terrain-slope corrections of WRF cal_deform_and_div (smag2d.cu:484-564,
  ``wrf_smag_deform`` :571-595);
* K_m = min(c_s^2 l^2 |S|, 10 l), l = sqrt(dx dy), divided by the slope
  factor alpha (or alpha^2 beyond the deformation limit), K_h = K_m / Pr,
  Pr = 1/3 (``wrf_smag2d_km`` :604-652);
* the momentum stress divergences of u, v, w (``wrf_smag_hd_u`` :765-795,
  ``wrf_smag_hd_v`` :798-824, ``wrf_smag_hd_w`` :1071-1093) from tau11,
  tau22, tau12, tau13, tau23 (:659-734, :826-899);
* the theta flux divergence with WRF's t_2 = theta - 300 (``wrf_smag_flux_s``
  / ``wrf_smag_hd_s`` :1236-1292, the flux form :1164-1225);
* with the dycore's own vertical mixing (bl_pbl_physics = 0,
  ``launch_wrf_smag2d_vertical`` dycore.py:1046-1144): the vertical stress
  divergences of u, v, w (``wrf_smag_vd_u/v/w`` :916-973) and the surface
  policy of isfflx (the constant drag :1608-1654 and heat flux
  :1658-1668; the surface-driver arms are inactive without a surface
  layer scheme).

Every product and sum is evaluated in the reference kernels' order in the
array's dtype, so the float32 results are bit-identical to a float32
device implementation that keeps that order.  Map factors are 1: every
factor msf enters as a multiplication or division by exactly 1.0, which
drops out without changing a rounding.
"""
from __future__ import annotations

import numpy as np

from .dycore_np import G, _b3

#: Smagorinsky Prandtl number (reference dycore.py:70)
PRANDTL = 1.0 / 3.0
#: WRF's reference theta of t_2 = theta - T0 (smag2d.cu:1103)
WRF_T0 = 300.0


def _r(a, shift, axis):
    """Value at index i + shift (periodic) placed at i."""
    return np.roll(a, -shift, axis=axis)


class Grid:
    """The on-demand metrics of the reference's WrfSmagGrid
    (smag2d.cu:284-420) for one time-t state: total geopotential phi,
    rdzw (mass levels), rdz (w levels), zx / zy at the u / v faces of every
    w level, and the dry density 1 / alt."""

    def __init__(self, s, dx, dy, php, alt, qv=None):
        t = s.dtype.type
        self.t = t
        self.rdx, self.rdy = t(1.0 / dx), t(1.0 / dy)
        self.dx, self.dy = t(dx), t(dy)
        self.g = t(G)
        ph = _b3(s.phb) + php                    # base + php
        self.ph = ph
        nz = ph.shape[0] - 1
        self.nz = nz
        self.rdzw = self.g / (ph[1:] - ph[:-1])
        g2 = t(2.0) * self.g
        rdz = np.empty_like(ph)
        rdz[0] = g2 / (ph[1] - ph[0])
        rdz[nz] = g2 / (ph[nz] - ph[nz - 1])
        rdz[1:nz] = g2 / (ph[2:] - ph[:-2])
        self.rdz = rdz
        self.zx = (self.rdx * (ph - _r(ph, -1, 2))) / self.g
        self.zy = (self.rdy * (ph - _r(ph, -1, 1))) / self.g
        one = t(1.0)
        self.rho = (one / alt) if qv is None else (one + qv) / alt
        self.fnm = s.fnm[:, None, None]
        self.fnp = s.fnp[:, None, None]
        self.dnw, self.dn = s.dnw, s.dn
        self.cf = (s.cf1, s.cf2, s.cf3)
        self.cft2 = (t(-0.5) * s.dnw[nz - 1]) / s.dn[nz - 1]

    def full_levels(self, P):
        """0.5 * WRF's w-level weights of per-mass-level sums P (nz, ...)
        (``wrf_full_weights`` smag2d.cu:420-431): the cf extrapolation at
        the surface, the one-sided cft2 form at the lid, fnm/fnp between."""
        t = self.t
        nz = self.nz
        cf1, cf2, cf3 = self.cf
        out = np.empty((nz + 1,) + P.shape[1:], dtype=P.dtype)
        out[0] = (cf1 * P[0] + cf2 * P[1]) + cf3 * P[2]
        out[nz] = (t(1.0) - self.cft2) * P[nz - 1] + self.cft2 * P[nz - 2]
        out[1:nz] = self.fnm[1:nz] * P[1:] + self.fnp[1:nz] * P[:-1]
        return t(0.5) * out

    def face_mavg(self, T, shift, axis):
        """0.5 (fnm (T_k(s) + T_k) + fnp (T_{k-1}(s) + T_{k-1})) at the
        interior w levels, 0 at the surface and the lid: the tau*_uavg /
        tau*_vavg / h*_mavg averages (smag2d.cu:688-733, :1206-1225).
        ``shift`` names the neighbour T_k(s) (index offset along axis)."""
        t = self.t
        nz = self.nz
        out = np.zeros((nz + 1,) + T.shape[1:], dtype=T.dtype)
        Ts = _r(T, shift, axis)
        out[1:nz] = t(0.5) * (self.fnm[1:nz] * (Ts[1:] + T[1:])
                              + self.fnp[1:nz] * (Ts[:-1] + T[:-1]))
        return out


def _sum4(a, b, c, d):
    return ((a + b) + c) + d


def deformation(gr, u, v):
    """(D11, D22, D12) of WRF cal_deform_and_div with its slope terms
    (reference smag2d.cu:484-564).  D12 sits at the corner shared by u face
    i and v face j of cell (j, i)."""
    t = gr.t
    q = t(0.25)
    uc, vc = u[..., :-1], v[:, :-1, :]
    zx, zy, rdzw = gr.zx, gr.zy, gr.rdzw
    # u_w_xcenter / v_w_ycenter: level sums of the two faces of a cell
    uwx = gr.full_levels(uc + _r(uc, 1, 2))
    vwy = gr.full_levels(vc + _r(vc, 1, 1))
    tmpzx = q * _sum4(zx[:-1], _r(zx[:-1], 1, 2), zx[1:], _r(zx[1:], 1, 2))
    slope = ((uwx[1:] - uwx[:-1]) * tmpzx) * rdzw
    d11 = t(2.0) * (gr.rdx * (_r(uc, 1, 2) - uc) - slope)
    tmpzy = q * _sum4(zy[:-1], _r(zy[:-1], 1, 1), zy[1:], _r(zy[1:], 1, 1))
    slope = ((vwy[1:] - vwy[:-1]) * tmpzy) * rdzw
    d22 = t(2.0) * (gr.rdy * (_r(vc, 1, 1) - vc) - slope)
    # corner: u_w_corner sums u(j-1), u(j); v_w_corner v(i-1), v(i)
    ucor = gr.full_levels(_r(uc, -1, 1) + uc)
    vcor = gr.full_levels(_r(vc, -1, 2) + vc)
    rr = _sum4(rdzw, _r(rdzw, -1, 2), _r(_r(rdzw, -1, 1), -1, 2),
               _r(rdzw, -1, 1))
    tzy = q * _sum4(_r(zy[:-1], -1, 2), zy[:-1], _r(zy[1:], -1, 2), zy[1:])
    uslope = (((ucor[1:] - ucor[:-1]) * q) * tzy) * rr
    tzx = q * _sum4(_r(zx[:-1], -1, 1), zx[:-1], _r(zx[1:], -1, 1), zx[1:])
    vslope = (((vcor[1:] - vcor[:-1]) * q) * tzx) * rr
    d12 = ((gr.rdy * (uc - _r(uc, -1, 1)) - uslope)
           + gr.rdx * (vc - _r(vc, -1, 2))) - vslope
    return d11, d22, d12


def smag_km(gr, d11, d22, d12, c_s):
    """K_m and K_h of WRF smag2d_km (reference smag2d.cu:604-652): the
    mixing length sqrt(dx dy), the 10 l cap, the terrain-slope factor
    alpha (squared where |S| exceeds max(10 / l, 1e-3)), K_h = K_m / Pr."""
    t = gr.t
    q = t(0.25)
    dc = q * _sum4(d12, _r(d12, 1, 1), _r(d12, 1, 2), _r(_r(d12, 1, 1), 1, 2))
    strain = np.sqrt((q * (d11 - d22)) * (d11 - d22) + dc * dc)
    mlen = np.sqrt(gr.dx * gr.dy)
    cs = t(c_s)
    km = np.minimum((((cs * cs) * mlen) * mlen) * strain, t(10.0) * mlen)
    zx, zy = np.abs(gr.zx), np.abs(gr.zy)
    sx = (q * _sum4(zx[:-1], _r(zx[:-1], 1, 2), zx[1:], _r(zx[1:], 1, 2))
          * gr.rdzw) * gr.dx
    sy = (q * _sum4(zy[:-1], _r(zy[:-1], 1, 1), zy[1:], _r(zy[1:], 1, 1))
          * gr.rdzw) * gr.dy
    alpha = np.maximum(np.sqrt(sx * sx + sy * sy), t(1.0))
    limit = max(t(10.0) / mlen, t(1.0e-3))
    km = km / np.where(strain > limit, alpha * alpha, alpha)
    return km, km / t(PRANDTL)


def stresses(gr, km, d11, d22, d12, u, v, w):
    """tau11, tau22 (mass), tau12 (corner), tau13 / tau23 (u / v faces of
    the w levels, 0 at the surface and the lid) and tau33 (mass)
    (reference smag2d.cu:659-686, :845-899, :945-956)."""
    t = gr.t
    q, h = t(0.25), t(0.5)
    rho = gr.rho
    nz = gr.nz
    t11 = ((-rho) * km) * d11
    t22 = ((-rho) * km) * d22
    ravg = q * _sum4(rho, _r(rho, -1, 2), _r(_r(rho, -1, 1), -1, 2),
                     _r(rho, -1, 1))
    kavg = q * _sum4(km, _r(km, -1, 2), _r(_r(km, -1, 1), -1, 2),
                     _r(km, -1, 1))
    t12 = ((-ravg) * kavg) * d12
    uc, vc = u[..., :-1], v[:, :-1, :]
    rdz = gr.rdz
    fnm, fnp = gr.fnm[1:nz], gr.fnp[1:nz]

    def wavg(axis):
        # w_xavg / w_yavg at mass levels k: w(k), w(k+1) and the cell behind
        return q * _sum4(w[:-1], w[1:], _r(w[:-1], -1, axis),
                         _r(w[1:], -1, axis))

    def tau_w(axis, vel, zface, rdz_):
        wa = wavg(axis)
        rz = h * (rdz_[1:nz] + _r(rdz_[1:nz], -1, axis))
        slope = ((wa[1:] - wa[:-1]) * zface[1:nz]) * rz
        rd = gr.rdx if axis == 2 else gr.rdy
        dw = rd * (w[1:nz] - _r(w[1:nz], -1, axis)) - slope
        d = dw + (vel[1:] - vel[:-1]) * rz
        ra = h * (fnm * (_r(rho[1:], -1, axis) + rho[1:])
                  + fnp * (_r(rho[:-1], -1, axis) + rho[:-1]))
        ka = h * (fnm * (_r(km[1:], -1, axis) + km[1:])
                  + fnp * (_r(km[:-1], -1, axis) + km[:-1]))
        out = np.zeros((nz + 1,) + rho.shape[1:], dtype=rho.dtype)
        out[1:nz] = ((-ra) * ka) * d
        return out

    t13 = tau_w(2, uc, gr.zx, rdz)
    t23 = tau_w(1, vc, gr.zy, rdz)
    t33 = ((-rho) * km) * ((t(2.0) * (w[1:] - w[:-1])) * gr.rdzw)
    return t11, t22, t12, t13, t23, t33


def _sum8(terms):
    acc = terms[0]
    for x in terms[1:]:
        acc = acc + x
    return acc


def momentum_tendencies(gr, tau, vertical, cd0, u, v):
    """The u, v, w stress divergences (reference smag2d.cu:765-824,
    :1071-1093), then with the dycore's vertical mixing the vertical
    divergences (:916-973) and, under isfflx = 0, the constant-drag wall
    stress (:1608-1654).  Returns (ru, rv, rw) shaped like u, v, w (the
    periodic duplicate face included)."""
    t = gr.t
    q, h, e = t(0.25), t(0.5), t(0.125)
    t11, t22, t12, t13, t23, t33 = tau
    nz = gr.nz
    zx, zy, rdzw, g = gr.zx, gr.zy, gr.rdzw, gr.g
    dnw = gr.dnw[:, None, None]
    dn = gr.dn[:, None, None]
    zero = t(0.0)
    one = t(1.0)
    # u faces
    tmpdz = h * (one / rdzw + one / _r(rdzw, -1, 2))
    zx_u = h * (zx[:-1] + zx[1:])
    zy_u = e * _sum8([_r(zy[:-1], -1, 2), zy[:-1],
                      _r(_r(zy[:-1], 1, 1), -1, 2), _r(zy[:-1], 1, 1),
                      _r(zy[1:], -1, 2), zy[1:],
                      _r(_r(zy[1:], 1, 1), -1, 2), _r(zy[1:], 1, 1)])
    divh = gr.rdx * (t11 - _r(t11, -1, 2)) + gr.rdy * (_r(t12, 1, 1) - t12)
    a11 = gr.face_mavg(t11, -1, 2)
    a12 = gr.face_mavg(t12, 1, 1)
    divz = (zx_u * (a11[1:] - a11[:-1])) / tmpdz \
        + (zy_u * (a12[1:] - a12[:-1])) / tmpdz
    ru = zero + ((g * tmpdz) / dnw) * (divh - divz)
    # v faces
    tmpdz = h * (one / rdzw + one / _r(rdzw, -1, 1))
    zx_v = e * _sum8([_r(zx[:-1], -1, 1), _r(_r(zx[:-1], -1, 1), 1, 2),
                      zx[:-1], _r(zx[:-1], 1, 2),
                      _r(zx[1:], -1, 1), _r(_r(zx[1:], -1, 1), 1, 2),
                      zx[1:], _r(zx[1:], 1, 2)])
    zy_v = h * (zy[:-1] + zy[1:])
    divh = gr.rdy * (t22 - _r(t22, -1, 1)) + gr.rdx * (_r(t12, 1, 2) - t12)
    a12v = gr.face_mavg(t12, 1, 2)
    a22 = gr.face_mavg(t22, -1, 1)
    divz = (zx_v * (a12v[1:] - a12v[:-1])) / tmpdz \
        + (zy_v * (a22[1:] - a22[:-1])) / tmpdz
    rv = zero + ((g * tmpdz) / dnw) * (divh - divz)
    # w levels 1..nz-1
    rw = np.zeros((nz + 1,) + rdzw.shape[1:], dtype=rdzw.dtype)
    zx_w = h * (zx[1:nz] + _r(zx[1:nz], 1, 2))
    zy_w = h * (zy[1:nz] + _r(zy[1:nz], 1, 1))
    rz = gr.rdz[1:nz]
    divh = gr.rdx * (_r(t13[1:nz], 1, 2) - t13[1:nz]) \
        + gr.rdy * (_r(t23[1:nz], 1, 1) - t23[1:nz])
    m13 = q * _sum4(t13[:-1], _r(t13[:-1], 1, 2), t13[1:], _r(t13[1:], 1, 2))
    m23 = q * _sum4(t23[:-1], _r(t23[:-1], 1, 1), t23[1:], _r(t23[1:], 1, 1))
    divz = rz * (zx_w * (m13[1:] - m13[:-1]) + zy_w * (m23[1:] - m23[:-1]))
    rw[1:nz] = zero + (g / (dn[1:nz] * rz)) * (divh - divz)
    if vertical:
        ru = ru + (g * (t13[1:] - t13[:-1])) / dnw
        rv = rv + (g * (t23[1:] - t23[:-1])) / dnw
        rw[1:nz] = rw[1:nz] + (g * (t33[1:] - t33[:-1])) / dn[1:nz]
        if cd0 is not None:
            cd = t(cd0)
            eps = t(1.0e-15)
            rho0 = gr.rho[0]
            uc, vc = u[0, :, :-1], v[0, :-1, :]
            vv = q * _sum4(vc, _r(vc, 1, 0), _r(vc, -1, 1),
                           _r(_r(vc, 1, 0), -1, 1))
            uu = uc
            speed = np.sqrt(uu * uu + vv * vv) + eps
            stress = (cd * speed) * uu
            ra = h * (rho0 + _r(rho0, -1, 1))
            ru[0] = ru[0] + ((g * stress) * ra) / gr.dnw[0]
            uu = q * _sum4(uc, _r(uc, 1, 1), _r(uc, -1, 0),
                           _r(_r(uc, -1, 0), 1, 1))
            vv = vc
            speed = np.sqrt(vv * vv + uu * uu) + eps
            stress = (cd * speed) * vv
            ra = h * (rho0 + _r(rho0, -1, 0))
            rv[0] = rv[0] + ((g * stress) * ra) / gr.dnw[0]
    ru = np.concatenate([ru, ru[..., :1]], axis=2)
    rv = np.concatenate([rv, rv[:, :1, :]], axis=1)
    return ru, rv, rw


def scalar_fluxes(gr, kh, S):
    """The x / y face fluxes H1, H2 of a scalar S at its u / v faces
    (reference smag2d.cu:1164-1200): -rho K_h times the slope-corrected
    gradient along constant height."""
    t = gr.t
    h = t(0.5)
    rho, rdzw = gr.rho, gr.rdzw
    one, two = t(1.0), t(2.0)

    def flux(axis, rd, zf):
        ra = h * (_r(rho, -1, axis) + rho)
        ka = h * (_r(kh, -1, axis) + kh)
        tz = h * (zf[:-1] + zf[1:])
        rdzu = two / (one / rdzw + one / _r(rdzw, -1, axis))
        sw = gr.full_levels(_r(S, -1, axis) + S)
        grad = rd * (S - _r(S, -1, axis)) - (tz * (sw[1:] - sw[:-1])) * rdzu
        return ((-ra) * ka) * grad

    return flux(2, gr.rdx, gr.zx), flux(1, gr.rdy, gr.zy)


def theta_tendency(gr, kh, thp, thb, heat_flux):
    """The theta flux divergence of WRF horizontal_diffusion_s_2 on
    t_2 = (thb + thp) - 300 (reference smag2d.cu:1095-1292), then the
    constant surface heat flux when ``heat_flux`` is given (isfflx 0 or 2
    with the dycore's vertical mixing, :1658-1668)."""
    t = gr.t
    return scalar_tendency(gr, kh, (_b3(thb) + thp) - t(WRF_T0), heat_flux)


def scalar_tendency(gr, kh, S, heat_flux=None):
    """The flux divergence of WRF horizontal_diffusion_s_2 for the scalar
    S (reference smag2d.cu:1254-1292): theta's t_2, or a moist species'
    mixing ratio as stored (dycore.py:1259-1264)."""
    t = gr.t
    zx, zy, rdzw, g = gr.zx, gr.zy, gr.rdzw, gr.g
    rho = gr.rho
    fx, fy = scalar_fluxes(gr, kh, S)
    q = t(0.25)
    zx_m = q * _sum4(zx[:-1], _r(zx[:-1], 1, 2), zx[1:], _r(zx[1:], 1, 2))
    zy_m = q * _sum4(zy[:-1], _r(zy[:-1], 1, 1), zy[1:], _r(zy[1:], 1, 1))
    divh = gr.rdx * (_r(fx, 1, 2) - fx) + gr.rdy * (_r(fy, 1, 1) - fy)
    h1 = gr.face_mavg(fx, 1, 2)
    h2 = gr.face_mavg(fy, 1, 1)
    divz = (zx_m * (h1[1:] - h1[:-1])) * rdzw \
        + (zy_m * (h2[1:] - h2[:-1])) * rdzw
    dnw = gr.dnw[:, None, None]
    out = t(0.0) + (g / (dnw * rdzw)) * (divh - divz)
    if heat_flux is not None:
        out[0] = out[0] - ((g * t(heat_flux)) * rho[0]) / gr.dnw[0]
    return out


def smag_held(s, cfg):
    """The held km_opt = 4 tendencies {ru_t, rv_t, rw_t, rth_t} of the
    time-t fields (u0 ... php0 and the refreshed alt), as
    ``_compute_wrf_smag_tendencies`` (reference dycore.py:1223-1343) leaves
    them in the carrying buffers for a periodic state; a moist state's
    density carries the vapour, rho = (1 + qv) / alt (smag2d.cu:394-404),
    and every transported species takes K_h's flux divergence under
    ``"q"`` ({name: tendency}, dycore.py:1193-1199)."""
    moist = getattr(s, "qv", None) is not None
    gr = Grid(s, cfg.dx, cfg.dy, s.php0, s.alt, s.qv0 if moist else None)
    d11, d22, d12 = deformation(gr, s.u0, s.v0)
    km, kh = smag_km(gr, d11, d22, d12, cfg.c_s)
    tau = stresses(gr, km, d11, d22, d12, s.u0, s.v0, s.w0)
    vertical = getattr(cfg, "bl_pbl_physics", 0) == 0
    isfflx = getattr(cfg, "isfflx", 1)
    cd0 = cfg.tke_drag_coefficient if (vertical and isfflx == 0) else None
    heat = cfg.tke_heat_flux if (vertical and isfflx in (0, 2)) else None
    ru, rv, rw = momentum_tendencies(gr, tau, vertical, cd0, s.u0, s.v0)
    rth = theta_tendency(gr, kh, s.thp0, s.thb, heat)
    out = {"ru_t": ru, "rv_t": rv, "rw_t": rw, "rth_t": rth}
    if moist:
        out["q"] = {n: scalar_tendency(gr, kh, getattr(s, n + "0"))
                    for n in ("qv", "qc", "qr") + tuple(getattr(s, "extra",
                                                              ()))}
    return out
