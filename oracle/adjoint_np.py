"""Adjoint of scalar transport for the oracle -- TEST INFRASTRUCTURE ONLY.

The flux-form scalar advection of the reference (advection.cu:206-416,
flux5 horizontal / Omega-signed flux3 vertical with 2nd-order faces one
level in) is linear in the scalar for given mass fluxes: each face flux is
sum_o w_o(vel) q[f + o], with weights that depend only on the face
velocity.  Its transpose gathers, at every cell, the face adjoints
Lambda_f = rdx (mu[f] - mu[f-1]) weighted by the coefficient with which
that cell entered face f.

The RK3 transport of one step (reference moist.py:434-698, unlimited
form) is q1 = (C0 q + dt0 D0 q) / C1s ..., linear in q for the stage fluxes
and masses of the step; tracer_step / tracer_step_adjoint are that map and
its transpose.  Pinned by the dot-product identity <D x, y> = <x, D^T y>
(float64) and by the forward matching flux_div_scalar.
"""
from __future__ import annotations

import numpy as np

from oracle import dycore_np as D


def flux5_weights(vel):
    """Coefficients of (qm3, qm2, qm1, q0, qp1, qp2) in WRF flux5
    (reference advection.cu:78-83)."""
    a = np.abs(vel)
    return ((vel + a) / 60.0, (-8.0 * vel - 5.0 * a) / 60.0,
            (37.0 * vel + 10.0 * a) / 60.0, (37.0 * vel - 10.0 * a) / 60.0,
            (-8.0 * vel + 5.0 * a) / 60.0, (vel - a) / 60.0)


def flux3_weights(vel):
    """Coefficients of (qm2, qm1, q0, qp1) in the Omega-signed flux3
    (reference advection.cu:86-90)."""
    a = np.abs(vel)
    return ((-vel + a) / 12.0, (7.0 * vel - 3.0 * a) / 12.0,
            (7.0 * vel + 3.0 * a) / 12.0, (-vel - a) / 12.0)


def flux_div_scalar_adjoint(lam, ru, rv, rw, rdnw, fnm, fnp, dx_inv,
                            dy_inv):
    """D^T lam for D = flux_div_scalar (periodic, msf = 1)."""
    nz, ny, nx = lam.shape
    out = np.zeros_like(lam)
    # x faces f = 0..nx-1 (face nx duplicates 0): between cells f-1 and f
    lx = dx_inv * (lam - np.roll(lam, 1, axis=2))        # Lambda at face f
    w = flux5_weights(ru[..., :nx])
    for o in range(6):                 # face f reads cell f + o - 3
        out += np.roll(w[o] * lx, o - 3, axis=2)
    ly = dy_inv * (lam - np.roll(lam, 1, axis=1))
    w = flux5_weights(rv[:, :ny, :])
    for o in range(6):
        out += np.roll(w[o] * ly, o - 3, axis=1)
    # eta faces kf = 1..nz-1: between cells kf-1 and kf
    rz = rdnw[:, None, None]
    for kf in range(1, nz):
        lz = rz[kf] * lam[kf] - rz[kf - 1] * lam[kf - 1]
        vel = rw[kf]
        if kf == 1 or kf == nz - 1:
            out[kf] += vel * fnm[kf] * lz
            out[kf - 1] += vel * fnp[kf] * lz
        else:
            w3 = flux3_weights(vel)
            for o in range(4):         # face kf reads cell kf + o - 2
                out[kf + o - 2] += w3[o] * lz
    return out


def tracer_step(q, ckpt, c1h, c2h, mub2d, rdnw, fnm, fnp, dx_inv, dy_inv,
                clamp=False):
    """One RK3 step of scalar transport with the checkpointed stage fluxes
    and masses: ckpt = {"mup_t", "stages": [(ru_m, rv_m, ww_m, mup_new,
    dt), ...]} (reference moist.py:385-398 per stage, unlimited form)."""
    mu0 = mub2d + ckpt["mup_t"]
    cur = q
    for s, (ru, rv, ww, mupn, dt) in enumerate(ckpt["stages"]):
        tend = D.flux_div_scalar(cur, ru, rv, ww, rdnw, fnm, fnp, dx_inv,
                                 dy_inv)
        mu = mub2d + mupn
        from oracle.moist_np import coupled_update
        cur = coupled_update(q, tend, c1h, c2h, mu0, mu, dt,
                             clamp=clamp and s == 2)
    return cur


def tracer_step_adjoint(lam, ckpt, c1h, c2h, mub2d, rdnw, fnm, fnp,
                        dx_inv, dy_inv):
    """Transpose of tracer_step (no clamp active): stages in reverse."""
    c1, c2 = c1h[:, None, None], c2h[:, None, None]
    chm0 = c1 * (mub2d + ckpt["mup_t"])[None] + c2
    g_q = np.zeros_like(lam)
    g = lam
    for s in (2, 1, 0):
        ru, rv, ww, mupn, dt = ckpt["stages"][s]
        chm = c1 * (mub2d + mupn)[None] + c2
        mu = g / chm
        g_q = g_q + chm0 * mu
        g = dt * flux_div_scalar_adjoint(mu, ru, rv, ww, rdnw, fnm, fnp,
                                         dx_inv, dy_inv)
    return g_q + g
