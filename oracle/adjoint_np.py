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


# ---------------------------------------------------------------------------
# Acoustic substep adjoint
# ---------------------------------------------------------------------------
#
# One acoustic substep (reference acoustic.cu:160-982; oracle
# dycore_np.acoustic_substep: advance_uv -> advance_mu_th -> advance_w_phi
# -> diagnose_p) maps the small-step state
#   X = (u_pp, v_pp, w_pp, th_pp, ph_pp, mu_pp, p_pp, p_pp_old, al_pp)
# and the slow tendencies (ru_t, rv_t, rw_t, rth_t, rph_t, rmu_t) to the new
# X, for a frozen stage reference state.  It is affine in X except for the
# new column mass mu_t + mu'' in three denominators (the theta'' term of the
# w equation, the phi'' update and the linearised EOS), so its Jacobian
# depends on X; acoustic_substep_adjoint applies J^T at the given X.
# Pinned by the dot-product identity against complex-step derivatives of
# the oracle's own forward substep (tests/test_oracle_adjoint.py).

SUBSTEP_STATE = ("u_pp", "v_pp", "w_pp", "th_pp", "ph_pp", "mu_pp", "p_pp",
                 "p_pp_old", "al_pp")
SUBSTEP_TEND = ("ru_t", "rv_t", "rw_t", "rth_t", "rph_t", "rmu_t")


def _bt(x, ax):
    """Transpose of the periodic shift B(a)[i] = a[i-1]."""
    return np.roll(x, -1, axis=ax)


def _forward_pieces(s, coefs, dtau, c, first):
    """The substep's intermediate states (run on a copy of s)."""
    import copy
    f = s.dtype.type
    smdiv = 0.0 if first else c.smdiv
    d32 = float(f(dtau))
    s1 = copy.deepcopy(s)
    D.advance_uv(s1, d32, smdiv, c.dx, c.dy, c.top_lid)
    U1, V1 = s1.u_pp.copy(), s1.v_pp.copy()
    D.advance_mu_th(s1, d32, c.dx, c.dy)
    M1, T1, WW1 = s1.mu_pp.copy(), s1.th_pp.copy(), s1.ww_pp.copy()
    dampmag = float(f(dtau * c.dampcoef)) if c.damp_opt == 3 else 0.0
    D.advance_w_phi(s1, coefs, d32, c.epssm, c.dx, c.dy, c.top_lid,
                    dampmag, c.zdamp)
    return {"smdiv": smdiv, "dt": d32, "dampmag": dampmag, "U1": U1,
            "V1": V1, "M1": M1, "T1": T1, "WW1": WW1, "W1": s1.w_pp,
            "Ph1": s1.ph_pp, "AL1": s1.al_pp, "PP1": s1.p_pp}


def acoustic_substep_adjoint(s, coefs, dtau, c, first, lam, ref=False):
    """J^T lam for one acoustic substep at the state s (its small-step
    fields are the substep's input X).  ``lam`` maps output names
    (SUBSTEP_STATE, optionally ww_pp) to adjoints; returns the adjoints of
    the inputs (SUBSTEP_STATE) and of the tendencies (SUBSTEP_TEND).
    ``ref``: also the adjoints of the stage reference state the substep
    reads (out["ref"], keyed u, v, w, thp, php, mup), through the stage EOS
    (p, alt) and the implicit-w coefficients (coefs must be calc_coefs of
    s, plain form)."""
    fw = _forward_pieces(s, coefs, dtau, c, first)
    f = s.dtype.type
    half, one = f(0.5), f(1.0)
    dt, eps, g = f(fw["dt"]), f(c.epssm), f(D.G)
    rdx, rdy = f(1.0 / c.dx), f(1.0 / c.dy)
    nz, ny, nx = s.thp.shape
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    c1f, c2f = s.c1f[:, None, None], s.c2f[:, None, None]
    rdnw = s.rdnw[:, None, None]
    mut = s.mub2d + s.mup
    thr = D._b3(s.thb) + s.thp
    c2a, a, alpha, gam = (coefs[k] for k in ("c2a", "a", "alpha", "gam"))
    U1, V1, M1, T1 = fw["U1"], fw["V1"], fw["M1"], fw["T1"]
    W1, AL1 = fw["W1"], fw["AL1"]
    M0, T0 = s.mu_pp, s.th_pp

    z = lambda x: np.zeros_like(x)
    L = {k: np.array(lam.get(k, z(getattr(s, k))), copy=True)
         for k in SUBSTEP_STATE}
    out = {k: z(getattr(s, k)) for k in SUBSTEP_STATE + SUBSTEP_TEND}
    lWW1 = np.array(lam.get("ww_pp", z(s.ww_pp)), copy=True)
    lU1, lV1 = L["u_pp"], L["v_pp"]
    lM1, lT1 = L["mu_pp"], L["th_pp"]
    lM0old, lT0old = z(M0), z(T0)
    lPh0 = out["ph_pp"]
    out["p_pp"] += L["p_pp_old"]                # output p_pp_old = input p_pp
    if ref:                       # reference-state adjoints, accumulated
        R = {"mut": z(mut), "alt": z(s.alt), "c2a": z(s.alt), "thr": z(thr),
             "ph": z(s.w_pp), "w": z(s.w), "u": z(s.u), "v": z(s.v),
             "chm": z(s.alt), "cfm": z(s.w_pp)}

    # ---- diagnose_p ------------------------------------------------------
    chmp = c1h * (mut + M1) + c2h
    q = s.alt / (chmp * thr)
    X = T1 - c1h * M1 * thr
    lPP1 = L["p_pp"]
    lT1 += c2a * q * lPP1
    lM1 += np.sum(c2a * (-q * c1h * thr - q * X * c1h / chmp) * lPP1, axis=0)
    lAL = L["al_pp"] - c2a * lPP1
    lM1 += np.sum((-(s.alt * c1h) / chmp - AL1 * c1h / chmp) * lAL, axis=0)
    if ref:
        R["c2a"] += lPP1 * (q * X - AL1)
        gq = lPP1 * c2a * X
        R["alt"] += gq / (chmp * thr) - lAL * c1h * M1 / chmp
        gchmp = -gq * q / chmp - lAL * AL1 / chmp
        R["thr"] += -gq * q / thr + lPP1 * c2a * q * (-c1h * M1)
        R["mut"] += np.sum(c1h * gchmp, axis=0)
    lPh1 = L["ph_pp"]
    lPh1[1:] += -rdnw / chmp * lAL
    lPh1[:-1] += rdnw / chmp * lAL

    # ---- phi'' update: Ph1[k] = rhs[k] + K W1[k] / cfs[k] ----------------
    K = half * dt * g * (one + eps)
    cfs = c1f * (mut + M1) + c2f
    lrhs = np.zeros_like(s.w_pp)
    lrhs[1:] += lPh1[1:]
    lW = np.array(L["w_pp"], copy=True)
    lW[1:] += K / cfs[1:] * lPh1[1:]
    lM1 += np.sum(-K * W1[1:] * c1f[1:] / cfs[1:] ** 2 * lPh1[1:], axis=0)
    if ref:
        R["mut"] += np.sum(-K * W1[1:] * c1f[1:] / cfs[1:] ** 2 * lPh1[1:],
                           axis=0)
    lPh0[0] += lPh1[0]

    # ---- damper (damp_opt = 3): w = (w - dw cfm w_ref) / (1 + dw) -------
    cfm = c1f * mut + c2f
    wsol = np.array(W1, copy=True)               # the solve's output
    if fw["dampmag"] > 0.0:
        phf = (s.phb if s.base3d else s.phb[:, None, None]) + s.php
        htop = phf[nz] / g
        hbot = htop - f(c.zdamp)
        for k in range(1, nz + 1):
            hk = phf[k] / g
            arg = f(1.5707963) * (hk - hbot) / f(c.zdamp)
            sn = np.sin(arg.astype(np.float64)).astype(s.dtype)
            dw = f(fw["dampmag"]) * sn * sn
            on = hk >= hbot
            if ref:
                Cw = cfm[k] * s.w[k]
                wsol[k] = np.where(on, W1[k] * (one + dw) + dw * Cw, W1[k])
                lk = np.where(on, lW[k], 0.0)
                R["w"][k] += -lk * dw * cfm[k] / (one + dw)
                R["cfm"][k] += -lk * dw * s.w[k] / (one + dw)
                gdw = -lk * (Cw + wsol[k]) / (one + dw) ** 2
                garg = gdw * f(fw["dampmag"]) * 2.0 * sn * np.cos(arg)
                ghk = garg * f(1.5707963) / f(c.zdamp)
                R["ph"][k] += ghk / g
                R["ph"][nz] += -ghk / g
            lW[k] = np.where(on, lW[k] / (one + dw), lW[k])

    # ---- Thomas solve (rows 1..nz, w[0] coupled through a[1]) ------------
    n = nz
    for k in range(1, n):
        lW[k + 1] += -gam[k] * lW[k]
    for k in range(n, 0, -1):
        lW[k] = alpha[k] * lW[k]
        lW[k - 1] += -a[k] * lW[k]
    lacc = lW                                   # adjoint of the pre-solve w
    if ref:
        # A x = r with rows k = 1..nz: d(x) = -A^-1 dA x, so the entries'
        # adjoints are -mu x (mu = lacc); then a, b, c -> c2a, chm, cfm
        cof = K * K
        chm_ = c1h * mut + c2h
        rdn, rdnw1 = s.rdn, s.rdnw
        mu_, x_ = lacc, wsol
        for k in range(1, nz + 1):
            if (2 <= k <= nz - 1) or (k == nz and not c.top_lid):
                ga = -mu_[k] * x_[k - 1]
                if k < nz:
                    av = -cof * rdn[k] * rdnw1[k - 1] * c2a[k - 1] \
                        / (chm_[k - 1] * cfm[k - 1])
                else:
                    av = f(-2.0) * cof * rdnw1[nz - 1] * rdnw1[nz - 1] \
                        * c2a[nz - 1] / (chm_[nz - 1] * cfm[nz - 1])
                R["c2a"][k - 1] += ga * av / c2a[k - 1]
                R["chm"][k - 1] += -ga * av / chm_[k - 1]
                R["cfm"][k - 1] += -ga * av / cfm[k - 1]
            gb = -mu_[k] * x_[k]
            if k <= nz - 1:
                t1 = cof * rdn[k] * rdnw1[k] * c2a[k] / (chm_[k] * cfm[k])
                t2b = cof * rdn[k] * rdnw1[k - 1] * c2a[k - 1] \
                    / (chm_[k - 1] * cfm[k])
                R["c2a"][k] += gb * t1 / c2a[k]
                R["chm"][k] += -gb * t1 / chm_[k]
                R["cfm"][k] += -gb * (t1 + t2b) / cfm[k]
                R["c2a"][k - 1] += gb * t2b / c2a[k - 1]
                R["chm"][k - 1] += -gb * t2b / chm_[k - 1]
                gc = -mu_[k] * x_[k + 1]
                cv = -cof * rdn[k] * rdnw1[k] * c2a[k] / (chm_[k] * cfm[k + 1])
                R["c2a"][k] += gc * cv / c2a[k]
                R["chm"][k] += -gc * cv / chm_[k]
                R["cfm"][k + 1] += -gc * cv / cfm[k + 1]
            else:
                t1 = f(2.0) * cof * rdnw1[nz - 1] * rdnw1[nz - 1] \
                    * c2a[nz - 1] / (chm_[nz - 1] * cfm[nz])
                R["c2a"][nz - 1] += gb * t1 / c2a[nz - 1]
                R["chm"][nz - 1] += -gb * t1 / chm_[nz - 1]
                R["cfm"][nz] += -gb * t1 / cfm[nz]
        # the phi'' right-hand side of the forward (for dph_up / dph_dn)
        ph_ = (s.phb if s.base3d else s.phb[:, None, None]) + s.php
        wd_ = half * (fw["WW1"][1:] + fw["WW1"][:-1]) * rdnw \
            * (ph_[1:] - ph_[:-1])
        r_ = np.zeros_like(s.w_pp)
        r_[1:] = dt * (s.rph_t[1:] + half * g * (one - eps) * s.w_pp[1:])
        for k in range(1, nz):
            r_[k] = r_[k] - dt * (s.fnm[k] * wd_[k] + s.fnp[k] * wd_[k - 1])
        rhs_ = np.zeros_like(s.w_pp)
        rhs_[1:] = s.ph_pp[1:] + r_[1:] / cfm[1:]
        if c.top_lid:
            rhs_[nz] = 0.0

    # ---- w accumulation ---------------------------------------------------
    chm = c1h * mut + c2h
    lt2 = np.zeros_like(s.thp)
    lmuave = np.zeros_like(mut)
    lW0 = out["w_pp"]
    s_rdn = s.rdn
    for k in range(1, nz):
        la = lacc[k]
        lW0[k] += la
        out["rw_t"][k] += dt * la
        cup = half * dt * g * s_rdn[k] * c2a[k] * s.rdnw[k] / chm[k]
        cdn = half * dt * g * s_rdn[k] * c2a[k - 1] * s.rdnw[k - 1] / chm[k - 1]
        lup, ldn = cup * la, -cdn * la
        lrhs[k + 1] += (one + eps) * lup
        lrhs[k] -= (one + eps) * lup
        lPh0[k + 1] += (one - eps) * lup
        lPh0[k] -= (one - eps) * lup
        lrhs[k] += (one + eps) * ldn
        lrhs[k - 1] -= (one + eps) * ldn
        lPh0[k] += (one - eps) * ldn
        lPh0[k - 1] -= (one - eps) * ldn
        lt2[k] += dt * g * s_rdn[k] * c2a[k] * s.alt[k] * la
        lt2[k - 1] -= dt * g * s_rdn[k] * c2a[k - 1] * s.alt[k - 1] * la
        lmuave += -dt * g * s.c1f[k] * la
        if ref:
            P0 = s.ph_pp
            dup = (one + eps) * (rhs_[k + 1] - rhs_[k]) \
                + (one - eps) * (P0[k + 1] - P0[k])
            ddn = (one + eps) * (rhs_[k] - rhs_[k - 1]) \
                + (one - eps) * (P0[k] - P0[k - 1])
            R["c2a"][k] += la * dup * cup / c2a[k]
            R["chm"][k] += -la * dup * cup / chm[k]
            R["c2a"][k - 1] += -la * ddn * cdn / c2a[k - 1]
            R["chm"][k - 1] += la * ddn * cdn / chm[k - 1]
            t2v = half * ((one + eps) * fw["T1"] + (one - eps) * T0) \
                / ((c1h * (mut + M1) + c2h) * thr)
            R["c2a"][k] += la * dt * g * s_rdn[k] * s.alt[k] * t2v[k]
            R["alt"][k] += la * dt * g * s_rdn[k] * c2a[k] * t2v[k]
            R["c2a"][k - 1] -= la * dt * g * s_rdn[k] * s.alt[k - 1] * t2v[k - 1]
            R["alt"][k - 1] -= la * dt * g * s_rdn[k] * c2a[k - 1] * t2v[k - 1]
    if not c.top_lid:
        la = lacc[nz]
        lW0[nz] += la
        out["rw_t"][nz] += dt * la
        # bterm = dt g c2a rdnw^2 / chm * dph_dn(nz)
        cb = dt * g * c2a[nz - 1] * s.rdnw[nz - 1] * s.rdnw[nz - 1] \
            / chm[nz - 1]
        ldn = -cb * la
        lrhs[nz] += (one + eps) * ldn
        lrhs[nz - 1] -= (one + eps) * ldn
        lPh0[nz] += (one - eps) * ldn
        lPh0[nz - 1] -= (one - eps) * ldn
        lt2[nz - 1] -= dt * g * f(2.0) * s.rdnw[nz - 1] * c2a[nz - 1] \
            * s.alt[nz - 1] * la
        lmuave += -dt * g * s.c1f[nz] * la
        if ref:
            P0 = s.ph_pp
            ddn = (one + eps) * (rhs_[nz] - rhs_[nz - 1]) \
                + (one - eps) * (P0[nz] - P0[nz - 1])
            bt = cb * ddn
            R["c2a"][nz - 1] += -la * bt / c2a[nz - 1]
            R["chm"][nz - 1] += la * bt / chm[nz - 1]
            t2v = half * ((one + eps) * fw["T1"] + (one - eps) * T0) \
                / ((c1h * (mut + M1) + c2h) * thr)
            R["c2a"][nz - 1] += -la * dt * g * f(2.0) * s.rdnw[nz - 1] \
                * s.alt[nz - 1] * t2v[nz - 1]
            R["alt"][nz - 1] += -la * dt * g * f(2.0) * s.rdnw[nz - 1] \
                * c2a[nz - 1] * t2v[nz - 1]
    lrhs[0] = 0.0                                # rhs[0] is the constant 0

    # ---- lower boundary w[0] from u'', v'' ---------------------------------
    l0 = lacc[0]
    ht = s.ht
    lvn = half * rdy * (np.roll(ht, -1, 0) - ht) * l0
    lvs = half * rdy * (ht - np.roll(ht, 1, 0)) * l0
    lue = half * rdx * (np.roll(ht, -1, 1) - ht) * l0
    luw = half * rdx * (ht - np.roll(ht, 1, 1)) * l0
    for lev, cf in enumerate((s.cf1, s.cf2, s.cf3)):
        lU1[lev, :, 1:] += cf * lue
        lU1[lev, :, :-1] += cf * luw
        lV1[lev, 1:, :] += cf * lvn
        lV1[lev, :-1, :] += cf * lvs

    # ---- t2 and muave -----------------------------------------------------
    chs = c1h * (mut + M1) + c2h
    t2 = half * ((one + eps) * T1 + (one - eps) * T0) / (chs * thr)
    lT1 += half * (one + eps) / (chs * thr) * lt2
    lT0old += half * (one - eps) / (chs * thr) * lt2
    lM1 += np.sum(-t2 * c1h / chs * lt2, axis=0)
    if ref:
        R["mut"] += np.sum(-t2 * c1h / chs * lt2, axis=0)
        R["thr"] += -t2 / thr * lt2
    lM1 += half * (one + eps) * lmuave
    lM0old += half * (one - eps) * lmuave

    # ---- rhs --------------------------------------------------------------
    if c.top_lid:
        lrhs[nz] = 0.0                            # rhs[nz] reset to 0
    lPh0[1:] += lrhs[1:]
    lr2 = lrhs[1:] / cfm[1:]                      # levels 1..nz
    out["rph_t"][1:] += dt * lr2
    lW0[1:] += dt * half * g * (one - eps) * lr2
    ph = (s.phb if s.base3d else s.phb[:, None, None]) + s.php
    lwd = np.zeros_like(s.thp)
    for k in range(1, nz):
        lwd[k] += -dt * s.fnm[k] * lr2[k - 1]
        lwd[k - 1] += -dt * s.fnp[k] * lr2[k - 1]
    cwd = half * rdnw * (ph[1:] - ph[:-1])
    lWW1[1:] += cwd * lwd
    lWW1[:-1] += cwd * lwd
    if ref:
        R["cfm"][1:] += -lrhs[1:] * r_[1:] / cfm[1:] ** 2
        gdp = lwd * half * (fw["WW1"][1:] + fw["WW1"][:-1]) * rdnw
        R["ph"][1:] += gdp
        R["ph"][:-1] -= gdp

    # ---- advance_mu_th ------------------------------------------------------
    lT0 = out["th_pp"]
    lT0 += lT1
    out["rth_t"] += dt * lT1
    gg = -dt * lT1
    th = thr
    lwdtn = np.zeros_like(s.w_pp)
    lwdtn[1:] += rdnw * gg
    lwdtn[:-1] -= rdnw * gg
    t_ip, t_im = np.roll(th, -1, axis=2), np.roll(th, 1, axis=2)
    t_jp, t_jm = np.roll(th, -1, axis=1), np.roll(th, 1, axis=1)
    lU1[..., 1:] += half * rdx * (t_ip + th) * gg
    lU1[..., :-1] -= half * rdx * (th + t_im) * gg
    lV1[:, 1:, :] += half * rdy * (t_jp + th) * gg
    lV1[:, :-1, :] -= half * rdy * (th + t_jm) * gg
    if ref:
        WW1_ = fw["WW1"]
        for k in range(1, nz):
            R["thr"][k] += lwdtn[k] * WW1_[k] * s.fnm[k]
            R["thr"][k - 1] += lwdtn[k] * WW1_[k] * s.fnp[k]
        ue_, uw_ = U1[..., 1:], U1[..., :-1]
        vn_, vs_ = V1[:, 1:, :], V1[:, :-1, :]
        R["thr"] += half * rdx * (ue_ - uw_) * gg \
            + half * rdy * (vn_ - vs_) * gg
        R["thr"] += np.roll(half * rdx * ue_ * gg, 1, axis=2) \
            - np.roll(half * rdx * uw_ * gg, -1, axis=2) \
            + np.roll(half * rdy * vn_ * gg, 1, axis=1) \
            - np.roll(half * rdy * vs_ * gg, -1, axis=1)
    lww = np.array(lWW1, copy=True)
    for k in range(1, nz):
        lww[k] += (s.fnm[k] * th[k] + s.fnp[k] * th[k - 1]) * lwdtn[k]
    lww[0] = 0.0
    lww[nz] = 0.0
    dnw = s.dnw
    linc = np.zeros_like(s.thp)
    acc = np.zeros_like(mut)
    for k in range(nz - 2, -1, -1):               # suffix sums of lww[k+1:nz]
        acc = acc + lww[k + 1]
        linc[k] = -acc
    ldmdt = np.zeros_like(mut)
    ldvp = np.zeros_like(s.thp)
    for k in range(nz):
        ldmdt += dnw[k] * s.c1h[k] * linc[k]
        out["rmu_t"] += dnw[k] * s.c1h[k] * linc[k]
        ldvp[k] += dnw[k] * linc[k]
    lM0 = out["mu_pp"]
    lM0 += lM1
    ldmdt += dt * lM1
    if ref:                       # dmdt_rf = sum_k dnw dvr (reference u, v)
        grf = dt * lM1
        for ax, rdd, xn in ((2, rdx, "u"), (1, rdy, "v")):
            x = getattr(s, xn)
            hi_ = (slice(None), slice(None), slice(1, None)) if ax == 2 \
                else (slice(None), slice(1, None), slice(None))
            lo_ = (slice(None), slice(None), slice(0, -1)) if ax == 2 \
                else (slice(None), slice(0, -1), slice(None))
            mhi = half * (mut + np.roll(mut, -1, axis=ax - 1))
            mlo = half * (np.roll(mut, 1, axis=ax - 1) + mut)
            gdv = dnw[:, None, None] * grf[None]
            R[xn][hi_] += rdd * (c1h * mhi + c2h) * gdv
            R[xn][lo_] -= rdd * (c1h * mlo + c2h) * gdv
            gmh = np.sum(rdd * c1h * x[hi_] * gdv, axis=0)
            gml = -np.sum(rdd * c1h * x[lo_] * gdv, axis=0)
            R["mut"] += half * (gmh + np.roll(gmh, 1, axis=ax - 1)) \
                + half * (gml + np.roll(gml, -1, axis=ax - 1))
    out["rmu_t"] += dt * lM1
    for k in range(nz):
        ldvp[k] += dnw[k] * ldmdt
    lU1[..., 1:] += rdx * ldvp
    lU1[..., :-1] -= rdx * ldvp
    lV1[:, 1:, :] += rdy * ldvp
    lV1[:, :-1, :] -= rdy * ldvp
    lM0 += lM0old
    lT0 += lT0old

    # ---- advance_uv -------------------------------------------------------
    smdiv = f(fw["smdiv"])
    lpd = np.zeros_like(s.p_pp)
    lal = out["al_pp"]
    for ax, rd, lpp, rt_name in ((-1, rdx, lU1, "ru_t"),
                                 (-2, rdy, lV1, "rv_t")):
        B = lambda x: np.roll(x, 1, axis=ax)
        out["u_pp" if ax == -1 else "v_pp"] += lpp
        out[rt_name] += dt * lpp
        if ax == -1:
            ld = -dt * lpp[..., :nx]
            ld[..., 0] += -dt * lpp[..., nx]
        else:
            ld = -dt * lpp[:, :ny, :]
            ld[:, 0, :] += -dt * lpp[:, ny, :]
        muf = half * (B(mut) + mut)
        cA = half * rd * (c1h * muf + c2h)
        lbrk = cA * ld
        phs = s.php[:-1] + s.php[1:]
        dphp = half * (phs - B(phs))
        if s.base3d:
            pbs = s.phb[:-1] + s.phb[1:]
            dphp = dphp + half * (pbs - B(pbs))
        t = rd * dphp * ld
        ldpn = np.zeros_like(s.w_pp)
        ldpn[1:] += rdnw * t
        ldpn[:-1] -= rdnw * t
        lms = np.sum(-half * c1h * t, axis=0)
        lM0 += lms + _bt(lms, ax)
        if ref:
            pd_ = D._pdmp(s.p_pp, s.p_pp_old, smdiv)
            dph_ = (s.ph_pp[1:] - B(s.ph_pp[1:])) + (s.ph_pp[:-1]
                                                     - B(s.ph_pp[:-1]))
            dpe_ = pd_ - B(pd_)
            brk_ = dph_ + (s.alt + B(s.alt)) * dpe_
            if s.base3d:
                brk_ = brk_ + (s.al_pp + B(s.al_pp)) * (s.pb - B(s.pb))
            gmf = np.sum(half * rd * c1h * brk_ * ld, axis=0)
            R["mut"] += half * (gmf + _bt(gmf, ax))
            xa = lbrk * dpe_
            R["alt"] += xa + _bt(xa, ax)
            dpn_ = D._dpn_face_levels(s, pd_, ax, c.top_lid)
            gdphp = rd * (rdnw * (dpn_[1:] - dpn_[:-1])
                          - half * c1h * (B(s.mu_pp) + s.mu_pp)) * ld
            gphs = half * (gdphp - _bt(gdphp, ax))
            R["ph"][:-1] += gphs
            R["ph"][1:] += gphs
        lPh0[1:] += lbrk - _bt(lbrk, ax)
        lPh0[:-1] += lbrk - _bt(lbrk, ax)
        ldpe = (s.alt + B(s.alt)) * lbrk
        lpd += ldpe - _bt(ldpe, ax)
        if s.base3d:
            x = (s.pb - B(s.pb)) * lbrk
            lal += x + _bt(x, ax)
        lS = np.zeros_like(s.p_pp)
        lS[0] += half * s.cf1 * ldpn[0]
        lS[1] += half * s.cf2 * ldpn[0]
        lS[2] += half * s.cf3 * ldpn[0]
        for kf in range(1, nz):
            lS[kf] += half * s.fnm[kf] * ldpn[kf]
            lS[kf - 1] += half * s.fnp[kf] * ldpn[kf]
        if c.top_lid:
            lS[nz - 1] += half * s.cf1 * ldpn[nz]
            lS[nz - 2] += half * s.cf2 * ldpn[nz]
            lS[nz - 3] += half * s.cf3 * ldpn[nz]
        lpd += lS + _bt(lS, ax)
    out["p_pp"] += (one + smdiv) * lpd
    out["p_pp_old"] += -smdiv * lpd
    if ref:
        # chm, cfm -> mut; c2a -> p, alt; the stage EOS -> thp, php, mup
        R["mut"] += np.sum(c1h * R["chm"], axis=0) \
            + np.sum(c1f * R["cfm"], axis=0)
        gp = R["c2a"] * GAMMA / s.alt
        galt = R["alt"] - R["c2a"] * c2a / s.alt
        gt, gph, gm = eos_adjoint(s, gp, np.zeros_like(s.al), galt)
        out["ref"] = {"u": R["u"], "v": R["v"], "w": R["w"],
                      "thp": R["thr"] + gt, "php": R["ph"] + gph,
                      "mup": R["mut"] + gm}
        if s.qv is not None:
            out["ref"]["qv"] = eos_q_adjoint(s, gp)
    return out


# ---------------------------------------------------------------------------
# Flux-form advection operators: brute-force Jacobians (tiny grids)
# ---------------------------------------------------------------------------
#
# The operators D(q; ru, rv, rw) of dycore_np (flux_div_scalar/u/v/w) are
# bilinear in the advected field and the mass fluxes except for |vel| in the
# upwind term.  Away from vel = 0 that is sign(vel) vel, analytic, so the
# complex-step derivative of D with respect to every input element (one
# forward call per element) builds the exact Jacobian on a tiny grid; its
# transpose is the reference for the hand-written device adjoints.

def _flux5_cs(qm3, qm2, qm1, q0, qp1, qp2, vel):
    sv = np.sign(np.real(vel))
    return (vel * (37.0 * (q0 + qm1) - 8.0 * (qp1 + qm2) + (qp2 + qm3))
            - sv * vel * (10.0 * (q0 - qm1) - 5.0 * (qp1 - qm2)
                          + (qp2 - qm3))) / 60.0


def _flux3_cs(qm2, qm1, q0, qp1, vel):
    sv = np.sign(np.real(vel))
    return (vel * (7.0 * (q0 + qm1) - (qp1 + qm2))
            + sv * vel * (3.0 * (q0 - qm1) - (qp1 - qm2))) / 12.0


def flux_div_jacobians(op, q, ru, rv, rw, rz, fnm, fnp, dx_inv, dy_inv,
                       h=1e-30):
    """{"q": J_q, "ru": J_ru, "rv": J_rv, "rw": J_rw}, each (n_tend, n_in),
    of D = dycore_np.flux_div_<op> at (q, ru, rv, rw), by complex steps."""
    fn = {"scalar": D.flux_div_scalar, "u": D.flux_div_u,
          "v": D.flux_div_v, "w": D.flux_div_w}[op]
    args = {"q": q, "ru": ru, "rv": rv, "rw": rw}
    saved = (D.flux5, D.flux3)
    D.flux5, D.flux3 = _flux5_cs, _flux3_cs
    try:
        base = fn(q, ru, rv, rw, rz, fnm, fnp, dx_inv, dy_inv)
        out = {}
        for name, arr in args.items():
            cols = []
            for idx in range(arr.size):
                x = {k: v.astype(np.complex128) for k, v in args.items()}
                x[name].reshape(-1)[idx] += 1j * h
                t = fn(x["q"], x["ru"], x["rv"], x["rw"], rz, fnm, fnp,
                       dx_inv, dy_inv)
                cols.append(np.imag(t).reshape(-1) / h)
            out[name] = np.stack(cols, axis=1)
    finally:
        D.flux5, D.flux3 = saved
    out["tend"] = base
    return out


# ---------------------------------------------------------------------------
# Hand adjoints of the advection operators (field and mass fluxes)
# ---------------------------------------------------------------------------
#
# Every operator is a sum over face fluxes F_f = sum_o w_o(vel_f) q[s_f + o]
# with w_o(v) = C_o v + S_o |v| (flux5 / flux3) or a 2nd-order average, and
# vel_f a fixed linear map of the mass fluxes.  With Lambda_f the adjoint of
# F_f (the difference of the adjacent cells' output adjoints, scaled), the
# transpose is
#   dL/dq[c]    += sum over faces reading c of w_o(vel_f) Lambda_f,
#   dL/dvel_f    = Lambda_f sum_o (C_o + S_o sign(vel_f)) q[s_f + o],
# and dL/dru etc. gather dL/dvel_f through the transpose of the velocity
# averages.  sign(0) = 0 (a subgradient of |v|).

C5 = np.array([1.0, -8.0, 37.0, 37.0, -8.0, 1.0]) / 60.0
S5 = np.array([1.0, -5.0, 10.0, -10.0, 5.0, -1.0]) / 60.0
C3 = np.array([-1.0, 7.0, 7.0, -1.0]) / 12.0
S3 = np.array([1.0, -3.0, 3.0, -1.0]) / 12.0


def _h_adjoint(q, vel, lam, ax, rd, dup):
    """One horizontal direction of flux5 transport along axis ax.  Cells
    c = 0..n-1 (periodic); face f lies between cells f-1 and f.  dup: faces
    0..n with face n a separate copy of face 0 (its own velocity, its flux
    entering only cell n-1); else n periodic faces.  Returns (dL/dq,
    dL/dvel)."""
    if dup:
        z = np.zeros_like(np.take(lam, [0], axis=ax))
        L = rd * (np.concatenate([lam, z], axis=ax)
                  - np.concatenate([z, lam], axis=ax))
    else:
        L = rd * (lam - np.roll(lam, 1, axis=ax))
    return _h_faces_adjoint(q, vel, L, ax, dup)


def _h_faces_adjoint(q, vel, L, ax, dup):
    """Transpose of the flux5 face fluxes along axis ax for the face
    adjoints L (same layout as vel): (dL/dq, dL/dvel)."""
    n = q.shape[ax]
    a, sv = np.abs(vel), np.sign(vel)
    gq = np.zeros_like(q)
    dv = np.zeros_like(vel)
    for o in range(6):                       # face f reads cell f + o - 3
        c = (C5[o] * vel + S5[o] * a) * L
        if dup:
            c = np.take(c, range(n), axis=ax) + np.where(
                np.arange(n).reshape([-1 if d == (ax % q.ndim) else 1
                                      for d in range(q.ndim)]) == 0,
                np.take(c, [n], axis=ax), 0.0)
        gq += np.roll(c, o - 3, axis=ax)
        qs = np.roll(q, 3 - o, axis=ax)
        if dup:
            qs = np.concatenate([qs, np.take(qs, [0], axis=ax)], axis=ax)
        dv += (C5[o] + S5[o] * sv) * qs
    return gq, dv * L


def _zhalf_adjoint(q, vel, lam, rdz, fnm, fnp):
    """Vertical transport of a half-level field (faces kf = 1..nz-1, 2nd
    order one face in from each boundary, flux3 inside)."""
    nz = q.shape[0]
    L = np.zeros_like(vel)
    for kf in range(1, nz):
        L[kf] = rdz[kf] * lam[kf] - rdz[kf - 1] * lam[kf - 1]
    return _zhalf_faces_adjoint(q, vel, L, fnm, fnp)


def _zhalf_faces_adjoint(q, vel, L, fnm, fnp):
    """Transpose of the half-level vertical face fluxes (dycore_np.
    _zfaces_half) for the face adjoints L (w levels; 0 and nz unused)."""
    nz = q.shape[0]
    gq = np.zeros_like(q)
    dv = np.zeros_like(vel)
    for kf in range(1, nz):
        v = vel[kf]
        if kf == 1 or kf == nz - 1:
            gq[kf] += v * fnm[kf] * L[kf]
            gq[kf - 1] += v * fnp[kf] * L[kf]
            dv[kf] = L[kf] * (fnm[kf] * q[kf] + fnp[kf] * q[kf - 1])
        else:
            for o in range(4):               # face kf reads cell kf + o - 2
                gq[kf + o - 2] += (C3[o] * v + S3[o] * np.abs(v)) * L[kf]
                dv[kf] += L[kf] * (C3[o] + S3[o] * np.sign(v)) * q[kf + o - 2]
    return gq, dv


def _zfull_adjoint(w, vel, lam, rdn):
    """Vertical transport of w: faces at mass levels m = 0..nz-1 between
    w[m] and w[m+1] (2nd order at m = 0 and nz-1, flux3 inside); the
    output rows 0 and nz are not part of the operator."""
    nz = w.shape[0] - 1
    gw = np.zeros_like(w)
    dv = np.zeros_like(vel)
    for m in range(nz):
        L = -(rdn[m] * lam[m]) if m >= 1 else np.zeros_like(lam[0])
        if m + 1 <= nz - 1:
            L = L + rdn[m + 1] * lam[m + 1]
        v = vel[m]
        if m == 0 or m == nz - 1:
            gw[m] += 0.5 * v * L
            gw[m + 1] += 0.5 * v * L
            dv[m] = L * 0.5 * (w[m + 1] + w[m])
        else:
            for o in range(4):               # face m reads w[m + o - 1]
                gw[m + o - 1] += (C3[o] * v + S3[o] * np.abs(v)) * L
                dv[m] += L * (C3[o] + S3[o] * np.sign(v)) * w[m + o - 1]
    return gw, dv


def flux_div_adjoint(op, lam, q, ru, rv, rw, rdz, fnm, fnp, dx_inv,
                     dy_inv):
    """Transpose of the Jacobian of dycore_np.flux_div_<op> at (q, ru, rv,
    rw) applied to lam (an array shaped like the operator's output):
    {"q", "ru", "rv", "rw"} gradients (reference advection.cu:206-1051)."""
    g = {"ru": np.zeros_like(ru), "rv": np.zeros_like(rv),
         "rw": np.zeros_like(rw)}
    half = 0.5
    if op == "scalar":
        gx, g["ru"] = _h_adjoint(q, ru, lam, -1, dx_inv, True)
        gy, g["rv"] = _h_adjoint(q, rv, lam, -2, dy_inv, True)
        gz, g["rw"] = _zhalf_adjoint(q, rw, lam, rdz, fnm, fnp)
        g["q"] = gx + gy + gz
        return g
    if op == "u":
        nx = q.shape[-1] - 1
        le = lam[..., :nx].copy()
        le[..., 0] += lam[..., nx]
        uc = q[..., :nx]
        velx = half * (ru[..., :nx] + ru[..., 1:])      # mass centre m
        gx, dvp = _h_adjoint(uc, np.roll(velx, 1, axis=-1), le, -1,
                             dx_inv, False)
        dvx = np.roll(dvp, -1, axis=-1)
        g["ru"][..., :nx] += half * dvx
        g["ru"][..., 1:] += half * dvx
        vely = half * (np.roll(rv, 1, axis=-1) + rv)
        gy, dvy = _h_adjoint(uc, vely, le, -2, dy_inv, True)
        g["rv"] += half * (dvy + np.roll(dvy, -1, axis=-1))
        velz = half * (np.roll(rw, 1, axis=-1) + rw)
        gz, dvz = _zhalf_adjoint(uc, velz, le, rdz, fnm, fnp)
        g["rw"] += half * (dvz + np.roll(dvz, -1, axis=-1))
        g["q"] = np.zeros_like(q)
        g["q"][..., :nx] = gx + gy + gz
        return g
    if op == "v":
        ny = q.shape[-2] - 1
        le = lam[:, :ny, :].copy()
        le[:, 0, :] += lam[:, ny, :]
        vc = q[:, :ny, :]
        velx = half * (np.roll(ru, 1, axis=-2) + ru)
        gx, dvx = _h_adjoint(vc, velx, le, -1, dx_inv, True)
        g["ru"] += half * (dvx + np.roll(dvx, -1, axis=-2))
        vely = half * (rv[:, :ny, :] + rv[:, 1:, :])
        gy, dvp = _h_adjoint(vc, np.roll(vely, 1, axis=-2), le, -2,
                             dy_inv, False)
        dvy = np.roll(dvp, -1, axis=-2)
        g["rv"][:, :ny] += half * dvy
        g["rv"][:, 1:] += half * dvy
        velz = half * (np.roll(rw, 1, axis=-2) + rw)
        gz, dvz = _zhalf_adjoint(vc, velz, le, rdz, fnm, fnp)
        g["rw"] += half * (dvz + np.roll(dvz, -1, axis=-2))
        g["q"] = np.zeros_like(q)
        g["q"][:, :ny] = gx + gy + gz
        return g
    if op == "w":
        nz = q.shape[0] - 1
        gq = np.zeros_like(q)
        for k in range(1, nz):
            velx = fnm[k] * ru[k] + fnp[k] * ru[k - 1]
            gx, dvx = _h_adjoint(q[k], velx, lam[k], -1, dx_inv, True)
            vely = fnm[k] * rv[k] + fnp[k] * rv[k - 1]
            gy, dvy = _h_adjoint(q[k], vely, lam[k], -2, dy_inv, True)
            gq[k] += gx + gy
            g["ru"][k] += fnm[k] * dvx
            g["ru"][k - 1] += fnp[k] * dvx
            g["rv"][k] += fnm[k] * dvy
            g["rv"][k - 1] += fnp[k] * dvy
        velz = half * (rw[:-1] + rw[1:])
        gz, dvz = _zfull_adjoint(q, velz, lam, rdz)
        g["rw"][:-1] += half * dvz
        g["rw"][1:] += half * dvz
        g["q"] = gq + gz
        return g
    raise ValueError(op)


# ---------------------------------------------------------------------------
# The large-step tendency pass: EOS -> coupled fluxes and Omega ->
# advection + PGF + buoyancy + rhs_ph  (dry, periodic, msf = 1, no
# diffusion), and its hand-written transpose
# ---------------------------------------------------------------------------

STAGE_IN = ("u", "v", "w", "thp", "php", "mup")
STAGE_OUT = ("ru_t", "rv_t", "rw_t", "rth_t", "rph_t")


def stage_tendencies(s, dx, dy, top_lid=True, diff=None):
    """The slow tendencies of one RK stage at the state s (reference
    dycore.py:2378-2389: calc_p_alpha, stage_fluxes, _add_slow_tendencies,
    and with ``diff`` = (khdif, kvdif) the constant-K diffusion).  Mutates
    s's diagnostics and tendencies; returns {name: tendency} and the stage
    fluxes (ru, rv, ww)."""
    for name in STAGE_OUT + ("rmu_t",):
        getattr(s, name)[...] = 0.0
    D.calc_p_alpha(s)
    ru, rv, ww = D.stage_fluxes(s, dx, dy)
    D.add_slow_tendencies(s, ru, rv, ww, dx, dy, top_lid)
    if diff is not None:
        D.add_diffusion_tendencies(s, diff[0], diff[1], dx, dy)
    return {n: getattr(s, n).copy() for n in STAGE_OUT}, (ru, rv, ww)


def diff2_adjoint(lam, kh, kv, dx, dy, zf, stagger):
    """Transpose of D.add_diff2 (the constant-K Laplacian on the periodic
    core, the duplicate face copying the first, w's end levels zeroed) for
    lam of its output: the horizontal 3-point operator is symmetric; the
    vertical one maps mu_k = kv lam_k rdzc_k to rdzf_(k-1) (mu_(k-1) -
    mu_k) + rdzf_k (mu_(k+1) - mu_k)."""
    nlev, nys, nxs = lam.shape
    nx = nxs - 1 if stagger == "x" else nxs
    ny = nys - 1 if stagger == "y" else nys
    rdzf, rdzc = D._diff2_spacings(zf, stagger)
    L = np.array(lam[:, :ny, :nx], copy=True)
    if nxs > nx:
        L[:, :, 0] += lam[:, :ny, nx]
    if nys > ny:
        L[:, 0, :] += lam[:, ny, :nx]
    if stagger == "z":
        L[0] = 0.0
        L[-1] = 0.0
    g = kh * ((1.0 / dx ** 2) * (np.roll(L, 1, -1) - 2.0 * L
                                 + np.roll(L, -1, -1))
              + (1.0 / dy ** 2) * (np.roll(L, 1, -2) - 2.0 * L
                                   + np.roll(L, -1, -2)))
    mu = kv * L * rdzc[:, None, None]
    g[1:] += rdzf[:, None, None] * (mu[:-1] - mu[1:])
    g[:-1] += rdzf[:, None, None] * (mu[1:] - mu[:-1])
    out = np.zeros_like(lam)
    out[:, :ny, :nx] = g
    return out


def diffusion_adjoint(s, lam, kh, kv, dx, dy):
    """(u, v, w, thp, mup) gradients of D.add_diffusion_tendencies (flat
    base; reference diffusion.py:100-130) for lam over ru_t, rv_t, rw_t,
    rth_t: tend += C(mu_face) diff2(field)."""
    zf = s.phb.astype(np.float64) / D.G
    mu = s.mub2d + s.mup
    mux, muy = D.mu_at_u_faces(mu), D.mu_at_v_faces(mu)
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    c1f, c2f = s.c1f[:, None, None], s.c2f[:, None, None]
    g = {}
    g_mu = np.zeros_like(mu)
    for name, tname, muf, c1, c2, stag, ax in (
            ("u", "ru_t", mux, c1h, c2h, "x", -1),
            ("v", "rv_t", muy, c1h, c2h, "y", -2),
            ("w", "rw_t", mu, c1f, c2f, "z", None),
            ("thp", "rth_t", mu, c1h, c2h, "", None)):
        lt = lam[tname]
        fld = getattr(s, name)
        tmp = D.add_diff2(fld, kh, kv, dx, dy, zf, stag, s.dtype)
        g[name] = diff2_adjoint((c1 * muf[None] + c2) * lt, kh, kv, dx, dy,
                                zf, stag)
        gm = np.sum(c1 * tmp * lt, axis=0)
        g_mu += gm if ax is None else _faces_to_mass(gm, ax)
    g["mup"] = g_mu
    return g


def stage_tendencies_cs(s, dx, dy, top_lid=True):
    """stage_tendencies with the complex-step-safe flux5/flux3 (|v| as
    sign(Re v) v), for complex states."""
    saved = (D.flux5, D.flux3)
    D.flux5, D.flux3 = _flux5_cs, _flux3_cs
    try:
        return stage_tendencies(s, dx, dy, top_lid)[0]
    finally:
        D.flux5, D.flux3 = saved


def _face_fold(x, ax):
    """A face array (n+1 along ax) folded onto the n periodic faces: the
    duplicate face n adds into face 0."""
    n = x.shape[ax] - 1
    out = np.take(x, range(n), axis=ax).copy()
    idx = [slice(None)] * x.ndim
    idx[ax] = 0
    src = [slice(None)] * x.ndim
    src[ax] = n
    out[tuple(idx)] += x[tuple(src)]
    return out


def eos_adjoint(s, g_p, g_al, g_alt):
    """(thp, php, mup) gradients of calc_p_alpha (dry, hypso = 1) given
    those of p, al, alt (reference diagnostics.cu:37-89)."""
    mu = s.mub2d + s.mup
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    th = D._b3(s.thb) + s.thp
    mh = c1h * mu + c2h
    ga = g_alt + g_al - g_p * (GAMMA * s.p / s.alt)
    g_thp = g_p * (GAMMA * s.p / th)
    g_dph = ga * (-s.rdnw[:, None, None] / mh)
    g_php = np.zeros_like(s.php)
    g_php[1:] += g_dph
    g_php[:-1] -= g_dph
    g_mup = np.sum(ga * (-s.alt * c1h / mh), axis=0)
    return g_thp, g_php, g_mup


GAMMA = D.GAMMA
SPECIES = ("qv", "qc", "qr")


def eos_q_adjoint(s, g_p):
    """qv gradient of the moist EOS (calc_p_alpha with theta_m = theta
    (1 + Rv/Rd qv), reference diagnostics.cu:71) given that of p: dp/dqv =
    gamma p / theta_m * theta Rv/Rd.  The theta' gradient of the moist EOS
    is eos_adjoint's gamma p / theta unchanged."""
    th = D._b3(s.thb) + s.thp
    thm = th * (1.0 + D.RVOVRD * s.qv)
    return g_p * (GAMMA * s.p / thm) * th * D.RVOVRD


def coupling_adjoint(s, g_ru, g_rv):
    """(u, v, mup) gradients of ru = (c1h mu_u + c2h) u, rv likewise
    (reference dycore.py:162-206, state.py:339-356)."""
    mu = s.mub2d + s.mup
    mux, muy = D.mu_at_u_faces(mu), D.mu_at_v_faces(mu)
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    g_u = (c1h * mux[None] + c2h) * g_ru
    g_v = (c1h * muy[None] + c2h) * g_rv
    gmx = _face_fold(np.sum(c1h * s.u * g_ru, axis=0), -1)   # (ny, nx)
    gmy = _face_fold(np.sum(c1h * s.v * g_rv, axis=0), -2)
    g_mu = 0.5 * (gmx + np.roll(gmx, -1, axis=1)) \
        + 0.5 * (gmy + np.roll(gmy, -1, axis=0))
    return g_u, g_v, g_mu


def omega_adjoint(s, g_ww, dx, dy):
    """(ru, rv) gradients of omega_ref's Omega (WRF calc_ww_cp, reference
    dycore.py:102-131)."""
    nz = s.thp.shape[0]
    rdx, rdy = 1.0 / dx, 1.0 / dy
    dnw = s.dnw[:, None, None]
    # ww[k+1] = -sum_{m<=k} inc[m], k = 0..nz-2
    g_inc = np.zeros((nz,) + g_ww.shape[1:], dtype=g_ww.dtype)
    run = np.zeros_like(g_ww[0])
    for k in range(nz - 2, -1, -1):
        run = run - g_ww[k + 1]
        g_inc[k] = run
    g_dmdt = np.sum((s.c1h * s.dnw)[:nz - 1, None, None] * g_inc[:nz - 1],
                    axis=0)
    g_divv = g_inc.copy()
    g_divv[nz - 1] = 0.0
    g_divv = g_divv + g_dmdt[None]
    gx = dnw * rdx * g_divv            # d divv / d ru[f] = +-dnw rdx
    gy = dnw * rdy * g_divv
    g_ru = np.zeros((nz,) + (g_ww.shape[1], g_ww.shape[2] + 1),
                    dtype=g_ww.dtype)
    g_ru[..., 1:] += gx
    g_ru[..., :-1] -= gx
    g_rv = np.zeros((nz, g_ww.shape[1] + 1, g_ww.shape[2]), dtype=g_ww.dtype)
    g_rv[:, 1:] += gy
    g_rv[:, :-1] -= gy
    return g_ru, g_rv


def pgf_adjoint(s, lam_u, lam_v, dx, dy, top_lid=True):
    """Gradients (p, al, alt, php, mup) of slow_pgf's subtraction from
    ru_t, rv_t (reference dycore.cu:98-211) for output adjoints lam_u,
    lam_v (shaped like ru_t, rv_t)."""
    nz = s.thp.shape[0]
    rdx, rdy = 1.0 / dx, 1.0 / dy
    hrdx, hrdy = 0.5 / dx, 0.5 / dy
    mut = s.mub2d + s.mup
    pp = s.p - D._b3(s.pb)
    pbd = D._b3(s.pb)
    dpn = D._dpn_levels(s, pp, s.cfn, s.cfn1, top_lid)
    if s.base3d:
        phh = 0.5 * ((s.phb[:-1] + s.phb[1:]) + (s.php[:-1] + s.php[1:]))
    else:
        phh = 0.5 * (s.php[:-1] + s.php[1:])
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    rdnw = s.rdnw[:, None, None]
    g = {"pp": np.zeros_like(pp), "al": np.zeros_like(s.al),
         "alt": np.zeros_like(s.alt), "php": np.zeros_like(s.php),
         "mup": np.zeros_like(s.mup), "dpn": np.zeros_like(dpn)}

    def face(lam, ax, rd, hrd):
        lt = -_face_fold(lam, ax)              # ru_t -= tu
        A = lambda a: a
        B = lambda a: np.roll(a, 1, axis=ax)
        BT = lambda a: np.roll(a, -1, axis=ax)  # transpose of B
        muf = 0.5 * (A(mut) + B(mut))
        dmu = 0.5 * (A(s.mup) + B(s.mup))
        layer = c1h * muf[None] + c2h
        bracket = (A(s.php[1:]) - B(s.php[1:])) + (A(s.php[:-1])
                                                   - B(s.php[:-1]))
        bracket = bracket + (A(s.alt) + B(s.alt)) * (A(pp) - B(pp))
        if s.base3d:
            bracket = bracket + (A(s.al) + B(s.al)) * (A(pbd) - B(pbd))
        dphp = A(phh) - B(phh)
        hi = 0.5 * (A(dpn[1:]) + B(dpn[1:]))
        lo = 0.5 * (A(dpn[:-1]) + B(dpn[:-1]))
        vertical = rdnw * (hi - lo) - c1h * dmu[None]
        g_br = lt * hrd * layer
        g_muf = np.sum(c1h * lt * hrd * bracket, axis=0)
        g_dphp = lt * rd * vertical
        g_vert = lt * rd * dphp
        both = lambda x: x + BT(x)
        anti = lambda x: x - BT(x)
        g["mup"] += 0.5 * both(g_muf)
        g["php"][1:] += anti(g_br)
        g["php"][:-1] += anti(g_br)
        g["alt"] += both(g_br * (A(pp) - B(pp)))
        g["pp"] += anti(g_br * (A(s.alt) + B(s.alt)))
        if s.base3d:
            g["al"] += both(g_br * (A(pbd) - B(pbd)))
        g["php"][:-1] += 0.5 * anti(g_dphp)
        g["php"][1:] += 0.5 * anti(g_dphp)
        g_hi = g_vert * rdnw
        g["dpn"][1:] += 0.5 * both(g_hi)
        g["dpn"][:-1] -= 0.5 * both(g_hi)
        g["mup"] += 0.5 * both(np.sum(-c1h * g_vert, axis=0))

    face(lam_u, -1, rdx, hrdx)
    face(lam_v, -2, rdy, hrdy)
    gd = g["dpn"]
    g["pp"][0] += s.cf1 * gd[0]
    g["pp"][1] += s.cf2 * gd[0]
    g["pp"][2] += s.cf3 * gd[0]
    for kf in range(1, nz):
        g["pp"][kf] += s.fnm[kf] * gd[kf]
        g["pp"][kf - 1] += s.fnp[kf] * gd[kf]
    if top_lid:
        g["pp"][nz - 1] += s.cfn * gd[nz]
        g["pp"][nz - 2] += s.cfn1 * gd[nz]
    return g


def buoyancy_adjoint_moist(s, lam_w):
    """(p, mup, q_total) gradients of slow_buoyancy with the moist loading
    (reference dycore.cu:231-308, 289-300): rw_t += G ((cq1 rdn) dp' - c1f
    mu' - cqw cq1 (c1f mub + c2f)) with cqw the face mean of qv + qc + qr
    and cq1 = 1 / (1 + cqw); d(term)/d(cqw) = -cq1^2 (rdn dp' + c1f mub +
    c2f)."""
    nz = s.thp.shape[0]
    pp = s.p - D._b3(s.pb)
    qt = (s.qv + s.qc) + s.qr
    c1f, c2f = s.c1f[:, None, None], s.c2f[:, None, None]
    g_pp = np.zeros_like(s.p)
    g_qt = np.zeros_like(s.p)
    lw = lam_w[1:nz]
    cqw = 0.5 * (qt[1:] + qt[:-1])
    cq1 = 1.0 / (1.0 + cqw)
    rdn = s.rdn[1:, None, None]
    g_pp[1:] += D.G * cq1 * rdn * lw
    g_pp[:-1] -= D.G * cq1 * rdn * lw
    dp = pp[1:] - pp[:-1]
    g_cqw = D.G * lw * (-cq1 * cq1) * (rdn * dp + (c1f[1:nz] * s.mub2d
                                                   + c2f[1:nz]))
    g_qt[1:] += 0.5 * g_cqw
    g_qt[:-1] += 0.5 * g_cqw
    g_mup = -D.G * np.sum(c1f[1:nz] * lw, axis=0)
    lt = lam_w[nz]
    cqw = 0.5 * (qt[nz - 1] + qt[nz - 2])
    cq1 = 1.0 / (1.0 + cqw)
    g_pp[nz - 1] -= D.G * cq1 * 2.0 * s.rdnw[nz - 1] * lt
    g_cqw = D.G * lt * (-cq1 * cq1) * (2.0 * s.rdnw[nz - 1] * (0.0 - pp[nz - 1])
                                       + (s.c1f[nz] * s.mub2d + s.c2f[nz]))
    g_qt[nz - 1] += 0.5 * g_cqw
    g_qt[nz - 2] += 0.5 * g_cqw
    g_mup = g_mup - D.G * s.c1f[nz] * lt
    return g_pp, g_mup, g_qt


def buoyancy_adjoint(s, lam_w):
    """(p, mup) gradients of slow_buoyancy (dry; reference dycore.cu:
    231-308)."""
    nz = s.thp.shape[0]
    g_pp = np.zeros_like(s.p)
    lw = lam_w[1:nz]
    g_pp[1:] += D.G * s.rdn[1:, None, None] * lw
    g_pp[:-1] -= D.G * s.rdn[1:, None, None] * lw
    g_mup = -D.G * np.sum(s.c1f[1:nz, None, None] * lw, axis=0)
    g_pp[nz - 1] -= D.G * 2.0 * s.rdnw[nz - 1] * lam_w[nz]
    g_mup = g_mup - D.G * s.c1f[nz] * lam_w[nz]
    return g_pp, g_mup


def geopotential_adjoint(s, ww, lam, dx, dy):
    """(ww, w, u, v, php, mup) gradients of slow_geopotential (reference
    dycore.cu:541-667), lam shaped like rph_t (rows 1..nz used)."""
    nz = s.thp.shape[0]
    qrx, qry = 0.25 / dx, 0.25 / dy
    mut = s.mub2d + s.mup
    ph = (s.phb + s.php) if s.base3d else (s.phb[:, None, None] + s.php)
    rdnw = s.rdnw[:, None, None]
    g = {"ww": np.zeros_like(ww), "w": np.zeros_like(s.w),
         "u": np.zeros_like(s.u), "v": np.zeros_like(s.v),
         "php": np.zeros_like(s.php), "mup": np.zeros_like(s.mup)}
    g_wd = np.zeros_like(s.thp)
    for k in range(1, nz):
        g_wd[k] -= s.fnm[k] * lam[k]
        g_wd[k - 1] -= s.fnp[k] * lam[k]
    dph = ph[1:] - ph[:-1]
    wavg = 0.5 * (ww[1:] + ww[:-1])
    gwa = g_wd * rdnw * dph
    g["ww"][1:] += 0.5 * gwa
    g["ww"][:-1] += 0.5 * gwa
    gd = g_wd * wavg * rdnw
    g["php"][1:] += gd
    g["php"][:-1] -= gd
    G = D.G
    for k in range(1, nz + 1):
        g["mup"] += s.c1f[k] * G * s.w[k] * lam[k]
        g["w"][k] += (s.c1f[k] * mut + s.c2f[k]) * G * lam[k]
    avgx = 0.5 * (mut + np.roll(mut, 1, axis=1))
    avgy = 0.5 * (mut + np.roll(mut, 1, axis=0))
    uc, vc = s.u[..., :-1], s.v[:, :-1, :]
    for kk in range(1, nz + 1):
        top = kk == nz
        qx = 2.0 * qrx if top else qrx
        qy = 2.0 * qry if top else qry
        lk = lam[kk]
        for ax, q, avg, xc, gname in ((1, qx, avgx, uc, "u"),
                                      (0, qy, avgy, vc, "v")):
            m = s.c1f[kk] * avg + s.c2f[kk]
            if top:
                wv = s.cfn * xc[kk - 1] + s.cfn1 * xc[kk - 2]
            else:
                wv = xc[kk] + xc[kk - 1]
            p = ph[kk]
            dp = p - np.roll(p, 1, axis=ax)
            g_f = -q * (lk + np.roll(lk, 1, axis=ax))   # f[i] in tend i, i-1
            g_dp = g_f * m * wv
            g["php"][kk] += g_dp - np.roll(g_dp, -1, axis=ax)
            g_wv = g_f * m * dp
            gx = g[gname]
            sl = (slice(None), slice(0, -1)) if gname == "u" else \
                (slice(0, -1), slice(None))
            if top:
                gx[(kk - 1,) + sl] += s.cfn * g_wv
                gx[(kk - 2,) + sl] += s.cfn1 * g_wv
            else:
                gx[(kk,) + sl] += g_wv
                gx[(kk - 1,) + sl] += g_wv
            g_avg = s.c1f[kk] * g_f * dp * wv
            g["mup"] += 0.5 * (g_avg + np.roll(g_avg, -1, axis=ax))
    return g


def stage_tendencies_adjoint(s, lam, dx, dy, top_lid=True, lam_flux=None,
                             diff=None):
    """Transpose of stage_tendencies' Jacobian at the state s applied to
    lam = {STAGE_OUT name: adjoint}: returns {STAGE_IN name: gradient}
    (and qv, qc, qr for a moist state).  ``lam_flux``: adjoints of the
    stage fluxes (ru, rv, ww) the stage also hands on (the moist scalars'
    sumflux).  s must hold the stage state (its diagnostics are
    recomputed)."""
    work = D.OState(**{k: (v.copy() if isinstance(v, np.ndarray) else v)
                       for k, v in vars(s).items()})
    _, (ru, rv, ww) = stage_tendencies(work, dx, dy, top_lid)
    z = lambda n: lam.get(n, np.zeros_like(getattr(s, n)))
    g_diff = None
    if diff is not None and (diff[0] > 0.0 or diff[1] > 0.0):
        g_diff = diffusion_adjoint(work, {n: z(n) for n in STAGE_OUT},
                                   diff[0], diff[1], dx, dy)
    l_ru, l_rv, l_rw, l_rth, l_rph = (z(n) for n in STAGE_OUT)
    dxi, dyi = 1.0 / dx, 1.0 / dy
    g = {n: np.zeros_like(getattr(s, n)) for n in STAGE_IN}
    g_ru, g_rv, g_ww = np.zeros_like(ru), np.zeros_like(rv), np.zeros_like(ww)
    if lam_flux is not None:
        g_ru += lam_flux["ru"]
        g_rv += lam_flux["rv"]
        g_ww += lam_flux["ww"]
    moist = work.qv is not None
    theta = D._b3(work.thb) + work.thp
    for op, q, l, rdz, name in (("scalar", theta, l_rth, work.rdnw, "thp"),
                                ("u", work.u, l_ru, work.rdnw, "u"),
                                ("v", work.v, l_rv, work.rdnw, "v"),
                                ("w", work.w, l_rw, work.rdn, "w")):
        a = flux_div_adjoint(op, l, q, ru, rv, ww, rdz, work.fnm, work.fnp,
                             dxi, dyi)
        g[name] += a["q"]
        g_ru += a["ru"]
        g_rv += a["rv"]
        g_ww += a["rw"]
    gp = pgf_adjoint(work, l_ru, l_rv, dx, dy, top_lid)
    if moist:
        g_pp, g_mup_b, g_qt = buoyancy_adjoint_moist(work, l_rw)
    else:
        g_pp, g_mup_b = buoyancy_adjoint(work, l_rw)
    gg = geopotential_adjoint(work, ww, l_rph, dx, dy)
    g_ww += gg["ww"]
    for n in ("w", "u", "v", "php", "mup"):
        g[n] += gg[n]
    g["php"] += gp["php"]
    g["mup"] += gp["mup"] + g_mup_b
    g_p = gp["pp"] + g_pp
    a_ru, a_rv = omega_adjoint(work, g_ww, dx, dy)
    g_ru += a_ru
    g_rv += a_rv
    gu, gv, gmu = coupling_adjoint(work, g_ru, g_rv)
    g["u"] += gu
    g["v"] += gv
    g["mup"] += gmu
    gt, gph, gm = eos_adjoint(work, g_p, gp["al"], gp["alt"])
    g["thp"] += gt
    g["php"] += gph
    g["mup"] += gm
    if moist:
        g["qv"] = eos_q_adjoint(work, g_p) + g_qt
        g["qc"] = g_qt.copy()
        g["qr"] = g_qt.copy()
    if g_diff is not None:
        for n, v in g_diff.items():
            g[n] += v
    return g


# ---------------------------------------------------------------------------
# Small-step initialisation and finish (plain form)
# ---------------------------------------------------------------------------

INIT_OUT = ("u_pp", "v_pp", "w_pp", "th_pp", "ph_pp", "mu_pp", "al_pp",
            "p_pp", "p_pp_old")
FINISH_IN = ("u_pp", "v_pp", "w_pp", "th_pp", "ph_pp", "mu_pp")


def _faces_to_mass(g, ax):
    """Transpose of a periodic face average with a duplicate face: g per
    face (n+1 along ax) of 0.5 (m[f] + m[f-1]) -> per mass column."""
    gf = _face_fold(g, ax)
    return 0.5 * (gf + np.roll(gf, -1, axis=ax))


def init_forward(s):
    """small_step_init at the state s (time-t fields in u0..mup0, stage
    fields in u..mup, the stage EOS recomputed first); returns copies of
    INIT_OUT."""
    D.calc_p_alpha(s)
    D.small_step_init(s)
    return {n: getattr(s, n).copy() for n in INIT_OUT}


def init_adjoint(s, lam):
    """Transpose of init_forward: gradients of the time-t fields
    (u0..mup0) and of the stage fields (u..mup, through the stage EOS's p
    and alt as well) for lam over INIT_OUT (reference dycore.cu:947-1085)."""
    w = D.OState(**{k: (v.copy() if isinstance(v, np.ndarray) else v)
                    for k, v in vars(s).items()})
    D.calc_p_alpha(w)
    z = lambda n: lam.get(n, np.zeros_like(getattr(s, n)))
    Lu, Lv, Lw, Lth, Lph, Lmu, Lal = (z(n) for n in INIT_OUT[:7])
    Lp = z("p_pp") + z("p_pp_old")
    c1h, c2h = w.c1h[:, None, None], w.c2h[:, None, None]
    c1f, c2f = w.c1f[:, None, None], w.c2f[:, None, None]
    rdnw = w.rdnw[:, None, None]
    mt = w.mub2d + w.mup0
    ms = w.mub2d + w.mup
    thb = D._b3(w.thb)
    tht, ths = thb + w.thp0, thb + w.thp
    cht, chs = c1h * mt + c2h, c1h * ms + c2h
    mupp = w.mup0 - w.mup
    phpp = w.php0 - w.php
    thpp = cht * tht - chs * ths
    c2a = GAMMA * w.p / w.alt
    al = -((w.alt * (c1h * mupp)) + rdnw * (phpp[1:] - phpp[:-1])) / cht
    N = thpp - (c1h * mupp) * ths
    p5 = (w.alt * N) / (cht * ths)
    g = {n: np.zeros_like(getattr(s, n)) for n in STAGE_IN}
    g.update({n + "0": np.zeros_like(getattr(s, n)) for n in STAGE_IN})
    # p'' = c2a (p5 - al)
    g_c2a = Lp * (p5 - al)
    g_p5 = Lp * c2a
    g_al = Lal - Lp * c2a
    g_alt = g_p5 * N / (cht * ths)
    g_N = g_p5 * w.alt / (cht * ths)
    g_cht = -g_p5 * p5 / cht
    g_ths = -g_p5 * p5 / ths - g_N * (c1h * mupp)
    g_thpp = g_N + Lth
    g_mupp3 = -g_N * c1h * ths
    # al'' = -(alt c1h mu'' + rdnw dph'') / cht
    g_alt = g_alt - g_al * c1h * mupp / cht
    g_mupp3 = g_mupp3 - g_al * w.alt * c1h / cht
    g_dph = -g_al * rdnw / cht
    g_phpp = Lph.copy()
    g_phpp[1:] += g_dph
    g_phpp[:-1] -= g_dph
    g_cht = g_cht - g_al * al / cht
    # c2a = gamma p / alt
    g_p = g_c2a * GAMMA / w.alt
    g_alt = g_alt - g_c2a * c2a / w.alt
    # theta'' = cht tht - chs ths
    g_cht = g_cht + g_thpp * tht
    g["thp0"] += g_thpp * cht
    g_chs = -g_thpp * ths
    g_ths = g_ths - g_thpp * chs
    g["thp"] += g_ths
    g_mt = np.sum(c1h * g_cht, axis=0)
    g_ms = np.sum(c1h * g_chs, axis=0)
    # ph'' and mu''
    g["php0"] += g_phpp
    g["php"] -= g_phpp
    g_mupp = Lmu + np.sum(g_mupp3, axis=0)
    g["mup0"] += g_mupp
    g["mup"] -= g_mupp
    # w''
    g["w0"] += (c1f * mt + c2f) * Lw
    g["w"] -= (c1f * ms + c2f) * Lw
    g_mt = g_mt + np.sum(c1f * w.w0 * Lw, axis=0)
    g_ms = g_ms - np.sum(c1f * w.w * Lw, axis=0)
    # u'', v''
    for L, n, ax, fac in ((Lu, "u", -1, D.mu_at_u_faces),
                          (Lv, "v", -2, D.mu_at_v_faces)):
        mtf, msf = fac(mt), fac(ms)
        g[n + "0"] += (c1h * mtf[None] + c2h) * L
        g[n] -= (c1h * msf[None] + c2h) * L
        g_mt = g_mt + _faces_to_mass(np.sum(c1h * getattr(w, n + "0") * L,
                                            axis=0), ax)
        g_ms = g_ms - _faces_to_mass(np.sum(c1h * getattr(w, n) * L, axis=0),
                                     ax)
    g["mup0"] += g_mt
    g["mup"] += g_ms
    gt, gph, gm = eos_adjoint(w, g_p, np.zeros_like(w.al), g_alt)
    g["thp"] += gt
    g["php"] += gph
    g["mup"] += gm
    if w.qv is not None:
        g["qv"] = eos_q_adjoint(w, g_p)
    return g


def finish_forward(s):
    """small_step_finish at s: the new stage fields from the stage fields
    (u..mup) and the small-step fields; returns copies of STAGE_IN."""
    D.small_step_finish(s)
    return {n: getattr(s, n).copy() for n in STAGE_IN}


def finish_adjoint(s, lam):
    """Transpose of finish_forward for lam over STAGE_IN (the new stage
    fields): gradients of the stage fields and of FINISH_IN (reference
    dycore.cu:1087-1198)."""
    z = lambda n: lam.get(n, np.zeros_like(getattr(s, n)))
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    c1f, c2f = s.c1f[:, None, None], s.c2f[:, None, None]
    ms = s.mub2d + s.mup
    mn = ms + s.mu_pp
    g = {n: np.zeros_like(getattr(s, n)) for n in STAGE_IN + FINISH_IN}
    g_ms = np.zeros_like(ms)
    g_mn = np.zeros_like(ms)
    for n, pn, ax, fac in (("u", "u_pp", -1, D.mu_at_u_faces),
                           ("v", "v_pp", -2, D.mu_at_v_faces)):
        L = z(n)
        msf, mnf = fac(ms), fac(mn)
        x, xpp = getattr(s, n), getattr(s, pn)
        den = c1h * mnf[None] + c2h
        new = ((c1h * msf[None] + c2h) * x + xpp) / den
        g_num = L / den
        g[n] += (c1h * msf[None] + c2h) * g_num
        g[pn] += g_num
        g_ms = g_ms + _faces_to_mass(np.sum(c1h * x * g_num, axis=0), ax)
        g_mn = g_mn + _faces_to_mass(np.sum(-c1h * g_num * new, axis=0), ax)
    L = z("w")
    den = c1f * mn + c2f
    new = ((c1f * ms + c2f) * s.w + s.w_pp) / den
    g["w"] += (c1f * ms + c2f) * L / den
    g["w_pp"] += L / den
    g_ms = g_ms + np.sum(c1f * s.w * L / den, axis=0)
    g_mn = g_mn - np.sum(c1f * L / den * new, axis=0)
    L = z("php")
    g["php"] += L
    g["ph_pp"] += L
    L = z("thp")
    thb = D._b3(s.thb)
    den = c1h * mn + c2h
    th_new = ((c1h * ms + c2h) * (thb + s.thp) + s.th_pp) / den
    g["thp"] += (c1h * ms + c2h) * L / den
    g["th_pp"] += L / den
    g_ms = g_ms + np.sum(c1h * (thb + s.thp) * L / den, axis=0)
    g_mn = g_mn - np.sum(c1h * L / den * th_new, axis=0)
    L = z("mup")
    g["mup"] += L + g_ms + g_mn
    g["mu_pp"] += L + g_mn
    return g


# ---------------------------------------------------------------------------
# The whole RK3 step (dry, plain form, km_opt = 0): the pieces above
# chained in reverse over the stages from per-stage checkpoints
# ---------------------------------------------------------------------------

def set_w_surface_adjoint(s, lam_w0, dx, dy):
    """(u, v) gradients of set_w_surface's w(sfc) = u . grad(ht)
    (reference dycore.py:1999-2026) for the adjoint lam_w0 of w[0]."""
    ht = s.ht
    dyn = np.roll(ht, -1, 0) - ht
    dys = ht - np.roll(ht, 1, 0)
    dxe = np.roll(ht, -1, 1) - ht
    dxw = ht - np.roll(ht, 1, 1)
    gvc_n = (0.5 / dy) * dyn * lam_w0           # vc[1:, :]
    gvc_s = (0.5 / dy) * dys * lam_w0           # vc[:-1, :]
    guc_e = (0.5 / dx) * dxe * lam_w0           # uc[:, 1:]
    guc_w = (0.5 / dx) * dxw * lam_w0           # uc[:, :-1]
    g_u, g_v = np.zeros_like(s.u), np.zeros_like(s.v)
    for lev, cf in enumerate((s.cf1, s.cf2, s.cf3)):
        g_u[lev, :, 1:] += cf * guc_e
        g_u[lev, :, :-1] += cf * guc_w
        g_v[lev, 1:, :] += cf * gvc_n
        g_v[lev, :-1, :] += cf * gvc_s
    return g_u, g_v


def _copy_state(s):
    return D.OState(**{k: (v.copy() if isinstance(v, np.ndarray) else v)
                       for k, v in vars(s).items()})


def scalar_update_adjoint(w, q, q0, fl, mu0, mu, dt_stage, clamp, lam, dx,
                          dy):
    """Transpose of one species' stage update (moist_np.advance_scalars_
    stage without the PD limiter: q_new = (C(mu0) q0 + dt flux_div_scalar
    (q; ru_m, rv_m, ww_m)) / C(mu), clamped at zero on the final stage;
    reference moist.py:385-398) for lam = adjoint of q_new.  Returns the
    gradients of q (the stage scalar), q0, the three averaged fluxes and
    the column masses mu0, mu (2-D).  The clamp is differentiated as
    1 where the update is positive, 0 elsewhere."""
    f = q.dtype.type
    ru_m, rv_m, ww_m = fl
    dxi, dyi = float(f(1.0 / dx)), float(f(1.0 / dy))
    c1, c2 = w.c1h[:, None, None], w.c2h[:, None, None]
    chm0 = c1 * mu0[None] + c2
    chm = c1 * mu[None] + c2
    tend = f(0.0) + D.flux_div_scalar(q, ru_m, rv_m, ww_m, w.rdnw, w.fnm,
                                      w.fnp, dxi, dyi)
    v = (chm0 * q0 + f(dt_stage) * tend) / chm
    lv = np.where(np.real(v) > 0.0, lam, 0.0) if clamp else lam
    g_tend = f(dt_stage) * lv / chm
    a = flux_div_adjoint("scalar", g_tend, q, ru_m, rv_m, ww_m, w.rdnw,
                         w.fnm, w.fnp, dxi, dyi)
    return {"q": a["q"], "q0": chm0 * lv / chm,
            "ru": a["ru"], "rv": a["rv"], "ww": a["rw"],
            "mu0": np.sum(c1 * q0 * lv / chm, axis=0),
            "mu": -np.sum(c1 * v * lv / chm, axis=0)}


def _upwind_adjoint(qm1, q0v, cr, g):
    """Transpose of moist_np.flux_upwind for g = adjoint of its output:
    (dqm1, dq0v, dcr); |cr| differentiated as sign(cr), the CFL caps at
    +-1/2 as constants."""
    a = cr + np.abs(cr)
    b = cr - np.abs(cr)
    sg = np.sign(cr)
    da = np.where(a < 1.0, 1.0 + sg, 0.0)
    db = np.where(b > -1.0, 1.0 - sg, 0.0)
    return (0.5 * np.fmin(1.0, a) * g, 0.5 * np.fmax(-1.0, b) * g,
            0.5 * g * (qm1 * da + q0v * db))


def _dup_fold(g, ax):
    """Transpose of appending a copy of index 0 along ax (n+1 -> n)."""
    n = g.shape[ax] - 1
    out = np.take(g, range(n), axis=ax).copy()
    idx = [slice(None)] * g.ndim
    idx[ax] = slice(0, 1)
    out[tuple(idx)] += np.take(g, [n], axis=ax)
    return out


def pd_update_adjoint(w, q, q0, fl, mu0, mu, dt_stage, lam, dx, dy):
    """Transpose of one species' final-stage update with WRF's
    positive-definite limiter (moist_np.advance_scalars_stage with
    moist_adv_opt = 1: pd_fluxes -> pd_scale -> pd_apply -> the clamped
    coupled update; reference pd_advection.cu, moist.py:385-398,
    :493-494) for lam = adjoint of q_new.  Returns the gradients of q,
    q0, the averaged fluxes ru, rv, ww and the column masses mu0 (time t)
    and mu (after the stage).  Piecewise: the clamp at zero, each
    correction's donor choice (the sign of the correction flux), the
    factor's clamps to [0, 1] and the upwind CFL caps are differentiated
    as the branch taken; |cr| as sign(cr)."""
    from oracle import moist_np as M
    f = q.dtype.type
    ru, rv, rw = fl
    nz, ny, nx = q.shape
    dxi = float(np.real(f(1.0 / dx)))
    dyi = float(np.real(f(1.0 / dy)))
    dt = float(dt_stage)
    c1, c2 = w.c1h[:, None, None], w.c2h[:, None, None]
    rz = w.rdnw[:, None, None]
    fx = M.pd_fluxes(q, q0, ru, rv, rw, mu, w.c1h, w.c2h, w.rdnw, w.fnm,
                     w.fnp, dx, dy, dt)
    fxl, fxc, fyl, fyc, fzl, fzc = fx
    sc = M.pd_scale(q0, mu0, w.c1h, w.c2h, w.rdnw, fx, dxi, dyi, dt)
    tend = f(0.0) + M.pd_apply(sc, fx, w.rdnw, dxi, dyi)
    chm0 = c1 * mu0[None] + c2
    chm = c1 * mu[None] + c2
    v = (chm0 * q0 + f(dt) * tend) / chm
    lv = np.where(v > 0.0, lam, 0.0)
    g_tend = dt * lv / chm
    g_q0 = chm0 * lv / chm
    g_mu0 = np.sum(c1 * q0 * lv / chm, axis=0)
    g_mu = -np.sum(c1 * v * lv / chm, axis=0)
    # pd_apply
    sx_l, sx_r, sy_l, sy_r, sz_b, sz_t = M.pd_donor_scales(sc, fx)
    gfxl, gfxc = np.zeros_like(fxl), np.zeros_like(fxc)
    gfyl, gfyc = np.zeros_like(fyl), np.zeros_like(fyc)
    gfzl, gfzc = np.zeros_like(fzl), np.zeros_like(fzc)
    ggx, ggy, ggz = -dxi * g_tend, -dyi * g_tend, -rz * g_tend
    gfxc[..., 1:] += ggx * sx_r
    gfxc[..., :-1] -= ggx * sx_l
    gfxl[..., 1:] += ggx
    gfxl[..., :-1] -= ggx
    gfyc[:, 1:] += ggy * sy_r
    gfyc[:, :-1] -= ggy * sy_l
    gfyl[:, 1:] += ggy
    gfyl[:, :-1] -= ggy
    gfzc[1:] += ggz * sz_t
    gfzc[:nz] -= ggz * sz_b
    gfzl[1:] += ggz
    gfzl[:nz] -= ggz
    # each factor's adjoint back to its donor cell
    g_sc = np.zeros_like(sc)
    gs = ggx * fxc[..., 1:]                  # sx_r
    g_sc += np.where(fxc[..., 1:] > 0.0, gs, 0.0)
    g_sc += M._pm(np.where(fxc[..., 1:] < 0.0, gs, 0.0), -1, 2)
    gs = -ggx * fxc[..., :-1]                # sx_l
    g_sc += M._pm(np.where(fxc[..., :-1] > 0.0, gs, 0.0), 1, 2)
    g_sc += np.where(fxc[..., :-1] < 0.0, gs, 0.0)
    gs = ggy * fyc[:, 1:]                    # sy_r
    g_sc += np.where(fyc[:, 1:] > 0.0, gs, 0.0)
    g_sc += M._pm(np.where(fyc[:, 1:] < 0.0, gs, 0.0), -1, 1)
    gs = -ggy * fyc[:, :-1]                  # sy_l
    g_sc += M._pm(np.where(fyc[:, :-1] > 0.0, gs, 0.0), 1, 1)
    g_sc += np.where(fyc[:, :-1] < 0.0, gs, 0.0)
    gs = ggz * fzc[1:]                       # sz_t: self, or above
    g_sc += np.where(fzc[1:] < 0.0, gs, 0.0)
    ga = np.where(fzc[1:] > 0.0, gs, 0.0)
    g_sc[1:] += ga[:-1]
    g_sc[-1] += ga[-1]
    gs = -ggz * fzc[:nz]                     # sz_b: self, or below
    g_sc += np.where(fzc[:nz] > 0.0, gs, 0.0)
    gb = np.where(fzc[:nz] < 0.0, gs, 0.0)
    g_sc[:-1] += gb[1:]
    g_sc[0] += gb[0]
    # pd_scale: the factor inside its clamps
    ph_low, fo = M.pd_budget(q0, mu0, w.c1h, w.c2h, w.rdnw, fx, dxi, dyi,
                             dt)
    den = fo + 1e-20
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = ph_low / den
        inside = (fo > ph_low) & (ratio > 0.0) & (ratio < 1.0)
        g_pl = np.where(inside, g_sc / den, 0.0)
        g_fo = np.where(inside, -g_sc * ratio / den, 0.0)
    g_q0 += chm0 * g_pl
    g_mu0 += np.sum(c1 * q0 * g_pl, axis=0)
    g_dl = -dt * g_pl
    gfxl[..., 1:] += dxi * g_dl
    gfxl[..., :-1] -= dxi * g_dl
    gfyl[:, 1:] += dyi * g_dl
    gfyl[:, :-1] -= dyi * g_dl
    gfzl[1:] += rz * g_dl
    gfzl[:-1] -= rz * g_dl
    a = dt * g_fo
    gfxc[..., 1:] += np.where(fxc[..., 1:] > 0.0, dxi * a, 0.0)
    gfxc[..., :-1] -= np.where(fxc[..., :-1] < 0.0, dxi * a, 0.0)
    gfyc[:, 1:] += np.where(fyc[:, 1:] > 0.0, dyi * a, 0.0)
    gfyc[:, :-1] -= np.where(fyc[:, :-1] < 0.0, dyi * a, 0.0)
    gfzc[1:] += np.where(fzc[1:] < 0.0, rz * a, 0.0)
    gfzc[:-1] -= np.where(fzc[:-1] > 0.0, rz * a, 0.0)
    # pd_fluxes: correction = high-order flux of q - upwind flux of q0
    g_q = np.zeros_like(q)
    gq_h, g_ru = _h_faces_adjoint(q, ru, gfxc, 2, True)
    g_q += gq_h
    gq_h, g_rv = _h_faces_adjoint(q, rv, gfyc, 1, True)
    g_q += gq_h
    gzc = gfzc.copy()
    gzc[0] = 0.0
    gzc[nz] = 0.0
    gq_h, g_rw = _zhalf_faces_adjoint(q, rw, gzc, w.fnm, w.fnp)
    g_q += gq_h
    glx, gly, glz = gfxl - gfxc, gfyl - gfyc, gfzl - gfzc
    for ax, rvel, gl, dd, m_ax in ((2, ru, glx, dx, 1), (1, rv, gly, dy, 0)):
        mb = (c1 * 0.5) * (mu + M._pm(mu, -1, m_ax))[None] + c2
        mf = np.concatenate([mb, np.take(mb, [0], axis=ax)], axis=ax)
        cr = ((rvel * dt) / dd) / mf
        qm1 = M._pm(q0, -1, ax)
        qm1 = np.concatenate([qm1, np.take(qm1, [0], axis=ax)], axis=ax)
        qc0 = np.concatenate([q0, np.take(q0, [0], axis=ax)], axis=ax)
        fu = M.flux_upwind(qm1, qc0, cr)
        g_fu = mf * (dd / dt) * gl
        dqm1, dqc0, dcr = _upwind_adjoint(qm1, qc0, cr, g_fu)
        g_m = (dd / dt) * fu * gl - dcr * cr / mf
        gvel = dcr * (dt / dd) / mf
        if ax == 2:
            g_ru = g_ru + gvel
        else:
            g_rv = g_rv + gvel
        g_q0 += M._pm(_dup_fold(dqm1, ax), 1, ax) + _dup_fold(dqc0, ax)
        sm = np.sum((c1 * 0.5) * _dup_fold(g_m, ax), axis=0)
        g_mu += sm + M._pm(sm, 1, m_ax)
    for k in range(1, nz):
        dz = 2.0 / (w.rdnw[k] + w.rdnw[k - 1])
        muf = w.c1h[k] * mu + w.c2h[k]
        crz = (rw[k] * dt) / (dz * muf)
        fu = M.flux_upwind(q0[k - 1], q0[k], crz)
        g_fu = muf * (dz / dt) * glz[k]
        dqm1, dqc0, dcr = _upwind_adjoint(q0[k - 1], q0[k], crz, g_fu)
        g_q0[k - 1] += dqm1
        g_q0[k] += dqc0
        g_rw[k] += dcr * dt / (dz * muf)
        g_mu += w.c1h[k] * ((dz / dt) * fu * glz[k] - dcr * crz / muf)
    return {"q": g_q, "q0": g_q0, "ru": g_ru, "rv": g_rv, "ww": g_rw,
            "mu0": g_mu0, "mu": g_mu}


def step_adjoint(s, cfg, lam):
    """J^T lam for one dycore step (D.step) at the state s: ``lam`` maps
    the new state's STAGE_IN fields (and qv, qc, qr on a moist state) to
    adjoints; returns the gradients of the step's input fields.  Plain
    form; constant-K diffusion on a flat base; a moist state's scalars with
    or without the final stage's PD limiter (moist_adv_opt 1 or 0).  The forward is re-run on a copy, checkpointing
    each stage's state and each substep's input; s is not modified.

    Moist: the scalars' stage update runs after the finish with the
    substep-averaged fluxes (sumflux: the mean of the substeps' u'', v'',
    Omega'' plus the stage fluxes), so its transpose comes first in the
    stage's reverse: it adds to the adjoint of the finish's new mu', to
    every substep's output u'', v'', Omega'' (1/nsub each) and to the stage
    fluxes, which the tendency-pass adjoint takes with the tendencies; the
    moist EOS and buoyancy loading add the stage scalars' share."""
    diffusion = cfg.km_opt == 1 and (cfg.khdif > 0.0 or cfg.kvdif > 0.0)
    moist = s.qv is not None
    if getattr(cfg, "precision_form", False):
        raise NotImplementedError("step adjoint: plain form")
    if diffusion and s.base3d:
        raise NotImplementedError("step adjoint: diffusion needs a flat base")
    diff = (cfg.khdif, cfg.kvdif) if diffusion else None
    pd_opt = getattr(cfg, "moist_adv_opt", 1)
    names = STAGE_IN + (SPECIES if moist else ())
    ns = cfg.time_step_sound
    stages = ((1, cfg.dt / 3.0), (max(ns // 2, 1), cfg.dt / ns),
              (ns, cfg.dt / ns))
    w = _copy_state(s)
    for n in D.PROGNOSTICS:
        getattr(w, n + "0")[...] = getattr(w, n)
    if moist:
        for q in SPECIES:
            getattr(w, q + "0")[...] = getattr(w, q)
    ck = []                                   # per stage: state, P_i list
    for istage, (nsub, dtau) in enumerate(stages):
        for n in STAGE_OUT + ("rmu_t",):
            getattr(w, n)[...] = 0.0
        D.calc_p_alpha(w)
        ru, rv, ww = D.stage_fluxes(w, cfg.dx, cfg.dy)
        D.add_slow_tendencies(w, ru, rv, ww, cfg.dx, cfg.dy, cfg.top_lid)
        if diff is not None:
            D.add_diffusion_tendencies(w, diff[0], diff[1], cfg.dx, cfg.dy)
        D.small_step_init(w)
        coefs = D.calc_coefs(w, dtau, cfg.epssm, cfg.top_lid)
        subs = []
        ex = None
        if moist:
            qst = {q: getattr(w, q).copy() for q in SPECIES}
            sm = [np.zeros_like(w.u), np.zeros_like(w.v), np.zeros_like(w.w)]
        for i in range(nsub):
            subs.append(_copy_state(w))       # substep i's input
            D.acoustic_substep(w, coefs, dtau, cfg, first=(i == 0))
            if moist:
                sm[0] += w.u_pp
                sm[1] += w.v_pp
                sm[2] += w.ww_pp
        after = _copy_state(w)
        D.small_step_finish(w)
        if moist:
            fl = D.sumflux_finish(sm[0], sm[1], sm[2], ru, rv, ww, nsub)
            ex = {"q": qst, "fl": fl, "mu": w.mub2d + w.mup,
                  "final": istage == 2, "dt": nsub * dtau}
            D.advance_scalars_stage(w, fl[0], fl[1], fl[2], nsub * dtau,
                                    istage == 2, cfg.dx, cfg.dy, pd_opt)
        ck.append((after, subs, coefs, dtau, nsub, ex))
    # reverse
    g0 = {n: np.zeros_like(getattr(s, n)) for n in names}   # time-t
    lx = {n: np.array(lam.get(n, np.zeros_like(getattr(s, n))), copy=True)
          for n in names}
    gu, gv = set_w_surface_adjoint(s, lx["w"][0], cfg.dx, cfg.dy)
    lx["u"] += gu
    lx["v"] += gv
    lx["w"][0] = 0.0
    mu0 = s.mub2d + s.mup
    for istage in range(2, -1, -1):
        after_subs, subs, coefs, dtau, nsub, ex = ck[istage]
        lam_flux = None
        inj = None
        gq = {}
        if moist:
            lam_flux = {"ru": 0.0, "rv": 0.0, "ww": 0.0}
            g_mu = 0.0
            for q in SPECIES:
                if ex["final"] and pd_opt == 1:
                    r = pd_update_adjoint(after_subs, ex["q"][q],
                                          getattr(s, q), ex["fl"], mu0,
                                          ex["mu"], ex["dt"], lx[q], cfg.dx,
                                          cfg.dy)
                else:
                    r = scalar_update_adjoint(after_subs, ex["q"][q],
                                              getattr(s, q), ex["fl"], mu0,
                                              ex["mu"], ex["dt"],
                                              ex["final"], lx[q], cfg.dx,
                                              cfg.dy)
                gq[q] = r["q"]
                g0[q] += r["q0"]
                g0["mup"] += r["mu0"]
                g_mu = g_mu + r["mu"]
                for k in ("ru", "rv", "ww"):
                    lam_flux[k] = lam_flux[k] + r[k]
            lx["mup"] = lx["mup"] + g_mu
            fn = s.dtype.type(nsub)
            inj = (lam_flux["ru"] / fn, lam_flux["rv"] / fn,
                   lam_flux["ww"] / fn)
        # the stage state Xs is the finish's input: after_subs holds it
        fin = finish_adjoint(after_subs, lx)
        gs = {n: fin[n] for n in STAGE_IN}
        gs.update(gq)
        lp = {n: fin[n] for n in FINISH_IN}
        lt = {n: np.zeros_like(getattr(s, n)) for n in SUBSTEP_TEND}
        for i in range(nsub - 1, -1, -1):
            st = subs[i]
            if moist:
                lp["u_pp"] = lp["u_pp"] + inj[0]
                lp["v_pp"] = lp["v_pp"] + inj[1]
                lp["ww_pp"] = inj[2]
            out = acoustic_substep_adjoint(st, coefs, dtau, cfg, i == 0, lp,
                                           ref=True)
            lp = {n: out[n] for n in SUBSTEP_STATE}
            for n in SUBSTEP_TEND:
                lt[n] += out[n]
            for n, v in out["ref"].items():
                gs[n] = gs[n] + v
        gi = init_adjoint(subs[0], lp)
        for n in STAGE_IN:
            g0[n] += gi[n + "0"]
            gs[n] += gi[n]
        if moist:
            gs["qv"] = gs["qv"] + gi["qv"]
        gt = stage_tendencies_adjoint(subs[0], {n: lt[n] for n in STAGE_OUT},
                                      cfg.dx, cfg.dy, cfg.top_lid, lam_flux,
                                      diff)
        for n, v in gt.items():
            gs[n] = gs[n] + v
        lx = gs                                # the previous stage's output
    for n in names:
        g0[n] += lx[n]
    return g0
