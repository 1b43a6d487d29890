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


def acoustic_substep_adjoint(s, coefs, dtau, c, first, lam):
    """J^T lam for one acoustic substep at the state s (its small-step
    fields are the substep's input X).  ``lam`` maps output names
    (SUBSTEP_STATE, optionally ww_pp) to adjoints; returns the adjoints of
    the inputs (SUBSTEP_STATE) and of the tendencies (SUBSTEP_TEND)."""
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

    # ---- diagnose_p ------------------------------------------------------
    chmp = c1h * (mut + M1) + c2h
    q = s.alt / (chmp * thr)
    X = T1 - c1h * M1 * thr
    lPP1 = L["p_pp"]
    lT1 += c2a * q * lPP1
    lM1 += np.sum(c2a * (-q * c1h * thr - q * X * c1h / chmp) * lPP1, axis=0)
    lAL = L["al_pp"] - c2a * lPP1
    lM1 += np.sum((-(s.alt * c1h) / chmp - AL1 * c1h / chmp) * lAL, axis=0)
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
    lPh0[0] += lPh1[0]

    # ---- damper (damp_opt = 3): w = (w - dw cfm w_ref) / (1 + dw) -------
    cfm = c1f * mut + c2f
    if fw["dampmag"] > 0.0:
        phf = (s.phb if s.base3d else s.phb[:, None, None]) + s.php
        htop = phf[nz] / g
        hbot = htop - f(c.zdamp)
        for k in range(1, nz + 1):
            hk = phf[k] / g
            arg = f(1.5707963) * (hk - hbot) / f(c.zdamp)
            sn = np.sin(arg.astype(np.float64)).astype(s.dtype)
            dw = f(fw["dampmag"]) * sn * sn
            lW[k] = np.where(hk >= hbot, lW[k] / (one + dw), lW[k])

    # ---- Thomas solve (rows 1..nz, w[0] coupled through a[1]) ------------
    n = nz
    for k in range(1, n):
        lW[k + 1] += -gam[k] * lW[k]
    for k in range(n, 0, -1):
        lW[k] = alpha[k] * lW[k]
        lW[k - 1] += -a[k] * lW[k]
    lacc = lW                                   # adjoint of the pre-solve w

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
    return out
