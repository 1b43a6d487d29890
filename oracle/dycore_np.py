"""NumPy restatement of the reference dycore step (test infrastructure).

One function per reference kernel / launcher, each citing the reference
file:line it follows.  Arrays use the reference layout ``(nz, ny, nx)``
x-fastest with C-grid staggering: ``u (nz,ny,nx+1)``, ``v (nz,ny+1,nx)``,
``w, php (nz+1,ny,nx)``; x and y are periodic and the duplicate face
column/row is carried like the reference does.

``dtype`` selects the arithmetic: float32 mirrors the device FP32 path
(every NumPy operation rounds, i.e. no fused multiply-add), float64 is the
"same discretisation in FP64" reference used to state true FP32 error.

Scope: periodic lateral boundaries, identity map factors, no Coriolis,
flat (1-D base) or terrain (3-D base) states, hybrid coordinates,
hypsometric_opt=1, h_sca_adv_order=2, optional constant-K diffusion
(km_opt=1), damp_opt 0/3, rigid or open top, optional moist qv/qc/qr tracer
transport (moist_adv_opt 0/1, mp_physics=0).
"""
from __future__ import annotations

from types import SimpleNamespace

import math

import numpy as np

G = 9.81
RD = 287.0
CP = 7.0 * RD / 2.0
CV = CP - RD
P0 = 1.0e5
GAMMA = CP / CV
RVOVRD = float(np.float32(np.float32(461.6) / np.float32(287.0)))

PROGNOSTICS = ("u", "v", "w", "thp", "php", "mup")


# ---------------------------------------------------------------------------
# State container
# ---------------------------------------------------------------------------

class OState(SimpleNamespace):
    """Oracle state: NumPy arrays named and shaped like the reference
    ``DomainState`` (gpuwm/core/state.py:389-695)."""

    @property
    def shape3(self):
        return self.p.shape


def make_state(nz, ny, nx, coord, base, *, dtype=np.float32, moist=False,
               cf=None, extra=(), n_load=0, h_diabatic=False) -> OState:
    """Allocate an at-rest oracle state and load coordinate/base arrays
    (reference state.py:359-771 ``DomainState.__init__`` + ``load_base``).

    ``coord``/``base`` are FP64 host objects with the WRF names (znw, dnw,
    c1h, ..., thb, pb, alb, phb, mub, p_top, terrain_z); ``cf`` holds the
    cf1..cf3/cfn/cfn1 extrapolation weights.  ``extra`` names a
    mixed-phase scheme's transported species beyond qv/qc/qr, of which the
    first ``n_load`` load the buoyancy (reference state.py:428-470,
    moist.py:107-240); ``h_diabatic`` allocates the retained microphysics
    heating that every stage adds to the theta tendency (reference
    dycore.py:572-592).
    """
    f = np.dtype(dtype).type
    z = lambda *sh: np.zeros(sh, dtype=dtype)
    s = OState(dtype=np.dtype(dtype))
    s.u, s.v, s.w = z(nz, ny, nx + 1), z(nz, ny + 1, nx), z(nz + 1, ny, nx)
    s.thp, s.php, s.mup = z(nz, ny, nx), z(nz + 1, ny, nx), z(ny, nx)
    s.p, s.al, s.alt = z(nz, ny, nx), z(nz, ny, nx), z(nz, ny, nx)
    for name in PROGNOSTICS:
        setattr(s, name + "0", np.zeros_like(getattr(s, name)))
    s.ru_t, s.rv_t, s.rw_t = (z(nz, ny, nx + 1), z(nz, ny + 1, nx),
                              z(nz + 1, ny, nx))
    s.rth_t, s.rph_t, s.rmu_t = z(nz, ny, nx), z(nz + 1, ny, nx), z(ny, nx)
    s.u_pp, s.v_pp, s.w_pp = (z(nz, ny, nx + 1), z(nz, ny + 1, nx),
                              z(nz + 1, ny, nx))
    s.th_pp, s.ph_pp, s.mu_pp = z(nz, ny, nx), z(nz + 1, ny, nx), z(ny, nx)
    s.p_pp, s.p_pp_old, s.al_pp = z(nz, ny, nx), z(nz, ny, nx), z(nz, ny, nx)
    s.ww_pp = z(nz + 1, ny, nx)
    s.pp = z(nz, ny, nx)       # precision form: p' of the last EOS
    if not moist and (extra or h_diabatic):
        raise ValueError("extra species and h_diabatic need moist=True")
    if moist:
        for q in ("qv", "qc", "qr") + tuple(extra):
            setattr(s, q, z(nz, ny, nx))
            setattr(s, q + "0", z(nz, ny, nx))
    else:
        s.qv = s.qc = s.qr = None
    s.extra = tuple(extra)
    s.n_load = int(n_load)
    s.h_diabatic = z(nz, ny, nx) if h_diabatic else None
    for name in ("dnw", "rdnw", "dn", "rdn", "fnp", "fnm", "znu", "znw",
                 "c1h", "c2h", "c1f", "c2f", "c3h", "c4h", "c3f", "c4f"):
        setattr(s, name, np.asarray(getattr(coord, name), dtype=dtype))
    for name in ("thb", "pb", "alb", "phb"):
        setattr(s, name, np.asarray(getattr(base, name), dtype=dtype))
    s.base3d = s.thb.ndim == 3
    s.mub2d = np.broadcast_to(np.asarray(base.mub, dtype=dtype),
                              (ny, nx)).copy()
    s.ht = (z(ny, nx) if base.terrain_z is None
            else np.asarray(base.terrain_z, dtype=dtype))
    s.p_top = f(base.p_top)
    cf = cf or {}
    for name in ("cf1", "cf2", "cf3", "cfn", "cfn1"):
        setattr(s, name, f(cf.get(name, 0.0)))
    s.elapsed_seconds = 0.0
    return s


def _b3(a):
    """Base profile as a broadcastable (n, 1, 1) or full 3-D array."""
    return a if a.ndim == 3 else a[:, None, None]


def _roll(a, shift, axis):
    """Value at index i+shift (periodic) placed at i."""
    return np.roll(a, -shift, axis=axis)


def _c(v, s):
    return np.asarray(v, dtype=s.dtype)


# ---------------------------------------------------------------------------
# EOS diagnostics
# ---------------------------------------------------------------------------

def calc_p_alpha(s, prec=False, hypso=1) -> None:
    """Full pressure, alpha', total alpha (reference diagnostics.cu:37-89;
    launcher diagnostics.py:14-57).  ``hypso`` is hypsometric_opt: 1 takes
    alpha from the layer's geopotential thickness, 2 from the log of its
    dry-pressure ratio (diagnostics.cu:75-80), the logarithm evaluated in
    double and rounded once.  ``prec``: the precision form's perturbation
    EOS (calc_p_alpha_pert, hypso 1)."""
    if prec:
        calc_p_alpha_pert(s)
        return
    mu = s.mub2d + s.mup
    th = _b3(s.thb) + s.thp
    if s.qv is not None:
        th = th * (_c(1.0, s) + _c(RVOVRD, s) * s.qv)
    ph = _b3(s.phb) + s.php
    dph = ph[1:] - ph[:-1]
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    if hypso == 2:
        lev = lambda a: a[:, None, None]
        pt = s.p_top
        pfu = (lev(s.c3f[1:]) * mu + lev(s.c4f[1:])) + pt
        pfd = (lev(s.c3f[:-1]) * mu + lev(s.c4f[:-1])) + pt
        phm = (lev(s.c3h) * mu + lev(s.c4h)) + pt
        lg = np.log((pfd / pfu).astype(np.float64)).astype(s.dtype)
        al = ((dph / phm) / lg) - _b3(s.alb)
        alt = al + _b3(s.alb)
        s.alt[...] = alt
        s.al[...] = al
    else:
        alt = (-dph * s.rdnw[:, None, None]) / (c1h * mu + c2h)
        s.alt[...] = alt
        s.al[...] = alt - _b3(s.alb)
    s.p[...] = _c(P0, s) * _pow_gamma((_c(RD, s) * th) / (_c(P0, s) * alt))


def calc_p_alpha_pert(s) -> None:
    """The EOS of calc_p_alpha in perturbation-from-base form (precision
    form).  The base state is discretely hydrostatic, alb = -rdnw dphb /
    (c1h mub + c2h), and in the EOS, pb = P0 (RD thb / (P0 alb))^GAMMA
    (grid.make_base_state), so with mh = c1h mu + c2h
      al  = -(alb c1h mu' + rdnw dphp) / mh         (WRF calc_p_rho's al)
      p'  = pb expm1(GAMMA (log1p(theta'_m / thb) - log1p(al / alb)))
    never difference the full geopotential or subtract pb from a full
    pressure; the float32 state keeps alt = alb + al and p = pb + p', and
    s.pp holds p' for the pressure-gradient terms.  The logarithms and the
    exponential run in float64 and p' is rounded once (the HIP kernels do
    the same)."""
    f64 = np.float64
    mu = s.mub2d + s.mup
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    alb = _b3(s.alb)
    mh = c1h * mu + c2h
    al = -((alb * (c1h * s.mup)) + s.rdnw[:, None, None]
           * (s.php[1:] - s.php[:-1])) / mh
    s.al[...] = al
    s.alt[...] = alb + al
    x = np.log1p(s.thp.astype(f64) / _b3(s.thb).astype(f64))
    if s.qv is not None:
        x = x + np.log1p(f64(np.float32(RVOVRD)) * s.qv.astype(f64))
    y = np.log1p(al.astype(f64) / alb.astype(f64))
    pb = _b3(s.pb)
    # the exponent the state's precision carries (as _pow_gamma)
    g = f64(np.float32(GAMMA)) if s.dtype == np.float32 else f64(GAMMA)
    s.pp[...] = (pb.astype(f64) * np.expm1(g * (x - y))).astype(s.dtype)
    s.p[...] = pb + s.pp


class _Kahan:
    """Compensated running sum (the precision form's column sums): s
    carries the sum, e the low-order part lost by the last addition."""

    def __init__(self, like):
        self.s = np.zeros_like(like)
        self.e = np.zeros_like(like)

    def add(self, x):
        y = x - self.e
        t = self.s + y
        self.e = (t - self.s) - y
        self.s = t
        return t


def _pow_gamma(x):
    """x**GAMMA.  For float32 the power is evaluated in float64 and rounded
    once (the HIP kernels do the same), so both sides agree to the bit."""
    if x.dtype == np.float32:
        g = np.float64(np.float32(GAMMA))
        return np.power(x.astype(np.float64), g).astype(np.float32)
    return np.power(x, GAMMA)


# ---------------------------------------------------------------------------
# Stage fluxes (couple_momentum, calc_ww_cp)
# ---------------------------------------------------------------------------

def mu_at_u_faces(mu):
    """(ny, nx) -> (ny, nx+1) periodic face average, face f between cells
    f-1 and f (reference state.py:339-346)."""
    mux = 0.5 * (mu + np.roll(mu, 1, axis=1))
    return np.concatenate([mux, mux[:, :1]], axis=1).astype(mu.dtype)


def mu_at_v_faces(mu):
    """(ny, nx) -> (ny+1, nx) (reference state.py:349-356)."""
    muy = 0.5 * (mu + np.roll(mu, 1, axis=0))
    return np.concatenate([muy, muy[:1, :]], axis=0).astype(mu.dtype)


def stage_fluxes(s, dx, dy, prec=False):
    """Coupled horizontal mass fluxes ru, rv and Omega ww at the stage
    reference state (reference dycore.py:162-206, 134-159, 102-131)."""
    mu = s.mub2d + s.mup
    mux, muy = mu_at_u_faces(mu), mu_at_v_faces(mu)
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    ru = (c1h * mux[None] + c2h) * s.u
    rv = (c1h * muy[None] + c2h) * s.v
    return ru, rv, omega_ref(s, ru, rv, dx, dy, prec)


def omega_ref(s, ru, rv, dx, dy, prec=False):
    """WRF calc_ww_cp (reference dycore.py:102-131).  ``prec``: the
    precision form's compensated column sums."""
    nz = s.thp.shape[0]
    rdx, rdy = _c(1.0 / dx, s), _c(1.0 / dy, s)
    dnw = s.dnw[:, None, None]
    divv = dnw * (rdx * (ru[:, :, 1:] - ru[:, :, :-1])
                  + rdy * (rv[:, 1:, :] - rv[:, :-1, :]))
    if prec:
        ks = _Kahan(divv[0])
        for k in range(nz):
            ks.add(divv[k])
        dmdt = ks.s
    else:
        dmdt = divv[0].copy()
        for k in range(1, nz):
            dmdt = dmdt + divv[k]
    ww = np.zeros((nz + 1,) + divv.shape[1:], dtype=s.dtype)
    inc = (s.c1h * s.dnw)[:nz - 1, None, None] * dmdt[None] + divv[:nz - 1]
    run = np.zeros_like(dmdt)
    kr = _Kahan(dmdt)
    for k in range(nz - 1):
        run = kr.add(inc[k]) if prec else run + inc[k]
        ww[k + 1] = -run
    return ww


# ---------------------------------------------------------------------------
# Advection (flux-form 5th/3rd order)
# ---------------------------------------------------------------------------

def flux5(qm3, qm2, qm1, q0, qp1, qp2, vel):
    """WRF flux5 (reference advection.cu:78-83)."""
    return (vel * (37.0 * (q0 + qm1) - 8.0 * (qp1 + qm2) + (qp2 + qm3))
            - np.abs(vel) * (10.0 * (q0 - qm1) - 5.0 * (qp1 - qm2)
                             + (qp2 - qm3))) / 60.0


def flux3(qm2, qm1, q0, qp1, vel):
    """Omega-signed vertical flux3 (reference advection.cu:86-90)."""
    return (vel * (7.0 * (q0 + qm1) - (qp1 + qm2))
            + np.abs(vel) * (3.0 * (q0 - qm1) - (qp1 - qm2))) / 12.0


def _xfaces(q, vel):
    """Periodic x-face flux5 of a cell field q (.., n) at faces f=0..n
    (face f between cells f-1 and f; vel has n+1 faces)."""
    st = [_roll(q, o, -1) for o in (-3, -2, -1, 0, 1, 2)]
    st = [np.concatenate([a, a[..., :1]], axis=-1) for a in st]
    return flux5(*st, vel)


def _yfaces(q, vel):
    st = [_roll(q, o, -2) for o in (-3, -2, -1, 0, 1, 2)]
    st = [np.concatenate([a, a[..., :1, :]], axis=-2) for a in st]
    return flux5(*st, vel)


def _zfaces_half(q, vel, fnm, fnp):
    """Vertical face fluxes of a half-level field at w-levels 0..nz
    (reference advection.cu:189-201 ``zface_half``)."""
    nz = q.shape[0]
    fz = np.zeros((nz + 1,) + q.shape[1:], dtype=q.dtype)
    for kf in range(1, nz):
        if kf == 1 or kf == nz - 1:
            fz[kf] = vel[kf] * (fnm[kf] * q[kf] + fnp[kf] * q[kf - 1])
        else:
            fz[kf] = flux3(q[kf - 2], q[kf - 1], q[kf], q[kf + 1], vel[kf])
    return fz


def _div_sum(dfx, dfy, dfz, rdx, rdy, rdz):
    t = -dfx * rdx
    t = t - dfy * rdy
    return t - dfz * rdz


def flux_div_scalar(q, ru, rv, rw, rdnw, fnm, fnp, dx_inv, dy_inv):
    """Periodic scalar flux divergence (reference advection.cu:206-416,
    periodic branch :384-415).  Returns the tendency (nz, ny, nx)."""
    fx = _xfaces(q, ru)
    fy = _yfaces(q, rv)
    fz = _zfaces_half(q, rw, fnm, fnp)
    return _div_sum(fx[..., 1:] - fx[..., :-1], fy[:, 1:] - fy[:, :-1],
                    fz[1:] - fz[:-1], dx_inv, dy_inv, rdnw[:, None, None])


def flux_div_u(u, ru, rv, rw, rdnw, fnm, fnp, dx_inv, dy_inv):
    """u-momentum flux divergence at u points (reference advection.cu:
    424-611, periodic branch :572-611).  Returns (nz, ny, nx+1)."""
    nx = u.shape[-1] - 1
    uc = u[..., :nx]
    # x faces at mass centres m (between u[m] and u[m+1])
    velx = 0.5 * (ru[..., :nx] + ru[..., 1:])
    FX = flux5(*[_roll(uc, o, -1) for o in (-2, -1, 0, 1, 2, 3)], velx)
    dfx = FX - np.roll(FX, 1, axis=-1)
    # y faces at corners g (rows), u column ic
    vely = 0.5 * (np.roll(rv, 1, axis=-1) + rv)             # (nz, ny+1, nx)
    fy = _yfaces(uc, vely)
    # z faces: corner Omega averaged over the two mass columns
    velz = 0.5 * (np.roll(rw, 1, axis=-1) + rw)
    fz = _zfaces_half(uc, velz, fnm, fnp)
    t = _div_sum(dfx, fy[:, 1:] - fy[:, :-1], fz[1:] - fz[:-1],
                 dx_inv, dy_inv, rdnw[:, None, None])
    return np.concatenate([t, t[..., :1]], axis=-1)


def flux_div_v(v, ru, rv, rw, rdnw, fnm, fnp, dx_inv, dy_inv):
    """v-momentum mirror of :func:`flux_div_u` (reference advection.cu:
    616-794, periodic branch :754-794).  Returns (nz, ny+1, nx)."""
    ny = v.shape[-2] - 1
    vc = v[:, :ny, :]
    velx = 0.5 * (np.roll(ru, 1, axis=-2) + ru)             # (nz, ny, nx+1)
    fx = _xfaces(vc, velx)
    vely = 0.5 * (rv[:, :ny, :] + rv[:, 1:, :])
    FY = flux5(*[_roll(vc, o, -2) for o in (-2, -1, 0, 1, 2, 3)], vely)
    dfy = FY - np.roll(FY, 1, axis=-2)
    velz = 0.5 * (np.roll(rw, 1, axis=-2) + rw)
    fz = _zfaces_half(vc, velz, fnm, fnp)
    t = _div_sum(fx[..., 1:] - fx[..., :-1], dfy, fz[1:] - fz[:-1],
                 dx_inv, dy_inv, rdnw[:, None, None])
    return np.concatenate([t, t[:, :1, :]], axis=-2)


def flux_div_w(w, ru, rv, rw, rdn, fnm, fnp, dx_inv, dy_inv):
    """w-momentum flux divergence at interior w levels (reference
    advection.cu:803-1051, periodic branch :1016-1050).  Returns
    (nz+1, ny, nx) with zero rows at k=0 and k=nz."""
    nz = w.shape[0] - 1
    out = np.zeros_like(w)
    FZ = np.zeros((nz,) + w.shape[1:], dtype=w.dtype)
    for m in range(nz):
        velz = 0.5 * (rw[m] + rw[m + 1])
        if m == 0 or m == nz - 1:
            FZ[m] = 0.5 * velz * (w[m + 1] + w[m])
        else:
            FZ[m] = flux3(w[m - 1], w[m], w[m + 1], w[m + 2], velz)
    for k in range(1, nz):
        velx = fnm[k] * ru[k] + fnp[k] * ru[k - 1]
        vely = fnm[k] * rv[k] + fnp[k] * rv[k - 1]
        fx = _xfaces(w[k], velx)
        fy = _yfaces(w[k], vely)
        out[k] = _div_sum(fx[..., 1:] - fx[..., :-1], fy[1:] - fy[:-1],
                          FZ[k] - FZ[k - 1], dx_inv, dy_inv, rdn[k])
    return out


# ---------------------------------------------------------------------------
# Large-step pressure gradient, buoyancy, geopotential RHS
# ---------------------------------------------------------------------------

def _pp(s, prec=False):
    """p' = p - pb; the precision form's EOS keeps it (s.pp)."""
    return s.pp if prec else s.p - _b3(s.pb)


def cq_mass_species(s, mp_physics):
    """The moist mass species WRF calc_cq sums (reference acoustic.py:
    61-110): qv for mp_physics 0, qv/qc/qr for 1, plus qi/qs/qg (and qh
    for NSSL) for the mixed-phase schemes; number moments never."""
    if mp_physics == 0:
        return ("qv",)
    return ("qv", "qc", "qr") + tuple(getattr(s, "extra", ())[
        :getattr(s, "n_load", 0)])


def calc_cq(s, mp_physics):
    """WRF calc_cq at the stage state (reference acoustic.cu:84-158):
    1 / (1 + 0.5 qtot) with qtot the species summed over the two cells
    around each u face, v face and interior w level (cqw = 1 at the
    surface and the lid).  Returns (cqu, cqv, cqw)."""
    f = s.dtype.type
    names = cq_mass_species(s, mp_physics)

    def pair(a_of, b_of):
        q = None
        for n in names:
            x = getattr(s, n)
            q = (f(0.0) + a_of(x)) if q is None else q + a_of(x)
            q = q + b_of(x)
        return f(1.0) / (f(1.0) + f(0.5) * q)

    cqu = pair(lambda x: x, lambda x: np.roll(x, 1, axis=2))
    cqu = np.concatenate([cqu, cqu[..., :1]], axis=2)
    cqv = pair(lambda x: x, lambda x: np.roll(x, 1, axis=1))
    cqv = np.concatenate([cqv, cqv[:, :1]], axis=1)
    nz = s.thp.shape[0]
    cqw = np.ones((nz + 1,) + s.thp.shape[1:], dtype=s.dtype)
    cqw[1:nz] = pair(lambda x: x[1:], lambda x: x[:-1])
    return cqu, cqv, cqw


def _dpn_levels(s, pp, cfn, cfn1, top_lid):
    """Perturbation pressure at full levels 0..nz per column (WRF dpn,
    reference dycore.cu:59-85)."""
    nz = pp.shape[0]
    d = np.zeros((nz + 1,) + pp.shape[1:], dtype=pp.dtype)
    d[0] = (s.cf1 * pp[0] + s.cf2 * pp[1]) + s.cf3 * pp[2]
    for kf in range(1, nz):
        d[kf] = s.fnm[kf] * pp[kf] + s.fnp[kf] * pp[kf - 1]
    if top_lid:
        d[nz] = cfn * pp[nz - 1] + cfn1 * pp[nz - 2]
    return d


def slow_pgf(s, dx, dy, top_lid=True, prec=False):
    """WRF horizontal_pressure_gradient on u and v faces, subtracted from
    ru_t, rv_t (reference dycore.cu:98-211; launcher dycore.py:374-394)."""
    f = s.dtype.type
    rdx, rdy = f(1.0 / dx), f(1.0 / dy)
    hrdx, hrdy = f(0.5 / dx), f(0.5 / dy)
    half = f(0.5)
    mut = s.mub2d + s.mup
    pp = _pp(s, prec)
    pbd = _b3(s.pb)
    dpn = _dpn_levels(s, pp, s.cfn, s.cfn1, top_lid)
    if s.base3d:
        phh = half * ((s.phb[:-1] + s.phb[1:]) + (s.php[:-1] + s.php[1:]))
    else:
        phh = half * (s.php[:-1] + s.php[1:])
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    rdnw = s.rdnw[:, None, None]

    def face(ax, rd, hrd):
        A = lambda a: a                                   # column i / j
        B = lambda a: np.roll(a, 1, axis=ax)              # column i-1 / j-1
        muf = half * (A(mut) + B(mut))
        dmu = half * (A(s.mup) + B(s.mup))
        layer = c1h * muf[None] + c2h
        bracket = (A(s.php[1:]) - B(s.php[1:])) + (A(s.php[:-1])
                                                   - B(s.php[:-1]))
        bracket = bracket + (A(s.alt) + B(s.alt)) * (A(pp) - B(pp))
        if s.base3d:
            bracket = bracket + (A(s.al) + B(s.al)) * (A(pbd) - B(pbd))
        else:
            bracket = bracket + (A(s.al) + B(s.al)) * f(0.0)
        left = (hrd * layer) * bracket
        dphp = A(phh) - B(phh)
        hi = half * (A(dpn[1:]) + B(dpn[1:]))
        lo = half * (A(dpn[:-1]) + B(dpn[:-1]))
        vertical = rdnw * (hi - lo) - c1h * dmu[None]
        return left + (rd * dphp) * vertical

    tu = face(-1, rdx, hrdx)
    tv = face(-2, rdy, hrdy)
    tu = np.concatenate([tu, tu[..., :1]], axis=-1)
    tv = np.concatenate([tv, tv[:, :1, :]], axis=-2)
    cq = getattr(s, "cq", None)
    if cq is not None:                   # moist_cq (dycore.cu:195, :208)
        tu = cq[0] * tu
        tv = cq[1] * tv
    s.ru_t -= tu
    s.rv_t -= tv


def _q_total(s):
    """qv + qc + qr, then the loading condensates in order (reference
    dycore.cu:213-230 q_total: qi, qs, qg and NSSL's qh)."""
    q = (s.qv + s.qc) + s.qr
    for name in getattr(s, "extra", ())[:getattr(s, "n_load", 0)]:
        q = q + getattr(s, name)
    return q


def slow_buoyancy(s, prec=False):
    """WRF pg_buoy_w: rw_t += g (rdn dp'/deta - c1f mu' [moist loading])
    at w levels 1..nz (reference dycore.cu:231-308)."""
    f = s.dtype.type
    nz = s.thp.shape[0]
    pp = _pp(s, prec)
    one, half, two = f(1.0), f(0.5), f(2.0)
    moist = s.qv is not None
    qt = _q_total(s) if moist else None
    c1f, c2f = s.c1f[:, None, None], s.c2f[:, None, None]
    dp = pp[1:] - pp[:-1]                                 # levels 1..nz-1
    if moist:
        cqw = half * (qt[1:] + qt[:-1])
        cq1 = one / (one + cqw)
        term = (cq1 * s.rdn[1:, None, None]) * dp - c1f[1:nz] * s.mup
        term = term - (cqw * cq1) * (c1f[1:nz] * s.mub2d + c2f[1:nz])
    else:
        term = s.rdn[1:, None, None] * dp - c1f[1:nz] * s.mup
    s.rw_t[1:nz] += f(G) * term
    cqw = half * (qt[nz - 1] + qt[nz - 2]) if moist else f(0.0)
    cq1 = one / (one + cqw)
    pressure = ((cq1 * two) * s.rdnw[nz - 1]) * (f(0.0) - pp[nz - 1])
    term = pressure - s.c1f[nz] * s.mup
    if moist:
        term = term - (cqw * cq1) * (s.c1f[nz] * s.mub2d + s.c2f[nz])
    s.rw_t[nz] += f(G) * term


def slow_geopotential(s, ww, dx, dy, prec=False, order=2):
    """WRF rhs_ph with the vertical Omega and g*w terms (reference
    dycore.cu:541-667; launcher dycore.py:443-462), horizontal branch of
    h_sca_adv_order ``order``: 2, or 5 (dycore.cu:668-748, periodic: the
    face-flux sum times the centred 6th-order difference
    (45 d1 - 9 d2 + d3) / 60 of the full geopotential).
    Writes rph_t at levels 1..nz (the top row is overwritten, not added).
    ``prec`` (precision form): the layer thickness of the full geopotential
    is dphb + dphp with dphb = -(dnw (c1h mub + c2h)) alb (the base
    state's hydrostatic relation), and on a flat base the horizontal
    differences are those of php (phb is horizontally uniform)."""
    f = s.dtype.type
    nz = s.thp.shape[0]
    half = f(0.5)
    qrx, qry = f(0.25 / dx), f(0.25 / dy)
    mut = s.mub2d + s.mup
    ph = (s.phb + s.php) if s.base3d else (s.phb[:, None, None] + s.php)
    rdnw = s.rdnw
    # vertical Omega term at interior levels 1..nz-1
    if prec:
        c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
        dphb = -(s.dnw[:, None, None] * (c1h * s.mub2d + c2h)) * _b3(s.alb)
        dph = dphb + (s.php[1:] - s.php[:-1])
    else:
        dph = ph[1:] - ph[:-1]
    wd = ((half * (ww[1:] + ww[:-1])) * rdnw[:, None, None]) \
        * dph                                               # levels 0..nz-1
    tend = s.rph_t.copy()
    tend[nz] = 0.0
    k = slice(1, nz)
    omega = s.fnm[k, None, None] * wd[1:] + s.fnp[k, None, None] * wd[:-1]
    tend[k] = tend[k] - omega
    mass = s.c1f[:, None, None] * mut + s.c2f[:, None, None]
    tend[1:] = tend[1:] + mass[1:] * (f(G) * s.w[1:])
    # horizontal advection of the full geopotential
    avgx = half * (mut + np.roll(mut, 1, axis=1))
    avgy = half * (mut + np.roll(mut, 1, axis=0))
    uc, vc = s.u[..., :-1], s.v[:, :-1, :]
    for kk in range(1, nz + 1):
        top = kk == nz
        qx = f(2.0) * qrx if top else qrx
        qy = f(2.0) * qry if top else qry
        mx = s.c1f[kk] * avgx + s.c2f[kk]
        my = s.c1f[kk] * avgy + s.c2f[kk]
        if top:
            wx = s.cfn * uc[kk - 1] + s.cfn1 * uc[kk - 2]
            wy = s.cfn * vc[kk - 1] + s.cfn1 * vc[kk - 2]
        else:
            wx = uc[kk] + uc[kk - 1]
            wy = vc[kk] + vc[kk - 1]
        fcx = mx * wx                       # face i (between i-1 and i)
        fcy = my * wy
        p = s.php[kk] if (prec and not s.base3d) else ph[kk]
        if order == 5:
            def cen(ax):
                d = [np.roll(p, -o, axis=ax) - np.roll(p, o, axis=ax)
                     for o in (1, 2, 3)]
                return ((f(45.0) * d[0] + f(-9.0) * d[1]) + f(1.0) * d[2]) \
                    / f(60.0)
            hx = (qx * (fcx + np.roll(fcx, -1, axis=1))) * cen(1)
            hy = (qy * (fcy + np.roll(fcy, -1, axis=0))) * cen(0)
            tend[kk] = tend[kk] - (hx + hy)
            continue
        fx = fcx * (p - np.roll(p, 1, axis=1))
        fxp = np.roll(fx, -1, axis=1)
        tend[kk] = tend[kk] - qx * (fxp + fx)
        fy = fcy * (p - np.roll(p, 1, axis=0))
        fyp = np.roll(fy, -1, axis=0)
        tend[kk] = tend[kk] - qy * (fyp + fy)
    s.rph_t[1:] = tend[1:]


def _diff2_spacings(zf, stagger):
    """Inverse vertical spacings (reference diffusion.py:34-61)."""
    zf = np.asarray(zf, dtype=np.float64)
    zh = 0.5 * (zf[:-1] + zf[1:])
    if stagger == "z":
        rdzf = 1.0 / np.diff(zf)
        rdzc = np.zeros(zf.size)
        rdzc[1:-1] = 1.0 / (zh[1:] - zh[:-1])
    else:
        rdzf = 1.0 / np.diff(zh)
        rdzc = 1.0 / np.diff(zf)
    return rdzf, rdzc


def add_diff2(fld, kh, kv, dx, dy, zf, stagger, dtype):
    """Constant-K Laplacian kh (fxx + fyy) + kv (fz)z (reference
    diffusion.cu:22-51; launcher diffusion.py:64-97).  Returns the
    uncoupled tendency array (same shape as ``fld``)."""
    f = np.dtype(dtype).type
    nlev, nys, nxs = fld.shape
    nx = nxs - 1 if stagger == "x" else nxs
    ny = nys - 1 if stagger == "y" else nys
    rdzf, rdzc = (np.asarray(a, dtype=dtype)
                  for a in _diff2_spacings(zf, stagger))
    core = fld[:, :ny, :nx]
    fc = core
    lap = f(kh) * (f(1.0 / dx ** 2) * (_roll(core, 1, -1) - f(2.0) * fc
                                       + _roll(core, -1, -1))
                   + f(1.0 / dy ** 2) * (_roll(core, 1, -2) - f(2.0) * fc
                                         + _roll(core, -1, -2)))
    fup = np.zeros_like(core)
    fdn = np.zeros_like(core)
    fup[:-1] = (core[1:] - core[:-1]) * rdzf[:, None, None]
    fdn[1:] = (core[1:] - core[:-1]) * rdzf[:, None, None]
    lap = lap + f(kv) * (fup - fdn) * rdzc[:, None, None]
    if stagger == "z":
        lap[0] = 0.0
        lap[-1] = 0.0
    out = np.zeros_like(fld)
    out[:, :ny, :nx] = lap
    if nxs > nx:
        out[:, :ny, nx] = lap[:, :, 0]
    if nys > ny:
        out[:, ny, :] = out[:, 0, :]
    return out


def add_diffusion_tendencies(s, khdif, kvdif, dx, dy):
    """Couple the constant-K tendencies with C(mu) and add them to the
    slow tendencies (reference diffusion.py:100-130)."""
    if khdif <= 0.0 and kvdif <= 0.0:
        return
    if s.base3d:
        raise NotImplementedError("constant-K diffusion needs a flat base")
    zf = np.real(s.phb).astype(np.float64) / G
    mu = s.mub2d + s.mup
    mux, muy = mu_at_u_faces(mu), mu_at_v_faces(mu)
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    c1f, c2f = s.c1f[:, None, None], s.c2f[:, None, None]
    for fld, tend, muf, c1, c2, stag in (
            (s.u, s.ru_t, mux, c1h, c2h, "x"),
            (s.v, s.rv_t, muy, c1h, c2h, "y"),
            (s.w, s.rw_t, mu, c1f, c2f, "z"),
            (s.thp, s.rth_t, mu, c1h, c2h, "")):
        tmp = add_diff2(fld, khdif, kvdif, dx, dy, zf, stag, s.dtype)
        tend += (c1 * muf[None] + c2) * tmp


def add_slow_tendencies(s, ru, rv, ww, dx, dy, top_lid=True, prec=False,
                        h_order=2):
    """Advection of theta/u/v/w plus PGF, buoyancy and rhs_ph (reference
    dycore.py:502-569)."""
    f = s.dtype.type
    dxi, dyi = f(1.0 / dx), f(1.0 / dy)
    theta = _b3(s.thb) + s.thp
    s.rth_t += flux_div_scalar(theta, ru, rv, ww, s.rdnw, s.fnm, s.fnp,
                               dxi, dyi)
    s.ru_t += flux_div_u(s.u, ru, rv, ww, s.rdnw, s.fnm, s.fnp, dxi, dyi)
    s.rv_t += flux_div_v(s.v, ru, rv, ww, s.rdnw, s.fnm, s.fnp, dxi, dyi)
    s.rw_t += flux_div_w(s.w, ru, rv, ww, s.rdn, s.fnm, s.fnp, dxi, dyi)
    slow_pgf(s, dx, dy, top_lid, prec)
    slow_buoyancy(s, prec)
    slow_geopotential(s, ww, dx, dy, prec, h_order)


def add_h_diabatic_tendency(s) -> None:
    """The retained microphysics heating enters the coupled theta tendency
    of every stage, rth_t += (c1h mu + c2h) h_diabatic with the stage
    state's column mass (reference dycore.py:572-592, after the slow
    tendencies and before diffusion, dycore.py:2381-2389)."""
    mu = s.mub2d + s.mup
    s.rth_t += (s.c1h[:, None, None] * mu[None]
                + s.c2h[:, None, None]) * s.h_diabatic


# ---------------------------------------------------------------------------
# Small-step preparation / acoustic substeps / finish
# ---------------------------------------------------------------------------

def small_step_init(s, prec=False) -> None:
    """WRF small_step_prep + calc_p_rho step 0 (reference dycore.cu:
    947-1085).  ``prec`` (precision form): theta'' = C(mu_t) theta_t -
    C(mu_s) theta_s is formed without the base theta's cancellation, as
    (c1h mu'') thb + (C(mu_t) thp_t - C(mu_s) thp_s)."""
    f = s.dtype.type
    half = f(0.5)
    mt = s.mub2d + s.mup0
    ms = s.mub2d + s.mup
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    c1f, c2f = s.c1f[:, None, None], s.c2f[:, None, None]
    for pp, x0, x, ax in ((s.u_pp, s.u0, s.u, -1), (s.v_pp, s.v0, s.v, -2)):
        mtf = half * (mt + np.roll(mt, 1, axis=ax))
        msf = half * (ms + np.roll(ms, 1, axis=ax))
        if ax == -1:
            mtf = np.concatenate([mtf, mtf[:, :1]], axis=1)
            msf = np.concatenate([msf, msf[:, :1]], axis=1)
        else:
            mtf = np.concatenate([mtf, mtf[:1]], axis=0)
            msf = np.concatenate([msf, msf[:1]], axis=0)
        pp[...] = (c1h * mtf + c2h) * x0 - (c1h * msf + c2h) * x
    mupp = s.mup0 - s.mup
    s.mu_pp[...] = mupp
    s.w_pp[...] = (c1f * mt + c2f) * s.w0 - (c1f * ms + c2f) * s.w
    s.ph_pp[...] = s.php0 - s.php
    cht = c1h * mt + c2h
    chs = c1h * ms + c2h
    thb = _b3(s.thb)
    tht = thb + s.thp0
    ths = thb + s.thp
    if prec:
        thpp = (c1h * mupp) * thb + (cht * s.thp0 - chs * s.thp)
    else:
        thpp = cht * tht - chs * ths
    s.th_pp[...] = thpp
    c2a = (f(GAMMA) * s.p) / s.alt
    al = -((s.alt * (c1h * mupp)) + s.rdnw[:, None, None]
           * (s.ph_pp[1:] - s.ph_pp[:-1])) / cht
    s.al_pp[...] = al
    p5 = (s.alt * (thpp - (c1h * mupp) * ths)) / (cht * ths)
    s.p_pp[...] = c2a * (p5 - al)
    s.p_pp_old[...] = s.p_pp


def calc_coefs(s, dtau, epssm, top_lid=True, prec=False):
    """c2a and the LU factors a/alpha/gam of the implicit w system
    (reference acoustic.cu:675-759, WRF calc_coef_w).  Returns a dict.
    ``prec``: the precision form -- inputs formed in the state's precision
    as usual, the factors evaluated in float64 (same expression order)."""
    f = s.dtype.type
    nz = s.thp.shape[0]
    mut = s.mub2d + s.mup
    cof = f(0.5) * f(dtau) * f(G) * (f(1.0) + f(epssm))
    cof = cof * cof
    c2a_s = f(GAMMA) * s.p / s.alt
    chm = [s.c1h[k] * mut + s.c2h[k] for k in range(nz)]
    cfm = [s.c1f[k] * mut + s.c2f[k] for k in range(nz + 1)]
    rdn, rdnw = s.rdn, s.rdnw
    if prec:
        f = np.float64
        cof = f(cof)
        c2a = c2a_s.astype(f)
        chm = [x.astype(f) for x in chm]
        cfm = [x.astype(f) for x in cfm]
        rdn, rdnw = rdn.astype(f), rdnw.astype(f)
    else:
        c2a = c2a_s
    shp = (nz + 1,) + mut.shape
    # moist_cq: row k's coefficient cof takes the factor cqw(k)
    # (acoustic.cu:700-744, (cq cof) then the same products)
    cq = getattr(s, "cq", None)
    qcof = [cof if cq is None else cq[2][k].astype(c2a.dtype) * cof
            for k in range(nz + 1)]
    a = np.zeros(shp, c2a.dtype)
    for k in range(2, nz):
        a[k] = (-qcof[k] * rdn[k] * rdnw[k - 1] * c2a[k - 1]
                / (chm[k - 1] * cfm[k - 1]))
    if not top_lid:
        a[nz] = (f(-2.0) * cof * rdnw[nz - 1] * rdnw[nz - 1]
                 * c2a[nz - 1] / (chm[nz - 1] * cfm[nz - 1]))
    b = np.zeros(shp, c2a.dtype)
    cc = np.zeros(shp, c2a.dtype)
    for k in range(1, nz):
        b[k] = f(1.0) + qcof[k] * rdn[k] * (
            rdnw[k] * c2a[k] / (chm[k] * cfm[k])
            + rdnw[k - 1] * c2a[k - 1] / (chm[k - 1] * cfm[k]))
        cc[k] = -qcof[k] * rdn[k] * rdnw[k] * c2a[k] / (chm[k] * cfm[k + 1])
    b[nz] = f(1.0) + f(2.0) * cof * rdnw[nz - 1] * rdnw[nz - 1] \
        * c2a[nz - 1] / (chm[nz - 1] * cfm[nz])
    alpha, gam = lu_factors(a, b, cc)
    return {"c2a": c2a_s, "a": a, "alpha": alpha, "gam": gam}


def lu_factors(a, b, c):
    """Thomas factors of the tridiagonal rows 1..n with sub-diagonal a,
    diagonal b, super-diagonal c (row n has no super-diagonal):
    alpha[k] = 1/(b[k] - a[k] gam[k-1]), gam[k] = c[k] alpha[k], gam[0] =
    gam[n] = 0 (reference acoustic.cu:739-758)."""
    n = a.shape[0] - 1
    one = np.asarray(1.0, dtype=a.dtype)
    alpha = np.zeros_like(a)
    gam = np.zeros_like(a)
    for k in range(1, n):
        alpha[k] = one / (b[k] - a[k] * gam[k - 1])
        gam[k] = c[k] * alpha[k]
    alpha[n] = one / (b[n] - a[n] * gam[n - 1])
    gam[n] = 0.0
    return alpha, gam


def thomas_solve(w, a, alpha, gam):
    """Solve the factored system in place for rows 1..n (row 0 holds the
    boundary value, coupled through a[1]) (reference acoustic.cu:941-948)."""
    n = w.shape[0] - 1
    for k in range(1, n + 1):
        w[k] = (w[k] - a[k] * w[k - 1]) * alpha[k]
    for k in range(n - 1, 0, -1):
        w[k] = w[k] - gam[k] * w[k + 1]
    return w


def _pdmp(p, pold, smdiv):
    return p + smdiv * (p - pold)


def _dpn_face_levels(s, pd, ax, top_lid):
    """Face-averaged damped p'' at full levels 0..nz (reference
    acoustic.cu:163-190 ``dpn_face``)."""
    nz = pd.shape[0]
    f = s.dtype.type
    half = f(0.5)
    S = pd + np.roll(pd, 1, axis=ax)           # A + B per level
    d = np.zeros((nz + 1,) + pd.shape[1:], pd.dtype)
    d[0] = half * (s.cf1 * S[0] + s.cf2 * S[1] + s.cf3 * S[2])
    for kf in range(1, nz):
        d[kf] = half * (s.fnm[kf] * S[kf] + s.fnp[kf] * S[kf - 1])
    if top_lid:
        d[nz] = half * (s.cf1 * S[nz - 1] + s.cf2 * S[nz - 2]
                        + s.cf3 * S[nz - 3])
    return d


def advance_uv(s, dtau, smdiv, dx, dy, top_lid=True):
    """Forward step of u''/v'' (reference acoustic.cu:205-314)."""
    f = s.dtype.type
    half = f(0.5)
    dtau = f(dtau)
    pd = _pdmp(s.p_pp, s.p_pp_old, f(smdiv))
    mut = s.mub2d + s.mup
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    rdnw = s.rdnw[:, None, None]
    for ax, rd, pp, rt in ((-1, f(1.0 / dx), s.u_pp, s.ru_t),
                           (-2, f(1.0 / dy), s.v_pp, s.rv_t)):
        B = lambda a: np.roll(a, 1, axis=ax)
        muf = half * (B(mut) + mut)
        dph = (s.ph_pp[1:] - B(s.ph_pp[1:])) + (s.ph_pp[:-1]
                                                - B(s.ph_pp[:-1]))
        dpe = pd - B(pd)
        brk = dph + (s.alt + B(s.alt)) * dpe
        if s.base3d:
            brk = brk + (s.al_pp + B(s.al_pp)) * (s.pb - B(s.pb))
        dpxy = half * rd * (c1h * muf + c2h) * brk
        dpn = _dpn_face_levels(s, pd, ax, top_lid)
        phs = s.php[:-1] + s.php[1:]
        dphp = half * (phs - B(phs))
        if s.base3d:
            pbs = s.phb[:-1] + s.phb[1:]
            dphp = dphp + half * (pbs - B(pbs))
        dpxy = dpxy + rd * dphp * (rdnw * (dpn[1:] - dpn[:-1])
                                   - half * c1h * (B(s.mu_pp) + s.mu_pp))
        if ax == -1:                     # duplicate face nx == face 0
            dpxy = np.concatenate([dpxy, dpxy[..., :1]], axis=-1)
        else:
            dpxy = np.concatenate([dpxy, dpxy[:, :1, :]], axis=-2)
        cq = getattr(s, "cq", None)
        if cq is not None:               # moist_cq (acoustic.cu:287, :307)
            dpxy = cq[0 if ax == -1 else 1] * dpxy
        pp += dtau * (rt - dpxy)


def advance_mu_th(s, dtau, dx, dy, prec=False):
    """mu'', Omega'' and (mu theta)'' forward step plus history saves
    (reference acoustic.cu:329-456)."""
    f = s.dtype.type
    half = f(0.5)
    dtau = f(dtau)
    rdx, rdy = f(1.0 / dx), f(1.0 / dy)
    nz = s.thp.shape[0]
    s.mu_pp_old = s.mu_pp.copy()
    s.th_pp_old = s.th_pp.copy()
    s.p_pp_old[...] = s.p_pp
    mut = s.mub2d + s.mup
    muw = half * (np.roll(mut, 1, axis=1) + mut)
    mue = half * (mut + np.roll(mut, -1, axis=1))
    mus = half * (np.roll(mut, 1, axis=0) + mut)
    mun = half * (mut + np.roll(mut, -1, axis=0))
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    uw, ue = s.u_pp[..., :-1], s.u_pp[..., 1:]
    vs, vn = s.v_pp[:, :-1, :], s.v_pp[:, 1:, :]
    dvp = rdx * (ue - uw) + rdy * (vn - vs)
    dvr = rdx * ((c1h * mue + c2h) * s.u[..., 1:]
                 - (c1h * muw + c2h) * s.u[..., :-1]) \
        + rdy * ((c1h * mun + c2h) * s.v[:, 1:, :]
                 - (c1h * mus + c2h) * s.v[:, :-1, :])
    dnw = s.dnw[:, None, None]
    dmdt_pp = np.zeros_like(mut)
    dmdt_rf = np.zeros_like(mut)
    kp, kr = _Kahan(mut), _Kahan(mut)
    for k in range(nz):
        if prec:
            dmdt_pp = kp.add(dnw[k] * dvp[k])
            dmdt_rf = kr.add(dnw[k] * dvr[k])
        else:
            dmdt_pp = dmdt_pp + dnw[k] * dvp[k]
            dmdt_rf = dmdt_rf + dnw[k] * dvr[k]
    rmu = s.rmu_t
    s.mu_pp[...] = s.mu_pp + dtau * ((dmdt_pp + dmdt_rf) + rmu)
    ww = np.zeros_like(s.ww_pp)
    wwk = np.zeros_like(mut)
    kw = _Kahan(mut)
    for kk in range(1, nz):
        k = kk - 1
        inc = dnw[k] * (c1h[k] * (dmdt_pp + rmu) + dvp[k])
        wwk = kw.add(-inc) if prec else wwk - inc
        ww[kk] = wwk
    s.ww_pp[...] = ww
    th = _b3(s.thb) + s.thp
    wdtn = np.zeros((nz + 1,) + mut.shape, ww.dtype)
    for k in range(1, nz):
        wdtn[k] = ww[k] * (s.fnm[k] * th[k] + s.fnp[k] * th[k - 1])
    t_ip = np.roll(th, -1, axis=2)
    t_im = np.roll(th, 1, axis=2)
    t_jp = np.roll(th, -1, axis=1)
    t_jm = np.roll(th, 1, axis=1)
    hx = half * rdx * (ue * (t_ip + th) - uw * (th + t_im))
    hy = half * rdy * (vn * (t_jp + th) - vs * (th + t_jm))
    s.th_pp[...] = s.th_pp + dtau * (
        s.rth_t - (hx + hy + s.rdnw[:, None, None] * (wdtn[1:] - wdtn[:-1])))


def advance_w_phi(s, coefs, dtau, epssm, dx, dy, top_lid=True,
                  dampmag=0.0, zdamp=5000.0):
    """Vertically implicit w''/phi'' solve, kinematic lower BC, damp_opt=3
    damper, phi'' update and the fused linearised EOS (reference
    acoustic.cu:774-982 with diagnose_p_column :636-665)."""
    f = s.dtype.type
    half, one = f(0.5), f(1.0)
    dtau, eps = f(dtau), f(epssm)
    g = f(G)
    nz = s.thp.shape[0]
    c2a, a, alpha, gam = (coefs[k] for k in ("c2a", "a", "alpha", "gam"))
    mut = s.mub2d + s.mup
    muts = mut + s.mu_pp
    muave = half * ((one + eps) * s.mu_pp + (one - eps) * s.mu_pp_old)
    cfm = s.c1f[:, None, None] * mut + s.c2f[:, None, None]
    rhs = np.zeros_like(s.w_pp)
    rhs[1:] = dtau * (s.rph_t[1:] + half * g * (one - eps) * s.w_pp[1:])
    ph = (s.phb if s.base3d else s.phb[:, None, None]) + s.php
    wd = half * (s.ww_pp[1:] + s.ww_pp[:-1]) * s.rdnw[:, None, None] \
        * (ph[1:] - ph[:-1])
    for k in range(1, nz):
        rhs[k] = rhs[k] - dtau * (s.fnm[k] * wd[k] + s.fnp[k] * wd[k - 1])
    rhs[1:] = s.ph_pp[1:] + rhs[1:] / cfm[1:]
    if top_lid:
        rhs[nz] = 0.0
    # kinematic terrain lower boundary on w''
    cf1, cf2, cf3 = s.cf1, s.cf2, s.cf3
    ue = cf1 * s.u_pp[0, :, 1:] + cf2 * s.u_pp[1, :, 1:] + cf3 * s.u_pp[2, :, 1:]
    uw = (cf1 * s.u_pp[0, :, :-1] + cf2 * s.u_pp[1, :, :-1]
          + cf3 * s.u_pp[2, :, :-1])
    vn = cf1 * s.v_pp[0, 1:] + cf2 * s.v_pp[1, 1:] + cf3 * s.v_pp[2, 1:]
    vs = cf1 * s.v_pp[0, :-1] + cf2 * s.v_pp[1, :-1] + cf3 * s.v_pp[2, :-1]
    ht = s.ht
    w = s.w_pp.copy()
    w[0] = half * f(1.0 / dy) * ((np.roll(ht, -1, 0) - ht) * vn
                                 + (ht - np.roll(ht, 1, 0)) * vs) \
        + half * f(1.0 / dx) * ((np.roll(ht, -1, 1) - ht) * ue
                                + (ht - np.roll(ht, 1, 1)) * uw)
    thr = _b3(s.thb) + s.thp
    chs = s.c1h[:, None, None] * muts + s.c2h[:, None, None]
    t2 = half * ((one + eps) * s.th_pp + (one - eps) * s.th_pp_old) \
        / (chs * thr)
    chm = s.c1h[:, None, None] * mut + s.c2h[:, None, None]
    cq = getattr(s, "cq", None)
    for k in range(1, nz):
        dph_up = (one + eps) * (rhs[k + 1] - rhs[k]) \
            + (one - eps) * (s.ph_pp[k + 1] - s.ph_pp[k])
        dph_dn = (one + eps) * (rhs[k] - rhs[k - 1]) \
            + (one - eps) * (s.ph_pp[k] - s.ph_pp[k - 1])
        pg = half * dtau * g * s.rdn[k] * (
            c2a[k] * s.rdnw[k] / chm[k] * dph_up
            - c2a[k - 1] * s.rdnw[k - 1] / chm[k - 1] * dph_dn)
        if cq is not None:               # moist_cq (acoustic.cu:899-905)
            pg = cq[2][k] * pg
        w[k] = w[k] + ((dtau * s.rw_t[k] + pg)
                       + dtau * g * (s.rdn[k] * (
                           c2a[k] * s.alt[k] * t2[k]
                           - c2a[k - 1] * s.alt[k - 1] * t2[k - 1])
                           - s.c1f[k] * muave))
    if top_lid:
        w[nz] = 0.0
    else:
        dph_dn = (one + eps) * (rhs[nz] - rhs[nz - 1]) \
            + (one - eps) * (s.ph_pp[nz] - s.ph_pp[nz - 1])
        bterm = dtau * g * c2a[nz - 1] * s.rdnw[nz - 1] * s.rdnw[nz - 1] \
            / chm[nz - 1] * dph_dn
        cterm = dtau * g * (f(2.0) * s.rdnw[nz - 1] * c2a[nz - 1]
                            * s.alt[nz - 1] * t2[nz - 1] + s.c1f[nz] * muave)
        w[nz] = w[nz] + ((dtau * s.rw_t[nz] - bterm) - cterm)
    if np.result_type(a, w) != w.dtype:   # precision form: FP64 tridiagonal
        w = thomas_solve(w.astype(a.dtype), a, alpha, gam).astype(w.dtype)
    else:
        thomas_solve(w, a, alpha, gam)
    if dampmag > 0.0:
        phf = (s.phb if s.base3d else s.phb[:, None, None]) + s.php
        htop = phf[nz] / g
        hbot = htop - f(zdamp)
        for k in range(1, nz + 1):
            hk = phf[k] / g
            dw = damper_weight(hk, hbot, zdamp, dampmag, s.dtype)
            new = (w[k] - dw * cfm[k] * s.w[k]) / (one + dw)
            w[k] = np.where(np.real(hk) >= np.real(hbot), new, w[k])
    s.w_pp[...] = w
    cfs = s.c1f[:, None, None] * muts + s.c2f[:, None, None]
    s.ph_pp[1:] = rhs[1:] + half * dtau * g * (one + eps) * w[1:] / cfs[1:]
    diagnose_p(s, c2a)


def damper_weight(hk, hbot, zdamp, dampmag, dtype):
    """damp_opt = 3 weight dtau*dampcoef*sin^2(pi/2 (z - z_bot)/zdamp) at
    height ``hk`` (reference acoustic.cu:950-965: the kernel's 1.5707963f
    literal; sin evaluated in double and rounded once)."""
    f = np.dtype(dtype).type
    arg = f(1.5707963) * (hk - hbot) / f(zdamp)
    if np.iscomplexobj(arg):        # complex-step derivatives
        sn = np.sin(arg)
    else:
        sn = np.sin(np.asarray(arg).astype(np.float64)).astype(dtype)
    return f(dampmag) * sn * sn


def diagnose_p(s, c2a):
    """Linearised EOS for p'' and alpha'' (reference acoustic.cu:636-665)."""
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    chm = c1h * (s.mub2d + s.mup + s.mu_pp) + c2h
    th_ref = _b3(s.thb) + s.thp
    al = -(s.alt * (c1h * s.mu_pp) + s.rdnw[:, None, None]
           * (s.ph_pp[1:] - s.ph_pp[:-1])) / chm
    s.al_pp[...] = al
    s.p_pp[...] = c2a * (s.alt * (s.th_pp - c1h * s.mu_pp * th_ref)
                         / (chm * th_ref) - al)


def acoustic_substep(s, coefs, dtau, cfg_like, first):
    """One substep: advance_uv -> advance_mu_th -> advance_w_phi
    (reference acoustic.py:330-479).  ``dtau`` is the host double; the
    kernels see float32(dtau), and the damper magnitude is formed in double
    first (acoustic.py:424) exactly like the reference launcher."""
    c = cfg_like
    smdiv = 0.0 if first else c.smdiv
    d32 = float(np.real(s.dtype.type(dtau)))
    advance_uv(s, d32, smdiv, c.dx, c.dy, c.top_lid)
    advance_mu_th(s, d32, c.dx, c.dy,
                  prec=bool(getattr(c, "precision_form", False))
                  and s.dtype == np.float32)
    dampmag = float(np.real(s.dtype.type(dtau * c.dampcoef))) \
        if c.damp_opt == 3 else 0.0
    advance_w_phi(s, coefs, d32, c.epssm, c.dx, c.dy, c.top_lid,
                  dampmag, c.zdamp)


def run_acoustic_only(s, cfg_like, dtau, n):
    """Drive n substeps about a frozen reference state (reference
    acoustic.py:570-583)."""
    if n <= 0:
        return
    coefs = calc_coefs(s, dtau, cfg_like.epssm, cfg_like.top_lid)
    for i in range(n):
        acoustic_substep(s, coefs, dtau, cfg_like, first=(i == 0))


def small_step_finish(s, prec=False, hdiab_dt=0.0) -> None:
    """WRF small_step_finish (reference dycore.cu:1087-1198).  ``prec``
    (precision form): theta' = (C(mu_s) theta_s + theta'') / C(mu_n) - thb
    is formed without the base theta's cancellation, as (C(mu_s) thp +
    theta'' - (c1h mu'') thb) / C(mu_n).  A nonzero ``hdiab_dt`` (the final
    stage of a step carrying h_diabatic) removes what the substeps
    integrated of the heating: the numerator drops
    (hdiab_dt C(mu_s)) h_diabatic (reference dycore.cu:1187-1195,
    dycore.py:1583-1600)."""
    f = s.dtype.type
    half = f(0.5)
    ms = s.mub2d + s.mup
    mn = ms + s.mu_pp
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    c1f, c2f = s.c1f[:, None, None], s.c2f[:, None, None]
    for x, pp, ax in ((s.u, s.u_pp, -1), (s.v, s.v_pp, -2)):
        msf = half * (ms + np.roll(ms, 1, axis=ax))
        mnf = half * (mn + np.roll(mn, 1, axis=ax))
        if ax == -1:
            msf = np.concatenate([msf, msf[:, :1]], axis=1)
            mnf = np.concatenate([mnf, mnf[:, :1]], axis=1)
        else:
            msf = np.concatenate([msf, msf[:1]], axis=0)
            mnf = np.concatenate([mnf, mnf[:1]], axis=0)
        x[...] = ((c1h * msf + c2h) * x + pp) / (c1h * mnf + c2h)
    s.w[...] = ((c1f * ms + c2f) * s.w + s.w_pp) / (c1f * mn + c2f)
    s.php[...] = s.php + s.ph_pp
    thb = _b3(s.thb)
    if prec:
        num = ((c1h * ms + c2h) * s.thp + s.th_pp) - (c1h * s.mu_pp) * thb
        s.thp[...] = num / (c1h * mn + c2h)
    else:
        cs = c1h * ms + c2h
        num = cs * (thb + s.thp) + s.th_pp
        if hdiab_dt:
            num = num - (f(hdiab_dt) * cs) * s.h_diabatic
        s.thp[...] = num / (c1h * mn + c2h) - thb
    s.mup[...] = s.mup + s.mu_pp


def set_w_surface(s, dx, dy) -> None:
    """Kinematic lower BC w(sfc) = u . grad(ht) (reference dycore.py:
    1999-2026)."""
    f = s.dtype.type
    ht = s.ht
    uc = s.cf1 * s.u[0] + s.cf2 * s.u[1] + s.cf3 * s.u[2]
    vc = s.cf1 * s.v[0] + s.cf2 * s.v[1] + s.cf3 * s.v[2]
    dyn = np.roll(ht, -1, 0) - ht
    dys = ht - np.roll(ht, 1, 0)
    dxe = np.roll(ht, -1, 1) - ht
    dxw = ht - np.roll(ht, 1, 1)
    s.w[0] = (f(0.5 / dy) * (dyn * vc[1:, :] + dys * vc[:-1, :])
              + f(0.5 / dx) * (dxe * uc[:, 1:] + dxw * uc[:, :-1]))


# ---------------------------------------------------------------------------
# Moist scalar transport (qv/qc/qr) with time-averaged mass fluxes
# ---------------------------------------------------------------------------

def sumflux_finish(ru_m, rv_m, ww_m, ru, rv, ww, nsub):
    """WRF sumflux at the last substep (reference acoustic.cu:1469-1489):
    the substep mean of the perturbation fluxes plus the stage-reference
    coupled fluxes."""
    fn = ru_m.dtype.type(nsub)
    return ru_m / fn + ru, rv_m / fn + rv, ww_m / fn + ww


def advance_scalars_stage(s, ru_m, rv_m, ww_m, dt_stage, final, dx, dy,
                          moist_adv_opt=1, fixed=None):
    """Coupled qv/qc/qr stage update with the substep-averaged fluxes
    (reference moist.py:434-698, _update_scalar_kernel :353-402) and the
    positive-definite limiter on the final stage (pd_advection.cu);
    ``fixed`` holds the held mixing tendencies per species."""
    from oracle import moist_np
    moist_np.advance_scalars_stage(s, ru_m, rv_m, ww_m, dt_stage, final,
                                   dx, dy, moist_adv_opt, fixed)


# ---------------------------------------------------------------------------
# The step
# ---------------------------------------------------------------------------

def diff6_factor(cfg) -> float:
    """Per-step 6th-order filter factor (reference dycore.py:1767-1786):
    the namelist factor, or, when the model clock step clock_dt differs
    from dt, the factor whose dt/clock_dt-fold composition removes the
    namelist fraction of a 2dx mode per clock step:
    1 - (1 - factor)^(dt/clock_dt)."""
    clock = cfg.clock_dt if cfg.clock_dt > 0.0 else cfg.dt
    factor = float(cfg.diff_6th_factor)
    if clock == cfg.dt or factor == 1.0:
        return factor
    return -math.expm1((cfg.dt / clock) * math.log1p(-factor))


# Xue (2000) eq. 3 flux weights of the 6th-order filter: 10, 5, 1
DIFF6_W = (10.0, 5.0)


def diff6(f, mut, c1, c2, coef, mono, variant, wstag, ny, nx, slope=False,
          phb=None, dzthr_x=1.0, dzthr_y=1.0):
    """WRF sixth_order_diffusion's coupled tendency of one periodic field
    (reference kernels/diff6.cu; Xue 2000 eq. 3 fluxes, the mono = 2
    up-gradient knock-down, the face masses of the u (variant 1), v (2)
    and mass/w (0) positions, the diff_6th_slopeopt taper when ``slope``;
    msf = 1).  f: (nlev, nys, nxs); returns 0 + (x part + y part), zero on
    a w field's boundary levels (wstag)."""
    t = f.dtype.type
    nlev, nys, nxs = f.shape
    jj = np.arange(nys) % ny
    ii = np.arange(nxs) % nx
    core = f[:, :ny, :nx]

    def fx(m):
        return core[:, jj][:, :, (ii + m) % nx]

    def fy(m):
        return core[:, (jj + m) % ny][:, :, ii]

    def mu(dj, di):
        return mut[((jj + dj) % ny)[:, None], ((ii + di) % nx)[None, :]][None]

    ten, five = t(DIFF6_W[0]), t(DIFF6_W[1])
    q, h, one, zero = t(0.25), t(0.5), t(1.0), t(0.0)
    C1, C2 = c1[:, None, None], c2[:, None, None]

    def fluxes(g):
        p0 = ten * (g(0) - g(-1)) - five * (g(1) - g(-2)) + (g(2) - g(-3))
        p1 = ten * (g(1) - g(0)) - five * (g(2) - g(-1)) + (g(3) - g(-2))
        if mono == 2:
            p0 = np.where(p0 * (g(0) - g(-1)) <= zero, zero, p0)
            p1 = np.where(p1 * (g(1) - g(0)) <= zero, zero, p1)
        return p0, p1

    def ph(dj, di):
        lev = np.minimum(np.arange(nlev), phb.shape[0] - 1)
        return phb[lev][:, ((jj + dj) % ny)[:, None], ((ii + di) % nx)[None, :]]

    def taper(dz, thr):
        return np.maximum(one - dz / t(thr), zero)

    px0, px1 = fluxes(fx)
    if variant == 1:
        mx0, mx1 = mu(0, -1), mu(0, 0)
    elif variant == 2:
        mx0 = q * (mu(-1, -1) + mu(-1, 0) + mu(0, -1) + mu(0, 0))
        mx1 = q * (mu(-1, 0) + mu(-1, 1) + mu(0, 0) + mu(0, 1))
    else:
        mx0 = h * (mu(0, -1) + mu(0, 0))
        mx1 = h * (mu(0, 0) + mu(0, 1))
    sx0 = sx1 = one
    sy0 = sy1 = one
    if slope:
        a0 = np.abs(ph(0, 0) - ph(0, -1))
        a1 = np.abs(ph(0, 1) - ph(0, 0))
        d0, d1 = a0, a1
        if variant == 1:
            d0 = np.fmax(a0, np.abs(ph(0, -1) - ph(0, -2)))
            d1 = np.fmax(a1, a0)
        elif variant == 2:
            d0 = np.fmax(a0, np.abs(ph(-1, 0) - ph(-1, -1)))
            d1 = np.fmax(a1, np.abs(ph(-1, 1) - ph(-1, 0)))
        sx0, sx1 = taper(d0, dzthr_x), taper(d1, dzthr_x)
        b0 = np.abs(ph(0, 0) - ph(-1, 0))
        b1 = np.abs(ph(1, 0) - ph(0, 0))
        d0, d1 = b0, b1
        if variant == 1:
            d0 = np.fmax(b0, np.abs(ph(0, -1) - ph(-1, -1)))
            d1 = np.fmax(b1, np.abs(ph(1, -1) - ph(0, -1)))
        elif variant == 2:
            d0 = np.fmax(b0, np.abs(ph(-1, 0) - ph(-2, 0)))
            d1 = np.fmax(b1, b0)
        sy0, sy1 = taper(d0, dzthr_y), taper(d1, dzthr_y)
    cf = t(coef)
    tx = cf * (sx1 * (C1 * mx1 + C2) * px1 - sx0 * (C1 * mx0 + C2) * px0)
    py0, py1 = fluxes(fy)
    if variant == 1:
        my0 = q * (mu(-1, -1) + mu(-1, 0) + mu(0, -1) + mu(0, 0))
        my1 = q * (mu(0, -1) + mu(0, 0) + mu(1, -1) + mu(1, 0))
    elif variant == 2:
        my0, my1 = mu(-1, 0), mu(0, 0)
    else:
        my0 = h * (mu(-1, 0) + mu(0, 0))
        my1 = h * (mu(0, 0) + mu(1, 0))
    ty = cf * (sy1 * (C1 * my1 + C2) * py1 - sy0 * (C1 * my0 + C2) * py0)
    out = zero + (tx + ty)
    if wstag:
        out[0] = zero
        out[-1] = zero
    return out


def diff6_held(s, cfg):
    """The once-per-step 6th-order tendencies of u, v, w, theta' from the
    time-t fields (reference dycore.py:1345-1420 with km_opt != 2..4:
    prepare_fixed_tendencies; launch_diff6 dycore.py:1647-1740 for the
    coefficient and the slope threshold), coupled with the time-t column
    mass; added to every RK stage's tendencies."""
    t = s.dtype.type
    ny, nx = s.mub2d.shape
    mut = s.mub2d + s.mup0
    coef = float(diff6_factor(cfg)) * 0.015625 / (2.0 * cfg.dt)
    slope = int(getattr(cfg, "diff_6th_slopeopt", 0)) >= 1 and s.base3d
    thr = float(getattr(cfg, "diff_6th_thresh", 0.10))
    kw = dict(slope=slope, phb=s.phb if slope else None,
              dzthr_x=float(t(thr * 9.81 * cfg.dx)),
              dzthr_y=float(t(thr * 9.81 * cfg.dy)))
    mono = int(cfg.diff_6th_opt)
    out = {"ru_t": diff6(s.u0, mut, s.c1h, s.c2h, coef, mono, 1, False,
                         ny, nx, **kw),
           "rv_t": diff6(s.v0, mut, s.c1h, s.c2h, coef, mono, 2, False,
                         ny, nx, **kw),
           "rw_t": diff6(s.w0, mut, s.c1f, s.c2f, coef, mono, 0, True,
                         ny, nx, **kw),
           "rth_t": diff6(s.thp0, mut, s.c1h, s.c2h, coef, mono, 0, False,
                          ny, nx, **kw)}
    if getattr(s, "qv", None) is not None:
        # every transported species, less WRF's moist array under
        # moist_mix6_off (reference dycore.py:1200-1220, :1391-1393)
        off = DIFF6_MOIST_ARRAY if getattr(cfg, "moist_mix6_off", False) \
            else ()
        out["q"] = {n: diff6(getattr(s, n + "0"), mut, s.c1h, s.c2h, coef,
                             mono, 0, False, ny, nx, **kw)
                    for n in ("qv", "qc", "qr") + tuple(getattr(s, "extra",
                                                              ()))
                    if n not in off}
    return out


#: WRF's moist array: the species moist_mix6_off exempts from the filter
#: (reference moist.py:172-173); number and volume tracers keep it
DIFF6_MOIST_ARRAY = ("qv", "qc", "qr", "qi", "qs", "qg", "qh")


def step(s, cfg) -> None:
    """One RK3 split-explicit step (reference dycore.py:2161-2503, dry or
    qv-tracer, periodic, msf = 1)."""
    if cfg.time_step_sound % 2 != 0:
        raise ValueError("time_step_sound must be even")
    f = s.dtype.type
    for name in PROGNOSTICS:
        getattr(s, name + "0")[...] = getattr(s, name)
    moist = s.qv is not None
    if moist:
        for q in ("qv", "qc", "qr") + getattr(s, "extra", ()):
            getattr(s, q + "0")[...] = getattr(s, q)
    hdiab = getattr(s, "h_diabatic", None)
    hypso = getattr(cfg, "hypsometric_opt", 1)
    ns = cfg.time_step_sound
    stages = ((1, cfg.dt / 3.0), (max(ns // 2, 1), cfg.dt / ns),
              (ns, cfg.dt / ns))
    prec = bool(getattr(cfg, "precision_form", False)) \
        and s.dtype == np.float32
    if hdiab is not None and prec:
        raise ValueError("the precision form carries no h_diabatic removal")
    held = {}
    if getattr(cfg, "km_opt", 1) == 4:
        # the Smagorinsky density 1/alt of the time-t state (reference
        # dycore.py:2353-2354 refreshes the EOS first)
        from .smag_np import smag_held
        calc_p_alpha(s, prec, hypso)
        held = smag_held(s, cfg)
    if getattr(cfg, "diff_6th_opt", 0) > 0:
        d6 = diff6_held(s, cfg)
        if held:
            for n, h in d6.items():
                if n == "q":
                    for q, hq in h.items():
                        held["q"][q] = held["q"][q] + hq
                else:
                    held[n] = held[n] + h
        else:
            held = d6
    held_q = held.pop("q", None)
    for istage, (nsub, dtau) in enumerate(stages):
        for name in ("ru_t", "rv_t", "rw_t", "rth_t", "rph_t", "rmu_t"):
            getattr(s, name)[...] = 0.0
        calc_p_alpha(s, prec, hypso)
        ru, rv, ww = stage_fluxes(s, cfg.dx, cfg.dy, prec)
        # moist_cq: WRF calc_cq once per stage from the stage state
        s.cq = (calc_cq(s, getattr(cfg, "mp_physics", 0))
                if moist and getattr(cfg, "moist_cq", False) else None)
        add_slow_tendencies(s, ru, rv, ww, cfg.dx, cfg.dy, cfg.top_lid, prec,
                            getattr(cfg, "h_sca_adv_order", 2))
        if hdiab is not None:
            add_h_diabatic_tendency(s)
        if cfg.km_opt == 1:
            add_diffusion_tendencies(s, cfg.khdif, cfg.kvdif, cfg.dx, cfg.dy)
        for name, h in held.items():       # held Smagorinsky / 6th-order
            getattr(s, name)[...] = getattr(s, name) + h
        small_step_init(s, prec)
        coefs = calc_coefs(s, dtau, cfg.epssm, cfg.top_lid, prec)
        if moist:
            ru_m = np.zeros_like(s.u)
            rv_m = np.zeros_like(s.v)
            ww_m = np.zeros_like(s.w)
        for i in range(nsub):
            acoustic_substep(s, coefs, dtau, cfg, first=(i == 0))
            if moist:
                ru_m += s.u_pp
                rv_m += s.v_pp
                ww_m += s.ww_pp
        # the final stage removes the heating its substeps integrated,
        # over its length nsub * dtau (= dt)
        small_step_finish(s, prec, nsub * dtau
                          if hdiab is not None and istage == 2 else 0.0)
        if moist:
            ru_m, rv_m, ww_m = sumflux_finish(ru_m, rv_m, ww_m, ru, rv, ww,
                                              nsub)
            advance_scalars_stage(s, ru_m, rv_m, ww_m, nsub * dtau,
                                  istage == 2, cfg.dx, cfg.dy,
                                  getattr(cfg, "moist_adv_opt", 1), held_q)
    s.cq = None
    set_w_surface(s, cfg.dx, cfg.dy)
    calc_p_alpha(s, prec, hypso)
    s.elapsed_seconds += cfg.dt


def advance_stage_advective(s, dt_eff) -> None:
    """Uncoupled update of theta, u, v, w from the time-t copies with the
    stage's coupled tendencies, mu' frozen (reference dycore.py:1602-1643,
    msf = 1)."""
    f = s.dtype.type
    dt_eff = f(dt_eff)
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    c1f, c2f = s.c1f[:, None, None], s.c2f[:, None, None]
    mu0 = s.mub2d + s.mup0
    mu = s.mub2d + s.mup
    thb = _b3(s.thb)
    s.thp[...] = (((c1h * mu0[None] + c2h) * (thb + s.thp0)
                   + dt_eff * s.rth_t) / (c1h * mu[None] + c2h)) - thb
    mux0, mux = mu_at_u_faces(mu0), mu_at_u_faces(mu)
    s.u[...] = (((c1h * mux0[None] + c2h) * s.u0 + dt_eff * s.ru_t)
                / (c1h * mux[None] + c2h))
    muy0, muy = mu_at_v_faces(mu0), mu_at_v_faces(mu)
    s.v[...] = (((c1h * muy0[None] + c2h) * s.v0 + dt_eff * s.rv_t)
                / (c1h * muy[None] + c2h))
    s.w[...] = (((c1f * mu0[None] + c2f) * s.w0 + dt_eff * s.rw_t)
                / (c1f * mu[None] + c2f))


def step_advective(s, cfg) -> None:
    """The reference's advection-only step, step(acoustic=False)
    (dycore.py:2318-2343): three stages of length dt/3, dt/2, dt, each the
    EOS, the coupled flux-form advection of theta, u, v, w with Omega = 0
    (advection.py:134-182; w must be zero) and the constant-K diffusion,
    advanced from the time-t fields with w, phi' and mu' frozen; then the
    kinematic lower boundary.  Dry periodic states only."""
    if s.qv is not None:
        raise NotImplementedError("the advection-only step is dry")
    f = s.dtype.type
    for name in PROGNOSTICS:
        getattr(s, name + "0")[...] = getattr(s, name)
    nz = s.thp.shape[0]
    dxi, dyi = f(1.0 / cfg.dx), f(1.0 / cfg.dy)
    c1h, c2h = s.c1h[:, None, None], s.c2h[:, None, None]
    for dt_eff in (cfg.dt / 3.0, cfg.dt / 2.0, cfg.dt):
        for name in ("ru_t", "rv_t", "rw_t", "rth_t", "rph_t", "rmu_t"):
            getattr(s, name)[...] = 0.0
        calc_p_alpha(s, False, getattr(cfg, "hypsometric_opt", 1))
        if np.any(s.w):
            raise ValueError("the advection-only step needs w == 0")
        mu = s.mub2d + s.mup
        ru = (c1h * mu_at_u_faces(mu)[None] + c2h) * s.u
        rv = (c1h * mu_at_v_faces(mu)[None] + c2h) * s.v
        rw = np.zeros((nz + 1,) + mu.shape, dtype=s.dtype)
        s.rth_t += flux_div_scalar(_b3(s.thb) + s.thp, ru, rv, rw, s.rdnw,
                                   s.fnm, s.fnp, dxi, dyi)
        s.ru_t += flux_div_u(s.u, ru, rv, rw, s.rdnw, s.fnm, s.fnp, dxi, dyi)
        s.rv_t += flux_div_v(s.v, ru, rv, rw, s.rdnw, s.fnm, s.fnp, dxi, dyi)
        s.rw_t += flux_div_w(s.w, ru, rv, rw, s.rdn, s.fnm, s.fnp, dxi, dyi)
        if cfg.km_opt == 1:
            add_diffusion_tendencies(s, cfg.khdif, cfg.kvdif, cfg.dx, cfg.dy)
        advance_stage_advective(s, dt_eff)
    set_w_surface(s, cfg.dx, cfg.dy)
    s.elapsed_seconds += cfg.dt


def run_steps(s, cfg, n) -> None:
    for _ in range(n):
        step(s, cfg)


def domain_mass(s) -> float:
    """FP64 sum of the total dry column mass (reference dycore.py:209-218,
    identity map factors)."""
    return float(np.sum((s.mub2d + s.mup).astype(np.float64)))


def stability_report(s, cfg=None) -> dict:
    """max |u|, |w|, |theta'| and the co-located CFL (reference
    dycore.py:2513-2582, without the boundary-width extras)."""
    u_max = float(np.max(np.abs(s.u)))
    w_max = float(np.max(np.abs(s.w)))
    th_max = float(np.max(np.abs(s.thp)))
    nan = not (np.isfinite(u_max) and np.isfinite(w_max)
               and np.isfinite(th_max))
    rep = {"u_max": u_max, "w_max": w_max, "th_max": th_max, "nan": nan,
           "cfl": None}
    if cfg is not None and not nan:
        ph = (s.phb if s.base3d else s.phb[:, None, None]) + s.php
        dz = (ph[1:] - ph[:-1]) / G
        wcfl = float(np.max(np.abs(s.w[1:]) / dz))
        rep["horizontal_cfl"] = cfg.dt * u_max / cfg.dx
        rep["vertical_cfl"] = cfg.dt * wcfl
        rep["cfl"] = max(rep["horizontal_cfl"], rep["vertical_cfl"])
    return rep


def from_case(case, dtype=np.float32) -> OState:
    """Oracle state from a host initial condition (``arwen_amd.cases``
    HostCase): allocate, load base, cast the initial fields."""
    cfg = case.cfg
    from oracle.moist_np import EXTRA_SPECIES, LOADING_EXTRA
    moist = "qv" in case.fields
    mp = getattr(cfg, "mp_physics", 0) if moist else 0
    s = make_state(cfg.nz, cfg.ny, cfg.nx, case.coord, case.base,
                   dtype=dtype, moist=moist, cf=case.cf,
                   extra=EXTRA_SPECIES[mp], n_load=LOADING_EXTRA[mp],
                   h_diabatic=mp != 0)
    for name, arr in case.fields.items():
        getattr(s, name)[...] = np.asarray(arr, dtype=dtype)
    return s
