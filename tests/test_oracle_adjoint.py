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


# ---------------------------------------------------------------------------
# Acoustic substep adjoint, pinned by complex-step derivatives of the
# oracle's own forward substep (exact to rounding: no cancellation)
# ---------------------------------------------------------------------------

def acoustic_case(seed=4, hill=0.0, damp=False, top_lid=True, nz=9, ny=6,
                  nx=10):
    """FP64 oracle state about a seeded random reference state with random
    small-step fields and tendencies (the pattern of the reference's
    test_acoustic.py:39-115), its coefficients and config."""
    import dataclasses
    from tests.helpers import random_state_case
    case, rng = random_state_case(seed=seed, nz=nz, ny=ny, nx=nx, hill=hill)
    cfg = dataclasses.replace(case.cfg, top_lid=top_lid,
                              damp_opt=3 if damp else 0, zdamp=5000.0,
                              dampcoef=0.2)
    o = D.from_case(case, np.float64)
    D.calc_p_alpha(o)
    amp = {"ru_t": 0.2, "rv_t": 0.2, "rw_t": 0.2, "rth_t": 0.2,
           "rph_t": 0.2, "rmu_t": 0.5, "u_pp": 0.2, "v_pp": 0.2,
           "w_pp": 0.2, "th_pp": 0.2, "mu_pp": 0.5, "ph_pp": 1.5,
           "p_pp": 1.0, "p_pp_old": 1.0, "al_pp": 3e-7}
    for n, a in amp.items():
        arr = getattr(o, n)
        arr[...] = a * rng.standard_normal(arr.shape)
    dtau = cfg.dt / cfg.time_step_sound
    coefs = D.calc_coefs(o, dtau, cfg.epssm, cfg.top_lid)
    return o, coefs, dtau, cfg, rng


def _complex_jvp(o, coefs, dtau, cfg, first, dx, h=1e-30):
    import copy
    sc = copy.deepcopy(o)
    for n in A.SUBSTEP_STATE + A.SUBSTEP_TEND + ("ww_pp",):
        setattr(sc, n, getattr(sc, n).astype(np.complex128))
    for n, d in dx.items():
        getattr(sc, n)[...] += 1j * h * d
    D.acoustic_substep(sc, coefs, dtau, cfg, first)
    return {n: np.imag(getattr(sc, n)) / h
            for n in A.SUBSTEP_STATE + ("ww_pp",)}


def _rand_like(o, names, rng):
    return {n: rng.standard_normal(getattr(o, n).shape) for n in names}


@pytest.mark.parametrize("hill,damp,top_lid,first", [
    (0.0, False, True, True), (0.0, False, True, False),
    (300.0, True, True, False), (0.0, True, False, False)])
def test_acoustic_substep_adjoint_dot_product(hill, damp, top_lid, first):
    o, coefs, dtau, cfg, rng = acoustic_case(hill=hill, damp=damp,
                                             top_lid=top_lid)
    dx = _rand_like(o, A.SUBSTEP_STATE + A.SUBSTEP_TEND, rng)
    y = _rand_like(o, A.SUBSTEP_STATE + ("ww_pp",), rng)
    jdx = _complex_jvp(o, coefs, dtau, cfg, first, dx)
    jty = A.acoustic_substep_adjoint(o, coefs, dtau, cfg, first, y)
    lhs = sum(np.vdot(jdx[n], y[n]) for n in y)
    rhs = sum(np.vdot(dx[n], jty[n]) for n in dx)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


def test_acoustic_adjoint_per_field_blocks():
    """Each input/output block of J^T on its own (a wrong block cannot
    hide behind the others)."""
    o, coefs, dtau, cfg, rng = acoustic_case(hill=200.0, damp=True)
    for nin in A.SUBSTEP_STATE + A.SUBSTEP_TEND:
        dx = _rand_like(o, (nin,), rng)
        jdx = _complex_jvp(o, coefs, dtau, cfg, False, dx)
        for nout in A.SUBSTEP_STATE + ("ww_pp",):
            y = _rand_like(o, (nout,), rng)
            jty = A.acoustic_substep_adjoint(o, coefs, dtau, cfg, False, y)
            lhs = np.vdot(jdx[nout], y[nout])
            rhs = np.vdot(dx[nin], jty[nin])
            scale = max(abs(lhs), 1e-300)
            assert abs(lhs - rhs) <= 1e-11 * scale + 1e-280, (nin, nout,
                                                              lhs, rhs)


def test_planted_acoustic_adjoint_mistake_fails(monkeypatch):
    """Dropping the mass-denominator term of the phi'' update (treating
    the substep as if it were affine) breaks the identity."""
    o, coefs, dtau, cfg, rng = acoustic_case()
    o.mu_pp[...] *= 1e3                     # make that term visible
    dx = {"mu_pp": rng.standard_normal(o.mu_pp.shape)}
    y = {"ph_pp": rng.standard_normal(o.ph_pp.shape)}
    jdx = _complex_jvp(o, coefs, dtau, cfg, False, dx)
    jty = A.acoustic_substep_adjoint(o, coefs, dtau, cfg, False, y)
    good = abs(np.vdot(jdx["ph_pp"], y["ph_pp"]) - np.vdot(dx["mu_pp"],
                                                            jty["mu_pp"]))
    o2 = o
    orig = A._forward_pieces

    def no_mass(*a, **k):
        fw = orig(*a, **k)
        fw["W1"] = fw["W1"] * 0.0           # kills -K W1 c1f / cfs^2 term
        return fw
    monkeypatch.setattr(A, "_forward_pieces", no_mass)
    jty2 = A.acoustic_substep_adjoint(o2, coefs, dtau, cfg, False, y)
    bad = abs(np.vdot(jdx["ph_pp"], y["ph_pp"]) - np.vdot(dx["mu_pp"],
                                                           jty2["mu_pp"]))
    assert good <= 1e-11 * abs(np.vdot(jdx["ph_pp"], y["ph_pp"]))
    assert bad > 1e3 * max(good, 1e-300)


# ---------------------------------------------------------------------------
# Brute-force Jacobians of the advection operators
# ---------------------------------------------------------------------------

def advection_case(op, seed=7, nz=5, ny=4, nx=6):
    """Tiny periodic inputs for flux_div_<op>: field, mass fluxes with
    consistent duplicate faces, coordinate factors (vel kept away from 0)."""
    rng = np.random.default_rng(seed)
    shape = {"scalar": (nz, ny, nx), "u": (nz, ny, nx + 1),
             "v": (nz, ny + 1, nx), "w": (nz + 1, ny, nx)}[op]
    q = rng.standard_normal(shape)
    sgn = lambda s: np.where(rng.uniform(size=s) < 0.5, -1.0, 1.0)
    ru = sgn((nz, ny, nx + 1)) * rng.uniform(0.5, 2.0, (nz, ny, nx + 1))
    ru[..., -1] = ru[..., 0]
    rv = sgn((nz, ny + 1, nx)) * rng.uniform(0.5, 2.0, (nz, ny + 1, nx))
    rv[:, -1] = rv[:, 0]
    rw = sgn((nz + 1, ny, nx)) * rng.uniform(0.5, 2.0, (nz + 1, ny, nx))
    rw[0] = 0.0
    rw[-1] = 0.0
    if op == "u":
        q[..., -1] = q[..., 0]
    if op == "v":
        q[:, -1] = q[:, 0]
    rz = -1.0 / rng.uniform(0.05, 0.2, nz + (1 if op == "w" else 0))
    fnm = np.r_[0.0, rng.uniform(0.4, 0.6, nz)]
    fnp = np.r_[0.0, 1.0 - fnm[1:]]
    return q, ru, rv, rw, rz, fnm, fnp, 0.37, 0.29


@pytest.mark.parametrize("op", ["scalar", "u", "v", "w"])
def test_complex_step_jacobian_matches_differences(op):
    args = advection_case(op)
    J = A.flux_div_jacobians(op, *args)
    rng = np.random.default_rng(1)
    fn = {"scalar": D.flux_div_scalar, "u": D.flux_div_u,
          "v": D.flux_div_v, "w": D.flux_div_w}[op]
    q, ru, rv, rw = args[:4]
    for name, x in (("q", q), ("ru", ru), ("rv", rv), ("rw", rw)):
        d = rng.standard_normal(x.shape) * 1e-6
        xs = {"q": q, "ru": ru, "rv": rv, "rw": rw}
        xp = dict(xs); xp[name] = x + d
        xm = dict(xs); xm[name] = x - d
        fd = (fn(xp["q"], xp["ru"], xp["rv"], xp["rw"], *args[4:])
              - fn(xm["q"], xm["ru"], xm["rv"], xm["rw"], *args[4:])) / 2
        jd = (J[name] @ d.reshape(-1)).reshape(fd.shape)
        assert np.allclose(fd, jd, rtol=1e-6, atol=1e-12), name


def test_scalar_jacobian_transpose_is_the_hand_adjoint():
    q, ru, rv, rw, rz, fnm, fnp, dxi, dyi = advection_case("scalar")
    J = A.flux_div_jacobians("scalar", q, ru, rv, rw, rz, fnm, fnp, dxi, dyi)
    lam = np.random.default_rng(2).standard_normal(q.shape)
    ref = A.flux_div_scalar_adjoint(lam, ru, rv, rw, rz, fnm, fnp, dxi, dyi)
    got = (J["q"].T @ lam.reshape(-1)).reshape(q.shape)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("op", ["scalar", "u", "v", "w"])
@pytest.mark.parametrize("shape", [(5, 4, 6), (6, 3, 7), (4, 1, 5)])
def test_advection_adjoint_is_jacobian_transpose(op, shape):
    """The hand adjoint of every advection operator equals the transpose
    of its brute-force Jacobian, for the field and all three mass fluxes,
    with independent (inconsistent) duplicate faces."""
    nz, ny, nx = shape
    args = list(advection_case(op, seed=11, nz=nz, ny=ny, nx=nx))
    rng = np.random.default_rng(3)
    for i in (1, 2):          # break the duplicate-face consistency
        args[i] = args[i] * rng.uniform(0.5, 1.5, args[i].shape)
    if op in ("u", "v"):
        args[0] = args[0] + rng.standard_normal(args[0].shape)
    J = A.flux_div_jacobians(op, *args)
    lam = rng.standard_normal(J["tend"].shape)
    got = A.flux_div_adjoint(op, lam, *args)
    for name in ("q", "ru", "rv", "rw"):
        ref = (J[name].T @ lam.reshape(-1)).reshape(got[name].shape)
        scale = np.abs(ref).max()
        assert np.abs(got[name] - ref).max() <= 1e-12 * scale, name


def test_planted_advection_adjoint_mistake_fails(monkeypatch):
    """A wrong upwind coefficient in the adjoint (S5 sign) is caught by the
    Jacobian comparison."""
    args = advection_case("u", seed=5)
    J = A.flux_div_jacobians("u", *args)
    lam = np.random.default_rng(4).standard_normal(J["tend"].shape)
    monkeypatch.setattr(A, "S5", -A.S5)
    got = A.flux_div_adjoint("u", lam, *args)
    ref = (J["ru"].T @ lam.reshape(-1)).reshape(got["ru"].shape)
    assert np.abs(got["ru"] - ref).max() > 1e-3 * np.abs(ref).max()


# ---------------------------------------------------------------------------
# The large-step tendency pass: hand adjoint vs complex-step JVPs
# ---------------------------------------------------------------------------

def _stage_state(name, seed, dtype=np.float64):
    from tests import pins
    case = pins.CASES[name]()
    s = D.from_case(case, dtype)
    rng = np.random.default_rng(seed)
    s.mup[...] += rng.standard_normal(s.mup.shape) * 20.0
    s.php[1:] += rng.standard_normal(s.php[1:].shape) * 30.0
    s.w[1:-1] += rng.standard_normal(s.w[1:-1].shape) * 0.3
    s.v[...] += rng.standard_normal(s.v.shape) * 0.5
    return case, s


def _cplx(s):
    c = D.OState(**{k: (v.astype(np.complex128) if isinstance(v, np.ndarray)
                        and v.dtype.kind == "f" else v)
                    for k, v in vars(s).items()})
    c.dtype = np.dtype(np.complex128)
    return c


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
@pytest.mark.parametrize("top_lid", [True, False])
def test_stage_tendency_adjoint_dot_product_vs_complex_step(name, top_lid):
    """<J v, lam> from complex-step JVPs of the oracle's own forward stage
    pass equals <v, J^T lam> from the hand adjoint, for random directions
    in all six stage inputs at once and one input at a time."""
    case, s = _stage_state(name, seed=13)
    dx, dy = case.cfg.dx, case.cfg.dy
    rng = np.random.default_rng(17)
    lam = {n: rng.standard_normal(getattr(s, n).shape)
           for n in A.STAGE_OUT}
    g = A.stage_tendencies_adjoint(s, lam, dx, dy, top_lid)
    h = 1e-20
    scale = {"u": 1.0, "v": 1.0, "w": 0.1, "thp": 0.5, "php": 20.0,
             "mup": 20.0}
    for which in (A.STAGE_IN,) + tuple((n,) for n in A.STAGE_IN):
        c = _cplx(s)
        dirs = {}
        for n in which:
            d = rng.standard_normal(getattr(s, n).shape) * scale[n]
            dirs[n] = d
            getattr(c, n)[...] = getattr(c, n) + 1j * h * d
        t = A.stage_tendencies_cs(c, dx, dy, top_lid)
        jv = sum(float(np.sum(np.imag(t[n]) / h * lam[n]))
                 for n in A.STAGE_OUT)
        vjt = sum(float(np.sum(dirs[n] * g[n])) for n in which)
        assert abs(jv - vjt) <= 1e-10 * max(abs(jv), 1e-30), (which, jv, vjt)


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
def test_init_and_finish_adjoints_vs_complex_step(name):
    """Hand adjoints of the small-step initialisation (through the stage
    EOS) and of the small-step finish against complex-step JVPs of the
    oracle's own forwards, for all inputs at once and one at a time."""
    case, s = _stage_state(name, seed=23)
    rng = np.random.default_rng(29)
    for n in A.STAGE_IN:
        x = getattr(s, n)
        getattr(s, n + "0")[...] = x + rng.standard_normal(x.shape) * (
            0.01 * (np.abs(x).max() + 1.0))
    for n in A.FINISH_IN:
        getattr(s, n)[...] = rng.standard_normal(getattr(s, n).shape) * 0.1
    s.mu_pp[...] *= 50.0
    h = 1e-20
    scale = {"u": 1.0, "v": 1.0, "w": 0.1, "thp": 0.5, "php": 20.0,
             "mup": 20.0}
    cases = (("init", A.init_forward, A.init_adjoint, A.INIT_OUT,
              A.STAGE_IN + tuple(n + "0" for n in A.STAGE_IN)),
             ("finish", A.finish_forward, A.finish_adjoint, A.STAGE_IN,
              A.STAGE_IN + A.FINISH_IN))
    for label, fwd, adj, outs, ins in cases:
        lam = {n: rng.standard_normal(getattr(s, n).shape) for n in outs}
        g = adj(s, lam)
        for which in (ins,) + tuple((n,) for n in ins):
            c = _cplx(s)
            dirs = {}
            for n in which:
                d = rng.standard_normal(getattr(s, n).shape) * \
                    scale.get(n.rstrip("0"), 0.1)
                dirs[n] = d
                getattr(c, n)[...] = getattr(c, n) + 1j * h * d
            t = fwd(c)
            jv = sum(float(np.sum(np.imag(t[n]) / h * lam[n])) for n in outs)
            vjt = sum(float(np.sum(dirs[n] * g[n])) for n in which)
            assert abs(jv - vjt) <= 1e-10 * max(abs(jv), 1e-30), \
                (label, which, jv, vjt)


def _complex_ref_jvp(o, dtau, cfg, first, dx, h=1e-30):
    """J_ref v by a complex step in the stage reference fields: the stage
    EOS and the implicit-w coefficients are recomputed from them."""
    sc = _cplx(o)
    for n, d in dx.items():
        getattr(sc, n)[...] += 1j * h * d
    D.calc_p_alpha(sc)
    coefs = D.calc_coefs(sc, dtau, cfg.epssm, cfg.top_lid)
    D.acoustic_substep(sc, coefs, dtau, cfg, first)
    return {n: np.imag(getattr(sc, n)) / h
            for n in A.SUBSTEP_STATE + ("ww_pp",)}


@pytest.mark.parametrize("hill,damp,top_lid,first", [
    (0.0, False, True, True), (0.0, False, True, False),
    (300.0, True, True, False), (0.0, True, False, False)])
def test_acoustic_substep_reference_adjoint_vs_complex_step(hill, damp,
                                                            top_lid, first):
    """The substep's adjoint with respect to the stage reference state (u,
    v, w, thp, php, mu' of the stage, through the stage EOS and the
    implicit-w coefficients) against complex-step JVPs, all six at once
    and one at a time."""
    o, coefs, dtau, cfg, rng = acoustic_case(hill=hill, damp=damp,
                                             top_lid=top_lid)
    y = _rand_like(o, A.SUBSTEP_STATE + ("ww_pp",), rng)
    out = A.acoustic_substep_adjoint(o, coefs, dtau, cfg, first, y, ref=True)
    g = out["ref"]
    scale = {"u": 1.0, "v": 1.0, "w": 0.1, "thp": 0.5, "php": 20.0,
             "mup": 20.0}
    for which in (A.STAGE_IN,) + tuple((n,) for n in A.STAGE_IN):
        dx = {n: rng.standard_normal(getattr(o, n).shape) * scale[n]
              for n in which}
        jv = _complex_ref_jvp(o, dtau, cfg, first, dx)
        lhs = sum(float(np.vdot(jv[n], y[n])) for n in y)
        rhs = sum(float(np.vdot(dx[n], g[n])) for n in which)
        assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1e-30), \
            (which, lhs, rhs)


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
def test_whole_step_adjoint_vs_complex_step(name):
    """The hand adjoint of one whole RK3 step (tendency pass, small-step
    init, every substep with its reference-state dependence, finish, over
    the three stages in reverse, and the kinematic lower boundary) against
    complex-step JVPs of the oracle's own step."""
    from tests import pins
    case = pins.CASES[name]()
    cfg = case.cfg
    s = D.from_case(case, np.float64)
    rng = np.random.default_rng(31)
    # a disturbed state: at rest, v and w sit at rounding noise about zero
    # after a stage, where the upwind switch sign(v) of the complex-arithmetic
    # forward can differ from that of the real forward the adjoint is
    # linearised about (the step is not differentiable there)
    s.u[...] += rng.standard_normal(s.u.shape) * 0.5
    s.v[...] += rng.standard_normal(s.v.shape) * 0.5
    s.w[1:-1] += rng.standard_normal(s.w[1:-1].shape) * 0.3
    s.mup[...] += rng.standard_normal(s.mup.shape) * 20.0
    s.php[1:] += rng.standard_normal(s.php[1:].shape) * 30.0
    D.set_w_surface(s, cfg.dx, cfg.dy)
    D.calc_p_alpha(s)
    lam = {n: rng.standard_normal(getattr(s, n).shape) for n in A.STAGE_IN}
    g = A.step_adjoint(s, cfg, lam)
    h = 1e-30
    scale = {"u": 1.0, "v": 1.0, "w": 0.1, "thp": 0.5, "php": 20.0,
             "mup": 20.0}
    saved = (D.flux5, D.flux3)
    for which in (A.STAGE_IN,) + tuple((n,) for n in A.STAGE_IN):
        dx = {n: rng.standard_normal(getattr(s, n).shape) * scale[n]
              for n in which}
        c = _cplx(s)
        for n, d in dx.items():
            getattr(c, n)[...] += 1j * h * d
        D.flux5, D.flux3 = A._flux5_cs, A._flux3_cs
        try:
            D.step(c, cfg)
        finally:
            D.flux5, D.flux3 = saved
        lhs = sum(float(np.sum(np.imag(getattr(c, n)) / h * lam[n]))
                  for n in A.STAGE_IN)
        rhs = sum(float(np.sum(dx[n] * g[n])) for n in which)
        assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1e-30), \
            (which, lhs, rhs)


@pytest.mark.parametrize("name,pd", [("dry_flat", 0), ("terrain", 0),
                                     ("dry_flat", 1), ("terrain", 1)])
def test_moist_step_adjoint_vs_complex_step(name, pd):
    """The hand adjoint of one moist RK3 step (qv/qc/qr through the moist
    EOS, the buoyancy loading and the sumflux-coupled scalar update,
    without (moist_adv_opt = 0) and with (1) the final stage's PD limiter)
    against complex-step JVPs of the oracle's own step.  With the limiter,
    qv is rough (a third of the cells nearly empty) so that it engages."""
    from tests import pins
    case = pins.CASES[name]()
    cfg = case.cfg.with_(moist=True, moist_adv_opt=pd)
    s = D.from_case(case, np.float64)
    rng = np.random.default_rng(59)
    s.u[...] += rng.standard_normal(s.u.shape) * 0.5
    s.v[...] += rng.standard_normal(s.v.shape) * 0.5
    s.w[1:-1] += rng.standard_normal(s.w[1:-1].shape) * 0.3
    s.mup[...] += rng.standard_normal(s.mup.shape) * 20.0
    s.php[1:] += rng.standard_normal(s.php[1:].shape) * 30.0
    shp = s.thp.shape
    s.qv = 0.008 + 0.002 * rng.standard_normal(shp)
    s.qc = 5e-4 + 1e-4 * rng.standard_normal(shp)
    s.qr = 5e-4 + 1e-4 * rng.standard_normal(shp)
    if pd:
        s.qv = np.abs(s.qv - 0.008) * 3.0
        s.qv[rng.random(shp) < 0.3] *= 1e-3
    for q in A.SPECIES:
        setattr(s, q + "0", np.zeros(shp))
    D.set_w_surface(s, cfg.dx, cfg.dy)
    D.calc_p_alpha(s)
    names = A.STAGE_IN + A.SPECIES
    lam = {n: rng.standard_normal(getattr(s, n).shape) for n in names}
    lam.update({q: lam[q] * 1e3 for q in A.SPECIES})
    g = A.step_adjoint(s, cfg, lam)
    h = 1e-30
    scale = {"u": 1.0, "v": 1.0, "w": 0.1, "thp": 0.5, "php": 20.0,
             "mup": 20.0, "qv": 1e-3, "qc": 1e-4, "qr": 1e-4}
    saved = (D.flux5, D.flux3)
    for which in (names,) + tuple((n,) for n in names):
        dx = {n: rng.standard_normal(getattr(s, n).shape) * scale[n]
              for n in which}
        c = _cplx(s)
        for n, d in dx.items():
            getattr(c, n)[...] += 1j * h * d
        D.flux5, D.flux3 = A._flux5_cs, A._flux3_cs
        try:
            D.step(c, cfg)
        finally:
            D.flux5, D.flux3 = saved
        lhs = sum(float(np.sum(np.imag(getattr(c, n)) / h * lam[n]))
                  for n in names)
        rhs = sum(float(np.sum(dx[n] * g[n])) for n in which)
        assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1e-30), \
            (which, lhs, rhs)


@pytest.mark.parametrize("moist", [False, True])
def test_step_adjoint_with_diffusion_vs_complex_step(moist):
    """The whole-step adjoint with constant-K diffusion (km_opt = 1, K =
    75, the Straka configuration's; flat base) against complex-step JVPs
    of the oracle's own step, dry and moist."""
    from tests import pins
    case = pins.CASES["dry_flat"]()
    cfg = case.cfg.with_(km_opt=1, khdif=75.0, kvdif=75.0, moist=moist,
                         moist_adv_opt=0)
    s = D.from_case(case, np.float64)
    rng = np.random.default_rng(79)
    s.u[...] += rng.standard_normal(s.u.shape) * 0.5
    s.v[...] += rng.standard_normal(s.v.shape) * 0.5
    s.w[1:-1] += rng.standard_normal(s.w[1:-1].shape) * 0.3
    s.mup[...] += rng.standard_normal(s.mup.shape) * 20.0
    s.php[1:] += rng.standard_normal(s.php[1:].shape) * 30.0
    s.thp[...] += rng.standard_normal(s.thp.shape) * 0.5
    names = A.STAGE_IN
    if moist:
        shp = s.thp.shape
        s.qv = 0.008 + 0.002 * rng.standard_normal(shp)
        s.qc = 5e-4 + 1e-4 * rng.standard_normal(shp)
        s.qr = 5e-4 + 1e-4 * rng.standard_normal(shp)
        for q in A.SPECIES:
            setattr(s, q + "0", np.zeros(shp))
        names = names + A.SPECIES
    D.calc_p_alpha(s)
    lam = {n: rng.standard_normal(getattr(s, n).shape) for n in names}
    g = A.step_adjoint(s, cfg, lam)
    h = 1e-30
    scale = {"u": 1.0, "v": 1.0, "w": 0.1, "thp": 0.5, "php": 20.0,
             "mup": 20.0, "qv": 1e-3, "qc": 1e-4, "qr": 1e-4}
    saved = (D.flux5, D.flux3)
    for which in (names,) + tuple((n,) for n in names):
        dx = {n: rng.standard_normal(getattr(s, n).shape) * scale[n]
              for n in which}
        c = _cplx(s)
        for n, d in dx.items():
            getattr(c, n)[...] += 1j * h * d
        D.flux5, D.flux3 = A._flux5_cs, A._flux3_cs
        try:
            D.step(c, cfg)
        finally:
            D.flux5, D.flux3 = saved
        lhs = sum(float(np.sum(np.imag(getattr(c, n)) / h * lam[n]))
                  for n in names)
        rhs = sum(float(np.sum(dx[n] * g[n])) for n in which)
        assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1e-30), \
            (which, lhs, rhs)


# ---------------------------------------------------------------------------
# The final stage's positive-definite limiter
# ---------------------------------------------------------------------------

def _pd_case(seed):
    """A flat-base state after its acoustic substeps, with a rough
    non-negative scalar (some cells nearly empty beside full ones) and
    averaged fluxes strong enough that the PD limiter scales some
    corrections (0 < factor < 1), zeroes others and leaves the rest."""
    from oracle import moist_np as M
    from tests import pins
    case = pins.CASES["dry_flat"]()
    s = D.from_case(case, np.float64)
    rng = np.random.default_rng(seed)
    shp = s.thp.shape
    q0 = np.abs(rng.standard_normal(shp)) * 4e-3
    q0[rng.random(shp) < 0.3] *= 1e-3
    s.qv0 = q0
    s.qv = q0 * (1.0 + 0.2 * rng.standard_normal(shp))
    s.qc = s.qc0 = np.zeros(shp)
    s.qr = s.qr0 = np.zeros(shp)
    s.mup0 = s.mup + rng.standard_normal(s.mup.shape) * 30.0
    mu = s.mub2d + s.mup
    cfl = 0.35
    cu = (np.concatenate([mu, mu[:, :1]], axis=1)[None] * s.c1h[:, None, None]
          + s.c2h[:, None, None])
    ru = rng.standard_normal(s.u.shape) * cfl * cu * case.cfg.dx / case.cfg.dt
    cv = (np.concatenate([mu, mu[:1]], axis=0)[None] * s.c1h[:, None, None]
          + s.c2h[:, None, None])
    rv = rng.standard_normal(s.v.shape) * cfl * cv * case.cfg.dy / case.cfg.dt
    ww = rng.standard_normal(s.w.shape) * 0.05 * s.mub2d.mean()
    ww[0] = 0.0
    ww[-1] = 0.0
    return case, s, (ru, rv, ww), M


def _pd_forward(s, fl, cfg, M):
    """q_new of the final stage with the limiter (moist_np's own forward
    over the three species; only qv is nonzero)."""
    c = D.OState(**{k: (v.copy() if isinstance(v, np.ndarray) else v)
                    for k, v in vars(s).items()})
    M.advance_scalars_stage(c, fl[0], fl[1], fl[2], cfg.dt, True, cfg.dx,
                            cfg.dy, 1)
    return c.qv


@pytest.mark.parametrize("seed", [3, 17])
def test_pd_update_adjoint_vs_complex_step(seed):
    """The hand transpose of the limited final-stage update (pd_fluxes ->
    pd_scale -> pd_apply -> clamped coupled update) against complex-step
    JVPs of moist_np's own forward, in every input at once and one at a
    time, on states where the limiter is partly active."""
    case, s, fl, M = _pd_case(seed)
    cfg = case.cfg
    mu0 = s.mub2d + s.mup0
    mu = s.mub2d + s.mup
    # the limiter is exercised: some factors strictly inside (0, 1)
    fx = M.pd_fluxes(s.qv, s.qv0, fl[0], fl[1], fl[2], mu, s.c1h, s.c2h,
                     s.rdnw, s.fnm, s.fnp, cfg.dx, cfg.dy, cfg.dt)
    sc = M.pd_scale(s.qv0, mu0, s.c1h, s.c2h, s.rdnw, fx, 1.0 / cfg.dx,
                    1.0 / cfg.dy, cfg.dt)
    inside = np.count_nonzero((sc > 0.0) & (sc < 1.0))
    assert inside > 10 and np.count_nonzero(sc == 1.0) > 10, inside
    qn = _pd_forward(s, fl, cfg, M)
    assert np.count_nonzero(qn > 0.0) > qn.size // 2
    rng = np.random.default_rng(seed + 1)
    lam = rng.standard_normal(s.qv.shape)
    g = A.pd_update_adjoint(s, s.qv, s.qv0, fl, mu0, mu, cfg.dt, lam,
                            cfg.dx, cfg.dy)
    h = 1e-30
    names = ("q", "q0", "ru", "rv", "ww", "mu0", "mu")
    scale = {"q": 1e-4, "q0": 1e-4, "ru": 1e3, "rv": 1e3, "ww": 30.0,
             "mu0": 20.0, "mu": 20.0}
    shapes = {"q": s.qv.shape, "q0": s.qv.shape, "ru": fl[0].shape,
              "rv": fl[1].shape, "ww": fl[2].shape, "mu0": s.mup.shape,
              "mu": s.mup.shape}
    saved = (D.flux5, D.flux3)
    for which in (names,) + tuple((n,) for n in names):
        dx = {n: rng.standard_normal(shapes[n]) * scale[n] for n in which}
        if "ww" in dx:
            dx["ww"][0] = dx["ww"][-1] = 0.0
        c = _cplx(s)
        cfl = [a.astype(np.complex128) for a in fl]
        for n, d in dx.items():
            if n == "q":
                c.qv = c.qv + 1j * h * d
            elif n == "q0":
                c.qv0 = c.qv0 + 1j * h * d
            elif n == "mu0":
                c.mup0 = c.mup0 + 1j * h * d
            elif n == "mu":
                c.mup = c.mup + 1j * h * d
            else:
                cfl[("ru", "rv", "ww").index(n)] += 1j * h * d
        D.flux5, D.flux3 = A._flux5_cs, A._flux3_cs
        try:
            qc = _pd_forward(c, cfl, cfg, M)
        finally:
            D.flux5, D.flux3 = saved
        lhs = float(np.sum(np.imag(qc) / h * lam))
        rhs = sum(float(np.sum(dx[n] * g[n])) for n in which)
        assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1e-30), \
            (which, lhs, rhs)


def test_pd_update_adjoint_planted_mistake_fails():
    """The check above is sharp: the same transpose with the upwind flux's
    Courant-number derivative dropped (its mass and velocity share)
    fails the dot-product identity."""
    case, s, fl, M = _pd_case(3)
    cfg = case.cfg
    mu0 = s.mub2d + s.mup0
    mu = s.mub2d + s.mup
    rng = np.random.default_rng(4)
    lam = rng.standard_normal(s.qv.shape)
    real = A._upwind_adjoint

    def dropped(qm1, q0v, cr, g):
        a, b, _ = real(qm1, q0v, cr, g)
        return a, b, np.zeros_like(g)
    A._upwind_adjoint = dropped
    try:
        g = A.pd_update_adjoint(s, s.qv, s.qv0, fl, mu0, mu, cfg.dt, lam,
                                cfg.dx, cfg.dy)
    finally:
        A._upwind_adjoint = real
    h = 1e-30
    d = rng.standard_normal(fl[0].shape) * 1e3
    c = _cplx(s)
    cfl = [a.astype(np.complex128) for a in fl]
    cfl[0] += 1j * h * d
    saved = (D.flux5, D.flux3)
    D.flux5, D.flux3 = A._flux5_cs, A._flux3_cs
    try:
        qc = _pd_forward(c, cfl, cfg, M)
    finally:
        D.flux5, D.flux3 = saved
    lhs = float(np.sum(np.imag(qc) / h * lam))
    rhs = float(np.sum(d * g["ru"]))
    assert abs(lhs - rhs) > 1e-6 * abs(lhs), (lhs, rhs)
