"""GPU checks of the whole-step adjoint (adjoint.step_adjoint,
adjoint.DycoreStep) and of its last two pieces, the substep adjoint's
stage-reference share (arw_acoustic_substep_adjoint_ref with
arw_acoustic_reference_gradients) and the kinematic lower boundary's
adjoint (arw_set_w_surface_adjoint):

- FP64: each equals the oracle's hand adjoint (oracle/adjoint_np.py:
  acoustic_substep_adjoint(ref=True), set_w_surface_adjoint, step_adjoint,
  themselves checked against complex-step JVPs of the oracle's forward) to
  1e-10 of the largest gradient, flat and terrain, rigid and open top,
  with and without the upper damper;
- FP64: <J v, lam> from central differences of the device's own fused
  step (arw_step) agrees with <v, J^T lam> (the step is nonlinear, so to
  the differencing error);
- autograd through DycoreStep gives the direct call's gradients and leaves
  the inputs untouched; FP32 agrees with FP64 to FP32 accuracy.
"""
import numpy as np
import pytest
import torch

from oracle import adjoint_np as A
from oracle import dycore_np as O
from tests.test_gpu_acoustic_adjoint import NAMES_OUT, _device_pair, _rel

pytestmark = pytest.mark.gpu

ORACLE_TOL = 1e-10


@pytest.mark.parametrize("hill,damp,top_lid,first", [
    (0.0, False, True, True), (0.0, False, True, False),
    (300.0, True, True, False), (0.0, True, False, False),
    (300.0, False, False, True)])
def test_substep_reference_share_matches_oracle(hill, damp, top_lid, first):
    from arwen_amd import adjoint
    o, coefs, dtau, cfg, rng, g = _device_pair(hill=hill, damp=damp,
                                               top_lid=top_lid)
    print(f"substep reference adjoint hill={hill} damp={damp} "
          f"lid={top_lid} first={first}", flush=True)
    lam = {n: rng.standard_normal(getattr(o, n).shape) for n in NAMES_OUT}
    ref = A.acoustic_substep_adjoint(o, coefs, dtau, cfg, first, lam,
                                     ref=True)
    tl = {n: torch.as_tensor(v, device=g.u.device) for n, v in lam.items()}
    acc = adjoint.reference_accumulators(g)
    got = adjoint.acoustic_substep_adjoint(g, cfg, dtau, first, tl, acc)
    plain = adjoint.acoustic_substep_adjoint(g, cfg, dtau, first, tl)
    for n in A.SUBSTEP_STATE + A.SUBSTEP_TEND:
        assert torch.equal(got[n], plain[n]), n
    rg = adjoint.reference_gradients(g, cfg, acc)
    bad = {n: _rel(rg[n], ref["ref"][n]) for n in A.STAGE_IN}
    assert max(bad.values()) < ORACLE_TOL, bad


def test_reference_accumulators_add_over_substeps():
    from arwen_amd import adjoint
    o, coefs, dtau, cfg, rng, g = _device_pair(hill=200.0, damp=True)
    print("substep reference accumulation", flush=True)
    dev = g.u.device
    lams = [{n: torch.as_tensor(rng.standard_normal(getattr(o, n).shape),
                                device=dev) for n in NAMES_OUT}
            for _ in range(2)]
    both = adjoint.reference_accumulators(g)
    single = []
    for lam in lams:
        adjoint.acoustic_substep_adjoint(g, cfg, dtau, False, lam, both)
        one = adjoint.reference_accumulators(g)
        adjoint.acoustic_substep_adjoint(g, cfg, dtau, False, lam, one)
        single.append(one)
    for n in adjoint.ACOUSTIC_REF:
        s = single[0][n] + single[1][n]
        scale = float(s.abs().max()) or 1.0
        assert float((both[n] - s).abs().max()) <= 1e-13 * scale, n


def test_given_substep_outputs_equal_recompute():
    """The reference-share adjoint given the forward substep's outputs
    (as the whole-step driver passes its checkpoints) is bit-identical to
    the one that recomputes them, with the inputs read through views."""
    from arwen_amd import _lib, adjoint, dycore
    o, coefs, dtau, cfg, rng, g = _device_pair(hill=300.0, damp=True)
    print("substep adjoint with given outputs", flush=True)
    dev = g.u.device
    x0 = {n: getattr(g, n).clone() for n in adjoint.ACOUSTIC_STATE}
    dycore.acoustic_substep(g, cfg, dtau, False)
    fwd = {n: getattr(g, n).clone() for n in _lib.ACOUSTIC_FWD}
    lam = {n: torch.as_tensor(rng.standard_normal(getattr(o, n).shape),
                              device=dev) for n in NAMES_OUT}
    a1 = adjoint.reference_accumulators(g)
    got = adjoint.acoustic_substep_adjoint(g, cfg, dtau, False, lam, a1,
                                           fwd=fwd, views=x0)
    for n, t in x0.items():
        getattr(g, n).copy_(t)
    a2 = adjoint.reference_accumulators(g)
    ref = adjoint.acoustic_substep_adjoint(g, cfg, dtau, False, lam, a2)
    for n in adjoint.ACOUSTIC_STATE + adjoint.ACOUSTIC_TEND:
        assert torch.equal(got[n], ref[n]), n
    for n in adjoint.ACOUSTIC_REF:
        assert torch.equal(a1[n], a2[n]), n


def test_stage_factors_and_flags_equal_plain_calls():
    """Two substep adjoints with the stage's precomputed Thomas factors,
    the first writing (clear) the accumulators and the tendency adjoints
    and the second adding to them, equal bit for bit the plain calls on
    zeroed accumulators with the tendency adjoints summed afterwards."""
    from arwen_amd import adjoint
    o, coefs, dtau, cfg, rng, g = _device_pair(hill=300.0, damp=True,
                                               top_lid=False)
    print("substep adjoint flags and stage factors", flush=True)
    dev = g.u.device
    lams = [{n: torch.as_tensor(rng.standard_normal(getattr(o, n).shape),
                                device=dev) for n in NAMES_OUT}
            for _ in range(2)]
    a_ref = adjoint.reference_accumulators(g)
    t_ref = {n: torch.zeros_like(getattr(g, n)) for n in adjoint.ACOUSTIC_TEND}
    outs_ref = []
    for lam in lams:
        out = adjoint.acoustic_substep_adjoint(g, cfg, dtau, False, lam, a_ref)
        for n in adjoint.ACOUSTIC_TEND:
            t_ref[n] += out[n]
        outs_ref.append(out)
    lu = adjoint.adjoint_factors(g, cfg, dtau)
    a_new = {n: torch.full_like(t, float("nan")) for n, t in a_ref.items()}
    t_new = {n: torch.full_like(t, float("nan")) for n, t in t_ref.items()}
    for q, lam in enumerate(lams):
        out = adjoint.acoustic_substep_adjoint(
            g, cfg, dtau, False, lam, a_new, lu=lu, clear_ref=(q == 0),
            tend_out=t_new, accumulate=(q > 0))
        for n in adjoint.ACOUSTIC_STATE:
            assert torch.equal(out[n], outs_ref[q][n]), (q, n)
    for n in adjoint.ACOUSTIC_TEND:
        assert torch.equal(t_new[n], t_ref[n]), n
    for n in adjoint.ACOUSTIC_REF:
        if n in ("u", "v"):          # written by reference_gradients only
            continue
        assert torch.equal(a_new[n], a_ref[n]), n


def test_set_w_surface_adjoint_matches_oracle():
    from arwen_amd import adjoint
    o, coefs, dtau, cfg, rng, g = _device_pair(hill=300.0)
    print("set_w_surface adjoint", flush=True)
    l0 = rng.standard_normal(o.mup.shape)
    gu0 = rng.standard_normal(o.u.shape)
    gv0 = rng.standard_normal(o.v.shape)
    eu, ev = A.set_w_surface_adjoint(o, l0, cfg.dx, cfg.dy)
    dev = g.u.device
    gu = torch.as_tensor(gu0.copy(), device=dev)
    gv = torch.as_tensor(gv0.copy(), device=dev)
    adjoint.set_w_surface_adjoint(g, cfg, torch.as_tensor(l0, device=dev),
                                  gu, gv)
    assert _rel(gu, gu0 + eu) < 1e-13
    assert _rel(gv, gv0 + ev) < 1e-13
    assert np.abs(eu).max() > 0.0


def _step_states(name, dtype=torch.float64, seed=31):
    """A disturbed oracle FP64 state (as the oracle's whole-step test) and a
    device state of dtype holding the same fields."""
    from arwen_amd import dycore
    from arwen_amd.state import DomainState
    from tests import pins
    case = pins.CASES[name]()
    cfg = case.cfg
    s = O.from_case(case, np.float64)
    rng = np.random.default_rng(seed)
    s.u[...] += rng.standard_normal(s.u.shape) * 0.5
    s.v[...] += rng.standard_normal(s.v.shape) * 0.5
    s.w[1:-1] += rng.standard_normal(s.w[1:-1].shape) * 0.3
    s.mup[...] += rng.standard_normal(s.mup.shape) * 20.0
    s.php[1:] += rng.standard_normal(s.php[1:].shape) * 30.0
    O.set_w_surface(s, cfg.dx, cfg.dy)
    if dtype == torch.float32:
        for n in A.STAGE_IN:
            x = getattr(s, n)
            x[...] = x.astype(np.float32).astype(np.float64)
    O.calc_p_alpha(s)
    g = DomainState.from_host(case, dtype=dtype)
    for n in A.STAGE_IN:
        getattr(g, n).copy_(torch.as_tensor(getattr(s, n)))
    dycore.update_diagnostics(g)
    return case, s, g, rng


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
def test_step_adjoint_matches_oracle(name):
    from arwen_amd import adjoint
    case, s, g, rng = _step_states(name)
    print(f"step adjoint vs oracle {name}", flush=True)
    lam = {n: rng.standard_normal(getattr(s, n).shape) for n in A.STAGE_IN}
    ref = A.step_adjoint(s, case.cfg, lam)
    before = {n: getattr(g, n).clone() for n in A.STAGE_IN}
    got = adjoint.step_adjoint(
        g, case.cfg, {n: torch.as_tensor(v, device=g.u.device)
                      for n, v in lam.items()})
    for n in A.STAGE_IN:
        assert torch.equal(getattr(g, n), before[n]), n
    bad = {n: _rel(got[n], ref[n]) for n in A.STAGE_IN}
    assert max(bad.values()) < ORACLE_TOL, bad


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
def test_step_adjoint_dot_product_vs_device_step(name):
    from arwen_amd import adjoint, dycore
    case, s, g, rng = _step_states(name, seed=37)
    cfg = case.cfg
    print(f"step adjoint FD dot product {name}", flush=True)
    dev = g.u.device
    base = {n: getattr(g, n).clone() for n in A.STAGE_IN}
    lam = {n: torch.as_tensor(rng.standard_normal(getattr(s, n).shape),
                              device=dev) for n in A.STAGE_IN}
    gt = adjoint.step_adjoint(g, cfg, lam)
    scale = {"u": 1.0, "v": 1.0, "w": 0.1, "thp": 0.5, "php": 20.0,
             "mup": 20.0}
    v = {n: torch.as_tensor(rng.standard_normal(getattr(s, n).shape)
                            * scale[n], device=dev) for n in A.STAGE_IN}
    h = 1e-6

    def run(sign):
        for n in A.STAGE_IN:
            getattr(g, n).copy_(base[n] + sign * h * v[n])
        dycore.update_diagnostics(g)
        dycore.step(g, cfg)
        return {n: getattr(g, n).clone() for n in A.STAGE_IN}

    fp, fm = run(1.0), run(-1.0)
    jv = sum(float(torch.sum((fp[n] - fm[n]) / (2 * h) * lam[n]))
             for n in A.STAGE_IN)
    vjt = sum(float(torch.sum(v[n] * gt[n])) for n in A.STAGE_IN)
    assert abs(jv - vjt) <= 1e-6 * abs(jv), (jv, vjt)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12),
                                       (torch.float32, 1e-5)])
def test_taped_step_equals_fused_step(dtype, tol):
    """The piecewise taped forward advances the state like the fused
    device step (to rounding: the stage EOS of stages 2 and 3 is formed by
    its own pass instead of inside the finish)."""
    from arwen_amd import adjoint, dycore
    case, s, g, rng = _step_states("terrain", dtype=dtype, seed=47)
    print(f"taped step vs fused step {dtype}", flush=True)
    x0 = {n: getattr(g, n).clone() for n in A.STAGE_IN}
    dycore.step(g, case.cfg)
    fused = {n: getattr(g, n).clone() for n in A.STAGE_IN}
    for n, t in x0.items():
        getattr(g, n).copy_(t)
    dycore.update_diagnostics(g)
    tape = adjoint.step_taped(g, case.cfg)
    for n in A.STAGE_IN:
        a = getattr(g, n).double()
        b = fused[n].double()
        assert torch.equal(tape.out[n], getattr(g, n)), n
        scale = float(b.abs().max()) or 1.0
        assert float((a - b).abs().max()) <= tol * scale, n


def test_dycore_step_autograd():
    from arwen_amd import adjoint
    case, s, g, rng = _step_states("terrain", seed=41)
    cfg = case.cfg
    print("dycore step autograd", flush=True)
    dev = g.u.device
    x0 = [getattr(g, n).clone() for n in A.STAGE_IN]
    lam = {n: torch.as_tensor(rng.standard_normal(getattr(s, n).shape),
                              device=dev) for n in A.STAGE_IN}
    direct = adjoint.step_adjoint(g, cfg, lam)
    xs = [x.clone().requires_grad_() for x in x0]
    outs = adjoint.DycoreStep.apply(g, cfg, *xs)
    loss = sum(torch.sum(o * lam[n]) for o, n in zip(outs, A.STAGE_IN))
    grads = torch.autograd.grad(loss, xs)
    for x, x_ in zip(xs, x0):
        assert torch.equal(x.detach(), x_)
    for n, gr in zip(A.STAGE_IN, grads):
        assert torch.allclose(gr, direct[n], rtol=1e-12, atol=0.0), n


def test_stage_checkpoint_tape_equals_substep_tape():
    """Keeping only the stage states and recomputing each stage in the
    reverse sweep gives the same gradients, bit for bit, as the full tape;
    and through DycoreStep with keep="stages"."""
    from arwen_amd import adjoint
    case, s, g, rng = _step_states("terrain", seed=53)
    cfg = case.cfg
    print("stage-checkpoint tape", flush=True)
    dev = g.u.device
    lam = {n: torch.as_tensor(rng.standard_normal(getattr(s, n).shape),
                              device=dev) for n in A.STAGE_IN}
    full = adjoint.step_adjoint(g, cfg, lam, keep="substeps")
    lean = adjoint.step_adjoint(g, cfg, lam, keep="stages")
    for n in A.STAGE_IN:
        assert torch.equal(full[n], lean[n]), n
    xs = [getattr(g, n).clone().requires_grad_() for n in A.STAGE_IN]
    outs = adjoint.DycoreStep.apply(g, cfg, "stages", *xs)
    loss = sum(torch.sum(o * lam[n]) for o, n in zip(outs, A.STAGE_IN))
    grads = torch.autograd.grad(loss, xs)
    for n, gr in zip(A.STAGE_IN, grads):
        assert torch.equal(gr, full[n]), n


def test_step_adjoint_fp32_close_to_fp64():
    from arwen_amd import adjoint
    case, s, g64, rng = _step_states("terrain", seed=43,
                                     dtype=torch.float32)
    _, _, g32, _ = _step_states("terrain", seed=43, dtype=torch.float32)
    from arwen_amd.state import DomainState
    g64 = DomainState.from_host(case, dtype=torch.float64)
    for n in A.STAGE_IN:
        getattr(g64, n).copy_(getattr(g32, n).double())
    from arwen_amd import dycore
    dycore.update_diagnostics(g64)
    print("step adjoint fp32 vs fp64", flush=True)
    lam = {n: torch.as_tensor(rng.standard_normal(getattr(s, n).shape),
                              device=g64.u.device) for n in A.STAGE_IN}
    a64 = adjoint.step_adjoint(g64, case.cfg, lam)
    a32 = adjoint.step_adjoint(g32, case.cfg,
                               {n: t.float() for n, t in lam.items()})
    bad = {n: _rel(a32[n].double(), a64[n].cpu().numpy())
           for n in A.STAGE_IN}
    print("fp32 vs fp64 relative errors", bad, flush=True)
    assert max(bad.values()) < 1e-3, bad


def test_step_adjoint_rejects_unsupported():
    from arwen_amd import adjoint
    from arwen_amd._lib import ArwenError
    case, s, g, rng = _step_states("dry_flat")
    lam = {"u": torch.ones_like(g.u)}
    with pytest.raises(ArwenError, match="plain form"):
        adjoint.step_adjoint(g, case.cfg.with_(precision_form=True), lam)


# ---------------------------------------------------------------------------
# Moist states: qv, qc, qr through the moist EOS, the buoyancy loading and
# the sumflux-coupled scalar update, without (moist_adv_opt = 0) and with
# (1) the final stage's positive-definite limiter
# ---------------------------------------------------------------------------

MOIST_NAMES = A.STAGE_IN + A.SPECIES


def _moist_states(name, dtype=torch.float64, seed=59, pd=0):
    """The oracle's moist whole-step test state and a device state of dtype
    holding the same fields (scalars kept positive).  pd = 1: the final
    stage's PD limiter on, with a rough qv (a third of the cells nearly
    empty) so that it engages."""
    from arwen_amd import dycore
    from arwen_amd.state import DomainState
    from tests import pins
    case = pins.CASES[name]()
    case.cfg = case.cfg.with_(moist=True, moist_adv_opt=pd)
    cfg = case.cfg
    s = O.from_case(case, np.float64)
    rng = np.random.default_rng(seed)
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
    O.set_w_surface(s, cfg.dx, cfg.dy)
    if dtype == torch.float32:
        for n in MOIST_NAMES:
            x = getattr(s, n)
            x[...] = x.astype(np.float32).astype(np.float64)
    O.calc_p_alpha(s)
    g = DomainState.from_host(case, dtype=dtype)
    assert g.qv is not None
    for n in MOIST_NAMES:
        getattr(g, n).copy_(torch.as_tensor(getattr(s, n)))
    dycore.update_diagnostics(g)
    return case, s, g, rng


def test_scalar_stage_and_adjoint_match_oracle():
    from arwen_amd import adjoint
    case, s, g, rng = _moist_states("terrain", seed=61)
    cfg = case.cfg
    print("scalar stage and its adjoint vs oracle", flush=True)
    dev = g.u.device
    nz, ny, nx = s.thp.shape
    q, q0 = s.qv, s.qv * (1.0 + 0.01 * rng.standard_normal(s.qv.shape))
    fl = (rng.standard_normal(s.u.shape) * 3e3,
          rng.standard_normal(s.v.shape) * 3e3,
          rng.standard_normal(s.w.shape) * 1e1)
    mup_t = s.mup * 0.9
    mup_new = s.mup * 1.1
    mu0, mu = s.mub2d + mup_t, s.mub2d + mup_new
    lam = rng.standard_normal(q.shape)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    for clamp in (False, True):
        r = A.scalar_update_adjoint(s, q, q0, fl, mu0, mu, 9.0, clamp, lam,
                                    cfg.dx, cfg.dy)
        g_q0 = torch.zeros_like(g.qv)
        g_fl = [torch.zeros_like(g.u), torch.zeros_like(g.v),
                torch.zeros_like(g.w)]
        g_mt = torch.zeros_like(g.mup)
        g_mn = torch.zeros_like(g.mup)
        g_q = adjoint.scalar_stage_adjoint(
            g, cfg, 9.0, clamp, t(q), t(q0), [t(x) for x in fl], t(mup_t),
            t(mup_new), t(lam), g_q0, g_fl, g_mt, g_mn)
        got = {"q": g_q, "q0": g_q0, "ru": g_fl[0], "rv": g_fl[1],
               "ww": g_fl[2], "mu0": g_mt, "mu": g_mn}
        bad = {k: _rel(v, r[k]) for k, v in got.items()}
        assert max(bad.values()) < 1e-10, (clamp, bad)


def _pd_inputs(s, cfg, rng):
    """A rough non-negative scalar pair and averaged fluxes strong enough
    (Courant numbers ~0.35) that the PD limiter scales some corrections."""
    shp = s.thp.shape
    q0 = np.abs(rng.standard_normal(shp)) * 4e-3
    q0[rng.random(shp) < 0.3] *= 1e-3
    q = q0 * (1.0 + 0.2 * rng.standard_normal(shp))
    mup_t = s.mup + rng.standard_normal(s.mup.shape) * 30.0
    mup_new = s.mup.copy()
    mu = s.mub2d + mup_new
    c1, c2 = s.c1h[:, None, None], s.c2h[:, None, None]
    cu = np.concatenate([mu, mu[:, :1]], axis=1)[None] * c1 + c2
    cv = np.concatenate([mu, mu[:1]], axis=0)[None] * c1 + c2
    fl = (rng.standard_normal(s.u.shape) * 0.35 * cu * cfg.dx / cfg.dt,
          rng.standard_normal(s.v.shape) * 0.35 * cv * cfg.dy / cfg.dt,
          rng.standard_normal(s.w.shape) * 0.05 * s.mub2d.mean())
    fl[2][0] = 0.0
    fl[2][-1] = 0.0
    return q, q0, fl, mup_t, mup_new


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
def test_pd_stage_and_adjoint_match_oracle(name):
    """arw_pd_stage equals the oracle's limited final-stage update (FP64 to
    round-off, FP32 bit for bit) and arw_pd_stage_adjoint its transpose
    (pinned by complex-step JVPs), on inputs where the limiter engages."""
    from arwen_amd import adjoint
    from arwen_amd.state import DomainState
    from oracle import moist_np as M
    case, s, g, rng = _moist_states(name, seed=89, pd=1)
    cfg = case.cfg
    print(f"pd stage and its adjoint vs oracle {name}", flush=True)
    dev = g.u.device
    q, q0, fl, mup_t, mup_new = _pd_inputs(s, cfg, rng)
    mu0, mu = s.mub2d + mup_t, s.mub2d + mup_new
    dt = cfg.dt
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)

    def oracle_forward(st, dtype):
        c = lambda a: np.asarray(a).astype(dtype)
        f = [c(x) for x in fl]
        m0, mn = c(st.mub2d) + c(mup_t), c(st.mub2d) + c(mup_new)
        fx = M.pd_fluxes(c(q), c(q0), f[0], f[1], f[2], mn, c(st.c1h),
                         c(st.c2h), c(st.rdnw), c(st.fnm), c(st.fnp),
                         cfg.dx, cfg.dy, dt)
        dxi = float(np.real(dtype(1.0 / cfg.dx)))
        dyi = float(np.real(dtype(1.0 / cfg.dy)))
        sc = M.pd_scale(c(q0), m0, c(st.c1h), c(st.c2h), c(st.rdnw), fx,
                        dxi, dyi, dt)
        assert np.count_nonzero((sc > 0) & (sc < 1)) > 10
        tend = dtype(0.0) + M.pd_apply(sc, fx, c(st.rdnw), dxi, dyi)
        return M.coupled_update(c(q0), tend, c(st.c1h), c(st.c2h), m0, mn,
                                dt, True)

    ref_q = oracle_forward(s, np.float64)
    got_q = adjoint.pd_stage(g, cfg, dt, t(q), t(q0), [t(x) for x in fl],
                             t(mup_t), t(mup_new))
    assert _rel(got_q, ref_q) < 1e-13
    # FP32: bit-identical to the oracle's float32 arithmetic
    g32 = DomainState.from_host(case, dtype=torch.float32)
    s32 = O.from_case(case, np.float32)
    f32 = lambda a: torch.as_tensor(np.ascontiguousarray(
        np.asarray(a).astype(np.float32)), device=dev)
    got32 = adjoint.pd_stage(g32, cfg, dt, f32(q), f32(q0),
                             [f32(x) for x in fl], f32(mup_t), f32(mup_new))
    ref32 = oracle_forward(s32, np.float32)
    assert np.array_equal(got32.cpu().numpy(), ref32)
    # the transpose
    lam = rng.standard_normal(q.shape)
    r = A.pd_update_adjoint(s, q, q0, fl, mu0, mu, dt, lam, cfg.dx, cfg.dy)
    g_q0 = torch.zeros_like(g.qv)
    g_fl = [torch.zeros_like(g.u), torch.zeros_like(g.v),
            torch.zeros_like(g.w)]
    g_mt = torch.zeros_like(g.mup)
    g_mn = torch.zeros_like(g.mup)
    g_q = adjoint.pd_stage_adjoint(
        g, cfg, dt, t(q), t(q0), [t(x) for x in fl], t(mup_t), t(mup_new),
        t(lam), g_q0, g_fl, g_mt, g_mn)
    got = {"q": g_q, "q0": g_q0, "ru": g_fl[0], "rv": g_fl[1],
           "ww": g_fl[2], "mu0": g_mt, "mu": g_mn}
    bad = {k: _rel(v, r[k]) for k, v in got.items()}
    print("pd adjoint relative errors", bad, flush=True)
    assert max(bad.values()) < 1e-10, bad


@pytest.mark.parametrize("name,pd", [("dry_flat", 0), ("terrain", 0),
                                     ("dry_flat", 1), ("terrain", 1)])
def test_moist_step_adjoint_matches_oracle(name, pd):
    from arwen_amd import adjoint
    case, s, g, rng = _moist_states(name, pd=pd)
    print(f"moist step adjoint vs oracle {name} pd={pd}", flush=True)
    lam = {n: rng.standard_normal(getattr(s, n).shape) for n in MOIST_NAMES}
    ref = A.step_adjoint(s, case.cfg, lam)
    before = {n: getattr(g, n).clone() for n in MOIST_NAMES}
    got = adjoint.step_adjoint(
        g, case.cfg, {n: torch.as_tensor(v, device=g.u.device)
                      for n, v in lam.items()})
    for n in MOIST_NAMES:
        assert torch.equal(getattr(g, n), before[n]), n
    bad = {n: _rel(got[n], ref[n]) for n in MOIST_NAMES}
    assert max(bad.values()) < ORACLE_TOL, bad


@pytest.mark.parametrize("pd", [0, 1])
def test_moist_step_adjoint_dot_product_vs_device_step(pd):
    from arwen_amd import adjoint, dycore
    case, s, g, rng = _moist_states("terrain", seed=67, pd=pd)
    cfg = case.cfg
    print(f"moist step adjoint FD dot product pd={pd}", flush=True)
    dev = g.u.device
    base = {n: getattr(g, n).clone() for n in MOIST_NAMES}
    lam = {n: torch.as_tensor(rng.standard_normal(getattr(s, n).shape),
                              device=dev) for n in MOIST_NAMES}
    lam.update({q: lam[q] * 1e3 for q in A.SPECIES})
    gt = adjoint.step_adjoint(g, cfg, lam)
    scale = {"u": 1.0, "v": 1.0, "w": 0.1, "thp": 0.5, "php": 20.0,
             "mup": 20.0, "qv": 1e-3, "qc": 1e-4, "qr": 1e-4}
    v = {n: torch.as_tensor(rng.standard_normal(getattr(s, n).shape)
                            * scale[n], device=dev) for n in MOIST_NAMES}
    h = 1e-6

    def run(sign):
        for n in MOIST_NAMES:
            getattr(g, n).copy_(base[n] + sign * h * v[n])
        dycore.update_diagnostics(g)
        dycore.step(g, cfg)
        return {n: getattr(g, n).clone() for n in MOIST_NAMES}

    fp, fm = run(1.0), run(-1.0)
    jv = sum(float(torch.sum((fp[n] - fm[n]) / (2 * h) * lam[n]))
             for n in MOIST_NAMES)
    vjt = sum(float(torch.sum(v[n] * gt[n])) for n in MOIST_NAMES)
    assert abs(jv - vjt) <= 1e-6 * abs(jv), (jv, vjt)


@pytest.mark.parametrize("pd", [0, 1])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12),
                                       (torch.float32, 1e-5)])
def test_taped_moist_step_equals_fused_step(dtype, tol, pd):
    from arwen_amd import adjoint, dycore
    case, s, g, rng = _moist_states("terrain", dtype=dtype, seed=71, pd=pd)
    print(f"taped moist step vs fused step {dtype} pd={pd}", flush=True)
    x0 = {n: getattr(g, n).clone() for n in MOIST_NAMES}
    dycore.step(g, case.cfg)
    fused = {n: getattr(g, n).clone() for n in MOIST_NAMES}
    for n, t in x0.items():
        getattr(g, n).copy_(t)
    dycore.update_diagnostics(g)
    tape = adjoint.step_taped(g, case.cfg)
    for n in MOIST_NAMES:
        a = getattr(g, n).double()
        b = fused[n].double()
        scale = float(b.abs().max()) or 1.0
        assert float((a - b).abs().max()) <= tol * scale, n


@pytest.mark.parametrize("pd", [0, 1])
def test_moist_dycore_step_autograd_and_stage_checkpoints(pd):
    from arwen_amd import adjoint
    case, s, g, rng = _moist_states("dry_flat", seed=73, pd=pd)
    cfg = case.cfg
    print(f"moist dycore step autograd pd={pd}", flush=True)
    dev = g.u.device
    lam = {n: torch.as_tensor(rng.standard_normal(getattr(s, n).shape),
                              device=dev) for n in MOIST_NAMES}
    full = adjoint.step_adjoint(g, cfg, lam)
    lean = adjoint.step_adjoint(g, cfg, lam, keep="stages")
    for n in MOIST_NAMES:
        assert torch.equal(full[n], lean[n]), n
    xs = [getattr(g, n).clone().requires_grad_() for n in MOIST_NAMES]
    outs = adjoint.DycoreStep.apply(g, cfg, *xs)
    loss = sum(torch.sum(o * lam[n]) for o, n in zip(outs, MOIST_NAMES))
    grads = torch.autograd.grad(loss, xs)
    for n, gr in zip(MOIST_NAMES, grads):
        assert torch.equal(gr, full[n]), n


@pytest.mark.parametrize("moist", [False, True])
def test_step_adjoint_with_diffusion_matches_oracle(moist):
    """Constant-K diffusion (K = 75, km_opt = 1, the Straka
    configuration's) in the whole-step adjoint, dry and moist, against the
    oracle's (pinned by complex-step JVPs); and central differences of the
    fused device step with diffusion."""
    from arwen_amd import adjoint, dycore
    if moist:
        case, s, g, rng = _moist_states("dry_flat", seed=83)
    else:
        case, s, g, rng = _step_states("dry_flat", seed=83)
    case.cfg = case.cfg.with_(km_opt=1, khdif=75.0, kvdif=75.0)
    cfg = case.cfg
    names = MOIST_NAMES if moist else A.STAGE_IN
    print(f"step adjoint with diffusion moist={moist}", flush=True)
    lam = {n: rng.standard_normal(getattr(s, n).shape) for n in names}
    ref = A.step_adjoint(s, cfg, lam)
    dev = g.u.device
    tl = {n: torch.as_tensor(v, device=dev) for n, v in lam.items()}
    got = adjoint.step_adjoint(g, cfg, tl)
    bad = {n: _rel(got[n], ref[n]) for n in names}
    assert max(bad.values()) < ORACLE_TOL, bad
    base = {n: getattr(g, n).clone() for n in names}
    scale = {"u": 1.0, "v": 1.0, "w": 0.1, "thp": 0.5, "php": 20.0,
             "mup": 20.0, "qv": 1e-3, "qc": 1e-4, "qr": 1e-4}
    v = {n: torch.as_tensor(rng.standard_normal(getattr(s, n).shape)
                            * scale[n], device=dev) for n in names}
    h = 1e-6

    def run(sign):
        for n in names:
            getattr(g, n).copy_(base[n] + sign * h * v[n])
        dycore.update_diagnostics(g)
        dycore.step(g, cfg)
        return {n: getattr(g, n).clone() for n in names}

    fp, fm = run(1.0), run(-1.0)
    jv = sum(float(torch.sum((fp[n] - fm[n]) / (2 * h) * tl[n]))
             for n in names)
    vjt = sum(float(torch.sum(v[n] * got[n])) for n in names)
    assert abs(jv - vjt) <= 1e-6 * abs(jv), (jv, vjt)
