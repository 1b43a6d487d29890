"""GPU checks of the tendency-pass adjoint (arw_stage_tendencies_adjoint,
arw64_*) and its autograd wrapper ``adjoint.StageTendencies``:

- FP64: every stage-state gradient equals the oracle's hand adjoint
  (oracle/adjoint_np.stage_tendencies_adjoint, itself checked against
  complex-step JVPs of the oracle's forward stage pass), on the flat and
  terrain pins, rigid and open top;
- FP64: <J v, lam> from central differences of the device's own forward
  stage pass agrees with <v, J^T lam> (the forward is nonlinear, so to
  the differencing error);
- FP32 agrees with FP64 to 1e-4;
- autograd through StageTendencies gives the same gradient as the direct
  call, and the state is left unmodified by the adjoint;
- the small-step initialisation and finish adjoints
  (arw64_small_step_init_adjoint / _finish_adjoint) equal the oracle's
  hand adjoints (checked against complex-step JVPs) in FP64.
"""
import numpy as np
import pytest
import torch

from oracle import adjoint_np as A
from oracle import dycore_np as O

pytestmark = pytest.mark.gpu

ORACLE_TOL = 1e-10
FD_TOL = 1e-6
FP32_TOL = 1e-4


def _states(name, seed, dtype=torch.float64, top_lid=True):
    """An oracle FP64 state and a device state of dtype holding the same
    disturbed stage state."""
    from arwen_amd.state import DomainState
    from tests import pins
    case = pins.CASES[name]()
    case.cfg = case.cfg.with_(top_lid=top_lid)
    o = O.from_case(case, np.float64)
    rng = np.random.default_rng(seed)
    o.mup[...] += rng.standard_normal(o.mup.shape) * 20.0
    o.php[1:] += rng.standard_normal(o.php[1:].shape) * 30.0
    o.w[1:-1] += rng.standard_normal(o.w[1:-1].shape) * 0.3
    o.v[...] += rng.standard_normal(o.v.shape) * 0.5
    g = DomainState.from_host(case, dtype=dtype)
    for n in A.STAGE_IN:
        x = getattr(o, n)
        if dtype == torch.float32:
            x = x.astype(np.float32)
            getattr(o, n)[...] = x.astype(np.float64)
        getattr(g, n).copy_(torch.as_tensor(x))
    return case, o, g


def _lam(g, seed):
    gen = torch.Generator(device=g.u.device).manual_seed(seed)
    return {n: torch.randn(getattr(g, n).shape, dtype=g.dtype,
                           device=g.u.device, generator=gen)
            for n in A.STAGE_OUT}


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
@pytest.mark.parametrize("top_lid", [True, False])
def test_stage_adjoint_matches_oracle(name, top_lid):
    from arwen_amd import adjoint
    case, o, g = _states(name, seed=3, top_lid=top_lid)
    print(f"stage adjoint vs oracle {name} top_lid={top_lid}", flush=True)
    lam = _lam(g, 5)
    got = adjoint.stage_tendencies_adjoint(g, case.cfg, lam)
    torch.cuda.synchronize()
    ref = A.stage_tendencies_adjoint(
        o, {n: t.cpu().numpy() for n, t in lam.items()}, case.cfg.dx,
        case.cfg.dy, top_lid)
    for n in A.STAGE_IN:
        r = ref[n]
        err = np.abs(got[n].cpu().numpy() - r).max()
        assert err <= ORACLE_TOL * np.abs(r).max(), (n, err, np.abs(r).max())


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
def test_stage_adjoint_dot_product_vs_device_differences(name):
    from arwen_amd import adjoint
    case, o, g = _states(name, seed=7)
    cfg = case.cfg
    print(f"stage adjoint FD dot product {name}", flush=True)
    base = {n: getattr(g, n).clone() for n in A.STAGE_IN}
    lam = _lam(g, 9)
    gt = adjoint.stage_tendencies_adjoint(g, cfg, lam)
    gen = torch.Generator(device=g.u.device).manual_seed(11)
    scale = {"u": 1.0, "v": 1.0, "w": 0.1, "thp": 0.5, "php": 20.0,
             "mup": 20.0}
    v = {n: torch.randn(base[n].shape, dtype=g.dtype, device=g.u.device,
                        generator=gen) * scale[n] for n in A.STAGE_IN}
    h = 1e-5

    def tend(sign):
        for n in A.STAGE_IN:
            getattr(g, n).copy_(base[n] + sign * h * v[n])
        return adjoint.stage_tendencies(g, cfg)

    tp, tm = tend(1.0), tend(-1.0)
    jv = sum(float(torch.sum((tp[n] - tm[n]) / (2 * h) * lam[n]))
             for n in A.STAGE_OUT)
    vjt = sum(float(torch.sum(v[n] * gt[n])) for n in A.STAGE_IN)
    assert abs(jv - vjt) <= FD_TOL * abs(jv), (jv, vjt)


def test_stage_adjoint_fp32_close_to_fp64():
    from arwen_amd import adjoint
    case, _, g64 = _states("terrain", seed=13)
    _, _, g32 = _states("terrain", seed=13, dtype=torch.float32)
    print("stage adjoint fp32 vs fp64", flush=True)
    lam = _lam(g64, 15)
    a64 = adjoint.stage_tendencies_adjoint(g64, case.cfg, lam)
    a32 = adjoint.stage_tendencies_adjoint(
        g32, case.cfg, {n: t.float() for n, t in lam.items()})
    for n in A.STAGE_IN:
        r = a64[n].cpu().numpy()
        err = np.abs(a32[n].double().cpu().numpy() - r).max()
        assert err <= FP32_TOL * np.abs(r).max(), (n, err)


def test_stage_autograd_and_state_untouched():
    from arwen_amd import adjoint
    case, _, g = _states("dry_flat", seed=17)
    cfg = case.cfg
    print("stage autograd", flush=True)
    before = {n: getattr(g, n).clone() for n in
              A.STAGE_IN + ("p", "al", "alt")}
    lam = _lam(g, 19)
    direct = adjoint.stage_tendencies_adjoint(g, cfg, lam)
    for n, t in before.items():
        assert torch.equal(getattr(g, n), t), n
    xs = [before[n].clone().requires_grad_() for n in A.STAGE_IN]
    outs = adjoint.StageTendencies.apply(g, cfg, *xs)
    loss = sum(torch.sum(o * lam[n]) for o, n in zip(outs, A.STAGE_OUT))
    grads = torch.autograd.grad(loss, xs)
    for n, gr in zip(A.STAGE_IN, grads):
        assert torch.allclose(gr, direct[n], rtol=1e-12, atol=0.0), n


def test_stage_adjoint_rejects_unsupported():
    from arwen_amd import adjoint
    from arwen_amd._lib import ArwenError
    case, _, g = _states("terrain", seed=1)
    lam = _lam(g, 2)
    # refused before device work (the state's diffusion spacings need a
    # flat base), as the forward refuses it
    with pytest.raises((ArwenError, NotImplementedError), match="flat base"):
        adjoint.stage_tendencies_adjoint(g, case.cfg.with_(khdif=75.0,
                                                           kvdif=75.0,
                                                           km_opt=1), lam)


def test_stage_adjoint_with_diffusion_matches_oracle():
    """Constant-K diffusion (K = 75, flat base) in the tendency pass:
    the device transpose equals the oracle's (itself pinned by complex-step
    JVPs of the whole step with diffusion)."""
    from arwen_amd import adjoint
    case, o, g = _states("dry_flat", seed=29)
    cfg = case.cfg.with_(km_opt=1, khdif=75.0, kvdif=75.0)
    print("stage adjoint with diffusion vs oracle", flush=True)
    lam = _lam(g, 31)
    got = adjoint.stage_tendencies_adjoint(g, cfg, lam)
    ref = A.stage_tendencies_adjoint(
        o, {n: t.cpu().numpy() for n, t in lam.items()}, cfg.dx, cfg.dy,
        cfg.top_lid, diff=(75.0, 75.0))
    for n in A.STAGE_IN:
        r = ref[n]
        err = np.abs(got[n].cpu().numpy() - r).max()
        assert err <= ORACLE_TOL * np.abs(r).max(), (n, err, np.abs(r).max())


def _with_small_step_fields(o, g, seed):
    """Time-t fields near the stage fields and small-step fields, the same
    on the oracle state and the device state."""
    rng = np.random.default_rng(seed)
    for n in A.STAGE_IN:
        x = getattr(o, n)
        getattr(o, n + "0")[...] = x + rng.standard_normal(x.shape) * (
            0.01 * (np.abs(x).max() + 1.0))
        getattr(g, n + "0").copy_(torch.as_tensor(getattr(o, n + "0")))
    for n in A.FINISH_IN:
        getattr(o, n)[...] = rng.standard_normal(getattr(o, n).shape) * 0.1
    o.mu_pp[...] *= 50.0
    for n in A.FINISH_IN:
        getattr(g, n).copy_(torch.as_tensor(getattr(o, n)))


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
def test_init_and_finish_adjoints_match_oracle(name):
    from arwen_amd import adjoint
    case, o, g = _states(name, seed=21)
    _with_small_step_fields(o, g, 23)
    print(f"init/finish adjoints vs oracle {name}", flush=True)
    gen = torch.Generator(device=g.u.device).manual_seed(25)
    lam_i = {n: torch.randn(getattr(g, n).shape, dtype=g.dtype,
                            device=g.u.device, generator=gen)
             for n in A.INIT_OUT}
    got = adjoint.small_step_init_adjoint(g, lam_i)
    ref = A.init_adjoint(o, {n: t.cpu().numpy() for n, t in lam_i.items()})
    for n, r in ref.items():
        err = np.abs(got[n].cpu().numpy() - r).max()
        assert err <= ORACLE_TOL * np.abs(r).max(), ("init", n, err)
    lam_f = {n: torch.randn(getattr(g, n).shape, dtype=g.dtype,
                            device=g.u.device, generator=gen)
             for n in A.STAGE_IN}
    got = adjoint.small_step_finish_adjoint(g, lam_f)
    ref = A.finish_adjoint(o, {n: t.cpu().numpy() for n, t in lam_f.items()})
    for n, r in ref.items():
        err = np.abs(got[n].cpu().numpy() - r).max()
        assert err <= ORACLE_TOL * np.abs(r).max(), ("finish", n, err)
