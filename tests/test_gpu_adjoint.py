"""GPU checks of the scalar-transport adjoint (arw_step_ckpt,
arw_tracer_step, arw_tracer_adjoint and the autograd wrapper):

- recording checkpoints leaves the step bit-identical to arw_step;
- the checkpointed transport reproduces the step's own qv/qc/qr update
  bit for bit (FP32, unlimited moist_adv_opt = 0);
- FP64: forward and adjoint agree with the oracle's float64 restatement
  (oracle/adjoint_np.py, itself pinned by its own dot-product tests);
- FP64 dot-product identity <J x, y> = <x, J^T y> to 1e-10 and a central
  finite-difference directional derivative of a scalar functional against
  the autograd gradient.
"""
import numpy as np
import pytest
import torch

from oracle import adjoint_np as A
from tests.helpers import PROG
from tests.test_gpu_moist import _case

pytestmark = pytest.mark.gpu

#: FP64 relative bars (the reference states 1e-10 for the dot product)
DOT_TOL = 1e-10
FD_TOL = 1e-8
ORACLE_TOL = 1e-12


def _state(case, dtype):
    from arwen_amd.state import DomainState
    return DomainState.from_host(case, dtype=dtype)


def _ckpt_np(ck, g):
    """Checkpoint and coefficients as float64 numpy for the oracle."""
    n = lambda t: t.detach().cpu().numpy().astype(np.float64)
    stages = [(n(ck.ru_m[s]), n(ck.rv_m[s]), n(ck.ww_m[s]), n(ck.mup_new[s]),
               ck.dt[s]) for s in range(3)]
    co = (n(g.c1h), n(g.c2h), n(g.mub2d), n(g.rdnw), n(g.fnm), n(g.fnp))
    return {"mup_t": n(ck.mup_t), "stages": stages}, co


def test_checkpointed_step_is_the_step():
    from arwen_amd import adjoint, dycore
    case = _case(moist_adv_opt=1)
    g1, g2 = _state(case, torch.float32), _state(case, torch.float32)
    print("ckpt step == step (moist)", flush=True)
    for _ in range(2):
        dycore.step(g1, case.cfg)
        adjoint.step_with_checkpoints(g2, case.cfg)
    for name in PROG + ("qv", "qc", "qr", "p", "alt"):
        assert torch.equal(getattr(g1, name), getattr(g2, name)), name


def test_checkpointed_dry_step_is_the_step():
    from arwen_amd import adjoint, dycore
    from tests import pins
    case = pins.CASES["dry_flat"]()
    g1, g2 = _state(case, torch.float32), _state(case, torch.float32)
    print("ckpt step == step (dry)", flush=True)
    dycore.step(g1, case.cfg)
    ck = adjoint.step_with_checkpoints(g2, case.cfg)
    for name in PROG + ("p", "alt"):
        assert torch.equal(getattr(g1, name), getattr(g2, name)), name
    assert ck.dt == pytest.approx((case.cfg.dt / 3.0,
                                   case.cfg.dt / 2.0, case.cfg.dt))


def test_tracer_step_reproduces_moist_update_bitwise():
    from arwen_amd import adjoint
    case = _case(moist_adv_opt=0)
    g = _state(case, torch.float32)
    print("tracer step == moist update", flush=True)
    q_t = {n: getattr(g, n).clone() for n in ("qv", "qc", "qr")}
    ck = adjoint.step_with_checkpoints(g, case.cfg)
    for n, q in q_t.items():
        out = adjoint.tracer_step(q, g, case.cfg, ck, clamp=True)
        assert torch.equal(out, getattr(g, n)), n


def test_fp64_forward_and_adjoint_match_oracle():
    from arwen_amd import adjoint
    case = _case(moist_adv_opt=0)
    g = _state(case, torch.float64)
    print("fp64 tracer forward/adjoint vs oracle", flush=True)
    ck = adjoint.step_with_checkpoints(g, case.cfg)
    ckn, co = _ckpt_np(ck, g)
    dxi, dyi = 1.0 / case.cfg.dx, 1.0 / case.cfg.dy
    rng = np.random.default_rng(11)
    x = rng.standard_normal(g.shape)
    y = rng.standard_normal(g.shape)
    fx = adjoint.tracer_step(torch.as_tensor(x, device=g.device), g,
                             case.cfg, ck).cpu().numpy()
    fty = adjoint.tracer_adjoint(torch.as_tensor(y, device=g.device), g,
                                 case.cfg, ck).cpu().numpy()
    ref_fx = A.tracer_step(x, ckn, *co, dxi, dyi)
    ref_fty = A.tracer_step_adjoint(y, ckn, *co, dxi, dyi)
    assert np.max(np.abs(fx - ref_fx)) <= ORACLE_TOL * np.max(np.abs(ref_fx))
    assert np.max(np.abs(fty - ref_fty)) <= \
        ORACLE_TOL * np.max(np.abs(ref_fty))


@pytest.mark.parametrize("moist", [True, False])
def test_fp64_dot_product_identity(moist):
    from arwen_amd import adjoint
    from tests import pins
    case = _case(moist_adv_opt=0) if moist else pins.CASES["terrain"]()
    g = _state(case, torch.float64)
    print(f"fp64 dot product moist={moist}", flush=True)
    ck = adjoint.step_with_checkpoints(g, case.cfg)
    gen = torch.Generator(device=g.device).manual_seed(3)
    x = torch.randn(g.shape, dtype=torch.float64, device=g.device,
                    generator=gen)
    y = torch.randn(g.shape, dtype=torch.float64, device=g.device,
                    generator=gen)
    lhs = float(torch.sum(adjoint.tracer_step(x, g, case.cfg, ck) * y))
    rhs = float(torch.sum(x * adjoint.tracer_adjoint(y, g, case.cfg, ck)))
    assert abs(lhs - rhs) <= DOT_TOL * abs(lhs), (lhs, rhs)


def test_fp64_autograd_gradient_matches_finite_differences():
    """Two checkpointed steps; L(q) = <w, J2 J1 q>^2 / 2 differentiated by
    autograd (hand-written adjoint) and by central differences along a
    random direction."""
    from arwen_amd import adjoint
    case = _case(moist_adv_opt=0)
    g = _state(case, torch.float64)
    print("fp64 autograd vs FD", flush=True)
    cks = [adjoint.step_with_checkpoints(g, case.cfg) for _ in range(2)]
    gen = torch.Generator(device=g.device).manual_seed(5)
    q = torch.rand(g.shape, dtype=torch.float64, device=g.device,
                   generator=gen) * 0.01
    w = torch.randn(g.shape, dtype=torch.float64, device=g.device,
                    generator=gen)
    dq = torch.randn(g.shape, dtype=torch.float64, device=g.device,
                     generator=gen)

    def loss(qq):
        return 0.5 * torch.sum(w * adjoint.transport(qq, g, case.cfg, cks)) ** 2

    qv = q.clone().requires_grad_(True)
    (grad,) = torch.autograd.grad(loss(qv), qv)
    eps = 1e-4
    with torch.no_grad():
        fd = (float(loss(q + eps * dq)) - float(loss(q - eps * dq))) / (2 * eps)
    ad = float(torch.sum(grad * dq))
    assert abs(fd - ad) <= FD_TOL * abs(ad), (fd, ad)


def test_tracer_entry_points_reject_bad_arguments():
    from arwen_amd import _lib, adjoint
    case = _case(moist_adv_opt=0)
    g = _state(case, torch.float32)
    ck = adjoint.TracerCheckpoints(g)
    with pytest.raises(ValueError):
        adjoint.tracer_step(torch.zeros(3, device=g.device), g, case.cfg, ck)
    ck.c.mup_t = None
    with pytest.raises(_lib.ArwenError):
        adjoint.tracer_step(g.qv, g, case.cfg, ck)
