"""The C4 adjoint configuration on the device: the 192x160x50 moist-tracer
case (dx = 3 km, dt = 18 s, qv seeded in 0-16 g/kg decaying with height,
the final stage's positive-definite limiter, moist_adv_opt = 1), FP64
adjoint with stage checkpoints only (each RK stage recomputed from its
checkpoint in the reverse sweep).

- On a full-height strip of the same case (192x16x50, the case's own
  columns, dx, dt and qv; the oracle's FP64 adjoint of one C4 step takes
  minutes at full size): the device step adjoint equals the oracle's hand
  adjoint to 1e-10 of each gradient's largest value, and the dot-product
  test <J dx, lam> = <dx, J^T lam> holds to 1e-10 with J dx a complex-step
  JVP of the oracle's own forward step (exact to rounding).  The case
  carries qc = qr = 0, where the final stage's clamp at zero and the
  limiter's budget sit exactly on their kinks (the step is not
  differentiable there), so the direction dx and the adjoint weights lam
  leave qc and qr out; their input gradients (through the buoyancy
  loading) are still compared with the oracle's.
- At full size: the stage-checkpoint reverse sweep gives the full tape's
  gradients bit for bit, and the FP32 adjoint agrees with the FP64 one to
  FP32 accuracy.

The state is disturbed first (winds, w, mu', phi'), as in the oracle's
own whole-step checks: the seeded case starts with Omega exactly zero,
where the upwind switches of the complex-arithmetic forward and of the
real forward the adjoint is taken about may disagree.
"""
import numpy as np
import pytest
import torch

from arwen_amd import cases
from oracle import adjoint_np as A
from oracle import dycore_np as O

pytestmark = pytest.mark.gpu

NAMES = A.STAGE_IN + A.SPECIES
TOL = 1e-10


def _rel(got, ref):
    got = got.detach().cpu().numpy() if torch.is_tensor(got) else got
    scale = float(np.max(np.abs(ref))) or 1.0
    return float(np.max(np.abs(got - ref))) / scale


def _disturbed(case, seed):
    """FP64 oracle state of the case with disturbed dynamics, and a device
    FP64 state holding the same fields."""
    from arwen_amd import dycore
    from arwen_amd.state import DomainState
    cfg = case.cfg
    s = O.from_case(case, np.float64)
    rng = np.random.default_rng(seed)
    s.u[...] += rng.standard_normal(s.u.shape) * 0.5
    s.v[...] += rng.standard_normal(s.v.shape) * 0.5
    s.w[1:-1] += rng.standard_normal(s.w[1:-1].shape) * 0.3
    s.mup[...] += rng.standard_normal(s.mup.shape) * 20.0
    s.php[1:] += rng.standard_normal(s.php[1:].shape) * 30.0
    s.u[..., -1] = s.u[..., 0]
    s.v[:, -1, :] = s.v[:, 0, :]
    O.set_w_surface(s, cfg.dx, cfg.dy)
    O.calc_p_alpha(s)
    g = DomainState.from_host(case, dtype=torch.float64)
    for n in NAMES:
        getattr(g, n).copy_(torch.as_tensor(getattr(s, n)))
    dycore.update_diagnostics(g)
    return s, g, rng


#: the fields C4 moves: the dynamics and qv (qc = qr = 0 stay zero)
ACTIVE = A.STAGE_IN + ("qv",)


def _lam(s, rng):
    lam = {n: rng.standard_normal(getattr(s, n).shape) for n in NAMES}
    lam["qv"] = lam["qv"] * 1e3
    for q in ("qc", "qr"):
        lam[q] = np.zeros_like(lam[q])
    return lam


def test_c4_strip_adjoint_matches_oracle_and_complex_step():
    from arwen_amd import adjoint
    case = cases.moist_tracer(ny=16)
    cfg = case.cfg
    assert cfg.moist_adv_opt == 1 and (cfg.nx, cfg.nz) == (192, 50)
    print("C4 full-height strip 192x16x50, FP64 adjoint vs oracle",
          flush=True)
    s, g, rng = _disturbed(case, 83)
    lam = _lam(s, rng)
    dev = g.u.device
    got = adjoint.step_adjoint(
        g, cfg, {n: torch.as_tensor(v, device=dev) for n, v in lam.items()},
        keep="stages")
    ref = A.step_adjoint(s, cfg, lam)
    bad = {n: _rel(got[n], ref[n]) for n in NAMES}
    print("device vs oracle adjoint", bad, flush=True)
    assert max(bad.values()) < TOL, bad
    # dot product: complex-step JVP of the oracle's forward step
    scale = {"u": 1.0, "v": 1.0, "w": 0.1, "thp": 0.5, "php": 20.0,
             "mup": 20.0, "qv": 1e-3, "qc": 1e-4, "qr": 1e-4}
    dx = {n: rng.standard_normal(getattr(s, n).shape) * scale[n]
          for n in ACTIVE}
    h = 1e-30
    c = O.OState(**{k: (v.astype(np.complex128) if isinstance(v, np.ndarray)
                        and v.dtype.kind == "f" else v)
                    for k, v in vars(s).items()})
    c.dtype = np.dtype(np.complex128)
    for n, d in dx.items():
        getattr(c, n)[...] += 1j * h * d
    saved = (O.flux5, O.flux3)
    O.flux5, O.flux3 = A._flux5_cs, A._flux3_cs
    try:
        O.step(c, cfg)
    finally:
        O.flux5, O.flux3 = saved
    lhs = sum(float(np.sum(np.imag(getattr(c, n)) / h * lam[n]))
              for n in NAMES)
    rhs = sum(float(torch.sum(torch.as_tensor(dx[n], device=dev) * got[n]))
              for n in ACTIVE)
    print(f"dot product <J dx, lam> {lhs!r} <dx, J^T lam> {rhs!r} "
          f"residual {abs(lhs - rhs) / abs(lhs):.3e}", flush=True)
    assert abs(lhs - rhs) <= TOL * abs(lhs), (lhs, rhs)


def test_c4_full_size_stage_checkpoints_and_fp32():
    from arwen_amd import adjoint
    from arwen_amd.state import DomainState
    case = cases.moist_tracer()
    cfg = case.cfg
    print("C4 192x160x50 full size: stage checkpoints vs full tape, FP32",
          flush=True)
    s, g, rng = _disturbed(case, 85)
    dev = g.u.device
    lam = {n: torch.as_tensor(v, device=dev)
           for n, v in _lam(s, rng).items()}
    full = adjoint.step_adjoint(g, cfg, lam, keep="substeps")
    stages = adjoint.step_adjoint(g, cfg, lam, keep="stages")
    for n in NAMES:
        assert torch.equal(full[n], stages[n]), n
        assert torch.isfinite(full[n]).all(), n
    g32 = DomainState.from_host(case, dtype=torch.float32)
    for n in NAMES:
        getattr(g32, n).copy_(getattr(g, n).float())
    from arwen_amd import dycore
    dycore.update_diagnostics(g32)
    got32 = adjoint.step_adjoint(
        g32, cfg, {n: t.float() for n, t in lam.items()}, keep="stages")
    bad = {n: _rel(got32[n].double(), full[n].cpu().numpy()) for n in NAMES}
    print("FP32 vs FP64 adjoint", bad, flush=True)
    assert max(bad.values()) < 1e-3, bad
