"""GPU checks of the acoustic-substep adjoint (arw_acoustic_substep_adjoint
and the AcousticSubstep autograd wrapper):

- FP64: J^T lam on the device equals the oracle's hand-written adjoint
  (oracle/adjoint_np.py, itself pinned by complex-step derivatives of the
  oracle's forward substep) on the same seeded state, every block;
- FP64: the dot product <J dx, y> from central differences of the device's
  own forward substep against <dx, J^T y> from the device adjoint;
- FP32 adjoint against the FP64 oracle adjoint at FP32 accuracy;
- autograd through two substeps against central differences.
"""
import numpy as np
import pytest
import torch

from oracle import adjoint_np as A
from oracle import dycore_np as O
from tests.test_oracle_adjoint import acoustic_case

pytestmark = pytest.mark.gpu

NAMES_IN = A.SUBSTEP_STATE + A.SUBSTEP_TEND
NAMES_OUT = A.SUBSTEP_STATE + ("ww_pp",)
REF = ("u", "v", "w", "thp", "php", "mup", "p", "al", "alt")


def _device_pair(dtype=torch.float64, **kw):
    """The oracle case and a device state holding the same values."""
    from arwen_amd import dycore
    from arwen_amd.state import DomainState
    from tests.helpers import random_state_case
    o, coefs, dtau, cfg, rng = acoustic_case(**kw)
    case, _ = random_state_case(seed=kw.get("seed", 4), nz=9, ny=6, nx=10,
                                hill=kw.get("hill", 0.0))
    g = DomainState.from_host(case, dtype=dtype)
    for n in REF:
        getattr(g, n).copy_(torch.as_tensor(getattr(o, n)))
    # the coefficient pass rebuilds Omega and the reference mass divergence
    # and clears rmu_t (the stage's own mass tendency is carried there), so
    # the small-step fields and tendencies go in afterwards
    dycore.prepare_acoustic_coefficients(g, cfg, dtau)
    for n in NAMES_IN + ("ww_pp",):
        getattr(g, n).copy_(torch.as_tensor(getattr(o, n)))
    return o, coefs, dtau, cfg, rng, g


def _rel(a, b):
    a = a.detach().cpu().numpy().astype(np.float64) if torch.is_tensor(a) else a
    s = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / s) if s > 0 else float(np.max(np.abs(a)))


@pytest.mark.parametrize("hill,damp,top_lid,first", [
    (0.0, False, True, True), (0.0, False, True, False),
    (300.0, True, True, False), (0.0, True, False, False)])
def test_fp64_adjoint_matches_oracle(hill, damp, top_lid, first):
    from arwen_amd import adjoint
    o, coefs, dtau, cfg, rng, g = _device_pair(hill=hill, damp=damp,
                                               top_lid=top_lid)
    print(f"acoustic adjoint fp64 hill={hill} damp={damp} lid={top_lid} "
          f"first={first}", flush=True)
    lam = {n: rng.standard_normal(getattr(o, n).shape) for n in NAMES_OUT}
    ref = A.acoustic_substep_adjoint(o, coefs, dtau, cfg, first, lam)
    got = adjoint.acoustic_substep_adjoint(
        g, cfg, dtau, first,
        {n: torch.as_tensor(v, device=g.u.device) for n, v in lam.items()})
    bad = {n: _rel(got[n], ref[n]) for n in NAMES_IN}
    assert max(bad.values()) < 1e-10, bad


def test_fp64_dot_product_against_device_forward():
    from arwen_amd import adjoint, dycore
    o, coefs, dtau, cfg, rng, g = _device_pair(hill=200.0, damp=True)
    print("acoustic adjoint fp64 dot product vs device FD", flush=True)
    dev = g.u.device
    x0 = {n: getattr(g, n).clone() for n in NAMES_IN}
    dx = {n: torch.as_tensor(rng.standard_normal(getattr(o, n).shape)
                             * max(float(x0[n].abs().max()), 1e-12), device=dev)
          for n in NAMES_IN}
    y = {n: torch.as_tensor(rng.standard_normal(getattr(o, n).shape),
                            device=dev) for n in NAMES_OUT}

    def F(eps):
        for n in NAMES_IN:
            getattr(g, n).copy_(x0[n] + eps * dx[n])
        dycore.acoustic_substep(g, cfg, dtau, False)
        return {n: getattr(g, n).clone() for n in NAMES_OUT}

    eps = 1e-5
    fp, fm = F(eps), F(-eps)
    lhs = sum(float(torch.sum((fp[n] - fm[n]) / (2 * eps) * y[n]))
              for n in NAMES_OUT)
    for n in NAMES_IN:
        getattr(g, n).copy_(x0[n])
    jty = adjoint.acoustic_substep_adjoint(g, cfg, dtau, False, y)
    rhs = sum(float(torch.sum(dx[n] * jty[n])) for n in NAMES_IN)
    assert abs(lhs - rhs) <= 1e-7 * abs(lhs), (lhs, rhs)


def test_fp32_adjoint_close_to_fp64_oracle():
    from arwen_amd import adjoint
    o, coefs, dtau, cfg, rng, g = _device_pair(dtype=torch.float32,
                                               hill=200.0, damp=True)
    print("acoustic adjoint fp32", flush=True)
    lam = {n: rng.standard_normal(getattr(o, n).shape) for n in NAMES_OUT}
    ref = A.acoustic_substep_adjoint(o, coefs, dtau, cfg, False, lam)
    got = adjoint.acoustic_substep_adjoint(
        g, cfg, dtau, False,
        {n: torch.as_tensor(v, dtype=torch.float32, device=g.u.device)
         for n, v in lam.items()})
    bad = {n: _rel(got[n], ref[n]) for n in NAMES_IN}
    assert max(bad.values()) < 1e-3, bad


def test_autograd_two_substeps_against_fd():
    from arwen_amd import adjoint
    o, coefs, dtau, cfg, rng, g = _device_pair(damp=True)
    print("acoustic autograd two substeps", flush=True)
    dev = g.u.device
    x0 = [getattr(g, n).clone() for n in NAMES_IN]
    w = [torch.as_tensor(rng.standard_normal(t.shape), device=dev)
         for t in x0[:9]]
    d = [torch.as_tensor(rng.standard_normal(t.shape), device=dev)
         * max(float(t.abs().max()), 1e-12) for t in x0]

    def loss(xs):
        st, tend = list(xs[:9]), list(xs[9:])
        out = adjoint.AcousticSubstep.apply(g, cfg, dtau, True, *st, *tend)
        out = adjoint.AcousticSubstep.apply(g, cfg, dtau, False,
                                            *out[:9], *tend)
        return sum(torch.sum(wi * oi) for wi, oi in zip(w, out[:9]))

    xs = [t.clone().requires_grad_(True) for t in x0]
    grads = torch.autograd.grad(loss(xs), xs)
    ad = sum(float(torch.sum(gi * di)) for gi, di in zip(grads, d))
    eps = 1e-5
    with torch.no_grad():
        lp = float(loss([t + eps * di for t, di in zip(x0, d)]))
        lm = float(loss([t - eps * di for t, di in zip(x0, d)]))
    fd = (lp - lm) / (2 * eps)
    assert abs(fd - ad) <= 1e-7 * abs(ad), (fd, ad)
