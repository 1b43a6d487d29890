"""Pin the CPU oracle against the reference's own fixtures and against
what the mathematics fixes (CPU only)."""
import os

import numpy as np
import pytest

import pins
from conftest import rel_err
from oracle import dycore_np as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

#: 5-step bound per field against the reference capture.  Two independent
#: FP32 implementations of this formulation drift apart by up to ~1e-2 of
#: the small fields (v, w) in 5 steps; the oracle's own FP32-vs-FP64
#: spread is the same size.  Planted mistakes exceed these bounds (see
#: test_planted_mistakes_are_caught).
PIN_TOL = {"u": 5e-4, "v": 3e-2, "w": 1e-2, "thp": 6e-4, "php": 1e-3,
           "mup": 5e-3}


def _run_pin(name, dtype, n=pins.PIN_STEPS):
    case = pins.CASES[name]()
    s = O.from_case(case, dtype)
    if name == "terrain":
        O.set_w_surface(s, case.cfg.dx, case.cfg.dy)
    O.run_steps(s, case.cfg, n)
    return s


@pytest.fixture(scope="module")
def golden_step():
    return np.load(os.path.join(GOLDEN, "phase2_step_regression.npz"))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("which", ["scalar", "u", "v", "w"])
def test_oracle_advection_matches_reference_fixture(which, dtype):
    """tests/data/advection_periodic_regression.npz holds the inputs and
    outputs of the reference's four periodic flux-divergence kernels."""
    from arwen_amd.grid import make_vertical_coord
    ref = np.load(os.path.join(GOLDEN, "advection_periodic_regression.npz"))
    nz = ref["q"].shape[0]
    c = make_vertical_coord(nz)
    d = {k: ref[k].astype(dtype) for k in ("q", "u", "v", "w", "ru", "rv",
                                           "rw")}
    co = {k: np.asarray(getattr(c, k), dtype) for k in
          ("rdnw", "rdn", "fnm", "fnp")}
    f = dtype(0.01)
    fn, field, rdz, key = {
        "scalar": (O.flux_div_scalar, "q", "rdnw", "tend_scalar"),
        "u": (O.flux_div_u, "u", "rdnw", "tend_u"),
        "v": (O.flux_div_v, "v", "rdnw", "tend_v"),
        "w": (O.flux_div_w, "w", "rdn", "tend_w")}[which]
    got = fn(d[field], d["ru"], d["rv"], d["rw"], co[rdz], co["fnm"],
             co["fnp"], f, f)
    assert rel_err(got, ref[key]) < 1e-6


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
def test_oracle_whole_step_matches_reference_capture(name, golden_step):
    s32 = _run_pin(name, np.float32)
    s64 = _run_pin(name, np.float64)
    for f in pins.PIN_FIELDS:
        ref = golden_step[f"{name}/{f}"]
        e32 = rel_err(getattr(s32, f), ref)
        assert e32 < PIN_TOL[f], (f, e32)
        # the capture sits within FP32 noise of the FP64 restatement: no
        # farther from it than 2.5x the oracle's own FP32 spread
        spread = rel_err(getattr(s32, f), getattr(s64, f))
        assert rel_err(ref, getattr(s64, f)) <= 2.5 * spread + 1e-6, f


# --- planted mistakes -------------------------------------------------------
# Each variant below plants one plausible transcription mistake into the
# oracle (on the CPU, via monkeypatch) and must push at least one field of
# the 5-step pin outside PIN_TOL -- evidence that the tolerances pin the
# restatement rather than admitting anything.  Two plausible mistakes are
# too small to resolve in 5 steps and are pinned elsewhere: a wrong index in
# the Thomas factorisation (test_oracle_closed_forms: dense solve), and the
# epssm off-centering weights (pinned by the reference source text only).

def _flux5_sign(qm3, qm2, qm1, q0, qp1, qp2, vel):
    return (vel * (37.0 * (q0 + qm1) - 8.0 * (qp1 + qm2) + (qp2 + qm3))
            + np.abs(vel) * (10.0 * (q0 - qm1) - 5.0 * (qp1 - qm2)
                             + (qp2 - qm3))) / 60.0


def _flux3_sign(qm2, qm1, q0, qp1, vel):
    return (vel * (7.0 * (q0 + qm1) - (qp1 + qm2))
            - np.abs(vel) * (3.0 * (q0 - qm1) - (qp1 - qm2))) / 12.0


def _lu_wrong_index(a, b, c):
    n = a.shape[0] - 1
    alpha = np.zeros_like(a)
    gam = np.zeros_like(a)
    for k in range(1, n):
        alpha[k] = 1.0 / (b[k] - a[k] * gam[k])          # gam[k-1] intended
        gam[k] = c[k] * alpha[k]
    alpha[n] = 1.0 / (b[n] - a[n] * gam[n - 1])
    return alpha, gam


MISTAKES = {
    "flux5 upwind sign": lambda mp: mp.setattr(O, "flux5", _flux5_sign),
    "flux3 dissipation sign": lambda mp: mp.setattr(O, "flux3", _flux3_sign),
    "Omega sign": lambda mp: mp.setattr(
        O, "omega_ref", lambda *a: -_ORIG["omega_ref"](*a)),
    "no divergence damping": lambda mp: mp.setattr(
        O, "advance_uv", lambda s, dtau, smdiv, *a: _ORIG["advance_uv"](
            s, dtau, 0.0, *a)),
    "buoyancy dropped": lambda mp: mp.setattr(O, "slow_buoyancy",
                                              lambda s, prec=False: None),
    "rhs_ph dropped": lambda mp: mp.setattr(
        O, "slow_geopotential", lambda s, ww, dx, dy, prec=False, order=2: None),
    "large-step PGF dropped": lambda mp: mp.setattr(
        O, "slow_pgf", lambda *a, **k: None),
    "theta'' update dropped": lambda mp: mp.setattr(
        O, "advance_mu_th", lambda s, dtau, dx, dy, prec=False:
        _ORIG["advance_mu_th"](s, 0.0, dx, dy, prec)),
}
_ORIG = {n: getattr(O, n) for n in ("omega_ref", "advance_uv",
                                    "advance_w_phi", "advance_mu_th")}


@pytest.mark.parametrize("mistake", sorted(MISTAKES))
def test_planted_mistakes_are_caught(mistake, monkeypatch, golden_step):
    MISTAKES[mistake](monkeypatch)
    worst = 0.0
    for name in ("dry_flat", "terrain"):
        with np.errstate(all="ignore"):
            s = _run_pin(name, np.float32)
        for f in pins.PIN_FIELDS:
            got = getattr(s, f)
            if not np.all(np.isfinite(got)):
                return
            worst = max(worst, rel_err(got, golden_step[f"{name}/{f}"])
                        / PIN_TOL[f])
    assert worst > 1.0, f"{mistake}: stays within the pin ({worst:.2f})"
