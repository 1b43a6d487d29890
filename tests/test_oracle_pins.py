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
