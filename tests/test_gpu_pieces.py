"""GPU parity of the step's passes one at a time (stage preparation,
acoustic substeps, small-step finish) against the float32 oracle, on a
stage reference state that differs from the time-t state."""
import numpy as np
import pytest
import torch

import pins
from helpers import compare, oracle_and_gpu
from oracle import dycore_np as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _stage_state(name):
    """Oracle + GPU states whose prognostics are one step ahead of their
    time-t (*0) copies, with consistent EOS diagnostics."""
    import torch
    from arwen_amd import dycore
    case = pins.CASES[name]()
    o, g = oracle_and_gpu(case, set_w=(name == "terrain"))
    for n in O.PROGNOSTICS:
        getattr(o, n + "0")[...] = getattr(o, n)
    O.step(o, case.cfg)
    for n in O.PROGNOSTICS:
        getattr(g, n).copy_(torch.as_tensor(getattr(o, n)))
        getattr(g, n + "0").copy_(torch.as_tensor(getattr(o, n + "0")))
    O.calc_p_alpha(o)
    dycore.update_diagnostics(g)
    return case, o, g


def _check(res, names):
    bad = {k: v for k, v in res.items() if k in names and v[0] > TOL}
    assert not bad, bad


@pytest.mark.parametrize("name", ["dry_flat", "terrain"])
def test_pieces(name):
    from arwen_amd import _lib, dycore
    case, o, g = _stage_state(name)
    cfg = case.cfg
    torch.cuda.synchronize()
    res = compare(o, g, ("p", "al", "alt"))
    print("eos", res)
    _check(res, ("p", "al", "alt"))

    # stage preparation (stage 2 of 3: dtau = dt/ns)
    dtau = cfg.dt / cfg.time_step_sound
    for n in ("ru_t", "rv_t", "rw_t", "rth_t", "rph_t", "rmu_t"):
        getattr(o, n)[...] = 0.0
    ru, rv, ww = O.stage_fluxes(o, cfg.dx, cfg.dy)
    O.add_slow_tendencies(o, ru, rv, ww, cfg.dx, cfg.dy, cfg.top_lid)
    O.small_step_init(o)
    lib = _lib.load()
    _lib.check(lib.arw_stage_prepare(g.domain_struct(), dycore.params(cfg),
                                     float(dtau), 0,
                                     torch.cuda.current_stream().cuda_stream),
               "stage_prepare")
    torch.cuda.synchronize()
    names = ("ru_t", "rv_t", "rw_t", "rth_t", "rph_t", "u_pp", "v_pp",
             "w_pp", "th_pp", "ph_pp", "mu_pp", "p_pp")
    res = compare(o, g, names)
    print("prepare", res)
    gww = g.scratch((cfg.nz + 1, cfg.ny, cfg.nx), "arw_ww").cpu().numpy()
    assert np.array_equal(gww, ww)
    _check(res, names)
    if g.thb.dim() == 3:
        _check(compare(o, g, ("al_pp",)), ("al_pp",))

    # two acoustic substeps
    coefs = O.calc_coefs(o, dtau, cfg.epssm, cfg.top_lid)
    dycore.prepare_acoustic_coefficients(g, cfg, dtau)
    for i in range(2):
        O.acoustic_substep(o, coefs, dtau, cfg, first=(i == 0))
        dycore.acoustic_substep(g, cfg, dtau, first=(i == 0))
        torch.cuda.synchronize()
        names = ("u_pp", "v_pp", "w_pp", "th_pp", "ph_pp", "mu_pp", "p_pp",
                 "al_pp", "ww_pp")
        res = compare(o, g, names)
        print("substep", i, res)
        _check(res, names)

    # finish
    O.small_step_finish(o)
    O.calc_p_alpha(o)
    _lib.check(lib.arw_small_step_finish(
        g.domain_struct(), torch.cuda.current_stream().cuda_stream), "finish")
    torch.cuda.synchronize()
    names = ("u", "v", "w", "thp", "php", "mup", "p", "al", "alt")
    res = compare(o, g, names)
    print("finish", res)
    _check(res, names)
