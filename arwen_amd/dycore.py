"""The dycore step: the reference's Python call surface over the HIP library.

``step(state, cfg, ...)`` keeps the signature of gpuwm/core/dycore.py:
2161-2163 and mutates ``state`` in place; ``run_steps`` follows
dycore.py:2506-2510.  All arithmetic runs in ``libarwen_dycore.so`` on the
current PyTorch stream; this module validates the configuration (raising
before any device work, like dycore.py:2238-2293), builds the raw-pointer
views and calls the C-ABI.
"""
from __future__ import annotations

import math

import torch

from . import _lib
from .config import (EXTRA_SPECIES, LOADING_EXTRA, RunConfig, diff6_factor,
                     validate_for_step)
from .state import DomainState, mu_at_u_faces, mu_at_v_faces


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _fn(state: DomainState, name: str):
    """The entry point of the state's precision: arw_* (float32) or the
    FP64 build arw64_* of the same discretisation."""
    prefix = "arw64_" if state.dtype == torch.float64 else "arw_"
    return prefix + name, getattr(_lib.load(), prefix + name)


def _dom(state: DomainState, cfg: RunConfig):
    """The state's C view with the arrays cfg's step options need
    (diffusion spacings, the precision form's pp and FP64 scratch)."""
    return state.domain_struct(diffusion=_uses_diffusion(cfg),
                               precision_form=bool(cfg.precision_form),
                               hypsometric_opt=cfg.hypsometric_opt)


def params(cfg: RunConfig) -> _lib.ArwParams:
    p = _lib.ArwParams()
    p.dx, p.dy, p.dt = float(cfg.dx), float(cfg.dy), float(cfg.dt)
    p.time_step_sound = int(cfg.time_step_sound)
    p.epssm, p.smdiv = float(cfg.epssm), float(cfg.smdiv)
    if cfg.km_opt == 1:
        p.khdif, p.kvdif = float(cfg.khdif), float(cfg.kvdif)
    else:
        p.khdif = p.kvdif = 0.0
    p.damp_opt = int(cfg.damp_opt)
    p.zdamp, p.dampcoef = float(cfg.zdamp), float(cfg.dampcoef)
    p.top_lid = int(bool(cfg.top_lid))
    p.moist_adv_opt = int(cfg.moist_adv_opt)
    p.precision_form = int(bool(cfg.precision_form))
    p.diff_6th_opt = int(cfg.diff_6th_opt)
    p.diff_6th_factor = float(diff6_factor(cfg)) if cfg.diff_6th_opt else 0.0
    p.diff_6th_slopeopt = int(cfg.diff_6th_slopeopt)
    p.diff_6th_thresh = float(cfg.diff_6th_thresh)
    p.h_sca_adv_order = int(cfg.h_sca_adv_order)
    p.moist_cq = moist_cq_species(cfg)
    p.km_opt = int(cfg.km_opt)
    p.c_s = float(cfg.c_s)
    p.bl_pbl_physics = int(cfg.bl_pbl_physics)
    p.isfflx = int(cfg.isfflx)
    p.tke_heat_flux = float(cfg.tke_heat_flux)
    p.tke_drag_coefficient = float(cfg.tke_drag_coefficient)
    p.diff6_species = diff6_species_mask(cfg)
    return p


#: WRF's moist array: moist_mix6_off exempts these species from the
#: 6th-order filter (reference moist.py:172-173, dycore.py:1200-1220)
MOIST_ARRAY = frozenset(("qv", "qc", "qr", "qi", "qs", "qg", "qh"))


def diff6_species_mask(cfg: RunConfig) -> int:
    """arw_params.diff6_species: bit e for species e (qv, qc, qr, then the
    scheme's extras) the 6th-order filter acts on."""
    if not (cfg.moist and cfg.diff_6th_opt):
        return 0
    names = ("qv", "qc", "qr") + EXTRA_SPECIES[cfg.mp_physics]
    mask = 0
    for e, n in enumerate(names):
        if not (cfg.moist_mix6_off and n in MOIST_ARRAY):
            mask |= 1 << e
    return mask


def moist_cq_species(cfg: RunConfig) -> int:
    """arw_params.moist_cq: 0 off, else how many moist mass species WRF
    calc_cq sums (reference acoustic.py:61-110): qv (mp_physics 0), qv qc
    qr (1), plus the ice and hail masses (the mixed-phase schemes)."""
    if not (cfg.moist and cfg.moist_cq):
        return 0
    if cfg.mp_physics == 0:
        return 1
    return 3 + LOADING_EXTRA[cfg.mp_physics]


def _uses_diffusion(cfg: RunConfig) -> bool:
    return cfg.km_opt == 1 and (cfg.khdif > 0.0 or cfg.kvdif > 0.0)


def _check_state(state: DomainState) -> None:
    if not state.u.is_cuda:
        raise RuntimeError("the dycore step runs on the GPU: state tensors "
                           "must live on a ROCm device")
    if state.has_msf or state.rotational:
        raise NotImplementedError(
            "map factors / Coriolis are not part of this port")


def step(state: DomainState, cfg: RunConfig, *, acoustic: bool = True,
         mass_flux_observer=None, mass_flux_accumulator=None,
         refl_10cm_due: bool = False, microphysics=None) -> None:
    """Advance ``state`` one RK3 split-explicit step of length cfg.dt.

    With ``cfg.mp_physics != 0`` the reference's step ends with its
    microphysics scheme and a diagnostics refresh (dycore.py:2491-2502).
    The scheme is not part of this library: the RK3 dynamics here carry
    the scheme's species and apply and remove the retained heating, and
    ``microphysics`` (the reference's ``apply_microphysics``, called as
    ``microphysics(state, cfg, cfg.dt, refl_10cm_due=...)``) runs in that
    slot when given, followed by the EOS refresh."""
    if mass_flux_observer is not None and mass_flux_accumulator is not None:
        raise ValueError(
            "mass_flux_observer and mass_flux_accumulator are mutually "
            "exclusive")
    if not acoustic:
        return _step_advective(state, cfg)
    if mass_flux_observer is not None or mass_flux_accumulator is not None:
        raise NotImplementedError(
            "boundary mass-flux observers apply to open boundaries, which "
            "are not part of this port (periodic domains conserve mass)")
    validate_for_step(cfg, acoustic=acoustic)
    if bool(cfg.moist) != (state.qv is not None):
        raise ValueError(
            "cfg.moist must match the state: a moist step needs a state "
            "allocated with the qv/qc/qr scalars, and a dry one none")
    _check_state(state)
    d = state.domain_struct(diffusion=_uses_diffusion(cfg),
                            precision_form=bool(cfg.precision_form),
                            diff6=cfg.diff_6th_opt > 0,
                            hypsometric_opt=cfg.hypsometric_opt,
                            moist_cq=bool(cfg.moist_cq),
                            smag=cfg.km_opt == 4)
    p = params(cfg)
    name, fn = _fn(state, "step")
    _lib.check(fn(d, p, _stream()), name)
    if microphysics is not None and cfg.mp_physics != 0:
        microphysics(state, cfg, cfg.dt, refl_10cm_due=refl_10cm_due)
        update_diagnostics(state, cfg.hypsometric_opt)
    state.elapsed_seconds += cfg.dt


def _step_advective(state: DomainState, cfg: RunConfig) -> None:
    """The reference's advection-only step, step(acoustic=False)
    (dycore.py:2318-2343): coupled flux-form advection of theta, u, v, w
    with Omega = 0 plus constant-K diffusion, three stages from the
    time-t fields with w, phi' and mu' frozen.  Dry states with w == 0,
    like the reference (advection.py:151-156)."""
    if state.qv is not None or cfg.moist:
        raise NotImplementedError(
            "moist transport requires the acoustic dycore path "
            "(step(acoustic=False) is the dry advection test path)")
    if cfg.km_opt in (2, 3, 4):
        raise NotImplementedError(
            f"km_opt={cfg.km_opt} Smagorinsky mixing requires the acoustic "
            "dycore path (step(acoustic=False) is the dry advection test "
            "path)")
    validate_for_step(cfg, acoustic=False)
    if cfg.precision_form or cfg.diff_6th_opt:
        raise NotImplementedError(
            "step(acoustic=False) is the plain-form advection test path "
            "(no precision form, no 6th-order filter)")
    _check_state(state)
    if bool(state.w.any()):
        raise ValueError(
            "step(acoustic=False) (advection-only path) requires w == 0: "
            "its eta mass flux is zero, which is the flux only for w == 0; "
            "the full step advects with the diagnosed Omega")
    d = state.domain_struct(diffusion=_uses_diffusion(cfg),
                            hypsometric_opt=cfg.hypsometric_opt)
    name, fn = _fn(state, "step_advective")
    _lib.check(fn(d, params(cfg), _stream()), name)
    state.elapsed_seconds += cfg.dt


def run_steps(state: DomainState, cfg: RunConfig, n: int, *,
              acoustic: bool = True) -> None:
    for _ in range(n):
        step(state, cfg, acoustic=acoustic)


def update_diagnostics(state: DomainState, hypsometric_opt: int = 1) -> None:
    """Recompute p, al, alt from (thp, php, mup) (reference
    diagnostics.py:14-57; hypsometric_opt 1 or 2)."""
    if hypsometric_opt not in (1, 2):
        raise ValueError(
            f"hypsometric_opt must be 1 or 2, got {hypsometric_opt}")
    _check_state(state)
    name, fn = _fn(state, "calc_p_alpha")
    _lib.check(fn(state.domain_struct(hypsometric_opt=hypsometric_opt),
                  _stream()), name)


def stage_fluxes(state: DomainState, cfg: RunConfig):
    """(ru, rv, ww) at the current state (reference dycore.py:162-206)."""
    _check_state(state)
    nz, ny, nx = state.shape
    ru = state.scratch((nz, ny, nx + 1), "rk_ru")
    rv = state.scratch((nz, ny + 1, nx), "rk_rv")
    d = _dom(state, cfg)
    name, fn = _fn(state, "stage_fluxes")
    _lib.check(fn(d, params(cfg), ru.data_ptr(), rv.data_ptr(), _stream()),
               name)
    return ru, rv, state.scratch((nz + 1, ny, nx), "arw_ww")


def stage_prepare(state: DomainState, cfg: RunConfig, dtau: float,
                  do_eos: bool = True) -> None:
    """One RK stage's preparation at the current state: EOS (when do_eos),
    Omega, the slow tendencies ru_t..rph_t and the small-step
    initialisation (reference dycore.py:2378-2389, :1502-1532)."""
    _check_state(state)
    name, fn = _fn(state, "stage_prepare")
    _lib.check(fn(_dom(state, cfg), params(cfg), float(dtau),
                  int(bool(do_eos)), _stream()), name)


def small_step_finish(state: DomainState) -> None:
    """Recover the stage's new u, v, w, theta', phi', mu' from the
    small-step fields (reference dycore.cu:1087-1198)."""
    _check_state(state)
    name, fn = _fn(state, "small_step_finish")
    _lib.check(fn(state.domain_struct(), _stream()), name)


def set_w_surface(state: DomainState, cfg: RunConfig) -> None:
    """w(sfc) = u . grad(ht) (reference dycore.py:1999-2026)."""
    _check_state(state)
    name, fn = _fn(state, "set_w_surface")
    _lib.check(fn(_dom(state, cfg), params(cfg), _stream()), name)


def prepare_acoustic_coefficients(state: DomainState, cfg: RunConfig,
                                  dtau: float) -> None:
    """Implicit-w coefficients for the current reference state (reference
    acoustic.py:299-327)."""
    _check_state(state)
    name, fn = _fn(state, "calc_coefs")
    _lib.check(fn(_dom(state, cfg), params(cfg), float(dtau), _stream()),
               name)


def acoustic_substep(state: DomainState, cfg: RunConfig, dtau: float,
                     first: bool) -> None:
    """One acoustic substep with the prepared coefficients (reference
    acoustic.py:482-567)."""
    _check_state(state)
    name, fn = _fn(state, "acoustic_substep")
    _lib.check(fn(_dom(state, cfg), params(cfg), float(dtau),
                  int(bool(first)), _stream()), name)


def run_acoustic_only(state: DomainState, cfg: RunConfig, dtau: float,
                      n: int) -> None:
    """n substeps about the frozen reference state (reference
    acoustic.py:570-583)."""
    if n <= 0:
        return
    prepare_acoustic_coefficients(state, cfg, dtau)
    for i in range(n):
        acoustic_substep(state, cfg, dtau, first=(i == 0))


def domain_mass_measure(state: DomainState) -> float:
    """FP64 sum of the total dry column mass (reference dycore.py:209-218;
    identity map factors)."""
    _check_state(state)
    out = torch.zeros(1, dtype=torch.float64, device=state.device)
    name, fn = _fn(state, "domain_mass")
    _lib.check(fn(state.domain_struct(), out.data_ptr(), _stream()), name)
    return float(out.item())


def stability_report(state: DomainState, cfg: RunConfig | None = None) -> dict:
    """max |u|, |w|, |theta'| and the co-located CFL with one device
    reduction and one readback (reference dycore.py:2513-2582)."""
    _check_state(state)
    if state.u.numel() == 0 or state.w.numel() == 0:
        raise ValueError("zero-size array to reduction operation maximum "
                         "which has no identity")
    out = torch.zeros(4, dtype=state.dtype, device=state.device)
    name, fn = _fn(state, "health")
    _lib.check(fn(state.domain_struct(), out.data_ptr(), _stream()), name)
    u_max, w_max, th_max, wcfl = (float(x) for x in out.cpu().tolist())
    nan = not (math.isfinite(u_max) and math.isfinite(w_max)
               and math.isfinite(th_max))
    rep = {"u_max": u_max, "w_max": w_max, "th_max": th_max, "cfl": None,
           "horizontal_cfl": None, "vertical_cfl": None, "nan": nan}
    if cfg is not None and not nan:
        rep["horizontal_cfl"] = cfg.dt * u_max / cfg.dx
        rep["vertical_cfl"] = cfg.dt * wcfl
        rep["cfl"] = max(rep["horizontal_cfl"], rep["vertical_cfl"])
    return rep


def stability_gate_failed(report: dict, *, max_cfl: float,
                          max_w_ms: float) -> bool:
    """True when a health sample crosses a safety limit (reference
    dycore.py:2585-2603)."""
    cfl = report.get("cfl")
    w_max = report.get("w_max")
    return (bool(report.get("nan")) or cfl is None
            or not math.isfinite(float(cfl)) or float(cfl) > max_cfl
            or w_max is None or not math.isfinite(float(w_max))
            or float(w_max) > max_w_ms)


__all__ = ["step", "run_steps", "update_diagnostics", "stage_fluxes",
           "set_w_surface", "stability_report", "stability_gate_failed",
           "domain_mass_measure", "mu_at_u_faces", "mu_at_v_faces",
           "run_acoustic_only", "acoustic_substep",
           "prepare_acoustic_coefficients"]
