"""Run configuration: the subset of the reference ``RunConfig`` the dycore
step reads (gpuwm/config.py:27-190), with identical field names and
defaults, plus the validation the step performs before any device work.

Fields for subsystems this build does not implement (physics schemes --
whose species the step still transports --, nesting, open lateral
boundaries, Smagorinsky) are
accepted with their reference defaults so a reference config constructs
unchanged; :func:`validate_for_step` refuses any non-default value with a
``NotImplementedError`` instead of silently ignoring it.
"""
from __future__ import annotations

import math

from dataclasses import dataclass, fields, replace


@dataclass(frozen=True)
class RunConfig:
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    ztop: float
    dt: float
    run_seconds: float = 0.0
    clock_dt: float = 0.0
    p_surf: float = 1.0e5
    time_step_sound: int = 4
    epssm: float = 0.1
    smdiv: float = 0.1
    khdif: float = 0.0
    kvdif: float = 0.0
    damp_opt: int = 0
    zdamp: float = 5000.0
    dampcoef: float = 0.2
    output_interval_s: float = 300.0
    case: str = ""
    hybrid_opt: int = 0
    etac: float = 0.2
    moist: bool = False
    mp_physics: int = 0
    moist_adv_opt: int = 1
    diff_6th_opt: int = 0
    diff_6th_factor: float = 0.12
    diff_6th_slopeopt: int = 0
    diff_6th_thresh: float = 0.10
    # WRF moist_mix6_off: the filter skips the moist array (qv, qc, qr,
    # qi, qs, qg, qh); number and volume tracers keep it (reference
    # config.py:529-551)
    moist_mix6_off: bool = False
    km_opt: int = 1
    # km_opt = 4 (2-D Smagorinsky): the Smagorinsky constant and the
    # surface policy of the dycore's vertical mixing (reference
    # config.py:83, :345, :380-381)
    c_s: float = 0.25
    isfflx: int = 1
    tke_heat_flux: float = 0.0
    tke_drag_coefficient: float = 0.0
    w_damping: int = 0
    open_x: bool = False
    open_y: bool = False
    specified: bool = False
    spec_zone: int = 1
    emdiv: float = 0.0
    h_sca_adv_order: int = 2
    terrain_opt: int = 0
    hill_height: float = 100.0
    hill_halfwidth: float = 10000.0
    hypsometric_opt: int = 1
    nested: bool = False
    top_lid: bool = True
    moist_cq: bool = False
    bl_pbl_physics: int = 0
    ra_physics: int = 0
    cu_physics: int = 0
    sf_sfclay_physics: int = 0
    sf_surface_physics: int = 0
    nwp_diagnostics: int = 0
    # --- extension of this build (no reference field): the precision form
    # of the float32 step -- Kahan-compensated column sums and an FP64
    # vertical implicit tridiagonal
    precision_form: bool = False

    def with_(self, **kw) -> "RunConfig":
        """Copy with fields replaced (frozen dataclass convenience)."""
        return replace(self, **kw)


def config_fields() -> tuple[str, ...]:
    return tuple(f.name for f in fields(RunConfig))


# ---------------------------------------------------------------------------
# Transported scalar species of a moist state (reference moist.py:107-240,
# allocation state.py:428-470).  The step advects every species of the
# state with the acoustic time-averaged fluxes (positive definite on the
# final stage); the microphysics scheme itself runs outside the step, in
# the caller's physics driver.  The condensate masses listed first in a
# scheme's extras also load the buoyancy (dycore.cu:213-230, q_total).
# ---------------------------------------------------------------------------

#: the qv/qc/qr base set of every moist state
BASE_SPECIES = ("qv", "qc", "qr")
_ICE_MASS = ("qi", "qs", "qg")
#: transported species beyond qv/qc/qr per mp_physics
EXTRA_SPECIES = {
    0: (), 1: (),
    6: _ICE_MASS,                                         # WSM6
    8: _ICE_MASS + ("nr", "ni"),                          # Thompson
    10: _ICE_MASS + ("nr", "ni", "ns", "ng"),             # Morrison
    18: ("qi", "qs", "qg", "qh", "qndrop", "qnr", "qni", "qns", "qng",
         "qnh", "qnn", "qvolg", "qvolh"),                 # NSSL
    28: _ICE_MASS + ("nr", "ni", "nc", "nwfa", "nifa"),   # Thompson aerosol
}
#: how many of a scheme's extra species (the first ones) load the
#: buoyancy with qv, qc, qr
LOADING_EXTRA = {0: 0, 1: 0, 6: 3, 8: 3, 10: 3, 18: 4, 28: 3}
#: the largest extra set any scheme carries (the C ABI's table size)
MAX_EXTRA_SPECIES = 13


def moist_species(cfg: RunConfig) -> tuple[str, ...]:
    """The transported species of a moist state under cfg.mp_physics."""
    if not cfg.moist:
        return ()
    return BASE_SPECIES + EXTRA_SPECIES[cfg.mp_physics]


def validate_for_step(cfg: RunConfig, *, acoustic: bool = True) -> None:
    """Fail-closed checks made before any device work.

    Mirrors the guards of gpuwm/core/dycore.py:2238-2293 for the features
    this build carries, and refuses (NotImplementedError) the reference
    features that are outside the ported path.
    """
    if acoustic and cfg.time_step_sound % 2 != 0:
        raise ValueError(
            f"time_step_sound must be even, got {cfg.time_step_sound}: RK3 "
            "stage 2 runs time_step_sound//2 acoustic substeps of "
            "dt/time_step_sound, which mis-times the dt/2 stage for odd "
            "values.")
    if cfg.time_step_sound < 1:
        raise ValueError("time_step_sound must be positive")
    if cfg.km_opt in (2, 3):
        raise NotImplementedError(
            f"km_opt={cfg.km_opt} (prognostic TKE / 3-D Smagorinsky) is not "
            "part of this port; km_opt 0, 1 and 4 are")
    if cfg.km_opt not in (0, 1, 4):
        raise ValueError(f"km_opt must be 0, 1, 2, 3 or 4, got {cfg.km_opt}")
    if cfg.km_opt == 4:
        if cfg.precision_form:
            raise NotImplementedError(
                "km_opt=4 is ported for the plain form")
        if cfg.nz < 3:
            raise ValueError("km_opt=4 needs nz >= 3 (the surface "
                             "extrapolation reads three levels)")
        if cfg.isfflx not in (0, 1, 2):
            raise ValueError(f"isfflx must be 0, 1 or 2, got {cfg.isfflx}")
        if not cfg.c_s >= 0.0:
            raise ValueError(f"c_s = {cfg.c_s!r} must be non-negative")
    if not 0.0 < cfg.epssm <= 1.0:
        raise ValueError(f"epssm must lie in (0, 1], got {cfg.epssm}")
    if cfg.smdiv < 0.0:
        raise ValueError(f"smdiv must be >= 0, got {cfg.smdiv}")
    if cfg.damp_opt not in (0, 3):
        raise ValueError(f"damp_opt must be 0 or 3, got {cfg.damp_opt}")
    if cfg.damp_opt == 3 and cfg.zdamp <= 0.0:
        raise ValueError("damp_opt=3 needs zdamp > 0")
    if not cfg.dampcoef >= 0.0:
        raise ValueError(f"dampcoef = {cfg.dampcoef!r}: a negative damping "
                         "rate would feed energy into the upper layer")
    if not cfg.diff_6th_factor >= 0.0:
        raise ValueError(f"diff_6th_factor = {cfg.diff_6th_factor!r}: a "
                         "negative factor makes the 6th-order filter "
                         "anti-diffusive")
    if cfg.h_sca_adv_order not in (2, 5):
        raise ValueError(
            f"h_sca_adv_order must be 2 or 5 (the rhs_ph horizontal "
            f"geopotential advection), got {cfg.h_sca_adv_order}")
    if cfg.hypsometric_opt not in (1, 2):
        raise ValueError(
            "hypsometric_opt must be 1 (d(phi)/d(eta) hydrostatic alpha) or "
            f"2 (log-pressure hypsometric alpha), got {cfg.hypsometric_opt}")
    if cfg.hypsometric_opt == 2 and cfg.precision_form:
        raise NotImplementedError(
            "the precision form's perturbation EOS is the hypsometric_opt=1 "
            "one")
    if cfg.moist_cq and cfg.precision_form:
        raise NotImplementedError(
            "moist_cq is ported for the plain form (the precision form's "
            "implicit solve runs in FP64)")
    for name in ("open_x", "open_y", "specified", "nested"):
        if getattr(cfg, name):
            raise NotImplementedError(f"{name}=True is not part of this port")
    if cfg.emdiv > 0.0 or cfg.w_damping:
        raise NotImplementedError(
            "emdiv / w_damping are not part of this port")
    if cfg.diff_6th_opt not in (0, 1, 2):
        raise ValueError(f"diff_6th_opt must be 0, 1 or 2, got "
                         f"{cfg.diff_6th_opt}")
    if cfg.diff_6th_opt and cfg.precision_form:
        raise NotImplementedError(
            "diff_6th_opt is ported for the plain form")
    if cfg.mp_physics not in EXTRA_SPECIES:
        raise NotImplementedError(
            f"mp_physics={cfg.mp_physics}: the step transports the species "
            f"of mp_physics {sorted(EXTRA_SPECIES)}")
    if cfg.mp_physics != 0 and not cfg.moist:
        raise ValueError("mp_physics != 0 needs moist=True")
    if cfg.mp_physics != 0 and cfg.precision_form:
        raise NotImplementedError(
            "the precision form is ported for mp_physics = 0 (no retained "
            "microphysics heating, no mixed-phase species)")
    # km_opt = 4 with a PBL scheme: the step mixes horizontally only and
    # leaves the vertical to the scheme, which runs outside the step
    # (reference dycore.py:1285)
    pbl = cfg.bl_pbl_physics and cfg.km_opt != 4
    if pbl or cfg.ra_physics or cfg.cu_physics \
            or cfg.sf_sfclay_physics or cfg.sf_surface_physics:
        raise NotImplementedError(
            "physics schemes are outside the dycore step port")
    if cfg.khdif < 0.0 or cfg.kvdif < 0.0:
        raise ValueError("khdif/kvdif must be non-negative")
    if cfg.moist_adv_opt not in (0, 1):
        raise NotImplementedError(
            f"moist_adv_opt={cfg.moist_adv_opt}: only 0 (unlimited) and 1 "
            "(positive definite) are ported")



def diff6_factor(cfg: RunConfig) -> float:
    """Per-step 6th-order filter factor (reference dycore.py:1767-1786):
    the namelist factor, or, when the model clock step ``clock_dt``
    differs from dt, 1 - (1 - factor)^(dt/clock_dt), so that the steps of
    one clock interval together remove the namelist fraction of a 2dx
    mode."""
    clock = cfg.clock_dt if cfg.clock_dt > 0.0 else cfg.dt
    factor = float(cfg.diff_6th_factor)
    if clock == cfg.dt:
        return factor
    if not 0.0 <= factor <= 1.0:
        raise ValueError("clock-scaled diff_6th_factor must lie in [0, 1], "
                         f"got {factor}")
    if factor == 1.0:
        return 1.0
    return -math.expm1((cfg.dt / clock) * math.log1p(-factor))
