"""ctypes binding of ``libarwen_dycore.so`` (see include/arwen_dycore.h).

The shared library is built in-tree by :mod:`arwen_amd.build`.  Loading
fails loudly when it is missing: there is no CPU fallback for the step.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_NAME = "libarwen_dycore.so"
LIB_PATH = Path(__file__).resolve().parent / LIB_NAME

ABI_VERSION = 9
#: transported species beyond qv/qc/qr a domain can carry (ARW_MAX_EXTRA)
MAX_EXTRA = 13

_P = ctypes.c_void_p
_I = ctypes.c_int32
_F = ctypes.c_float
_D = ctypes.c_double

DOMAIN_ARRAYS = (
    "u", "v", "w", "thp", "php", "mup",
    "p", "al", "alt",
    "u0", "v0", "w0", "thp0", "php0", "mup0",
    "ru_t", "rv_t", "rw_t", "rth_t", "rph_t", "rmu_t",
    "u_pp", "v_pp", "w_pp", "th_pp", "ph_pp", "mu_pp",
    "p_pp", "p_pp_old", "ww_pp", "al_pp",
    "thb", "pb", "alb", "phb", "mub2d", "ht",
    "c1h", "c2h", "c1f", "c2f", "dnw", "rdnw", "dn", "rdn", "fnm", "fnp",
)
SCRATCH_ARRAYS = ("ww", "gam", "dmdt_rf", "mup_next")
DIFF_ARRAYS = ("rdzf_h", "rdzc_h", "rdzf_w", "rdzc_w")
DIFF6_ARRAYS = ("d6_u", "d6_v", "d6_w", "d6_t")
#: moist scalars: state arrays, then the third stage buffer (scratch)
MOIST_ARRAYS = ("qv", "qc", "qr", "qv0", "qc0", "qr0", "qsv", "qsc", "qsr")
#: moist scratch: WRF sumflux averages and the PD limiter's buffers
MOIST_SCRATCH = ("ru_m", "rv_m", "ww_m", "pd_fxl", "pd_fxc", "pd_fyl",
                 "pd_fyc", "pd_fzl", "pd_fzc", "pd_sc")


class ArwDomain(ctypes.Structure):
    _fields_ = ([("nz", _I), ("ny", _I), ("nx", _I), ("base3d", _I)]
                + [(n, _P) for n in DOMAIN_ARRAYS]
                + [(n, _F) for n in ("cf1", "cf2", "cf3", "cfn", "cfn1")]
                + [(n, _P) for n in SCRATCH_ARRAYS]
                + [(n, _P) for n in DIFF_ARRAYS]
                + [(n, _P) for n in MOIST_ARRAYS]
                + [(n, _P) for n in MOIST_SCRATCH]
                + [("tri64", _P), ("pp", _P)]
                + [(n, _P) for n in DIFF6_ARRAYS]
                + [("n_extra", _I), ("n_load", _I)]
                + [(n, _P * MAX_EXTRA) for n in ("qx", "qx0", "qxs")]
                + [("h_diabatic", _P), ("hypsometric_opt", _I)]
                + [(n, _P) for n in ("c3h", "c4h", "c3f", "c4f")]
                + [("p_top", _D)]
                + [(n, _P) for n in ("cqu", "cqv", "cqw", "sm_buf",
                                     "held_q")])


class ArwDomain64(ctypes.Structure):
    """arw64_domain: the same fields with FP64 arrays and scalars."""
    _fields_ = [(n, _D if t is _F else t) for n, t in ArwDomain._fields_]


class ArwParams(ctypes.Structure):
    _fields_ = [("dx", _D), ("dy", _D), ("dt", _D),
                ("time_step_sound", _I),
                ("epssm", _D), ("smdiv", _D),
                ("khdif", _D), ("kvdif", _D),
                ("damp_opt", _I), ("zdamp", _D), ("dampcoef", _D),
                ("top_lid", _I), ("moist_adv_opt", _I),
                ("precision_form", _I), ("diff_6th_opt", _I),
                ("diff_6th_factor", _D), ("diff_6th_slopeopt", _I),
                ("diff_6th_thresh", _D), ("h_sca_adv_order", _I),
                ("moist_cq", _I), ("km_opt", _I), ("c_s", _D),
                ("bl_pbl_physics", _I), ("isfflx", _I),
                ("tke_heat_flux", _D), ("tke_drag_coefficient", _D),
                ("diff6_species", ctypes.c_uint32)]


class ArwTracerCkpt(ctypes.Structure):
    """arw_tracer_ckpt: per-stage sumflux averages, post-acoustic mu' and
    stage lengths of one step, plus the time-t mu'."""
    _fields_ = [("ru_m", _P * 3), ("rv_m", _P * 3), ("ww_m", _P * 3),
                ("mup_new", _P * 3), ("mup_t", _P), ("dt", _D * 3)]


#: arw64_tracer_ckpt has the same layout (pointers and doubles)
ArwTracerCkpt64 = ArwTracerCkpt

#: small-step state and tendencies of the acoustic substep adjoint
ACOUSTIC_STATE = ("u_pp", "v_pp", "w_pp", "th_pp", "ph_pp", "mu_pp", "p_pp",
                  "p_pp_old", "al_pp")
ACOUSTIC_TEND = ("ru_t", "rv_t", "rw_t", "rth_t", "rph_t", "rmu_t")


class ArwAcousticAdj(ctypes.Structure):
    """arw_acoustic_adj (and arw64_acoustic_adj: same layout)."""
    _fields_ = ([("l_" + n, _P) for n in ACOUSTIC_STATE + ("ww_pp",)]
                + [("g_" + n, _P) for n in ACOUSTIC_STATE + ACOUSTIC_TEND])


ArwAcousticAdj64 = ArwAcousticAdj

#: stage state and slow tendencies of the tendency-pass adjoint
STAGE_STATE = ("u", "v", "w", "thp", "php", "mup")
STAGE_TEND = ("ru_t", "rv_t", "rw_t", "rth_t", "rph_t")


class ArwStageAdj(ctypes.Structure):
    """arw_stage_adj (and arw64_stage_adj: same layout)."""
    _fields_ = ([("l_" + n, _P) for n in STAGE_TEND]
                + [("g_" + n, _P) for n in STAGE_STATE]
                + [("g_qv", _P), ("g_qc", _P), ("g_qr", _P),
                   ("l_ru", _P), ("l_rv", _P), ("l_ww", _P),
                   ("flags", ctypes.c_int32)])


ArwStageAdj64 = ArwStageAdj

#: outputs of the small-step initialisation, inputs of the finish
INIT_OUT = ("u_pp", "v_pp", "w_pp", "th_pp", "ph_pp", "mu_pp", "al_pp",
            "p_pp", "p_pp_old")
FINISH_PP = ("u_pp", "v_pp", "w_pp", "th_pp", "ph_pp", "mu_pp")


class ArwInitAdj(ctypes.Structure):
    """arw_init_adj (and arw64_init_adj: same layout)."""
    _fields_ = ([("l_" + n, _P) for n in INIT_OUT]
                + [("g_" + n + "0", _P) for n in STAGE_STATE]
                + [("g_" + n, _P) for n in STAGE_STATE]
                + [("g_qv", _P), ("flags", ctypes.c_int32)])


class ArwFinishAdj(ctypes.Structure):
    """arw_finish_adj (and arw64_finish_adj: same layout)."""
    _fields_ = ([("l_" + n, _P) for n in STAGE_STATE]
                + [("g_" + n, _P) for n in STAGE_STATE]
                + [("g_" + n, _P) for n in FINISH_PP])


ArwInitAdj64 = ArwInitAdj
ArwFinishAdj64 = ArwFinishAdj

#: accumulators of the substep adjoint's stage-reference share
ACOUSTIC_REF = ("mut", "c2a", "alt", "thr", "ph", "w", "grf")
#: flags of the accumulating adjoint entry points
ADJ_CLEAR = 1
ADJ_ACCUMULATE = 2
#: outputs of a substep the reference-share adjoint can take as given
ACOUSTIC_FWD = ("u_pp", "v_pp", "w_pp", "th_pp", "mu_pp", "al_pp", "ww_pp")


class ArwAcousticRef(ctypes.Structure):
    """arw_acoustic_ref (and arw64_acoustic_ref: same layout)."""
    _fields_ = [(n, _P) for n in ACOUSTIC_REF]


ArwAcousticRef64 = ArwAcousticRef


class ArwAcousticFwd(ctypes.Structure):
    """arw_acoustic_fwd (and arw64_acoustic_fwd: same layout)."""
    _fields_ = [(n, _P) for n in ACOUSTIC_FWD]


ArwAcousticFwd64 = ArwAcousticFwd


#: flags of arw_health_field
HEALTH_LOWER, HEALTH_UPPER, HEALTH_STRICT_LOWER = 1, 2, 4
HEALTH_ADD_AUX, HEALTH_AUX_LEVEL = 8, 16


class ArwHealthField(ctypes.Structure):
    """arw_health_field: one descriptor of the full-state health gate."""
    _fields_ = [("values", _P), ("aux", _P), ("size", ctypes.c_uint64),
                ("lower", _D), ("upper", _D), ("flags", ctypes.c_uint32),
                ("plane", ctypes.c_uint32), ("status_bit", ctypes.c_uint64)]


class ArwenError(RuntimeError):
    """A library entry point returned a nonzero status."""


_lib = None


def load():
    """Load (once) and return the library handle."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ArwenError(
            f"{LIB_NAME} is not built ({LIB_PATH}); run "
            "`python -c 'import arwen_amd.build as b; b.build()'` first -- "
            "the dycore step has no CPU fallback")
    lib = ctypes.CDLL(os.fspath(LIB_PATH))
    dp, pp = ctypes.POINTER(ArwDomain), ctypes.POINTER(ArwParams)
    sigs = {
        "arw_abi_version": ([], _I),
        "arw_last_error": ([], ctypes.c_char_p),
        "arw_step": ([dp, pp, _P], _I),
        "arw_step_noalloc": ([dp, pp, _P], _I),
        "arw_step_advective": ([dp, pp, _P], _I),
        "arw_health_gate": ([_P, _I, ctypes.c_uint64, ctypes.c_uint32, _P,
                             _P], _I),
        "arw_calc_p_alpha": ([dp, _P], _I),
        "arw_stage_fluxes": ([dp, pp, _P, _P, _P], _I),
        "arw_stage_prepare": ([dp, pp, _D, _I, _P], _I),
        "arw_acoustic_substep": ([dp, pp, _D, _I, _P], _I),
        "arw_calc_coefs": ([dp, pp, _D, _P], _I),
        "arw_small_step_finish": ([dp, _P], _I),
        "arw_set_w_surface": ([dp, pp, _P], _I),
        "arw_health": ([dp, _P, _P], _I),
        "arw_domain_mass": ([dp, _P, _P], _I),
        "arw_profile_begin": ([_I], _I),
        "arw_profile_end": ([_P, _P, _I], _I),
        "arw_kernel_name": ([_I], ctypes.c_char_p),
        "arw_kernel_count": ([], _I),
        "arw_step_ckpt": ([dp, pp, ctypes.POINTER(ArwTracerCkpt), _P], _I),
        "arw_tracer_step": ([dp, pp, ctypes.POINTER(ArwTracerCkpt), _P, _P,
                             _P, _I, _P], _I),
        "arw_tracer_adjoint": ([dp, pp, ctypes.POINTER(ArwTracerCkpt), _P,
                                _P, _P, _P], _I),
        "arw_acoustic_substep_adjoint": ([dp, pp, _D, _I,
                                          ctypes.POINTER(ArwAcousticAdj), _P,
                                          _P], _I),
        "arw_acoustic_substep_adjoint_ref": ([dp, pp, _D, _I,
                                              ctypes.POINTER(ArwAcousticAdj),
                                              ctypes.POINTER(ArwAcousticRef),
                                              ctypes.POINTER(ArwAcousticFwd),
                                              _P, _I, _P, _P], _I),
        "arw_acoustic_adjoint_factors": ([dp, pp, _D, _P, _P], _I),
        "arw_acoustic_reference_gradients": ([dp, pp,
                                              ctypes.POINTER(ArwAcousticRef),
                                              _P, _P, _P, _P, _P, _P, _P, _I,
                                              _P, _P], _I),
        "arw_set_w_surface_adjoint": ([dp, pp, _P, _P, _P, _P], _I),
        "arw_scalar_stage": ([dp, pp, _D, _I] + [_P] * 9, _I),
        "arw_scalar_stage_adjoint": ([dp, pp, _D, _I] + [_P] * 17, _I),
        "arw_pd_stage": ([dp, pp, _D] + [_P] * 10, _I),
        "arw_pd_stage_adjoint": ([dp, pp, _D] + [_P] * 17, _I),
        "arw_acoustic_adjoint_work": ([_I, _I, _I,
                                       ctypes.POINTER(ctypes.c_int64)], _I),
        "arw_stage_tendencies_adjoint": ([dp, pp, ctypes.POINTER(ArwStageAdj),
                                          _P, _P], _I),
        "arw_stage_adjoint_work": ([_I, _I, _I,
                                    ctypes.POINTER(ctypes.c_int64)], _I),
        "arw_small_step_init_adjoint": ([dp, ctypes.POINTER(ArwInitAdj), _P,
                                         _P], _I),
        "arw_small_step_finish_adjoint": ([dp, ctypes.POINTER(ArwFinishAdj),
                                           _P, _P], _I),
        "arw_small_step_adjoint_work": ([_I, _I, _I,
                                         ctypes.POINTER(ctypes.c_int64)], _I),
        "arw_ensemble_create": ([dp, _I, pp, _P, _I, _I,
                                 ctypes.POINTER(ctypes.c_void_p)], _I),
        "arw_ensemble_step": ([_P, _I, _P], _I),
        "arw_ensemble_capture": ([_P, _I, _I], _I),
        "arw_ensemble_launch": ([_P, _P], _I),
        "arw_ensemble_destroy": ([_P], _I),
    }
    flux = [_P, _P, _P, _P, _P, _P, _P, _P, _F, _F, _I, _I, _I, _P]
    for name in ("arw_flux_div_scalar", "arw_flux_div_u", "arw_flux_div_v",
                 "arw_flux_div_w"):
        sigs[name] = (flux, _I)
    flux_adj = [_P] * 12 + [_F, _F, _I, _I, _I, _P]
    for name in FLUX_DIV_ADJOINTS:
        sigs["arw_" + name] = (flux_adj, _I)
    # the FP64 build of the same entry points
    d64 = ctypes.POINTER(ArwDomain64)
    for name in FP64_ENTRY_POINTS:
        base = "arw_" + name[len("arw64_"):]
        args, res = sigs[base]
        args = [d64 if a is dp else a for a in args]
        if name.startswith("arw64_flux_div_"):
            args = [_D if a is _F else a for a in args]
        sigs[name] = (args, res)
    for name, (args, res) in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.arw_abi_version() != ABI_VERSION \
            or lib.arw64_abi_version() != ABI_VERSION:
        raise ArwenError(f"{LIB_NAME} ABI {lib.arw_abi_version()} != "
                         f"expected {ABI_VERSION}; rebuild the library")
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        lib = load()
        msg = (lib.arw64_last_error() if what.startswith("arw64_")
               else lib.arw_last_error())
        raise ArwenError(f"{what} failed ({rc}): "
                         f"{msg.decode() if msg else ''}")


#: adjoints of the advection kernels (arw_ / arw64_ prefixed)
FLUX_DIV_ADJOINTS = ("flux_div_scalar_adjoint", "flux_div_u_adjoint",
                     "flux_div_v_adjoint", "flux_div_w_adjoint")

#: FP64 entry points (same discretisation in double precision)
FP64_ENTRY_POINTS = (
    "arw64_abi_version", "arw64_last_error", "arw64_step",
    "arw64_calc_p_alpha", "arw64_stage_fluxes", "arw64_stage_prepare",
    "arw64_acoustic_substep", "arw64_calc_coefs", "arw64_small_step_finish",
    "arw64_set_w_surface", "arw64_health", "arw64_domain_mass",
    "arw64_flux_div_scalar", "arw64_flux_div_u", "arw64_flux_div_v",
    "arw64_flux_div_w", "arw64_step_ckpt", "arw64_tracer_step",
    "arw64_tracer_adjoint", "arw64_acoustic_substep_adjoint",
    "arw64_acoustic_adjoint_work", "arw64_ensemble_create",
    "arw64_ensemble_step",
    "arw64_ensemble_capture", "arw64_ensemble_launch",
    "arw64_ensemble_destroy", "arw64_stage_tendencies_adjoint",
    "arw64_stage_adjoint_work", "arw64_small_step_init_adjoint",
    "arw64_small_step_finish_adjoint", "arw64_small_step_adjoint_work",
    "arw64_acoustic_substep_adjoint_ref", "arw64_acoustic_adjoint_factors",
    "arw64_acoustic_reference_gradients",
    "arw64_set_w_surface_adjoint", "arw64_scalar_stage",
    "arw64_scalar_stage_adjoint", "arw64_pd_stage",
    "arw64_pd_stage_adjoint", "arw64_step_noalloc",
    "arw64_health_gate", "arw64_step_advective") + tuple(
        "arw64_" + n for n in FLUX_DIV_ADJOINTS)


def pd_work_elems(nz: int, ny: int, nx: int) -> int:
    """Work elements of arw_pd_stage / arw_pd_stage_adjoint
    (ARW_PD_WORK in the header)."""
    return 26 * (nz + 1) * (ny + 1) * (nx + 1)


def exported_symbols() -> list[str]:
    """Names of the C entry points this binding declares."""
    return list(FP64_ENTRY_POINTS) + ["arw_abi_version", "arw_last_error", "arw_step",
            "arw_calc_p_alpha", "arw_stage_fluxes", "arw_stage_prepare",
            "arw_acoustic_substep", "arw_calc_coefs", "arw_small_step_finish",
            "arw_set_w_surface", "arw_health", "arw_domain_mass",
            "arw_flux_div_scalar", "arw_flux_div_u", "arw_flux_div_v",
            "arw_flux_div_w", "arw_profile_begin", "arw_profile_end",
            "arw_kernel_name", "arw_kernel_count", "arw_step_ckpt",
            "arw_tracer_step", "arw_tracer_adjoint",
            "arw_acoustic_substep_adjoint", "arw_acoustic_adjoint_work",
            "arw_ensemble_create",
            "arw_ensemble_step", "arw_ensemble_capture",
            "arw_ensemble_launch", "arw_ensemble_destroy",
            "arw_stage_tendencies_adjoint", "arw_stage_adjoint_work",
            "arw_small_step_init_adjoint", "arw_small_step_finish_adjoint",
            "arw_small_step_adjoint_work",
            "arw_acoustic_substep_adjoint_ref",
            "arw_acoustic_adjoint_factors",
            "arw_acoustic_reference_gradients",
            "arw_set_w_surface_adjoint", "arw_scalar_stage",
            "arw_scalar_stage_adjoint", "arw_pd_stage",
            "arw_pd_stage_adjoint", "arw_step_noalloc",
            "arw_health_gate", "arw_step_advective"] + [
            "arw_" + n for n in FLUX_DIV_ADJOINTS]


class KernelTimer:
    """Context manager collecting per-kernel HIP-event time of arw_step."""

    def __init__(self, max_events: int = 1 << 16):
        self.max_events = max_events
        self.ms = {}
        self.launches = {}

    def __enter__(self):
        check(load().arw_profile_begin(self.max_events), "arw_profile_begin")
        return self

    def __exit__(self, *exc):
        lib = load()
        n = lib.arw_kernel_count()
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int32 * n)()
        check(lib.arw_profile_end(ctypes.cast(ms, _P), ctypes.cast(cnt, _P),
                                  n), "arw_profile_end")
        for i in range(n):
            name = lib.arw_kernel_name(i).decode()
            self.ms[name] = ms[i]
            self.launches[name] = cnt[i]
        return False
