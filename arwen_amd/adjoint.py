"""Adjoints of the step's transport and acoustic kernels.

Scalar transport, with per-RK-stage checkpoints:

The dycore step moves a passive scalar (qv and friends; WRF
rk_scalar_tend + rk_update_scalar, reference moist.py:385-398 and
:434-698) with the substep-averaged mass fluxes of each RK stage.  For
given fluxes and column masses that transport is linear in the scalar, so
one step is q_{n+1} = J q_n with J fixed by the step's dynamics.

``step_with_checkpoints`` runs the ordinary step and records what J needs:
per stage the averaged fluxes ru_m, rv_m, ww_m, the post-acoustic mu' and
the stage length, plus the time-t mu'.  ``TracerStep`` is the
``torch.autograd.Function`` over J: forward launches the HIP transport
(``arw_tracer_step``, bit-identical to the update inside the step),
backward the hand-written transpose (``arw_tracer_adjoint``).

Acoustic substep: ``acoustic_substep_adjoint`` applies the transpose of
one acoustic substep's Jacobian (``arw_acoustic_substep_adjoint``) at the
state's current small-step fields; ``AcousticSubstep`` is the
``torch.autograd.Function`` over the substep (inputs: the nine small-step
fields and six slow tendencies; outputs: the nine fields and Omega'').
The reference has no adjoint; these entry points are new.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .config import RunConfig, validate_for_step
from .dycore import (_check_state, _dom, _fn, _stream, _uses_diffusion,
                     params)
from .state import DomainState


class TracerCheckpoints:
    """Device buffers of one step's stage checkpoints (about 9 fields of
    the state's size) and the ctypes view the library fills."""

    def __init__(self, state: DomainState):
        nz, ny, nx = state.shape
        dev, dt = state.device, state.dtype

        def z(*shape):
            return torch.zeros(shape, dtype=dt, device=dev)

        self.ru_m = [z(nz, ny, nx + 1) for _ in range(3)]
        self.rv_m = [z(nz, ny + 1, nx) for _ in range(3)]
        self.ww_m = [z(nz + 1, ny, nx) for _ in range(3)]
        self.mup_new = [z(ny, nx) for _ in range(3)]
        self.mup_t = z(ny, nx)
        self.shape = (nz, ny, nx)
        self.dtype = dt
        c = _lib.ArwTracerCkpt()
        for s in range(3):
            c.ru_m[s] = self.ru_m[s].data_ptr()
            c.rv_m[s] = self.rv_m[s].data_ptr()
            c.ww_m[s] = self.ww_m[s].data_ptr()
            c.mup_new[s] = self.mup_new[s].data_ptr()
        c.mup_t = self.mup_t.data_ptr()
        self.c = c

    @property
    def dt(self) -> tuple[float, float, float]:
        """Stage lengths nsub * dtau (dt/3, dt/2, dt for even ns)."""
        return tuple(self.c.dt)


def step_with_checkpoints(state: DomainState, cfg: RunConfig,
                          ckpt: TracerCheckpoints | None = None
                          ) -> TracerCheckpoints:
    """``dycore.step(state, cfg)`` that also records the stage checkpoints
    of the scalar transport.  The state advances exactly as in ``step``."""
    validate_for_step(cfg, acoustic=True)
    if bool(cfg.moist) != (state.qv is not None):
        raise ValueError("cfg.moist must match the state")
    _check_state(state)
    if ckpt is None:
        ckpt = TracerCheckpoints(state)
    if ckpt.shape != state.shape or ckpt.dtype != state.dtype:
        raise ValueError("checkpoint buffers do not match the state")
    d = state.domain_struct(diffusion=_uses_diffusion(cfg),
                            precision_form=bool(cfg.precision_form),
                            diff6=cfg.diff_6th_opt > 0,
                            hypsometric_opt=cfg.hypsometric_opt,
                            moist_cq=bool(cfg.moist_cq),
                            smag=cfg.km_opt == 4)
    name, fn = _fn(state, "step_ckpt")
    _lib.check(fn(d, params(cfg), ctypes.byref(ckpt.c), _stream()), name)
    state.elapsed_seconds += cfg.dt
    return ckpt


def _tracer_call(state, cfg, ckpt, kind, src, dst, clamp=False):
    n = src.numel()
    work = state.scratch((3 * n,), "arw_tracer_work")   # J uses 2n, J^T 3n
    name, fn = _fn(state, kind)
    d, p = _dom(state, cfg), params(cfg)
    if kind == "tracer_step":
        rc = fn(d, p, ctypes.byref(ckpt.c), src.data_ptr(), dst.data_ptr(),
                work.data_ptr(), int(bool(clamp)), _stream())
    else:
        rc = fn(d, p, ctypes.byref(ckpt.c), src.data_ptr(), dst.data_ptr(),
                work.data_ptr(), _stream())
    _lib.check(rc, name)
    return dst


def _as_field(x: torch.Tensor, state: DomainState) -> torch.Tensor:
    if tuple(x.shape) != state.shape:
        raise ValueError(f"scalar field shape {tuple(x.shape)} != "
                         f"{state.shape}")
    if x.dtype != state.dtype or x.device != state.u.device:
        raise ValueError("scalar field must match the state's dtype/device")
    return x.contiguous()


def tracer_step(q: torch.Tensor, state: DomainState, cfg: RunConfig,
                ckpt: TracerCheckpoints, clamp: bool = False) -> torch.Tensor:
    """J q: one step of the unlimited scalar transport recorded in
    ``ckpt`` (the final-stage clamp at zero when ``clamp``)."""
    _check_state(state)
    q = _as_field(q, state)
    return _tracer_call(state, cfg, ckpt, "tracer_step", q,
                        torch.empty_like(q), clamp)


def tracer_adjoint(lam: torch.Tensor, state: DomainState, cfg: RunConfig,
                   ckpt: TracerCheckpoints) -> torch.Tensor:
    """J^T lam (the clamp, a no-op on positive fields, is not
    differentiated)."""
    _check_state(state)
    lam = _as_field(lam, state)
    return _tracer_call(state, cfg, ckpt, "tracer_adjoint", lam,
                        torch.empty_like(lam))


class TracerStep(torch.autograd.Function):
    """q -> J q with the hand-written adjoint as its backward."""

    @staticmethod
    def forward(ctx, q, state, cfg, ckpt):
        ctx.state, ctx.cfg, ctx.ckpt = state, cfg, ckpt
        return tracer_step(q.detach(), state, cfg, ckpt)

    @staticmethod
    def backward(ctx, grad):
        return (tracer_adjoint(grad, ctx.state, ctx.cfg, ctx.ckpt), None,
                None, None)


def transport(q: torch.Tensor, state: DomainState, cfg: RunConfig,
              ckpts: list[TracerCheckpoints]) -> torch.Tensor:
    """Differentiable multi-step transport: J_n ... J_1 q."""
    for c in ckpts:
        q = TracerStep.apply(q, state, cfg, c)
    return q


# ---------------------------------------------------------------------------
# Acoustic substep
# ---------------------------------------------------------------------------

ACOUSTIC_STATE = _lib.ACOUSTIC_STATE
ACOUSTIC_TEND = _lib.ACOUSTIC_TEND


def _dom_views(state: DomainState, cfg: RunConfig, views: dict | None):
    """The state's C view with some arrays replaced by other tensors of the
    same shape (checkpoints the call reads in place of the state's)."""
    d = _dom(state, cfg)
    nz, ny, nx = state.shape
    scratch = {"ww": (nz + 1, ny, nx), "dmdt_rf": (ny, nx)}
    for n, t in (views or {}).items():
        want = scratch[n] if n in scratch else tuple(getattr(state, n).shape)
        if tuple(t.shape) != want or t.dtype != state.dtype \
                or t.device != state.u.device or not t.is_contiguous():
            raise ValueError(f"view of {n}: expected a contiguous {want} "
                             f"{state.dtype} tensor on {state.u.device}")
        setattr(d, n, t.data_ptr())
    return d


def acoustic_substep_adjoint(state: DomainState, cfg: RunConfig,
                             dtau: float, first: bool,
                             lam: dict, ref: dict | None = None,
                             fwd: dict | None = None,
                             views: dict | None = None,
                             lu: torch.Tensor | None = None,
                             clear_ref: bool = False,
                             tend_out: dict | None = None,
                             accumulate: bool = False) -> dict:
    """J^T lam for one acoustic substep taken from the state's current
    small-step fields (coefficients prepared as for the forward substep,
    ``dycore.prepare_acoustic_coefficients``).  ``lam`` maps the outputs
    (ACOUSTIC_STATE, optionally ``ww_pp``) to adjoints (missing = 0);
    returns the adjoints of ACOUSTIC_STATE + ACOUSTIC_TEND.  ``ref``
    (from ``reference_accumulators``): also add the share of the stage
    reference state to those accumulators (dry states; map them with
    ``reference_gradients`` once the stage's substeps are done); with
    ``ref``, ``fwd`` may give the substep's outputs (ACOUSTIC_FWD names:
    the small-step fields and Omega'' after the forward substep), which
    are then not recomputed, ``lu`` the stage's Thomas factors
    (``adjoint_factors``), ``clear_ref`` makes this call write the
    accumulators instead of adding (a stage's first call), ``tend_out``
    gives the tensors the tendency adjoints go to and ``accumulate`` adds
    them there.  ``views``: tensors read in place of the state's arrays of
    those names (e.g. a checkpointed substep input).  The state is not
    modified."""
    _check_state(state)
    lib = _lib.load()
    pre = "arw64_" if state.dtype == torch.float64 else "arw_"
    nz, ny, nx = state.shape
    n = ctypes.c_int64()
    _lib.check(getattr(lib, pre + "acoustic_adjoint_work")(
        nz, ny, nx, ctypes.byref(n)), pre + "acoustic_adjoint_work")
    # the work area (about 30 fields) persists in the state's scratch pool
    work = state.scratch((n.value,), "arw_acoustic_adjoint_work")
    keep = []
    a = _lib.ArwAcousticAdj()
    for name in ACOUSTIC_STATE + ("ww_pp",):
        t = lam.get(name)
        if t is None:
            if name == "ww_pp":
                setattr(a, "l_ww_pp", None)
                continue
            t = torch.zeros_like(getattr(state, name))
        t = _as_like(t, getattr(state, name), name)
        keep.append(t)
        setattr(a, "l_" + name, t.data_ptr())
    out = {}
    for name in ACOUSTIC_STATE + ACOUSTIC_TEND:
        if tend_out is not None and name in ACOUSTIC_TEND:
            out[name] = _as_like(tend_out[name], getattr(state, name), name)
            if out[name] is not tend_out[name]:
                raise ValueError(f"tend_out[{name}] must be contiguous")
        else:
            out[name] = torch.empty_like(getattr(state, name))
        setattr(a, "g_" + name, out[name].data_ptr())
    d = _dom_views(state, cfg, views)
    if ref is None:
        if fwd is not None or lu is not None or clear_ref or accumulate:
            raise ValueError("fwd, lu, clear_ref and accumulate need ref")
        fn = getattr(lib, pre + "acoustic_substep_adjoint")
        _lib.check(fn(d, params(cfg), float(dtau), int(bool(first)),
                      ctypes.byref(a), work.data_ptr(), _stream()),
                   pre + "acoustic_substep_adjoint")
        return out
    r = _ref_struct(state, ref)
    fp = None
    if fwd is not None:
        fs = _lib.ArwAcousticFwd()
        for n in _lib.ACOUSTIC_FWD:
            t = fwd[n]
            like = getattr(state, n)
            if tuple(t.shape) != tuple(like.shape) or t.dtype != like.dtype \
                    or t.device != like.device or not t.is_contiguous():
                raise ValueError(f"substep output {n}: expected a contiguous "
                                 f"{tuple(like.shape)} {like.dtype} tensor")
            setattr(fs, n, t.data_ptr())
        fp = ctypes.byref(fs)
    if lu is not None:
        nz, ny, nx = state.shape
        if tuple(lu.shape) != (3, nz + 1, ny, nx) or lu.dtype != state.dtype \
                or not lu.is_contiguous():
            raise ValueError("lu: expected adjoint_factors' tensor")
    flags = (_lib.ADJ_CLEAR if clear_ref else 0) \
        | (_lib.ADJ_ACCUMULATE if accumulate else 0)
    fn = getattr(lib, pre + "acoustic_substep_adjoint_ref")
    _lib.check(fn(d, params(cfg), float(dtau), int(bool(first)),
                  ctypes.byref(a), ctypes.byref(r), fp,
                  None if lu is None else lu.data_ptr(), flags,
                  work.data_ptr(), _stream()),
               pre + "acoustic_substep_adjoint_ref")
    return out


def adjoint_factors(state: DomainState, cfg: RunConfig, dtau: float,
                    views: dict | None = None) -> torch.Tensor:
    """The stage's implicit-w Thomas factors (3, nz+1, ny, nx) the substep
    adjoints of a stage can share (they depend on the stage state, its EOS
    and dtau only)."""
    _check_state(state)
    lib = _lib.load()
    pre = "arw64_" if state.dtype == torch.float64 else "arw_"
    nz, ny, nx = state.shape
    lu = torch.empty((3, nz + 1, ny, nx), dtype=state.dtype,
                     device=state.u.device)
    _lib.check(getattr(lib, pre + "acoustic_adjoint_factors")(
        _dom_views(state, cfg, views), params(cfg), float(dtau),
        lu.data_ptr(), _stream()), pre + "acoustic_adjoint_factors")
    return lu


ACOUSTIC_REF = _lib.ACOUSTIC_REF


def _ref_shapes(state: DomainState) -> dict:
    nz, ny, nx = state.shape
    return {"mut": (ny, nx), "c2a": (nz, ny, nx), "alt": (nz, ny, nx),
            "thr": (nz, ny, nx), "ph": (nz + 1, ny, nx),
            "w": (nz + 1, ny, nx), "grf": (ny, nx)}


def reference_accumulators(state: DomainState) -> dict:
    """Zeroed accumulators of the substep adjoint's stage-reference
    share: mut, grf (ny,nx); c2a, alt, thr (mass points); ph, w (w
    levels)."""
    return {n: torch.zeros(shp, dtype=state.dtype, device=state.u.device)
            for n, shp in _ref_shapes(state).items()}


def _ref_struct(state, ref):
    want = _ref_shapes(state)
    r = _lib.ArwAcousticRef()
    for n in ACOUSTIC_REF:
        t = ref[n]
        if tuple(t.shape) != want[n] or t.dtype != state.dtype \
                or t.device != state.u.device or not t.is_contiguous():
            raise ValueError(f"reference accumulator {n}: expected a "
                             f"contiguous {want[n]} {state.dtype} tensor on "
                             f"{state.u.device}")
        setattr(r, n, t.data_ptr())
    return r


def _out_like(state, out, names, accumulate):
    """Output tensors for names: the caller's (to write or add to) or
    fresh ones."""
    if out is None:
        if accumulate:
            raise ValueError("accumulate needs out")
        return {n: torch.empty_like(getattr(state, n)) for n in names}
    res = {}
    for n in names:
        t = _as_like(out[n], getattr(state, n), n)
        if t is not out[n]:
            raise ValueError(f"out[{n}] must be contiguous")
        res[n] = t
    return res


def reference_gradients(state: DomainState, cfg: RunConfig, ref: dict,
                        views: dict | None = None, out: dict | None = None,
                        accumulate: bool = False) -> dict:
    """The stage-state gradients (STAGE_STATE names) of the substep
    adjoints accumulated in ``ref`` over one stage: dmdt_rf's share to u,
    v, ref's w to w, and c2a and alpha_t through the stage EOS (the
    state's p and alt, or those of ``views``, must hold it, as after
    ``dycore.stage_prepare``) to thp, php, mup.  Written to ``out`` (or
    fresh tensors), or added to ``out`` with ``accumulate``."""
    _check_state(state)
    lib = _lib.load()
    pre = "arw64_" if state.dtype == torch.float64 else "arw_"
    r = _ref_struct(state, ref)
    nz, ny, nx = state.shape
    work = state.scratch((2 * (nz + 1) * (ny + 1) * (nx + 1),),
                         "arw_reference_gradients_work")
    o = _out_like(state, out, STAGE_STATE + (("qv",) if _moist(state)
                                            else ()), accumulate)
    _lib.check(getattr(lib, pre + "acoustic_reference_gradients")(
        _dom_views(state, cfg, views), params(cfg), ctypes.byref(r),
        *[o[n].data_ptr() for n in STAGE_STATE],
        o["qv"].data_ptr() if _moist(state) else None,
        _lib.ADJ_ACCUMULATE if accumulate else 0, work.data_ptr(),
        _stream()), pre + "acoustic_reference_gradients")
    return o


def set_w_surface_adjoint(state: DomainState, cfg: RunConfig,
                          lam_w0: torch.Tensor, g_u: torch.Tensor,
                          g_v: torch.Tensor) -> None:
    """Add the (u, v) gradients of the kinematic lower boundary w(sfc) =
    u . grad(ht) (reference dycore.py:1999-2026) for the adjoint ``lam_w0``
    (ny, nx) of w(sfc) to ``g_u``, ``g_v`` in place."""
    _check_state(state)
    lib = _lib.load()
    pre = "arw64_" if state.dtype == torch.float64 else "arw_"
    l0 = _as_like(lam_w0, state.mup, "w(sfc)")
    for t, ref_t, n in ((g_u, state.u, "u"), (g_v, state.v, "v")):
        if tuple(t.shape) != tuple(ref_t.shape) or t.dtype != ref_t.dtype \
                or t.device != ref_t.device or not t.is_contiguous():
            raise ValueError(f"gradient of {n}: expected a contiguous "
                             f"{tuple(ref_t.shape)} {ref_t.dtype} tensor")
    _lib.check(getattr(lib, pre + "set_w_surface_adjoint")(
        _dom(state, cfg), params(cfg), l0.data_ptr(), g_u.data_ptr(),
        g_v.data_ptr(), _stream()), pre + "set_w_surface_adjoint")


def _as_like(t, ref, name):
    if tuple(t.shape) != tuple(ref.shape) or t.dtype != ref.dtype \
            or t.device != ref.device:
        raise ValueError(f"adjoint of {name}: expected {tuple(ref.shape)} "
                         f"{ref.dtype} on {ref.device}")
    return t.contiguous()


class AcousticSubstep(torch.autograd.Function):
    """(X, tendencies) -> (X_new, Omega'') for one acoustic substep, with
    the hand-written adjoint as its backward.  ``apply(state, cfg, dtau,
    first, *inputs)``: ``inputs`` are the 9 ACOUSTIC_STATE fields then the
    6 ACOUSTIC_TEND fields; the call writes them into the state (the
    substep runs in the state's arrays) and returns new tensors.  The
    stage coefficients must be prepared (prepare_acoustic_coefficients)."""

    @staticmethod
    def forward(ctx, state, cfg, dtau, first, *inputs):
        from .dycore import acoustic_substep
        names = ACOUSTIC_STATE + ACOUSTIC_TEND
        if len(inputs) != len(names):
            raise ValueError(f"expected {len(names)} input tensors")
        for name, x in zip(names, inputs):
            getattr(state, name).copy_(x)
        acoustic_substep(state, cfg, dtau, first)
        ctx.state, ctx.cfg, ctx.dtau, ctx.first = state, cfg, dtau, first
        ctx.save_for_backward(*[x.detach() for x in inputs])
        return tuple(getattr(state, n).clone()
                     for n in ACOUSTIC_STATE + ("ww_pp",))

    @staticmethod
    def backward(ctx, *grads):
        st = ctx.state
        names = ACOUSTIC_STATE + ACOUSTIC_TEND
        for name, x in zip(names, ctx.saved_tensors):
            getattr(st, name).copy_(x)
        lam = {n: g for n, g in zip(ACOUSTIC_STATE + ("ww_pp",), grads)
               if g is not None}
        out = acoustic_substep_adjoint(st, ctx.cfg, ctx.dtau, ctx.first, lam)
        return (None, None, None, None) + tuple(out[n] for n in names)


# ---------------------------------------------------------------------------
# Large-step tendency pass
# ---------------------------------------------------------------------------

STAGE_STATE = _lib.STAGE_STATE
STAGE_TEND = _lib.STAGE_TEND


def stage_tendencies(state: DomainState, cfg: RunConfig) -> dict:
    """The slow tendencies of one RK stage at the state (EOS, Omega,
    advection, PGF, buoyancy, rhs_ph; reference dycore.py:2378-2389) as
    new tensors.  Runs the stage preparation in the state (its
    diagnostics, Omega and small-step fields are overwritten)."""
    from .dycore import stage_prepare
    stage_prepare(state, cfg, cfg.dt / cfg.time_step_sound, True)
    return {n: getattr(state, n).clone() for n in STAGE_TEND}


SPECIES = ("qv", "qc", "qr")


def _moist(state):
    return state.qv is not None


def stage_tendencies_adjoint(state: DomainState, cfg: RunConfig,
                             lam: dict, out: dict | None = None,
                             accumulate: bool = False,
                             lam_flux: dict | None = None) -> dict:
    """J^T lam for the tendency pass at the state's stage fields: ``lam``
    maps STAGE_TEND names to adjoints (missing = 0); returns the
    gradients of STAGE_STATE, and of qv, qc, qr on a moist state (the
    moist EOS and buoyancy loading), written to ``out`` or fresh tensors,
    or added to ``out`` with ``accumulate``.  ``lam_flux``: adjoints of
    the stage fluxes ru, rv, ww handed on besides the tendencies (the
    scalars' sumflux).  Constant-K diffusion (km_opt = 1) on a flat base;
    plain form.  The state is not modified."""
    _check_state(state)
    lib = _lib.load()
    pre = "arw64_" if state.dtype == torch.float64 else "arw_"
    nz, ny, nx = state.shape
    n = ctypes.c_int64()
    _lib.check(getattr(lib, pre + "stage_adjoint_work")(
        nz, ny, nx, ctypes.byref(n)), pre + "stage_adjoint_work")
    work = state.scratch((n.value,), "arw_stage_adjoint_work")
    a = _lib.ArwStageAdj()
    keep = []
    for name in STAGE_TEND:
        t = lam.get(name)
        if t is None:
            t = torch.zeros_like(getattr(state, name))
        t = _as_like(t, getattr(state, name), name)
        keep.append(t)
        setattr(a, "l_" + name, t.data_ptr())
    names = STAGE_STATE + (SPECIES if _moist(state) else ())
    out = _out_like(state, out, names, accumulate)
    for name in names:
        setattr(a, "g_" + name, out[name].data_ptr())
    if lam_flux is not None:
        nz, ny, nx = state.shape
        want = {"ru": (nz, ny, nx + 1), "rv": (nz, ny + 1, nx),
                "ww": (nz + 1, ny, nx)}
        for n, shp in want.items():
            t = lam_flux[n]
            if tuple(t.shape) != shp or t.dtype != state.dtype \
                    or not t.is_contiguous():
                raise ValueError(f"lam_flux[{n}]: expected contiguous {shp}")
            keep.append(t)
            setattr(a, "l_" + n, t.data_ptr())
    a.flags = _lib.ADJ_ACCUMULATE if accumulate else 0
    fn = getattr(lib, pre + "stage_tendencies_adjoint")
    _lib.check(fn(_dom(state, cfg), params(cfg), ctypes.byref(a),
                  work.data_ptr(), _stream()),
               pre + "stage_tendencies_adjoint")
    return out


class StageTendencies(torch.autograd.Function):
    """(u, v, w, thp, php, mu') -> (ru_t, rv_t, rw_t, rth_t, rph_t) of
    one RK stage, with the hand-written adjoint as its backward.
    ``apply(state, cfg, *fields)`` writes the fields into the state and
    runs the stage preparation there."""

    @staticmethod
    def forward(ctx, state, cfg, *fields):
        if len(fields) != len(STAGE_STATE):
            raise ValueError(f"expected {len(STAGE_STATE)} fields")
        for name, x in zip(STAGE_STATE, fields):
            getattr(state, name).copy_(x)
        ctx.state, ctx.cfg = state, cfg
        ctx.save_for_backward(*[x.detach() for x in fields])
        t = stage_tendencies(state, cfg)
        return tuple(t[n] for n in STAGE_TEND)

    @staticmethod
    def backward(ctx, *grads):
        st = ctx.state
        for name, x in zip(STAGE_STATE, ctx.saved_tensors):
            getattr(st, name).copy_(x)
        lam = {n: (g if g is not None else torch.zeros_like(getattr(st, n)))
               for n, g in zip(STAGE_TEND, grads)}
        g = stage_tendencies_adjoint(st, ctx.cfg, lam)
        return (None, None) + tuple(g[n] for n in STAGE_STATE)


# ---------------------------------------------------------------------------
# Small-step initialisation and finish
# ---------------------------------------------------------------------------

INIT_OUT = _lib.INIT_OUT
FINISH_PP = _lib.FINISH_PP


def _small_step_work(state):
    lib = _lib.load()
    pre = "arw64_" if state.dtype == torch.float64 else "arw_"
    nz, ny, nx = state.shape
    n = ctypes.c_int64()
    _lib.check(getattr(lib, pre + "small_step_adjoint_work")(
        nz, ny, nx, ctypes.byref(n)), pre + "small_step_adjoint_work")
    return lib, pre, state.scratch((n.value,), "arw_small_step_adjoint_work")


def small_step_init_adjoint(state: DomainState, lam: dict,
                            out: dict | None = None,
                            accumulate: bool = False) -> dict:
    """J^T lam of the small-step initialisation (reference dycore.cu:
    947-1085) at the state's time-t fields (u0..mup0) and stage fields
    (u..mup): ``lam`` maps INIT_OUT names to adjoints (missing = 0);
    returns gradients named u0..mup0 and u..mup (the latter including the
    stage EOS's share), written to ``out`` or fresh tensors, or added to
    ``out`` with ``accumulate``.  Dry, plain form; the state is not
    modified."""
    _check_state(state)
    lib, pre, work = _small_step_work(state)
    a = _lib.ArwInitAdj()
    keep = []
    for name in INIT_OUT:
        t = lam.get(name)
        if t is None:
            t = torch.zeros_like(getattr(state, name))
        t = _as_like(t, getattr(state, name), name)
        keep.append(t)
        setattr(a, "l_" + name, t.data_ptr())
    names = [n for name in STAGE_STATE for n in (name + "0", name)]
    if _moist(state):
        names.append("qv")
    if out is None:
        if accumulate:
            raise ValueError("accumulate needs out")
        out = {n: torch.empty_like(getattr(state, n.rstrip("0")))
               for n in names}
    else:
        out = {n: _as_like(out[n], getattr(state, n.rstrip("0")), n)
               if out[n].is_contiguous() else None for n in names}
        if any(t is None for t in out.values()):
            raise ValueError("out tensors must be contiguous")
    for n in names:
        setattr(a, "g_" + n, out[n].data_ptr())
    a.flags = _lib.ADJ_ACCUMULATE if accumulate else 0
    _lib.check(getattr(lib, pre + "small_step_init_adjoint")(
        state.domain_struct(), ctypes.byref(a), work.data_ptr(), _stream()),
        pre + "small_step_init_adjoint")
    return out


def small_step_finish_adjoint(state: DomainState, lam: dict,
                              views: dict | None = None) -> dict:
    """J^T lam of the small-step finish (reference dycore.cu:1087-1198) at
    the state's stage fields and small-step fields (or the tensors of
    ``views`` in their place): ``lam`` maps the new stage fields
    (STAGE_STATE names) to adjoints; returns gradients of STAGE_STATE and
    FINISH_PP."""
    _check_state(state)
    lib, pre, work = _small_step_work(state)
    a = _lib.ArwFinishAdj()
    keep = []
    for name in STAGE_STATE:
        t = lam.get(name)
        if t is None:
            t = torch.zeros_like(getattr(state, name))
        t = _as_like(t, getattr(state, name), name)
        keep.append(t)
        setattr(a, "l_" + name, t.data_ptr())
    out = {}
    for name in STAGE_STATE + FINISH_PP:
        out[name] = torch.empty_like(getattr(state, name))
        setattr(a, "g_" + name, out[name].data_ptr())
    d = state.domain_struct()
    for n, t in (views or {}).items():
        _as_like(t, getattr(state, n), n)
        setattr(d, n, t.data_ptr())
    _lib.check(getattr(lib, pre + "small_step_finish_adjoint")(
        d, ctypes.byref(a), work.data_ptr(), _stream()),
        pre + "small_step_finish_adjoint")
    return out


# ---------------------------------------------------------------------------
# One whole step
# ---------------------------------------------------------------------------

def _step_stages(cfg: RunConfig):
    """(substeps, dtau) of the three RK stages (reference dycore.py:
    2290-2300)."""
    ns = cfg.time_step_sound
    return ((1, cfg.dt / 3.0), (max(ns // 2, 1), cfg.dt / ns),
            (ns, cfg.dt / ns))


#: what a stage's preparation leaves that its substeps read
_PREP = ("p", "al", "alt", "ru_t", "rv_t", "rw_t", "rth_t", "rph_t", "rmu_t")


class StepTape:
    """What the reverse sweep of one step needs, recorded by ``step_taped``:
    per RK stage the stage state and, with ``keep="substeps"``, what its
    preparation leaves (EOS, slow tendencies, Omega, dmdt_rf) and every
    substep's small-step fields and Omega'' (``keep="stages"`` keeps only
    the stage states, and the reverse sweep recomputes each stage from
    its checkpoint); and the step's output fields."""

    def __init__(self, cfg: RunConfig, keep: str):
        self.cfg = cfg
        self.keep = keep
        self.stages = []
        self.out = None


def _check_adjoint_cfg(state, cfg):
    if getattr(cfg, "precision_form", False):
        raise _lib.ArwenError("step adjoint: plain form only")
    if getattr(cfg, "diff_6th_opt", 0) or getattr(cfg, "km_opt", 1) == 4:
        raise _lib.ArwenError("step adjoint: no held mixing tendencies "
                              "(diff_6th_opt = 0, km_opt != 4)")
    if getattr(cfg, "mp_physics", 0) != 0:
        raise _lib.ArwenError("step adjoint: qv/qc/qr states only "
                              "(mp_physics = 0, no h_diabatic)")
    if getattr(cfg, "h_sca_adv_order", 2) != 2 \
            or getattr(cfg, "hypsometric_opt", 1) != 1 \
            or getattr(cfg, "moist_cq", False):
        raise _lib.ArwenError("step adjoint: h_sca_adv_order = 2, "
                              "hypsometric_opt = 1, moist_cq off only")
    if _uses_diffusion(cfg) and state.thb.dim() == 3:
        raise _lib.ArwenError("step adjoint: constant-K diffusion needs a "
                              "flat base")


def _snapshot(tensors: dict) -> dict:
    """Copies of the named tensors, made by one batched copy launch."""
    out = {n: torch.empty_like(t) for n, t in tensors.items()}
    torch._foreach_copy_(list(out.values()), list(tensors.values()))
    return out


def _restore(state, names, src: dict, suffix: str = "") -> None:
    """state.<name><suffix> = src[name] for every name, one batched copy."""
    torch._foreach_copy_([getattr(state, n + suffix) for n in names],
                         [src[n] for n in names])


def _run_stage(state, cfg, nsub, dtau, record, sums=True):
    """One stage's preparation and substeps from the state's stage and
    time-t fields; with ``record``, returns what the stage's adjoint reads
    (the preparation's outputs and every substep's small-step fields).  On
    a moist state, with ``sums``, also the stage fluxes and the substep sums
    of u'', v'', Omega'' (sumflux; the reverse sweep's recompute has them
    on the tape already)."""
    from . import dycore as dyn
    nz, ny, nx = state.shape
    dyn.stage_prepare(state, cfg, dtau, True)
    prep = subs = None
    flx = None
    if sums and _moist(state):
        ru, rv, ww = dyn.stage_fluxes(state, cfg)
        flx = _snapshot({"ru": ru, "rv": rv, "ww": ww})
    if record:
        src = {n: getattr(state, n) for n in _PREP}
        src["ww"] = state.scratch((nz + 1, ny, nx), "arw_ww")
        src["dmdt_rf"] = state.scratch((ny, nx), "arw_dmdt_rf")
        src.update({"sub0_" + n: getattr(state, n) for n in ACOUSTIC_STATE})
        snap = _snapshot(src)
        subs = [{n: snap.pop("sub0_" + n) for n in ACOUSTIC_STATE}]
        prep = snap
    for i in range(nsub):
        dyn.acoustic_substep(state, cfg, dtau, first=(i == 0))
        if flx is not None:
            # the substep sums of u'', v'', Omega'' (the first one copies)
            pp = [state.u_pp, state.v_pp, state.ww_pp]
            if i == 0:                   # 0 + x, as the sums start at zero
                flx.update(zip(("su", "sv", "sw"),
                               torch._foreach_add(pp, 0.0)))
            else:
                torch._foreach_add_([flx["su"], flx["sv"], flx["sw"]], pp)
        if record:
            src = {n: getattr(state, n) for n in ACOUSTIC_STATE}
            src["ww_pp"] = state.ww_pp
            subs.append(_snapshot(src))
    return prep, subs, flx


def scalar_stage(state: DomainState, cfg: RunConfig, dt_stage: float,
                 clamp: bool, q, q0, fl, mup_t, mup_new) -> torch.Tensor:
    """One stage of a scalar's unlimited update with explicit fluxes fl =
    (ru, rv, ww) and masses (``arw_scalar_stage``)."""
    lib = _lib.load()
    pre = "arw64_" if state.dtype == torch.float64 else "arw_"
    q_new = torch.empty_like(q)
    _lib.check(getattr(lib, pre + "scalar_stage")(
        state.domain_struct(), params(cfg), float(dt_stage), int(bool(clamp)),
        q.data_ptr(), q0.data_ptr(), *[t.data_ptr() for t in fl],
        mup_t.data_ptr(), mup_new.data_ptr(), q_new.data_ptr(), _stream()),
        pre + "scalar_stage")
    return q_new


def scalar_stage_adjoint(state: DomainState, cfg: RunConfig,
                         dt_stage: float, clamp: bool, q, q0, fl, mup_t,
                         mup_new, lam, g_q0, g_fl, g_mup_t, g_mup_new):
    """Transpose of ``scalar_stage`` for lam: returns the stage scalar's
    gradient; adds to g_q0, g_fl (ru, rv, ww), g_mup_t, g_mup_new."""
    lib = _lib.load()
    pre = "arw64_" if state.dtype == torch.float64 else "arw_"
    g_q = torch.empty_like(q)
    work = state.scratch((3 * q.numel(),), "arw_scalar_adjoint_work")
    _lib.check(getattr(lib, pre + "scalar_stage_adjoint")(
        state.domain_struct(), params(cfg), float(dt_stage), int(bool(clamp)),
        q.data_ptr(), q0.data_ptr(), *[t.data_ptr() for t in fl],
        mup_t.data_ptr(), mup_new.data_ptr(), lam.data_ptr(), g_q.data_ptr(),
        g_q0.data_ptr(), *[t.data_ptr() for t in g_fl], g_mup_t.data_ptr(),
        g_mup_new.data_ptr(), work.data_ptr(), _stream()),
        pre + "scalar_stage_adjoint")
    return g_q


def _pd_work(state):
    nz, ny, nx = state.shape
    return state.scratch((_lib.pd_work_elems(nz, ny, nx),), "arw_pd_work")


def pd_stage(state: DomainState, cfg: RunConfig, dt_stage: float, q, q0,
             fl, mup_t, mup_new) -> torch.Tensor:
    """The final stage's update of a scalar with the positive-definite
    limiter, explicit fluxes fl = (ru, rv, ww) and masses
    (``arw_pd_stage``; the step's moist_adv_opt = 1 arithmetic)."""
    lib = _lib.load()
    pre = "arw64_" if state.dtype == torch.float64 else "arw_"
    q_new = torch.empty_like(q)
    _lib.check(getattr(lib, pre + "pd_stage")(
        state.domain_struct(), params(cfg), float(dt_stage), q.data_ptr(),
        q0.data_ptr(), *[t.data_ptr() for t in fl], mup_t.data_ptr(),
        mup_new.data_ptr(), q_new.data_ptr(), _pd_work(state).data_ptr(),
        _stream()), pre + "pd_stage")
    return q_new


def pd_stage_adjoint(state: DomainState, cfg: RunConfig, dt_stage: float,
                     q, q0, fl, mup_t, mup_new, lam, g_q0, g_fl, g_mup_t,
                     g_mup_new):
    """Transpose of ``pd_stage`` for lam: returns the stage scalar's
    gradient; adds to g_q0, g_fl (ru, rv, ww), g_mup_t, g_mup_new."""
    lib = _lib.load()
    pre = "arw64_" if state.dtype == torch.float64 else "arw_"
    g_q = torch.empty_like(q)
    _lib.check(getattr(lib, pre + "pd_stage_adjoint")(
        state.domain_struct(), params(cfg), float(dt_stage), q.data_ptr(),
        q0.data_ptr(), *[t.data_ptr() for t in fl], mup_t.data_ptr(),
        mup_new.data_ptr(), lam.data_ptr(), g_q.data_ptr(), g_q0.data_ptr(),
        *[t.data_ptr() for t in g_fl], g_mup_t.data_ptr(),
        g_mup_new.data_ptr(), _pd_work(state).data_ptr(), _stream()),
        pre + "pd_stage_adjoint")
    return g_q


def _limited(cfg, ex):
    return ex["final"] and getattr(cfg, "moist_adv_opt", 1) == 1


def step_taped(state: DomainState, cfg: RunConfig,
               keep: str = "substeps") -> StepTape:
    """One dycore step of the state (as ``dycore.step``; reference
    dycore.py:2161-2503), run stage by stage through the piecewise entry
    points so that it records the tape its adjoint needs.  ``keep``:
    "substeps" records every stage's preparation and substep states (the
    faster reverse sweep; at 1000x1000x50 FP32 about 30 GB), "stages" only
    the three stage states (3.6 GB there; the reverse sweep recomputes each
    stage from its checkpoint).  Plain form; constant-K diffusion on a flat
    base; moist states with or without the final stage's PD limiter."""
    from . import dycore as dyn
    _check_state(state)
    _check_adjoint_cfg(state, cfg)
    if keep not in ("substeps", "stages"):
        raise ValueError("keep must be 'substeps' or 'stages'")
    tape = StepTape(cfg, keep)
    names = _step_fields(state)
    _restore(state, names, {n: getattr(state, n) for n in names}, "0")
    for istage, (nsub, dtau) in enumerate(_step_stages(cfg)):
        xs = _snapshot({n: getattr(state, n) for n in names})
        prep, subs, flx = _run_stage(state, cfg, nsub, dtau,
                                     keep == "substeps")
        dyn.small_step_finish(state)
        ex = None
        if flx is not None:
            # sumflux, then the scalars' stage update from the stage's
            # scalars and the time-t ones
            fl = tuple(torch._foreach_add(
                torch._foreach_div([flx["su"], flx["sv"], flx["sw"]],
                                   float(nsub)),
                [flx["ru"], flx["rv"], flx["ww"]]))
            ex = {"fl": fl, "mu_new": state.mup.clone(),
                  "dt": nsub * dtau, "final": istage == 2}
            for q in SPECIES:
                if _limited(cfg, ex):
                    q_new = pd_stage(state, cfg, ex["dt"], getattr(state, q),
                                     getattr(state, q + "0"), fl,
                                     state.mup0, state.mup)
                else:
                    q_new = scalar_stage(state, cfg, ex["dt"], ex["final"],
                                         getattr(state, q),
                                         getattr(state, q + "0"), fl,
                                         state.mup0, state.mup)
                getattr(state, q).copy_(q_new)
        tape.stages.append((xs, prep, subs, nsub, dtau, ex))
    dyn.set_w_surface(state, cfg)
    dyn.update_diagnostics(state)
    state.elapsed_seconds += cfg.dt
    tape.out = _snapshot({n: getattr(state, n) for n in names})
    return tape


def _step_fields(state):
    return STAGE_STATE + (SPECIES if _moist(state) else ())


def step_adjoint_taped(state: DomainState, tape: StepTape,
                       lam: dict) -> dict:
    """The reverse sweep of one step: J^T lam from the tape ``step_taped``
    recorded (``lam`` maps the step's output STAGE_STATE fields to
    adjoints, missing = 0; returns the gradients of its input fields).
    Per stage, last first (recomputing the stage's preparation and
    substeps from its checkpoint when the tape kept only stage states):
    the finish adjoint, the substep adjoints in
    reverse order (given each substep's outputs, adding the stage-reference
    share), the small-step initialisation adjoint (to the time-t and stage
    fields) and the tendency-pass adjoint; the stages chain through their
    states and the kinematic lower boundary of the step's end comes first.
    The state is the workspace: its prognostic fields end as the step's
    output, its other fields are overwritten."""
    _check_state(state)
    cfg = tape.cfg
    _check_adjoint_cfg(state, cfg)
    names = _step_fields(state)
    moist = _moist(state)
    lx = {}
    for n in names:
        t = lam.get(n)
        lx[n] = (torch.zeros_like(getattr(state, n)) if t is None
                 else _as_like(t, getattr(state, n), n).clone())
    set_w_surface_adjoint(state, cfg, lx["w"][0], lx["u"], lx["v"])
    lx["w"][0].zero_()
    x0 = tape.stages[0][0]                     # the step's input
    _restore(state, names, x0, "0")
    g0 = {n: torch.zeros_like(getattr(state, n)) for n in names}
    for xs, prep, subs, nsub, dtau, ex in reversed(tape.stages):
        _restore(state, names, xs)
        if prep is None:                 # recompute the stage's tape
            prep, subs, _ = _run_stage(state, cfg, nsub, dtau, True,
                                       sums=False)
        gq, inj = {}, None
        if moist:
            # the scalars' stage update (last in the stage): its adjoint
            # feeds the finish's new mu', the stage fluxes and, through
            # sumflux, every substep's output u'', v'', Omega''
            lf = (torch.zeros_like(ex["fl"][0]), torch.zeros_like(ex["fl"][1]),
                  torch.zeros_like(ex["fl"][2]))
            g_mu = torch.zeros_like(state.mup)
            for q in SPECIES:
                if _limited(cfg, ex):
                    gq[q] = pd_stage_adjoint(
                        state, cfg, ex["dt"], xs[q], x0[q], ex["fl"],
                        x0["mup"], ex["mu_new"], lx[q], g0[q], lf,
                        g0["mup"], g_mu)
                else:
                    gq[q] = scalar_stage_adjoint(
                        state, cfg, ex["dt"], ex["final"], xs[q], x0[q],
                        ex["fl"], x0["mup"], ex["mu_new"], lx[q], g0[q], lf,
                        g0["mup"], g_mu)
            lx["mup"] = lx["mup"] + g_mu
            inj = [t / float(nsub) for t in lf]
            lam_flux = {"ru": lf[0], "rv": lf[1], "ww": lf[2]}
        fin = small_step_finish_adjoint(
            state, {n: lx[n] for n in STAGE_STATE},
            views={n: subs[nsub][n] for n in FINISH_PP})
        gs = {n: fin[n] for n in STAGE_STATE}
        gs.update(gq)
        lp = {n: fin[n] for n in FINISH_PP}
        # the stage's substep adjoints write (first) then add to one set of
        # tendency adjoints and reference accumulators, and share the
        # stage's Thomas factors
        lt = {n: torch.empty_like(getattr(state, n)) for n in ACOUSTIC_TEND}
        acc = {n: torch.empty(shp, dtype=state.dtype, device=state.u.device)
               for n, shp in _ref_shapes(state).items()}
        lu = adjoint_factors(state, cfg, dtau, views=prep)
        for i in range(nsub - 1, -1, -1):
            views = dict(prep)
            views.update(subs[i])
            views.pop("ww_pp", None)
            fwd = {n: subs[i + 1][n] for n in _lib.ACOUSTIC_FWD}
            last = i == nsub - 1
            if inj is not None:
                lp["u_pp"] = lp["u_pp"] + inj[0]
                lp["v_pp"] = lp["v_pp"] + inj[1]
                lp["ww_pp"] = inj[2]
            out = acoustic_substep_adjoint(
                state, cfg, dtau, i == 0, lp, acc, fwd=fwd, views=views,
                lu=lu, clear_ref=last, tend_out=lt, accumulate=not last)
            lp = {n: out[n] for n in ACOUSTIC_STATE}
        # every share adds into the stage's gradients gs (the finish's)
        # and the time-t gradients g0
        reference_gradients(state, cfg, acc, views=prep, out=gs,
                            accumulate=True)
        gio = {n: gs[n] for n in STAGE_STATE + (("qv",) if moist else ())}
        gio.update({n + "0": g0[n] for n in STAGE_STATE})
        small_step_init_adjoint(state, lp, out=gio, accumulate=True)
        stage_tendencies_adjoint(state, cfg, lt, out=gs, accumulate=True,
                                 lam_flux=lam_flux if moist else None)
        lx = gs
    torch._foreach_add_([g0[n] for n in names], [lx[n] for n in names])
    _restore(state, names, tape.out)
    return g0


def step_adjoint(state: DomainState, cfg: RunConfig, lam: dict,
                 keep: str = "substeps") -> dict:
    """J^T lam of one dycore step (``dycore.step``; reference dycore.py:
    2161-2503) at the state's prognostic fields: ``lam`` maps the new
    state's STAGE_STATE fields (and qv, qc, qr on a moist state) to
    adjoints (missing = 0); returns the gradients of the step's input
    fields.  Plain form; constant-K diffusion on a flat base; moist states
    with or without the final stage's PD limiter (moist_adv_opt 1 or 0,
    each branch of the limiter differentiated as taken).  ``step_taped`` then
    ``step_adjoint_taped``; the state's prognostic fields are left as they
    were (their EOS recomputed), its other fields are overwritten."""
    from . import dycore as dyn
    _check_state(state)
    _check_adjoint_cfg(state, cfg)
    x_in = {n: getattr(state, n).clone() for n in _step_fields(state)}
    t0 = state.elapsed_seconds
    tape = step_taped(state, cfg, keep)
    g = step_adjoint_taped(state, tape, lam)
    for n, t in x_in.items():
        getattr(state, n).copy_(t)
    state.elapsed_seconds = t0
    dyn.update_diagnostics(state)
    return g


class DycoreStep(torch.autograd.Function):
    """(u, v, w, thp, php, mu'[, qv, qc, qr]) -> the same fields one dycore
    step later (the scalars on a moist state),
    with the hand-written adjoint as its backward.  ``apply(state, cfg,
    *fields)`` (or ``apply(state, cfg, keep, *fields)``, keep as for
    ``step_taped``) writes the fields into the state and steps it there
    with ``step_taped``; the backward is ``step_adjoint_taped`` on that
    tape (the state is its workspace)."""

    @staticmethod
    def forward(ctx, state, cfg, *fields):
        ctx.keep_arg = bool(fields) and isinstance(fields[0], str)
        keep = fields[0] if ctx.keep_arg else "substeps"
        if ctx.keep_arg:
            fields = fields[1:]
        names = _step_fields(state)
        if len(fields) != len(names):
            raise ValueError(f"expected {len(names)} fields {names}")
        for name, x in zip(names, fields):
            getattr(state, name).copy_(x)
        ctx.state, ctx.names = state, names
        ctx.tape = step_taped(state, cfg, keep)
        return tuple(getattr(state, n).clone() for n in names)

    @staticmethod
    def backward(ctx, *grads):
        lam = {n: g for n, g in zip(ctx.names, grads) if g is not None}
        g = step_adjoint_taped(ctx.state, ctx.tape, lam)
        ctx.tape = None
        lead = (None, None, None) if ctx.keep_arg else (None, None)
        return lead + tuple(g[n] for n in ctx.names)
