"""Batched ensemble: many independent domains advanced together.

The reference runs ensemble members one after another, each with its own
launches (gpuwm/ensemble/engine.py:1-14).  On MI355X a member of
100-250 columns a side fills a fraction of the 256 CUs, and its step is
dominated by launch gaps.  ``Ensemble`` instead issues every pass of the
RK3 step once for all members: each workgroup of a launch looks up its
member and its workgroup within that member in a device work table
(``arw_ensemble_*``), so one launch fills the GPU with columns of all
members.  ``capture`` records N such steps as one HIP graph; ``launch``
replays it.  Each member's arithmetic is that of ``dycore.step`` -- the
results are bit-identical to stepping the members one by one.

``autotune`` picks the launch layout -- the member order in the work
tables and the number of member groups that run as parallel graph
branches -- by timing the candidates on the ensemble itself (state saved
and restored around the trials).
"""
from __future__ import annotations

import ctypes
import sys
import time

import torch

from . import _lib
from .config import RunConfig, validate_for_step
from .dycore import _check_state, _uses_diffusion, params
from .state import DomainState

#: RunConfig fields every member must share (they enter arw_params)
SHARED = ("dx", "dy", "dt", "time_step_sound", "epssm", "smdiv", "km_opt",
          "khdif", "kvdif", "damp_opt", "zdamp", "dampcoef", "top_lid",
          "precision_form")

#: prognostic and diagnostic arrays that carry a member's state between
#: steps (everything else is rebuilt inside the step)
STATE_ARRAYS = ("u", "v", "w", "thp", "php", "mup", "p", "al", "alt")


def _prefix(states):
    return "arw64_" if states[0].dtype == torch.float64 else "arw_"


class Ensemble:
    """Device work tables for ``states`` stepped with ``cfg``'s step
    parameters (members may differ in nx, ny, nz)."""

    def __init__(self, states: list[DomainState], cfg: RunConfig,
                 order: list[int] | None = None, groups: int = 1,
                 join_every: int = 0, col_tile: int = 256):
        if not states:
            raise ValueError("an ensemble needs at least one member")
        validate_for_step(cfg, acoustic=True)
        dt0 = states[0].dtype
        for i, st in enumerate(states):
            _check_state(st)
            if st.dtype != dt0:
                raise ValueError("members must share one dtype")
            if st.qv is not None:
                raise NotImplementedError("ensemble members must be dry")
            for f in SHARED:
                if getattr(st.cfg, f) != getattr(cfg, f):
                    raise ValueError(f"member {i}: cfg.{f} differs from the "
                                     "ensemble's step parameters")
        self.states = states
        self.cfg = cfg
        self.order = list(order) if order is not None else None
        self.groups = int(groups)
        self.join_every = int(join_every)
        self.col_tile = int(col_tile)
        n = len(states)
        Dom = _lib.ArwDomain64 if dt0 == torch.float64 else _lib.ArwDomain
        arr = (Dom * n)()
        for i, st in enumerate(states):
            arr[i] = st.domain_struct(diffusion=_uses_diffusion(cfg),
                                      precision_form=bool(cfg.precision_form))
        self._doms = arr
        ordp = None
        if self.order is not None:
            if sorted(self.order) != list(range(n)):
                raise ValueError("order must be a permutation of the members")
            self._ord = (ctypes.c_int32 * n)(*self.order)
            ordp = ctypes.cast(self._ord, ctypes.c_void_p)
        self._pre = _prefix(states)
        lib = _lib.load()
        h = ctypes.c_void_p()
        fn = getattr(lib, self._pre + "ensemble_create")
        _lib.check(fn(arr, n, params(cfg), ordp, self.groups, self.col_tile,
                      ctypes.byref(h)), self._pre + "ensemble_create")
        self._h = h
        self.graph_steps = 0

    def _fn(self, name):
        if self._h is None:
            raise RuntimeError("ensemble is closed")
        return self._pre + name, getattr(_lib.load(), self._pre + name)

    def step(self, n: int = 1) -> None:
        """n steps of every member, launched on the current stream."""
        name, fn = self._fn("ensemble_step")
        _lib.check(fn(self._h, int(n), torch.cuda.current_stream().cuda_stream),
                   name)
        for st in self.states:
            st.elapsed_seconds += n * self.cfg.dt

    def capture(self, nsteps: int, join_every: int | None = None) -> None:
        """Record nsteps steps as one HIP graph (nothing runs yet).  With
        several groups the parallel branches rejoin every ``join_every``
        steps (default: the layout's own, else only at the end)."""
        if join_every is None:
            join_every = self.join_every
        name, fn = self._fn("ensemble_capture")
        _lib.check(fn(self._h, int(nsteps), int(join_every)), name)
        self.graph_steps = int(nsteps)

    def launch(self) -> None:
        """Run the captured graph on the current stream."""
        name, fn = self._fn("ensemble_launch")
        _lib.check(fn(self._h, torch.cuda.current_stream().cuda_stream), name)
        for st in self.states:
            st.elapsed_seconds += self.graph_steps * self.cfg.dt

    def close(self) -> None:
        if self._h is not None:
            torch.cuda.synchronize()
            name, fn = self._fn("ensemble_destroy")
            _lib.check(fn(self._h), name)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sequential_steps(states: list[DomainState], cfg: RunConfig,
                     n: int) -> None:
    """The reference's launch model: each member advanced n steps with
    the single-domain step, one member after another."""
    from .dycore import run_steps
    for st in states:
        run_steps(st, cfg, n)


def save_state(states):
    return [({a: getattr(st, a).clone() for a in STATE_ARRAYS},
             st.elapsed_seconds) for st in states]


def restore_state(states, saved):
    for st, (sv, el) in zip(states, saved):
        for a, t in sv.items():
            getattr(st, a).copy_(t)
        st.elapsed_seconds = el


def member_orders(states) -> dict[str, list[int]]:
    """Candidate work-table orders: as given, largest member first,
    smallest first."""
    size = [st.u.numel() for st in states]
    idx = list(range(len(states)))
    return {"given": idx,
            "largest_first": sorted(idx, key=lambda i: -size[i]),
            "smallest_first": sorted(idx, key=lambda i: size[i])}


_TUNED: dict = {}


def _time_layout(states, cfg, saved, order, groups, join_every, steps, reps,
                 col_tile=256):
    with Ensemble(states, cfg, order=order, groups=groups,
                  join_every=join_every, col_tile=col_tile) as ens:
        ens.capture(steps)
        ens.launch()                               # warm-up replay
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ens.launch()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / (reps * steps)
    restore_state(states, saved)
    return ms


def autotune(states: list[DomainState], cfg: RunConfig, *, steps: int = 10,
             reps: int = 2, groups=(1, 2, 4, 8), joins=(1, 5, 0),
             col_tiles=(256, 128, 64), orders=None,
             verbose: bool = False) -> dict:
    """Pick the launch layout by timing it on the ensemble itself: first
    the member order of the work tables (one group), then the number of
    groups that run as parallel graph branches and how often the branches
    rejoin (the substep schedule across the branches), then the column
    solve's tile (columns per workgroup) for that layout.  Returns
    {"order", "order_name", "groups", "join_every", "col_tile",
    "ms_per_step", "trials"}; the members' state is restored afterwards
    and the result is cached per member shapes and step parameters."""
    key = (tuple(st.shape for st in states), states[0].dtype,
           tuple(getattr(cfg, f) for f in SHARED), tuple(groups),
           tuple(joins), tuple(col_tiles), steps)
    if key in _TUNED:
        return _TUNED[key]
    saved = save_state(states)
    cands = member_orders(states)
    if orders is not None:
        cands = {k: v for k, v in cands.items() if k in orders}
    trials = []

    def run(oname, g, j, t=256):
        ms = _time_layout(states, cfg, saved, cands[oname], g, j, steps,
                          reps, t)
        trials.append({"order": oname, "groups": g, "join_every": j,
                       "col_tile": t, "ms_per_step": ms})
        if verbose:
            print(f"autotune order={oname} groups={g} join_every={j} "
                  f"col_tile={t}: {ms:.3f} ms/step", file=sys.stderr,
                  flush=True)
        return ms

    best_order = min(cands, key=lambda o: run(o, 1, 0))
    for g in groups:
        if g == 1 or g > len(states):
            continue
        for j in joins:
            run(best_order, g, j)
    lay = min(trials, key=lambda t: t["ms_per_step"])
    for t in col_tiles:
        if t != 256:
            run(lay["order"], lay["groups"], lay["join_every"], t)
    best = min(trials, key=lambda t: t["ms_per_step"])
    # a trial graph of `steps` steps with join_every = 0 rejoins at its end,
    # i.e. every `steps` steps: that is what was measured, and what a longer
    # capture should do (never rejoining over 100 steps is slower)
    join = best["join_every"] if best["join_every"] > 0 else steps
    out = {"order": cands[best["order"]], "order_name": best["order"],
           "groups": best["groups"], "join_every": join,
           "col_tile": best["col_tile"], "ms_per_step": best["ms_per_step"],
           "trials": trials}
    _TUNED[key] = out
    return out
