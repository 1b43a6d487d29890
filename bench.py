#!/usr/bin/env python3
"""Benchmark of the MI355X dycore step (one JSON line on stdout).

Workload (default): the plain FP32 RK3 split-explicit step on a
1000x1000x50 dry stratified domain (dx = 1 km, dt = 6 s, 4 acoustic
substeps; seeded +3 K warm bubbles), the grid the throughput/roofline
target is stated on.  Metric: Mpoint-steps/s (mass points x steps / s).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c2]

The timed region runs K steps with the state resident in HBM, bracketed
by torch.cuda.synchronize().  Per-kernel durations over the same region
come from HIP events the library records on its launch stream.  The CPU
baseline (rank 0 only) times the float32 NumPy oracle on a bounded strip
of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0

#: Algorithmic FP32 words per mass point of one pass of the algorithm
#: (each distinct 3-D input read once, each output written once, one HBM
#: pass per required global synchronisation): stage pass S (19 words on
#: stage 1, 24 on stages 2-3), acoustic substep pass U (24), finish F (15).
WORDS_S1, WORDS_S23, WORDS_U, WORDS_F = 19, 24, 24, 15


def step_words(ns: int) -> int:
    """Algorithmic words per point-step: 67 + 24 (1 + ns/2 + ns) + 45."""
    return WORDS_S1 + 2 * WORDS_S23 + WORDS_U * (1 + ns // 2 + ns) \
        + 3 * WORDS_F


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=("c3", "c2", "c5", "c3-precision",
                                           "c4-adjoint"),
                    default="c3",
                    help="c5: the batched 32-member ensemble against the "
                         "same members stepped one after another; "
                         "c3-precision: FP32 and the precision form against "
                         "the FP64 build after --prec-steps steps")
    ap.add_argument("--prec-steps", type=int, default=600)
    ap.add_argument("--ens-steps", type=int, default=100,
                    help="c5: steps captured in the one graph launch")
    ap.add_argument("--ens-members", type=int, default=32)
    ap.add_argument("--no-fp64", action="store_true",
                    help="c5: skip the FP64 error reference")
    ap.add_argument("--diff6", type=int, default=0, choices=(0, 1, 2),
                    help="c3/c2: run with WRF's held 6th-order filter "
                         "(diff_6th_opt; factor 0.12) to time its cost")
    ap.add_argument("--km-opt", type=int, default=1, choices=(1, 4),
                    help="c3/c2: 4 runs with WRF's held 2-D Smagorinsky "
                         "mixing (km_opt=4, c_s=0.25) to time its cost")
    ap.add_argument("--species", type=int, default=-1,
                    choices=(-1, 0, 6, 8, 10, 18, 28),
                    help="c3/c2: run moist, carrying the species of this "
                         "mp_physics scheme (and, for mp > 0, a retained "
                         "heating); a throughput line for the scalar "
                         "family, without extras or CPU baseline")
    ap.add_argument("--no-extras", action="store_true",
                    help="c3/c2: leave out the compact adjoint (C4) and "
                         "ensemble (C5) objects measured after the timed "
                         "region")
    ap.add_argument("--extra-ens-steps", type=int, default=20,
                    help="steps of the compact C5 ensemble comparison")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=160,
                    help="y rows of the CPU-baseline strip")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also launch one 1 GiB device copy (for PMC "
                         "byte-counter calibration under rocprofv3)")
    ap.add_argument("--traffic-file", default=None,
                    help="PMC summary giving HBM bytes per launch (default "
                         "for c3: the committed rocprofv3 summary)")
    return ap.parse_args()


def build_case(workload: str, ny_override: int | None = None):
    from arwen_amd import cases
    if workload == "c3":
        kw = {} if ny_override is None else {"ny": ny_override}
        return cases.warm_bubbles(**kw), (
            "dry stratified 1000x1000x50, dx=1 km, dt=6 s, 4 acoustic "
            "substeps, 16 seeded +3 K warm bubbles, u=5 v=2 m/s")
    kw = {} if ny_override is None else {"ny": ny_override}
    return cases.stratified_hill(**kw), (
        "3-D dry stratified flow 250x200x49 over a 1 km Gaussian hill, "
        "dx=12 km, dt=60 s, 4 acoustic substeps, seeded smooth u/v/theta'")


def read_traffic(path, kernel):
    """Bytes per launch of ``kernel`` from a PMC summary text file with
    lines ``<kernel> fetch_bytes=<F> write_bytes=<W>``."""
    if not path or not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if parts and parts[0] == kernel:
                kv = dict(p.split("=", 1) for p in parts[1:] if "=" in p)
                if "fetch_bytes" in kv and "write_bytes" in kv:
                    return float(kv["fetch_bytes"]) + float(kv["write_bytes"])
    return None


def host_cpu() -> str:
    """The host's CPU model and the cores this process may use."""
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count()
    return f"{model}, {ncpu} cores visible"


def cpu_baseline(workload: str, rows: int, steps: int, dtype_name="float32"):
    """NumPy oracle (float32, or float64) on a (nz, rows, nx) strip of the
    workload: one untimed warm-up step, then ``steps`` timed steps.  A
    strip, not the full grid: a full-size step of the oracle takes about
    half a minute (float32) to a minute (float64) per step, and the
    baseline is bounded to a few tens of seconds so that the default run
    stays within minutes.  Every column of the strip is a full-height
    column of the workload, and NumPy's ufuncs are single-threaded, so the
    rate per point is the full grid's."""
    import numpy as np
    from oracle import dycore_np as O
    case, _ = build_case(workload, ny_override=rows)
    s = O.from_case(case, np.dtype(dtype_name).type)
    cfg = case.cfg
    O.step(s, cfg)                              # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        O.step(s, cfg)
    dt = time.perf_counter() - t0
    pts = cfg.nx * cfg.ny * cfg.nz
    return {"value": pts * steps / dt / 1e6, "unit": "Mpoint-steps/s",
            "cores": 1, "kind": "port",
            "sample": f"{dtype_name} NumPy oracle, 1 warm-up + {steps} "
                      f"timed steps on a {cfg.nx}x{cfg.ny}x{cfg.nz} strip "
                      f"of the same workload ({dt:.1f} s timed, "
                      f"single-threaded ufuncs; host: {host_cpu()})"}


def extras_adjoint(reps: int = 5) -> dict:
    """Compact C4 adjoint record: the 192x160x50 moist-tracer case with the
    final stage's PD limiter (the reference's default moist_adv_opt = 1),
    FP32 and FP64.  The fused step, the taped step and the reverse sweep
    in both tape modes (every substep kept, or stage checkpoints only with
    each stage recomputed in the reverse sweep); the FP64 dot-product
    residual of the scalar transport (exactly linear); and the FP64
    central-difference agreement of the whole step's adjoint (the step is
    nonlinear, so this is bounded by the differencing error; the exact
    whole-step check is against complex-step JVPs of the oracle, in the
    tests)."""
    import torch
    from arwen_amd import adjoint, cases, dycore
    from arwen_amd.state import DomainState

    case = cases.moist_tracer()
    cfg = case.cfg
    names = adjoint.STAGE_STATE + adjoint.SPECIES
    out = {"workload": "C4 moist tracer 192x160x50 + qv/qc/qr, dx=3 km, "
                       "dt=18 s, 4 substeps, PD limiter on the final stage",
           "ratio": "reverse sweep ms / fused forward step ms"}
    for tag, dtype in (("fp32", torch.float32), ("fp64", torch.float64)):
        st = DomainState.from_host(case, dtype=dtype)
        for _ in range(3):                  # leave the state at rest
            dycore.step(st, cfg)
        saved = {a: getattr(st, a).clone() for a in names}

        def restore():
            for a, t in saved.items():
                getattr(st, a).copy_(t)
            dycore.update_diagnostics(st)

        gen = torch.Generator(device=st.device).manual_seed(11)
        lam = {n: torch.randn(getattr(st, n).shape, dtype=dtype,
                              device=st.device, generator=gen) for n in names}
        t_rest = _time_ms(restore, reps)

        def fused():
            restore()
            dycore.step(st, cfg)

        t_f = _time_ms(fused, reps) - t_rest
        rec = {"forward_step_ms": round(t_f, 4)}
        for keep in ("stages", "substeps"):
            holder = {}

            def taped():
                restore()
                holder["tape"] = adjoint.step_taped(st, cfg, keep=keep)

            t_t = _time_ms(taped, reps) - t_rest
            t_r = _time_ms(lambda: adjoint.step_adjoint_taped(
                st, holder["tape"], lam), reps)
            holder.clear()
            rec[keep] = {"taped_step_ms": round(t_t, 4),
                         "reverse_ms": round(t_r, 4),
                         "reverse_to_forward": round(t_r / t_f, 3),
                         "taped_plus_reverse_to_forward":
                             round((t_t + t_r) / t_f, 3)}
        if dtype == torch.float64:
            ck = adjoint.TracerCheckpoints(st)
            restore()
            adjoint.step_with_checkpoints(st, cfg, ck)
            x = torch.randn(st.qv.shape, dtype=dtype, device=st.device,
                            generator=gen)
            y = torch.randn(st.qv.shape, dtype=dtype, device=st.device,
                            generator=gen)
            lhs = float(torch.sum(adjoint.tracer_step(x, st, cfg, ck) * y))
            rhs = float(torch.sum(x * adjoint.tracer_adjoint(y, st, cfg, ck)))
            rec["transport_dot_product_rel_residual"] = abs(lhs - rhs) / abs(lhs)
            restore()
            gm = adjoint.step_adjoint(st, cfg, lam, keep="stages")
            scl = {"u": 1.0, "v": 1.0, "w": 0.1, "thp": 0.5, "php": 20.0,
                   "mup": 20.0, "qv": 1e-4, "qc": 1e-5, "qr": 1e-5}
            vv = {n: torch.randn(getattr(st, n).shape, dtype=dtype,
                                 device=st.device, generator=gen) * scl[n]
                  for n in names}
            for n in adjoint.SPECIES:
                vv[n] = vv[n] * (saved[n] > 0)
            h = 1e-6

            def moved(hh):
                restore()
                for n in names:
                    getattr(st, n).add_(hh * vv[n])
                dycore.update_diagnostics(st)
                dycore.step(st, cfg)
                return {n: getattr(st, n).clone() for n in names}

            fp, fm = moved(h), moved(-h)
            jv = sum(float(torch.sum((fp[n] - fm[n]) / (2 * h) * lam[n]))
                     for n in names)
            vjt = sum(float(torch.sum(vv[n] * gm[n])) for n in names)
            rec["step_fd_rel_err_h1e-6"] = abs(jv - vjt) / abs(jv)
            del fp, fm, gm, ck
        out[tag] = rec
        del st, saved
        torch.cuda.empty_cache()
    return out


def extras_ensemble(steps: int) -> dict:
    """Compact C5 record: the 32 uneven members (nx, ny ~ U{96..256}, nz ~
    U{40..60}, dx = 3 km, dt = 18 s) advanced ``steps`` steps in one
    captured graph (4 member groups rejoining every 5 steps, column tile
    256: the autotuner's choice on this ensemble) against the same members
    stepped one after another (eagerly, the reference's launch model);
    bit-identity of the two; and the worst member's error against the FP64
    build of the same discretisation after the same steps."""
    import torch
    from arwen_amd import cases, dycore
    from arwen_amd import ensemble as E
    from arwen_amd.state import DomainState

    shapes = cases.ensemble_shapes()
    st32, st64 = [], []
    for i, (nx, ny, nz) in enumerate(shapes):
        h = cases.ensemble_member(nx, ny, nz, 20260731 + 1 + i)
        st32.append(DomainState.from_host(h))
        st64.append(DomainState.from_host(h, dtype=torch.float64))
        del h
    cfg = st32[0].cfg
    pts = sum(a * b * c for a, b, c in shapes)
    print(f"extras: C5 {len(st32)} members built", file=sys.stderr,
          flush=True)
    init = E.save_state(st32)
    with E.Ensemble(st32, cfg, groups=4, join_every=5, col_tile=256) as ens:
        ens.capture(steps)
        ens.launch()                        # first replay: one-off costs
        torch.cuda.synchronize()
        E.restore_state(st32, init)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ens.launch()
        torch.cuda.synchronize()
        t_b = time.perf_counter() - t0
    batched = E.save_state(st32)
    E.restore_state(st32, init)
    for st in st32:                         # warm each member's launch path
        dycore.step(st, cfg)
    E.restore_state(st32, init)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E.sequential_steps(st32, cfg, steps)
    torch.cuda.synchronize()
    t_s = time.perf_counter() - t0
    same = all(torch.equal(getattr(st, a), sv[0][a])
               for st, sv in zip(st32, batched) for a in E.STATE_ARRAYS)
    del batched, init
    with E.Ensemble(st64, cfg, groups=1) as e64:
        e64.step(steps)
    torch.cuda.synchronize()
    worst = {}
    for a, b in zip(st32, st64):
        for f in ("u", "v", "w", "thp", "php", "mup", "p"):
            x, y = getattr(a, f).double(), getattr(b, f)
            e = float((x - y).abs().max()) / float(y.abs().max())
            worst[f] = max(worst.get(f, 0.0), e)
    return {"workload": f"C5 32 members, {pts} points, {steps} steps",
            "batched_graph_ms_per_step": round(t_b * 1e3 / steps, 3),
            "sequential_ms_per_step": round(t_s * 1e3 / steps, 3),
            "speedup": round(t_s / t_b, 3),
            "bitwise_equal_to_sequential": same,
            "max_rel_err_vs_fp64_worst_member": worst}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus "
              f"{args.gpus}); the step is a single-GPU kernel set and the "
              "multi-GPU sharding is design-only", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c5":
        return main_ensemble(args)
    if args.workload == "c3-precision":
        return main_precision(args)
    if args.workload == "c4-adjoint":
        return main_adjoint(args)
    import torch
    from arwen_amd import _lib, dycore
    from arwen_amd.state import DomainState

    case, desc = build_case(args.workload)
    if args.species >= 0:
        from arwen_amd import cases as _cases
        from arwen_amd.config import moist_species
        _cases.add_species(case, args.species, qv_max=1.0e-3)
        n_q = len(moist_species(case.cfg))
        desc += (f", moist: mp_physics={args.species} species ({n_q} "
                 "transported, PD on the final stage)"
                 + (" and a retained heating" if args.species else ""))
        args.no_extras = True
        args.no_cpu_baseline = True
    cfg = case.cfg
    if args.diff6:
        cfg = cfg.with_(diff_6th_opt=args.diff6)
        desc += f", held 6th-order filter diff_6th_opt={args.diff6}"
    if args.km_opt == 4:
        cfg = cfg.with_(km_opt=4)
        desc += ", held 2-D Smagorinsky mixing km_opt=4"
    state = DomainState.from_host(case)
    del case
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        dycore.step(state, cfg)
    torch.cuda.synchronize()

    timer = _lib.KernelTimer(max_events=320 * max(args.steps, 1) + 64)
    t0 = time.perf_counter()
    with timer:
        for _ in range(args.steps):
            dycore.step(state, cfg)
        torch.cuda.synchronize()
    t1 = time.perf_counter()
    if args.copy_calibration:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
    rep = dycore.stability_report(state, cfg)
    if rep["nan"]:
        raise SystemExit("bench.py: the state went NaN")
    # the full-state health gate on the same state (outside the timed
    # region): one pass over every checked field
    from arwen_amd.health import StateHealthValidator
    gate = StateHealthValidator(state)
    gate.require_healthy(phase="bench")
    t_gate = _time_ms(lambda: gate.require_healthy(phase="bench"), 5)

    pts = cfg.nx * cfg.ny * cfg.nz
    sec = (t1 - t0) / args.steps
    value = pts / sec / 1e6
    ns = cfg.time_step_sound
    wstep = step_words(ns)
    step_gbs = wstep * 4 * pts / sec / 1e9

    # the acoustic substep pass (U) dominates: 7 launches per step at ns=4
    n_sub = timer.launches.get("k_column", 0)
    u_ms = timer.ms.get("k_uv", 0.0) + timer.ms.get("k_column", 0.0)
    u_avg_ms = u_ms / n_sub if n_sub else float("nan")
    u_bytes = WORDS_U * 4 * pts
    u_gbs = u_bytes / (u_avg_ms * 1e-3) / 1e9 if n_sub else None
    traffic = None
    tfile = args.traffic_file
    if tfile is None and args.workload == "c3" and args.species < 0:
        tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                             "profiles",
                             "dycore_step_1000x1000x50_pmc_traffic.txt")
    t_uv = read_traffic(tfile, "k_uv")
    t_col = read_traffic(tfile, "k_column")
    if t_uv is not None and t_col is not None:
        traffic = t_uv + t_col

    out = {
        "metric": "RK3 dycore Mpoint-steps/s, % HBM roofline; "
                  "adjoint:forward cost at stated err",
        "value": round(value, 3),
        "unit": "Mpoint-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded initial state)",
        "config": {"workload": desc, "grid": [cfg.nx, cfg.ny, cfg.nz],
                   "dt": cfg.dt, "time_step_sound": ns,
                   "parallelism": "single GPU"},
        "step_roofline": {"alg_bytes_per_point_step": wstep * 4,
                          "achieved_GBs": round(step_gbs, 1),
                          "frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "roofline": {
            "bound": "hbm",
            "kernel": "acoustic substep pass (k_uv + k_column)",
            "avg_launch_ms": round(u_avg_ms, 4),
            "alg_bytes_per_launch": u_bytes,
            "achieved": None if u_gbs is None else round(u_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": None if u_gbs is None else round(u_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "kernel_ms_per_step": {k: round(v / args.steps, 4)
                               for k, v in timer.ms.items() if v > 0},
        "health_gate_ms": round(t_gate, 4),
    }
    del state
    torch.cuda.empty_cache()
    if not args.no_extras:
        out["adjoint"] = extras_adjoint()
        torch.cuda.empty_cache()
        out["ensemble"] = extras_ensemble(args.extra_ens_steps)
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_rows,
                                           args.cpu_steps)
        out["cpu_baseline_fp64"] = cpu_baseline(
            args.workload, max(args.cpu_rows // 2, 8), 1, "float64")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def _rel_err(a, b):
    """max |a - b| / max |b| over one field (b the FP64 reference)."""
    import torch
    a64, b64 = a.double(), b.double()
    den = float(b64.abs().max())
    return float((a64 - b64).abs().max()) / den if den > 0 else 0.0


def main_precision(args):
    """C3 precision form: the warm-bubble 1000x1000x50 state advanced
    ``--prec-steps`` steps in plain FP32, in the precision form (FP32 bulk,
    Kahan column sums, FP64 tridiagonal) and in the FP64 build of the same
    discretisation; per-field max and RMS error of both FP32 runs against
    FP64 after 1, 10 and the last step, and the three step times.  The
    FP64 snapshots stay in HBM."""
    import dataclasses

    import torch
    from arwen_amd import dycore
    from arwen_amd.state import DomainState

    case, desc = build_case("c3")
    n = args.prec_steps
    marks = sorted({1, min(10, n), n})
    fields = ("u", "v", "w", "thp", "php", "mup", "p")
    ref = {}                                    # step -> field -> FP64 tensor
    runs, err = {}, {}
    for name, dtype, prec in (("fp64", torch.float64, False),
                              ("fp32", torch.float32, False),
                              ("precision_form", torch.float32, True)):
        cfg = dataclasses.replace(case.cfg, precision_form=prec)
        st = DomainState.from_host(case, dtype=dtype)
        m0 = dycore.domain_mass_measure(st)
        torch.cuda.synchronize()
        t_step = 0.0
        e_run = {}
        for k in range(n):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            dycore.step(st, cfg)
            torch.cuda.synchronize()
            t_step += time.perf_counter() - t0
            if k + 1 in marks:
                if name == "fp64":
                    ref[k + 1] = {f: getattr(st, f).clone() for f in fields}
                else:
                    e = {}
                    for f in fields:
                        b = ref[k + 1][f]
                        dlt = getattr(st, f).double() - b
                        scale = float(b.abs().max())
                        e[f] = {"max_rel": float(dlt.abs().max()) / scale,
                                "rms_rel": float(dlt.pow(2).mean().sqrt())
                                / scale}
                        del dlt
                    e_run[f"step_{k + 1}"] = e
            if (k + 1) % 100 == 0:
                print(f"{name}: {k + 1}/{n} steps", file=sys.stderr,
                      flush=True)
        m1 = dycore.domain_mass_measure(st)
        if dycore.stability_report(st, cfg)["nan"]:
            raise SystemExit(f"bench.py: {name} went NaN")
        runs[name] = {"ms_per_step": round(t_step * 1e3 / n, 3),
                      "mass_rel_change": (m1 - m0) / m0}
        if name != "fp64":
            err[name] = e_run
        del st
        torch.cuda.empty_cache()
    out = {"workload": desc + f", {n} steps", "steps": n,
           "ms_per_step": {k: v["ms_per_step"] for k, v in runs.items()},
           "mass_rel_change": {k: v["mass_rel_change"]
                               for k, v in runs.items()},
           "error_vs_fp64": err,
           "scale": "max/RMS of |field - fp64| over max |fp64 field|"}
    print(json.dumps(out), flush=True)


def _time_ms(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def main_adjoint(args):
    """C4 adjoint form: the 192x160x50 moist-tracer case (dx = 3 km, dt =
    18 s).  Times one checkpointed step, the checkpointed scalar transport
    J q and its hand-written adjoint J^T lam (FP32 and FP64), and states
    the FP64 dot-product residual and the central-difference agreement of
    the autograd gradient."""
    import dataclasses

    import torch
    from arwen_amd import adjoint, cases, dycore
    from arwen_amd import advection as A_
    from arwen_amd.state import DomainState

    case = cases.moist_tracer(moist_adv_opt=0)
    cfg = case.cfg
    reps = max(args.steps, 5)
    out = {"workload": "C4 moist tracer 192x160x50, dx=3 km, dt=18 s, "
                       "4 substeps, qv U(0,16) g/kg exp(-z/2.5 km)",
           "reps": reps}
    for name, dtype in (("fp32", torch.float32), ("fp64", torch.float64)):
        st = DomainState.from_host(case, dtype=dtype)
        ck = adjoint.TracerCheckpoints(st)
        q0 = st.qv.clone()
        saved = {a: getattr(st, a).clone() for a in
                 ("u", "v", "w", "thp", "php", "mup", "p", "al", "alt",
                  "qv", "qc", "qr")}

        def fwd_step():
            for a, t in saved.items():
                getattr(st, a).copy_(t)
            adjoint.step_with_checkpoints(st, cfg, ck)

        def plain_step():
            for a, t in saved.items():
                getattr(st, a).copy_(t)
            dycore.step(st, cfg)

        t_copy = _time_ms(lambda: [getattr(st, a).copy_(t)
                                   for a, t in saved.items()], reps)
        t_step = _time_ms(plain_step, reps) - t_copy
        t_ckpt = _time_ms(fwd_step, reps) - t_copy
        lam = torch.randn_like(q0)
        t_fwd = _time_ms(lambda: adjoint.tracer_step(q0, st, cfg, ck), reps)
        t_adj = _time_ms(lambda: adjoint.tracer_adjoint(lam, st, cfg, ck),
                         reps)
        rec = {"step_ms": round(t_step, 4),
               "checkpointed_step_ms": round(t_ckpt, 4),
               "tracer_forward_ms": round(t_fwd, 4),
               "tracer_adjoint_ms": round(t_adj, 4),
               "adjoint_to_forward": round(t_adj / t_fwd, 3),
               "adjoint_to_step": round(t_adj / t_step, 3)}
        if dtype == torch.float64:
            gen = torch.Generator(device=st.device).manual_seed(7)
            x = torch.randn(q0.shape, dtype=dtype, device=st.device,
                            generator=gen)
            y = torch.randn(q0.shape, dtype=dtype, device=st.device,
                            generator=gen)
            lhs = float(torch.sum(adjoint.tracer_step(x, st, cfg, ck) * y))
            rhs = float(torch.sum(x * adjoint.tracer_adjoint(y, st, cfg, ck)))
            rec["dot_product_rel_residual"] = abs(lhs - rhs) / abs(lhs)
            w = torch.randn(q0.shape, dtype=dtype, device=st.device,
                            generator=gen)

            def loss(qq):
                return 0.5 * torch.sum(
                    w * adjoint.transport(qq, st, cfg, [ck])) ** 2

            qv = q0.clone().requires_grad_(True)
            (grad,) = torch.autograd.grad(loss(qv), qv)
            eps = 1e-4
            with torch.no_grad():
                fd = (float(loss(q0 + eps * x))
                      - float(loss(q0 - eps * x))) / (2 * eps)
            ad = float(torch.sum(grad * x))
            rec["fd_rel_err"] = abs(fd - ad) / abs(ad)
        # the whole moist step (qv/qc/qr through the moist EOS, buoyancy
        # loading and the sumflux-coupled scalar update), without and with
        # ("_pd" keys; the reference's default moist_adv_opt = 1) the final
        # stage's positive-definite limiter, against its adjoint: the taped
        # step, then the reverse sweep
        cfg_pd = dataclasses.replace(cfg, moist_adv_opt=1)
        for tag, mcfg in (("", cfg), ("_pd", cfg_pd)):
            names = adjoint.STAGE_STATE + adjoint.SPECIES
            holder = {}

            def moist_taped():
                for a, t in saved.items():
                    getattr(st, a).copy_(t)
                holder["tape"] = adjoint.step_taped(st, mcfg)

            gen = torch.Generator(device=st.device).manual_seed(7)
            lam_m = {n: torch.randn(getattr(st, n).shape, dtype=dtype,
                                    device=st.device, generator=gen)
                     for n in names}

            def moist_plain():
                for a, t in saved.items():
                    getattr(st, a).copy_(t)
                dycore.step(st, mcfg)

            t_ms = _time_ms(moist_plain, reps) - t_copy
            t_mt = _time_ms(moist_taped, reps) - t_copy
            t_mr = _time_ms(lambda: adjoint.step_adjoint_taped(
                st, holder["tape"], lam_m), reps)
            holder.clear()
            rec.update({f"moist{tag}_step_ms": round(t_ms, 4),
                        f"moist{tag}_taped_step_ms": round(t_mt, 4),
                        f"moist{tag}_reverse_sweep_ms": round(t_mr, 4),
                        f"moist{tag}_reverse_to_forward": round(t_mr / t_ms,
                                                                3)})
            if dtype == torch.float64:
                # <J v, lam> by central differences of the fused step against
                # <v, J^T lam>, v scaled per field, about the state after three
                # steps (the initial state is at rest vertically, where the
                # vertical upwind switch sign(Omega) is not differentiable)
                for a, t in saved.items():
                    getattr(st, a).copy_(t)
                for _ in range(3):
                    dycore.step(st, mcfg)
                base = {a: getattr(st, a).clone() for a in saved}
                gm = adjoint.step_adjoint(st, mcfg, lam_m)
                scl = {"u": 1.0, "v": 1.0, "w": 0.1, "thp": 0.5, "php": 20.0,
                       "mup": 20.0, "qv": 1e-4, "qc": 1e-5, "qr": 1e-5}
                vv = {n: torch.randn(getattr(st, n).shape, dtype=dtype,
                                     device=st.device, generator=gen) * scl[n]
                      for n in names}
                for n in adjoint.SPECIES:      # keep the scalars non-negative
                    vv[n] = vv[n] * (base[n] > 0)

                def moved(hh):
                    for a, t in base.items():
                        getattr(st, a).copy_(t)
                    for n in names:
                        getattr(st, n).add_(hh * vv[n])
                    dycore.update_diagnostics(st)
                    dycore.step(st, mcfg)
                    return {n: getattr(st, n).clone() for n in names}

                vjt = sum(float(torch.sum(vv[n] * gm[n])) for n in names)
                fd = {}
                for hh in (1e-3, 1e-4, 1e-5, 1e-6):
                    fp, fm = moved(hh), moved(-hh)
                    jv = sum(float(torch.sum((fp[n] - fm[n]) / (2 * hh)
                                             * lam_m[n])) for n in names)
                    fd[f"{hh:g}"] = abs(jv - vjt) / abs(jv)
                    del fp, fm
                rec[f"moist{tag}_step_fd_rel_err_by_h"] = fd
                del gm, base
        for a, t in saved.items():
            getattr(st, a).copy_(t)
        dycore.update_diagnostics(st)
        # one acoustic substep (stage-3 shape, state after a step) and its
        # adjoint at the same small-step state
        dtau = cfg.dt / cfg.time_step_sound
        dycore.prepare_acoustic_coefficients(st, cfg, dtau)
        xs = {n: getattr(st, n).clone() for n in adjoint.ACOUSTIC_STATE}

        def restore():
            for n, t in xs.items():
                getattr(st, n).copy_(t)

        def sub_fwd():
            restore()
            dycore.acoustic_substep(st, cfg, dtau, False)

        lam = {n: torch.randn_like(getattr(st, n))
               for n in adjoint.ACOUSTIC_STATE}
        t_rest = _time_ms(restore, reps)
        t_sub = _time_ms(sub_fwd, reps) - t_rest
        restore()
        t_sub_adj = _time_ms(lambda: adjoint.acoustic_substep_adjoint(
            st, cfg, dtau, False, lam), reps)
        rec.update({"acoustic_substep_ms": round(t_sub, 4),
                    "acoustic_adjoint_ms": round(t_sub_adj, 4),
                    "acoustic_adjoint_to_forward": round(t_sub_adj / t_sub,
                                                         3)})
        # the large-step tendency pass (a dry copy of the state: the
        # pass's adjoint covers dry states) against its adjoint
        cfg_dry = dataclasses.replace(cfg, moist=False)
        dry_case = dataclasses.replace(
            case, cfg=cfg_dry, fields={k: v for k, v in case.fields.items()
                                       if k not in ("qv", "qc", "qr")})
        sd = DomainState.from_host(dry_case, dtype=dtype)
        for n in adjoint.STAGE_STATE:
            getattr(sd, n).copy_(getattr(st, n))
        lam_t = {n: torch.randn_like(getattr(sd, n))
                 for n in adjoint.STAGE_TEND}
        t_sp = _time_ms(lambda: dycore.stage_prepare(
            sd, cfg_dry, cfg.dt / cfg.time_step_sound, True), reps)
        t_sa = _time_ms(lambda: adjoint.stage_tendencies_adjoint(
            sd, cfg_dry, lam_t), reps)
        rec.update({"stage_pass_ms": round(t_sp, 4),
                    "stage_pass_adjoint_ms": round(t_sa, 4),
                    "stage_pass_adjoint_to_forward": round(t_sa / t_sp, 3)})
        # one whole dry step against its adjoint (the forward re-run with
        # checkpoints, then the reverse sweep over the three stages)
        saved_d = {n: getattr(sd, n).clone() for n in adjoint.STAGE_STATE}

        def restore_d():
            for n, t in saved_d.items():
                getattr(sd, n).copy_(t)

        def dry_step():
            restore_d()
            dycore.step(sd, cfg_dry)

        t_rd = _time_ms(restore_d, reps)
        t_ds = _time_ms(dry_step, reps) - t_rd
        restore_d()
        dycore.update_diagnostics(sd)
        lam_s = {n: torch.randn_like(getattr(sd, n))
                 for n in adjoint.STAGE_STATE}
        t_da = _time_ms(lambda: adjoint.step_adjoint(sd, cfg_dry, lam_s),
                        reps)
        holder = {}

        def taped():
            restore_d()
            holder["tape"] = adjoint.step_taped(sd, cfg_dry)

        t_tp = _time_ms(taped, reps) - t_rd
        t_rev = _time_ms(lambda: adjoint.step_adjoint_taped(
            sd, holder["tape"], lam_s), reps)

        def taped_lean():
            restore_d()
            holder["tape"] = adjoint.step_taped(sd, cfg_dry, keep="stages")

        t_tl = _time_ms(taped_lean, reps) - t_rd
        t_rl = _time_ms(lambda: adjoint.step_adjoint_taped(
            sd, holder["tape"], lam_s), reps)
        holder.clear()
        rec.update({"dry_step_ms": round(t_ds, 4),
                    "step_adjoint_ms": round(t_da, 4),
                    "step_adjoint_to_forward": round(t_da / t_ds, 3),
                    "taped_step_ms": round(t_tp, 4),
                    "reverse_sweep_ms": round(t_rev, 4),
                    "reverse_to_forward": round(t_rev / t_ds, 3),
                    "stage_ckpt_taped_step_ms": round(t_tl, 4),
                    "stage_ckpt_reverse_ms": round(t_rl, 4),
                    "stage_ckpt_reverse_to_forward": round(t_rl / t_ds, 3)})
        del sd
        # the four advection kernels at the stage mass fluxes of the
        # state: forward tendency against the adjoint (field and all three
        # mass-flux gradients), both writing fresh zeroed outputs
        ru, rv, ww = dycore.stage_fluxes(st, cfg)
        fields = {"scalar": st.thp, "u": st.u, "v": st.v, "w": st.w}
        for op, q in fields.items():
            lam_q = torch.randn_like(q)
            t_f = _time_ms(lambda: A_.flux_div(op, q, ru, rv, ww, st, cfg.dx,
                                               cfg.dy), reps)
            t_a = _time_ms(lambda: A_.flux_div_adjoint(
                op, lam_q, q, ru, rv, ww, st, cfg.dx, cfg.dy), reps)
            rec[f"advect_{op}_forward_ms"] = round(t_f, 4)
            rec[f"advect_{op}_adjoint_ms"] = round(t_a, 4)
            rec[f"advect_{op}_adjoint_to_forward"] = round(t_a / t_f, 3)
        out[name] = rec
        del st, ck
        torch.cuda.empty_cache()
    print(json.dumps(out), flush=True)


def main_ensemble(args):
    """C5: 32 dry members of uneven extents (nx, ny ~ U{96..256}, nz ~
    U{40..60}, dx = 3 km, dt = 18 s, 4 substeps).  Timed: one HIP graph
    launch of ``--ens-steps`` batched steps (layout from the autotuner)
    against the same members stepped one after another -- eagerly through
    dycore.step, and as one captured graph per member."""
    import torch
    from arwen_amd import cases, dycore
    from arwen_amd import ensemble as E
    from arwen_amd.state import DomainState

    t_init = time.perf_counter()
    hosts = cases.ensemble_members(args.ens_members)
    cfg = hosts[0].cfg
    states = [DomainState.from_host(h) for h in hosts]
    print(f"c5: {len(states)} members built in "
          f"{time.perf_counter() - t_init:.1f} s", file=sys.stderr, flush=True)
    pts = sum(st.shape[0] * st.shape[1] * st.shape[2] for st in states)
    n = args.ens_steps
    init = E.save_state(states)

    tune = E.autotune(states, cfg, steps=10, reps=2, verbose=True)
    ens = E.Ensemble(states, cfg, order=tune["order"], groups=tune["groups"],
                     join_every=tune["join_every"], col_tile=tune["col_tile"])
    ens.capture(n)
    # warm-up: the first replay of a fresh graph pays one-off costs
    ens.launch()
    torch.cuda.synchronize()
    E.restore_state(states, init)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.launch()
    torch.cuda.synchronize()
    t_batched = time.perf_counter() - t0
    batched = E.save_state(states)
    ens.close()
    print(f"c5: batched graph {t_batched * 1e3:.1f} ms for {n} steps",
          file=sys.stderr, flush=True)

    # sequential, eager: the reference's launch model
    E.restore_state(states, init)
    for st in states:                     # warm every member's launch path
        dycore.step(st, cfg)
    E.restore_state(states, init)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E.sequential_steps(states, cfg, n)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    same = all(torch.equal(getattr(st, a), sv[0][a])
               for st, sv in zip(states, batched) for a in E.STATE_ARRAYS)
    print(f"c5: sequential eager {t_seq * 1e3:.1f} ms; bitwise equal to "
          f"batched: {same}", file=sys.stderr, flush=True)

    # sequential, one captured graph per member (launch gaps removed)
    E.restore_state(states, init)
    singles = []
    for st in states:
        e1 = E.Ensemble([st], cfg)
        e1.capture(n)
        singles.append(e1)
    for e1 in singles:
        e1.launch()
    torch.cuda.synchronize()
    E.restore_state(states, init)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e1 in singles:
        e1.launch()
    torch.cuda.synchronize()
    t_seq_graph = time.perf_counter() - t0
    for e1 in singles:
        e1.close()
    del singles

    errs = None
    if not args.no_fp64:
        # the same members in FP64 (batched; bit-identical to FP64
        # sequential), then per-field max relative error of the FP32 run
        s64 = [DomainState.from_host(h, dtype=torch.float64) for h in hosts]
        with E.Ensemble(s64, cfg, groups=1) as e64:
            e64.step(n)
        torch.cuda.synchronize()
        errs = {}
        for a in ("u", "v", "w", "thp", "php", "mup", "p"):
            errs[a] = max(_rel_err(sv[0][a], getattr(s64i, a))
                          for sv, s64i in zip(batched, s64))
        del s64

    value = pts * n / t_batched / 1e6
    out = {
        "metric": "RK3 dycore Mpoint-steps/s, % HBM roofline; "
                  "adjoint:forward cost at stated err",
        "value": round(value, 3),
        "unit": "Mpoint-steps/s",
        "n_gpus": 1,
        "steps": n,
        "warmup": 1,
        "ms_per_step": round(t_batched * 1e3 / n, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded initial states)",
        "config": {"workload": "C5 batched ensemble: 32 dry members, nx,ny "
                               "~U{96..256}, nz ~U{40..60}, dx=3 km, dt=18 s, "
                               "4 substeps, one HIP graph launch",
                   "members": [list(st.shape) for st in states],
                   "points": pts, "parallelism": "single GPU"},
        "ensemble": {
            "layout": {"order": tune["order_name"], "groups": tune["groups"],
                       "join_every": tune["join_every"],
                       "col_tile": tune["col_tile"]},
            "autotune_trials": tune["trials"],
            "batched_ms": round(t_batched * 1e3, 2),
            "sequential_eager_ms": round(t_seq * 1e3, 2),
            "sequential_graph_ms": round(t_seq_graph * 1e3, 2),
            "speedup_vs_sequential_eager": round(t_seq / t_batched, 3),
            "speedup_vs_sequential_graph": round(t_seq_graph / t_batched, 3),
            "bitwise_equal_to_sequential": bool(same),
            "max_rel_err_vs_fp64": errs,
        },
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
