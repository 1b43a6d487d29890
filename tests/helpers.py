"""Shared builders for the GPU-vs-oracle parity tests."""
from __future__ import annotations

import numpy as np

from oracle import dycore_np as O

PROG = ("u", "v", "w", "thp", "php", "mup")
DIAG = ("p", "al", "alt")


def oracle_and_gpu(case, *, set_w=False, dtype=np.float32):
    """An FP32 oracle state and a GPU DomainState holding the identical
    FP32 initial condition (set_w applies the kinematic lower boundary on
    the FP32 state first, as the reference case builders do on device)."""
    import torch
    from arwen_amd.state import DomainState
    o = O.from_case(case, dtype)
    if set_w:
        O.set_w_surface(o, case.cfg.dx, case.cfg.dy)
    g = DomainState.from_host(case)
    for name in PROG:
        src = getattr(o, name)
        if dtype != np.float32:
            src = src.astype(np.float32)
        getattr(g, name).copy_(torch.as_tensor(src))
    return o, g


def compare(o, g, names=PROG + DIAG):
    """{name: (normwise rel err, fraction of bit-identical values)}."""
    out = {}
    for n in names:
        a = getattr(g, n).detach().cpu().numpy()
        b = getattr(o, n)
        scale = np.max(np.abs(b.astype(np.float64)))
        err = np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)))
        out[n] = (err / scale if scale > 0 else err,
                  float(np.mean(a == b)))
    return out


def random_state_case(seed=0, nz=8, ny=6, nx=12, hybrid_opt=0, hill=0.0):
    """Small consistent state like the reference's acoustic test builder
    (gpuwm/verify/npref.py:2217-2324): isothermal 300 K base, seeded
    reference fields; returns (HostCase, rng) for further filling."""
    from arwen_amd.cases import HostCase, host_at_rest
    from arwen_amd.config import RunConfig
    from arwen_amd.grid import bell_hill, make_base_state, make_vertical_coord
    cfg = RunConfig(nx=nx, ny=ny, nz=nz, dx=500.0, dy=500.0, ztop=8000.0,
                    dt=3.0, hybrid_opt=hybrid_opt,
                    terrain_opt=1 if hill > 0 else 0)
    coord = make_vertical_coord(nz, hybrid_opt=hybrid_opt)
    tz = bell_hill(nx, ny, 500.0, hill, 1500.0) if hill > 0 else None
    base = make_base_state(coord, lambda z: np.full_like(z, 300.0),
                           cfg.p_surf, cfg.ztop, terrain_z=tz)
    case = host_at_rest(cfg, coord, base)
    rng = np.random.default_rng(seed)
    f = case.fields
    f["u"] = 0.05 * rng.standard_normal(f["u"].shape)
    f["u"][..., -1] = f["u"][..., 0]
    f["v"] = 0.05 * rng.standard_normal(f["v"].shape)
    f["v"][:, -1, :] = f["v"][:, 0, :]
    f["w"] = 0.05 * rng.standard_normal(f["w"].shape)
    f["w"][0] = 0.0
    f["w"][-1] = 0.0
    f["thp"] = 0.5 * rng.standard_normal(f["thp"].shape)
    f["php"] = 20.0 * rng.standard_normal(f["php"].shape)
    f["php"][0] = 0.0
    f["mup"] = 20.0 * rng.standard_normal(f["mup"].shape)
    return case, rng
