"""Initial-state builders of the reference's whole-step regression pin.

The reference captured ``tests/golden/phase2_step_regression.npz`` by
building these states and running 5 ``dycore.step`` calls (reference
tests/_phase2_pin.py:22-127).  The builders are restated here so the same
initial states feed the oracle and the HIP step.
"""
from __future__ import annotations

import numpy as np

from arwen_amd.cases import HostCase, cell_x, height_half, host_theta_perturbation
from arwen_amd.config import RunConfig
from arwen_amd.grid import bell_hill, make_base_state, make_vertical_coord

PIN_FIELDS = ("u", "v", "w", "thp", "php", "mup")
PIN_STEPS = 5


def _bubble3(cfg, amp=2.0, zc=2000.0, rz=1500.0, rx=2000.0, ry=2000.0):
    y = (np.arange(cfg.ny) + 0.5) * cfg.dy - 0.5 * cfg.ny * cfg.dy

    def thp(x, z):
        zz = z[:, None, None] if np.ndim(z) == 1 else z
        L = np.sqrt((x[None, None, :] / rx) ** 2
                    + (y[None, :, None] / ry) ** 2 + ((zz - zc) / rz) ** 2)
        return np.where(L < 1.0, amp * np.cos(np.pi * L / 2) ** 2, 0.0) \
            * np.ones((cfg.nz, cfg.ny, cfg.nx))
    return thp


def _shear_u(case: HostCase, u0=5.0):
    cfg = case.cfg
    z = height_half(case.base)
    zz = z if np.ndim(z) == 3 else np.broadcast_to(
        np.asarray(z)[:, None, None], (cfg.nz, cfg.ny, cfg.nx))
    prof = u0 * np.tanh(zz / 3000.0)
    jvar = 1.0 + 0.1 * np.sin(2 * np.pi * np.arange(cfg.ny) / cfg.ny)
    u = prof * jvar[None, :, None]
    u = np.concatenate([u, u[:, :, :1]], axis=2)
    # the reference fills the device array in FP32
    case.fields["u"] = u.astype(np.float32).astype(np.float64)


def _theta(z):
    return 300.0 * np.exp(1e-4 * np.asarray(z, float) / 9.81)


def build_dry_flat() -> HostCase:
    cfg = RunConfig(nx=14, ny=12, nz=16, dx=500.0, dy=500.0, ztop=8000.0,
                    dt=2.0, run_seconds=0.0)
    vc = make_vertical_coord(cfg.nz)
    b = make_base_state(vc, _theta, cfg.p_surf, cfg.ztop)
    case = host_theta_perturbation(cfg, vc, b, _bubble3(cfg))
    _shear_u(case)
    return case


def build_terrain() -> HostCase:
    """Bell ridge + hybrid coordinate; w(sfc) is set by the caller in FP32
    (the reference runs set_w_surface on the device state)."""
    cfg = RunConfig(nx=16, ny=10, nz=16, dx=1000.0, dy=1000.0, ztop=10000.0,
                    dt=2.0, run_seconds=0.0, terrain_opt=1, hybrid_opt=2,
                    hill_height=300.0, hill_halfwidth=3000.0)
    vc = make_vertical_coord(cfg.nz, hybrid_opt=2, etac=cfg.etac)
    tz = bell_hill(cfg.nx, cfg.ny, cfg.dx, cfg.hill_height,
                   cfg.hill_halfwidth)
    b = make_base_state(vc, _theta, cfg.p_surf, cfg.ztop, terrain_z=tz)
    case = host_theta_perturbation(cfg, vc, b, _bubble3(cfg, zc=3000.0))
    _shear_u(case, u0=3.0)
    return case


CASES = {"dry_flat": build_dry_flat, "terrain": build_terrain}
