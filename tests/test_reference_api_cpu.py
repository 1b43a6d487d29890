"""The reference's own case-building pattern runs unchanged against this
package (CPU only: the state is built on CPU tensors, no step is taken).

Restated from the reference's Straka builder (gpuwm/verify/cases/straka.py
:42-75): RunConfig -> make_vertical_coord -> make_base_state ->
init_theta_perturbation, which returns a DomainState.
"""
import numpy as np
import pytest
import torch

from arwen_amd import constants as c
from arwen_amd.config import RunConfig
from arwen_amd.grid import make_base_state, make_vertical_coord
from arwen_amd.state import DomainState, init_at_rest, init_theta_perturbation


def _straka_cfg(nx=512):
    return RunConfig(nx=nx, ny=1, nz=64, dx=100.0, dy=100.0, ztop=6400.0,
                     dt=0.5, run_seconds=900.0, khdif=75.0, kvdif=75.0,
                     case="straka")


def _sounding(z):
    return np.full_like(np.asarray(z, dtype=np.float64), 300.0)


def _build(cfg, coord, base):
    exner_b = (np.asarray(base.pb, dtype=np.float64) / c.P0) ** c.RCP

    def thp_func(x, z):
        L = np.sqrt((x[None, None, :] / 4000.0) ** 2
                    + ((z[:, None, None] - 3000.0) / 2000.0) ** 2)
        t_pert = np.where(L <= 1.0,
                          -15.0 * (np.cos(np.pi * L) + 1.0) / 2.0, 0.0)
        thp = t_pert / exner_b[:, None, None]
        return np.broadcast_to(thp, (cfg.nz, cfg.ny, cfg.nx))

    return init_theta_perturbation(cfg, coord, base, thp_func, device="cpu")


def test_straka_build_pattern_returns_a_domain_state():
    cfg = _straka_cfg()
    coord = make_vertical_coord(cfg.nz)
    base = make_base_state(coord, _sounding, p_surf=cfg.p_surf,
                           ztop=cfg.ztop)
    s = _build(cfg, coord, base)
    assert isinstance(s, DomainState)
    assert tuple(s.u.shape) == (64, 1, 513)
    assert tuple(s.php.shape) == (65, 1, 512)
    assert s.thp.dtype == torch.float32
    # the cold blob: -15 K of T' at its centre, theta' = T'/Exner there
    assert -17.5 < float(s.thp.min()) < -15.0
    assert float(s.thp.max()) == 0.0
    # rebalanced: phi' is non-zero where theta' is, winds and mu' are zero
    assert float(s.php.abs().max()) > 0.0
    assert float(s.u.abs().max()) == 0.0 and float(s.mup.abs().max()) == 0.0
    # the same state as this package's own Straka workload builder
    from arwen_amd.cases import straka
    case = straka(nx=512, nz=64, dx=100.0, dt=0.5)
    for name in ("thp", "php"):
        want = np.asarray(case.fields[name], dtype=np.float32)
        assert np.array_equal(getattr(s, name).numpy(), want), name
    # and it reads back as a host case for the oracle
    back = s.host_case()
    assert np.array_equal(back.fields["thp"].astype(np.float32),
                          s.thp.numpy())


def test_terrain_z_is_a_consistency_check():
    from arwen_amd.grid import gaussian_hill
    cfg = RunConfig(nx=16, ny=12, nz=10, dx=1000.0, dy=1000.0, ztop=10000.0,
                    dt=6.0, terrain_opt=1)
    coord = make_vertical_coord(cfg.nz)
    tz = gaussian_hill(cfg.nx, cfg.ny, cfg.dx, cfg.dy, 300.0, 3000.0)
    base = make_base_state(coord, _sounding, cfg.p_surf, cfg.ztop,
                           terrain_z=tz)
    s = init_at_rest(cfg, coord, base, terrain_z=tz, device="cpu")
    assert np.array_equal(s.ht.numpy(), tz.astype(np.float32))
    with pytest.raises(ValueError, match="terrain_z"):
        init_at_rest(cfg, coord, base, terrain_z=tz + 1.0, device="cpu")
    flat = make_base_state(coord, _sounding, cfg.p_surf, cfg.ztop)
    with pytest.raises(ValueError, match="terrain_z"):
        init_at_rest(cfg.with_(terrain_opt=0), coord, flat, terrain_z=tz,
                     device="cpu")
