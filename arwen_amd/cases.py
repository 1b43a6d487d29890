"""Host-side (FP64 NumPy) initial conditions for the dycore step.

A :class:`HostCase` bundles a :class:`~arwen_amd.config.RunConfig`, the
vertical coordinate, the base state and the initial prognostic fields.
``arwen_amd.state.DomainState.from_host`` uploads it to the GPU; the test
oracle builds its own state from the same object.

The reference's own initial-state API, :func:`init_at_rest` and
:func:`init_theta_perturbation` (gpuwm/core/state.py:922-993), returns a
device :class:`~arwen_amd.state.DomainState`, as the reference's does; it
is built from the host-side :func:`host_at_rest` /
:func:`host_theta_perturbation`, which the oracle and the workload builders
below share.

Builders:

* :func:`host_theta_perturbation` -- at-rest state plus a hydrostatically
  rebalanced theta perturbation (reference state.py:961-993);
* :func:`straka` -- Straka et al. (1993) density current (reference
  verify/cases/straka.py:42-75), at the resolution and time step given;
* :func:`stratified_hill` -- 3-D dry stratified flow over a Gaussian hill
  with seeded smooth random u/v/theta' (250x200x49, dx=12 km);
* :func:`warm_bubbles` -- stratified dry state with seeded +3 K cosine^2
  warm bubbles (1000x1000x50, dx=1 km);
* :func:`init_moist_balanced` -- discretely balanced moist at-rest state
  (reference moist.py:700-785, WRF quarter_ss construction);
* :func:`moist_tracer` -- 3-D stratified flow carrying a seeded qv field,
  0-16 g/kg decaying with height (192x160x50, dx=3 km).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import constants as c
from .config import RunConfig
from .grid import (BaseState, VerticalCoord, gaussian_hill, make_base_state,
                   make_vertical_coord, rebalance_hydrostatic,
                   rebalance_hydrostatic_terrain, surface_extrapolation)


@dataclass
class HostCase:
    cfg: RunConfig
    coord: VerticalCoord
    base: BaseState
    fields: dict = field(default_factory=dict)   # u v w thp php mup [qv..]

    @property
    def cf(self) -> dict:
        return surface_extrapolation(self.coord)

    def zeros_like_state(self):
        nz, ny, nx = self.cfg.nz, self.cfg.ny, self.cfg.nx
        return {"u": np.zeros((nz, ny, nx + 1)), "v": np.zeros((nz, ny + 1, nx)),
                "w": np.zeros((nz + 1, ny, nx)), "thp": np.zeros((nz, ny, nx)),
                "php": np.zeros((nz + 1, ny, nx)), "mup": np.zeros((ny, nx))}


def height_half(base: BaseState) -> np.ndarray:
    """Base-state half-level heights (m): (nz,) flat, (nz,ny,nx) terrain
    (reference state.py:314-316)."""
    return 0.5 * (base.phb[:-1] + base.phb[1:]) / c.G


def cell_x(cfg: RunConfig) -> np.ndarray:
    """Domain-centred cell-centre x coordinates (nx,)."""
    return (np.arange(cfg.nx) + 0.5) * cfg.dx - 0.5 * cfg.nx * cfg.dx


def cell_y(cfg: RunConfig) -> np.ndarray:
    return (np.arange(cfg.ny) + 0.5) * cfg.dy - 0.5 * cfg.ny * cfg.dy


def _same_terrain(base: BaseState, terrain_z) -> None:
    """An explicit ``terrain_z`` is a consistency check against the
    terrain the base state was built over, never a second source of it
    (reference state.py:900-919)."""
    if terrain_z is None:
        return
    if base.terrain_z is None:
        raise ValueError("terrain_z was passed for a base state built "
                         "without terrain; pass the terrain to "
                         "make_base_state instead")
    if np.shape(terrain_z) != np.shape(base.terrain_z) or not np.array_equal(
            np.asarray(terrain_z, np.float64),
            np.asarray(base.terrain_z, np.float64)):
        raise ValueError("terrain_z differs from base.terrain_z, the "
                         "terrain the base state was built over")


def host_at_rest(cfg: RunConfig, coord: VerticalCoord, base: BaseState,
                 terrain_z=None) -> HostCase:
    """Host-side at-rest state: zero perturbations on ``base``."""
    _same_terrain(base, terrain_z)
    case = HostCase(cfg, coord, base)
    case.fields = case.zeros_like_state()
    return case


def init_at_rest(cfg: RunConfig, coord: VerticalCoord, base: BaseState,
                 terrain_z=None, *, device="cuda", dtype=None):
    """Device state with zero perturbations on ``base`` (reference
    state.py:922-932)."""
    from .state import DTYPE, DomainState
    return DomainState.from_host(host_at_rest(cfg, coord, base, terrain_z),
                                 device=device, dtype=dtype or DTYPE)


def init_theta_perturbation(cfg: RunConfig, coord: VerticalCoord,
                            base: BaseState, thp_func, terrain_z=None, *,
                            device="cuda", dtype=None):
    """Device state at rest plus theta' = thp_func(x, z), hydrostatically
    rebalanced; winds zero, mu' = 0 (reference state.py:961-993).  ``x`` is
    the domain-centred cell-centre coordinate (nx,), ``z`` the base
    half-level heights, (nz,) flat or (nz, ny, nx) over terrain."""
    from .state import DTYPE, DomainState
    case = host_theta_perturbation(cfg, coord, base, thp_func, terrain_z)
    return DomainState.from_host(case, device=device, dtype=dtype or DTYPE)


def host_theta_perturbation(cfg: RunConfig, coord: VerticalCoord,
                            base: BaseState, thp_func,
                            terrain_z=None) -> HostCase:
    """At-rest state plus theta' = thp_func(x, z), with phi' chosen so
    every column is discretely hydrostatic (reference state.py:961-993);
    winds zero, mu' = 0."""
    case = host_at_rest(cfg, coord, base, terrain_z)
    z = height_half(base)
    thp = np.asarray(thp_func(cell_x(cfg), z), dtype=np.float64)
    thp = np.broadcast_to(thp, (cfg.nz, cfg.ny, cfg.nx)).copy()
    if base.terrain_z is None:
        th_total = base.thb[:, None, None] + thp
        ph = rebalance_hydrostatic(th_total, base.mub, coord, cfg.p_surf)
        php = ph - base.phb[:, None, None]
    else:
        th_total = base.thb + thp
        php = rebalance_hydrostatic_terrain(th_total, base, coord) - base.phb
    case.fields["thp"] = thp
    case.fields["php"] = php
    return case


def set_periodic_duplicates(fields: dict) -> None:
    """Make the staggered duplicate face column/row equal to face 0."""
    fields["u"][..., -1] = fields["u"][..., 0]
    fields["v"][:, -1, :] = fields["v"][:, 0, :]


def stratified_sounding(n_bv: float = 0.01, theta0: float = 300.0):
    """theta(z) = theta0 exp(N^2 z / g): constant Brunt-Vaisala N."""
    return lambda z: theta0 * np.exp(n_bv * n_bv * np.asarray(z, float) / c.G)


def isentropic_sounding(theta0: float = 300.0):
    return lambda z: np.full_like(np.asarray(z, dtype=np.float64), theta0)


# ---------------------------------------------------------------------------
# Benchmark workloads
# ---------------------------------------------------------------------------

def straka(nx=512, nz=64, dx=100.0, dt=1.0, time_step_sound=6,
           k=75.0) -> HostCase:
    """Straka density current: 300 K isentropic, cold blob T' = -15 K
    (cos(pi L)+1)/2 inside L = sqrt((x/4 km)^2 + ((z-3 km)/2 km)^2) <= 1,
    theta' = T'/Exner_b (reference verify/cases/straka.py:42-75)."""
    cfg = RunConfig(nx=nx, ny=1, nz=nz, dx=dx, dy=dx, ztop=nz * dx, dt=dt,
                    run_seconds=900.0, khdif=k, kvdif=k,
                    time_step_sound=time_step_sound, case="straka")
    coord = make_vertical_coord(nz)
    base = make_base_state(coord, isentropic_sounding(), cfg.p_surf,
                           cfg.ztop)
    exner = (np.asarray(base.pb, dtype=np.float64) / c.P0) ** c.RCP

    def thp(x, z):
        L = np.sqrt((x[None, None, :] / 4000.0) ** 2
                    + ((z[:, None, None] - 3000.0) / 2000.0) ** 2)
        tp = np.where(L <= 1.0, -15.0 * (np.cos(np.pi * L) + 1.0) / 2.0, 0.0)
        return tp / exner[:, None, None]

    return host_theta_perturbation(cfg, coord, base, thp)


def smooth_noise(rng, shape, dx, dy, scale, sigma, dz_levels=None):
    """Seeded white noise smoothed with a Gaussian spectral filter of
    length ``scale`` (m) in x/y (and ``dz_levels`` levels in z), then
    scaled to standard deviation ``sigma``.  Periodic in x and y."""
    noise = rng.standard_normal(shape)
    ny, nx = shape[-2], shape[-1]
    kx = 2.0 * np.pi * np.fft.fftfreq(nx, d=dx)
    ky = 2.0 * np.pi * np.fft.fftfreq(ny, d=dy)
    filt = np.exp(-0.5 * scale ** 2 * (ky[:, None] ** 2 + kx[None, :] ** 2))
    spec = np.fft.fft2(noise, axes=(-2, -1)) * filt
    out = np.real(np.fft.ifft2(spec, axes=(-2, -1)))
    if dz_levels and out.ndim == 3 and out.shape[0] > 1:
        kz = 2.0 * np.pi * np.fft.fftfreq(out.shape[0])
        fz = np.exp(-0.5 * (dz_levels * kz) ** 2)
        out = np.real(np.fft.ifft(np.fft.fft(out, axis=0)
                                  * fz[:, None, None], axis=0))
    out -= out.mean()
    std = out.std()
    return out * (sigma / std if std > 0 else 0.0)


def stratified_hill(nx=250, ny=200, nz=49, dx=12000.0, dt=60.0,
                    time_step_sound=4, seed=20260731, ztop=20000.0,
                    u0=10.0, n_bv=0.01, hill=1000.0, radius=50000.0,
                    damp_opt=3) -> HostCase:
    """3-D dry stratified flow over a Gaussian hill: theta = 300
    exp(N^2 z/g), u = u0 + u', v = v', theta' smooth seeded noise (sigma
    1 m/s, 1 m/s, 0.5 K at a 200 km scale), hybrid coordinate, terrain
    base state, rigid lid with the w damper over the top 5 km."""
    cfg = RunConfig(nx=nx, ny=ny, nz=nz, dx=dx, dy=dx, ztop=ztop, dt=dt,
                    time_step_sound=time_step_sound, terrain_opt=1,
                    hybrid_opt=2, damp_opt=damp_opt, zdamp=5000.0,
                    dampcoef=0.2, case="stratified_hill")
    coord = make_vertical_coord(nz, hybrid_opt=2, etac=0.2)
    tz = gaussian_hill(nx, ny, dx, dx, hill, radius)
    base = make_base_state(coord, stratified_sounding(n_bv), cfg.p_surf,
                           ztop, terrain_z=tz)
    rng = np.random.default_rng(seed)
    th_noise = smooth_noise(rng, (nz, ny, nx), dx, dx, 200e3, 0.5, 3)
    case = host_theta_perturbation(cfg, coord, base, lambda x, z: th_noise)
    u = u0 + smooth_noise(rng, (nz, ny, nx), dx, dx, 200e3, 1.0, 3)
    v = smooth_noise(rng, (nz, ny, nx), dx, dx, 200e3, 1.0, 3)
    case.fields["u"][..., :nx] = u
    case.fields["v"][:, :ny, :] = v
    set_periodic_duplicates(case.fields)
    set_w_surface_host(case)
    return case


def warm_bubbles(nx=1000, ny=1000, nz=50, dx=1000.0, dt=6.0,
                 time_step_sound=4, seed=20260731, ztop=15000.0,
                 n_bubbles=16, amp=3.0, n_bv=0.01, u0=5.0,
                 v0=2.0) -> HostCase:
    """Flat stratified dry state with ``n_bubbles`` seeded +3 K cosine^2
    bubbles (horizontal radius 10 km, vertical 1.5 km, centred at 2 km)
    and a uniform background wind (u0, v0)."""
    cfg = RunConfig(nx=nx, ny=ny, nz=nz, dx=dx, dy=dx, ztop=ztop, dt=dt,
                    time_step_sound=time_step_sound, case="warm_bubbles")
    coord = make_vertical_coord(nz)
    base = make_base_state(coord, stratified_sounding(n_bv), cfg.p_surf, ztop)
    rng = np.random.default_rng(seed)
    xs = rng.uniform(-0.5, 0.5, n_bubbles) * nx * dx
    ys = rng.uniform(-0.5, 0.5, n_bubbles) * ny * dx
    x, y = cell_x(cfg), cell_y(cfg)
    Lx, Lz = 10000.0, 1500.0

    def thp(_x, z):
        out = np.zeros((nz, ny, nx))
        zz = z[:, None, None]
        hw = int(np.ceil(Lx / dx)) + 1          # half-width of the window
        for bx, by in zip(xs, ys):
            ii = np.arange(nx) if 2 * hw + 1 >= nx else (
                np.round((bx + 0.5 * nx * dx) / dx).astype(int)
                + np.arange(-hw, hw + 1)) % nx
            jj = np.arange(ny) if 2 * hw + 1 >= ny else (
                np.round((by + 0.5 * ny * dx) / dx).astype(int)
                + np.arange(-hw, hw + 1)) % ny
            ddx = (x[ii][None, None, :] - bx + 0.5 * nx * dx) % (nx * dx) \
                - 0.5 * nx * dx
            ddy = (y[jj][None, :, None] - by + 0.5 * ny * dx) % (ny * dx) \
                - 0.5 * ny * dx
            L = np.sqrt((ddx / Lx) ** 2 + (ddy / Lx) ** 2
                        + ((zz - 2000.0) / Lz) ** 2)
            out[:, jj[:, None], ii[None, :]] += np.where(
                L < 1.0, amp * np.cos(0.5 * np.pi * L) ** 2, 0.0)
        return out

    case = host_theta_perturbation(cfg, coord, base, thp)
    case.fields["u"][...] = u0
    case.fields["v"][...] = v0
    return case


def ensemble_member(nx, ny, nz, seed, dx=3000.0, dt=18.0,
                    time_step_sound=4, ztop=20000.0, n_bv=0.01, u0=10.0,
                    scale=30000.0) -> HostCase:
    """One ensemble member: flat stratified dry state (theta = 300
    exp(N^2 z/g)), u = u0 + u', v = v', hydrostatically balanced theta',
    seeded smooth noise (sigma 1 m/s, 1 m/s, 0.5 K at ``scale``), rigid
    lid with the w damper over the top 5 km."""
    cfg = RunConfig(nx=nx, ny=ny, nz=nz, dx=dx, dy=dx, ztop=ztop, dt=dt,
                    time_step_sound=time_step_sound, damp_opt=3,
                    zdamp=5000.0, dampcoef=0.2, case="ensemble_member")
    coord = make_vertical_coord(nz)
    base = make_base_state(coord, stratified_sounding(n_bv), cfg.p_surf,
                           ztop)
    rng = np.random.default_rng(seed)
    th = smooth_noise(rng, (nz, ny, nx), dx, dx, scale, 0.5, 3)
    case = host_theta_perturbation(cfg, coord, base, lambda x, z: th)
    case.fields["u"][..., :nx] = u0 + smooth_noise(
        rng, (nz, ny, nx), dx, dx, scale, 1.0, 3)
    case.fields["v"][:, :ny, :] = smooth_noise(
        rng, (nz, ny, nx), dx, dx, scale, 1.0, 3)
    set_periodic_duplicates(case.fields)
    return case


def ensemble_shapes(n=32, seed=20260731, nxy=(96, 256), nzr=(40, 60)):
    """Member extents (nx, ny, nz): nx, ny ~ U{96..256}, nz ~ U{40..60}
    from ``seed``."""
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(nxy[0], nxy[1] + 1)),
             int(rng.integers(nxy[0], nxy[1] + 1)),
             int(rng.integers(nzr[0], nzr[1] + 1))) for _ in range(n)]


def ensemble_members(n=32, seed=20260731, **kw) -> list:
    """The uneven 32-member ensemble: shapes from ensemble_shapes, member
    i seeded with seed + 1 + i."""
    return [ensemble_member(nx, ny, nz, seed + 1 + i, **kw)
            for i, (nx, ny, nz) in enumerate(ensemble_shapes(n, seed))]


def init_moist_balanced(cfg: RunConfig, coord: VerticalCoord,
                        base: BaseState, qv3: np.ndarray,
                        thp: np.ndarray | None = None) -> HostCase:
    """Discretely balanced moist at-rest state on a flat dry base
    (reference moist.py:700-785, WRF module_initialize_ideal.F quarter_ss):
    with mu' = 0, integrate p' down the column by inverting the moist
    pg_buoy_w balance row by row, invert the moist EOS for the dry specific
    volume, then integrate phi' up with the discrete hydrostatic
    recurrence.  ``qv3`` is (nz, ny, nx); ``thp`` (optional) rides on top
    without entering the pressure construction, as in WRF.  FP64."""
    if not cfg.moist:
        raise ValueError("init_moist_balanced requires cfg.moist=True")
    if base.terrain_z is not None:
        raise NotImplementedError("moist balance is built on flat bases only")
    nz, ny, nx = cfg.nz, cfg.ny, cfg.nx
    case = host_at_rest(cfg, coord, base)
    qv3 = np.broadcast_to(np.asarray(qv3, np.float64), (nz, ny, nx))
    thp = np.zeros((nz, ny, nx)) if thp is None else np.broadcast_to(
        np.asarray(thp, np.float64), (nz, ny, nx))
    th = base.thb[:, None, None] + thp
    mub, mup = float(base.mub), 0.0
    c1f, c2f, c1h, c2h = coord.c1f, coord.c2f, coord.c1h, coord.c2h
    p = np.zeros((nz, ny, nx))
    q1 = qv3[nz - 1]
    q2 = 1.0 / (1.0 + q1)
    q1 = q1 * q2
    p[nz - 1] = (-0.5 * (c1f[nz] * mup + q1 * (c1f[nz] * mub + c2f[nz]))
                 / coord.rdnw[nz - 1] / q2)
    for kk in range(nz - 2, -1, -1):
        kw = kk + 1
        q1 = 0.5 * (qv3[kk] + qv3[kk + 1])
        q2 = 1.0 / (1.0 + q1)
        q1 = q1 * q2
        p[kk] = p[kk + 1] - (c1f[kw] * mup + q1 * (c1f[kw] * mub + c2f[kw])) \
            / q2 / coord.rdn[kw]
    qvf = 1.0 + c.RVOVRD * qv3
    alt = (c.RD / c.P0) * th * qvf * ((p + base.pb[:, None, None]) / c.P0) \
        ** (-c.CV / c.CP)
    al = alt - base.alb[:, None, None]
    php = np.zeros((nz + 1, ny, nx))
    for k in range(nz):
        php[k + 1] = php[k] - coord.dnw[k] * (
            ((c1h[k] * mub + c2h[k]) + c1h[k] * mup) * al[k]
            + (c1h[k] * mup) * base.alb[k])
    case.fields["thp"] = np.array(thp)
    case.fields["php"] = php
    for q in ("qv", "qc", "qr"):
        case.fields[q] = np.zeros((nz, ny, nx))
    case.fields["qv"] = np.array(qv3)
    return case


def moist_tracer(nx=192, ny=160, nz=50, dx=3000.0, dt=18.0,
                 time_step_sound=4, seed=20260802, ztop=20000.0,
                 qv_max=0.016, qv_scale_height=2500.0, n_bv=0.01,
                 u0=10.0, v0=5.0, moist_adv_opt=1) -> HostCase:
    """3-D stratified flow carrying a seeded water-vapour field: qv drawn
    uniformly in [0, qv_max] per cell and decaying as exp(-z/H), balanced
    by :func:`init_moist_balanced`, with a uniform wind (u0, v0).
    dt = 18 s follows WRF's 6 s per km of dx."""
    cfg = RunConfig(nx=nx, ny=ny, nz=nz, dx=dx, dy=dx, ztop=ztop, dt=dt,
                    time_step_sound=time_step_sound, moist=True,
                    mp_physics=0, moist_adv_opt=moist_adv_opt,
                    case="moist_tracer")
    coord = make_vertical_coord(nz)
    base = make_base_state(coord, stratified_sounding(n_bv), cfg.p_surf, ztop)
    rng = np.random.default_rng(seed)
    z = height_half(base)
    qv = rng.uniform(0.0, qv_max, (nz, ny, nx)) \
        * np.exp(-z / qv_scale_height)[:, None, None]
    case = init_moist_balanced(cfg, coord, base, qv)
    case.fields["u"][...] = u0
    case.fields["v"][...] = v0
    return case


def add_species(case: HostCase, mp_physics: int, seed: int = 20261203,
                heating: float = 2.0e-3, qv_max: float = 0.0) -> HostCase:
    """Give ``case`` a mixed-phase scheme's species and a retained heating
    (in place; returns the case).  Mass species are drawn in [0, 1e-3]
    kg/kg, number moments in [0, 1e6] /kg (NSSL's CCN qnn about its
    408163264 /kg start), particle volumes in [0, 1e-9] m3/kg; a third of
    the cells of every species hold zero, so the final stage's
    positive-definite limiter acts.  h_diabatic is a seeded field in
    [-heating, heating] K/s.  A dry case becomes moist with qv drawn in
    [0, qv_max] (its balance is the dry one; for throughput runs)."""
    from .config import EXTRA_SPECIES
    cfg = case.cfg
    shape = (cfg.nz, cfg.ny, cfg.nx)
    rng = np.random.default_rng(seed + mp_physics)
    if "qv" not in case.fields:
        case.fields["qv"] = rng.uniform(0.0, qv_max, shape)
    case.cfg = cfg.with_(moist=True, mp_physics=mp_physics)
    for name in ("qc", "qr") + EXTRA_SPECIES[mp_physics]:
        if name == "qnn":
            x = 408163264.0 * rng.uniform(0.5, 1.5, shape)
        elif name.startswith("qvol"):
            x = rng.uniform(0.0, 1.0e-9, shape)
        elif name.startswith("n") or name.startswith("qn"):
            x = rng.uniform(0.0, 1.0e6, shape)
        else:
            x = rng.uniform(0.0, 1.0e-3, shape)
        x[rng.uniform(size=shape) < 1.0 / 3.0] = 0.0
        case.fields[name] = x
    if mp_physics != 0:
        case.fields["h_diabatic"] = rng.uniform(-heating, heating, shape)
    return case


def mixed_phase(mp_physics=10, nx=64, ny=48, nz=30, dx=3000.0, dt=18.0,
                time_step_sound=4, seed=20261203, ztop=20000.0,
                moist_adv_opt=1, heating=2.0e-3) -> HostCase:
    """:func:`moist_tracer` carrying a mixed-phase scheme's species and a
    retained microphysics heating (:func:`add_species`)."""
    case = moist_tracer(nx=nx, ny=ny, nz=nz, dx=dx, dt=dt,
                        time_step_sound=time_step_sound, seed=seed, ztop=ztop,
                        moist_adv_opt=moist_adv_opt)
    case = add_species(case, mp_physics, seed=seed, heating=heating)
    case.cfg = case.cfg.with_(case="mixed_phase")
    return case


def set_w_surface_host(case: HostCase) -> None:
    """w(sfc) = u . grad(ht) on the host (reference dycore.py:1999-2026)."""
    cfg, f = case.cfg, case.fields
    cf = case.cf
    ht = (np.zeros((cfg.ny, cfg.nx)) if case.base.terrain_z is None
          else case.base.terrain_z)
    uc = cf["cf1"] * f["u"][0] + cf["cf2"] * f["u"][1] + cf["cf3"] * f["u"][2]
    vc = cf["cf1"] * f["v"][0] + cf["cf2"] * f["v"][1] + cf["cf3"] * f["v"][2]
    dyn = np.roll(ht, -1, 0) - ht
    dys = ht - np.roll(ht, 1, 0)
    dxe = np.roll(ht, -1, 1) - ht
    dxw = ht - np.roll(ht, 1, 1)
    f["w"][0] = ((0.5 / cfg.dy) * (dyn * vc[1:] + dys * vc[:-1])
                 + (0.5 / cfg.dx) * (dxe * uc[:, 1:] + dxw * uc[:, :-1]))
