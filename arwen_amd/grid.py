"""Host (FP64) setup of the eta coordinate and the discretely hydrostatic
base state.

Restates, in this package's own code, the reference setup path
(gpuwm/core/grid.py:100-147 ``compute_hybrid_coeffs``, :351-409
``make_vertical_coord``, :435-549 ``make_base_state``, :552-569
``rebalance_hydrostatic``).  The arrays it produces are part of the drop-in
contract (names and meanings of WRF's ``znw, dnw, rdnw, dn, rdn, fnm, fnp,
c1h..c4f``), so the construction steps follow WRF exactly; everything is
float64 and is cast to float32 once, when loaded into a state.

Conventions: ``znw[0] = 1`` at the surface, ``znw[nz] = 0`` at the top,
``dnw[k] = znw[k+1] - znw[k] < 0``; ``dn, rdn, fnm, fnp`` are defined for
k >= 1 and zero at k = 0.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable

import numpy as np

from . import constants as c

Sounding = Callable[[np.ndarray], np.ndarray]


@dataclass
class VerticalCoord:
    znw: np.ndarray
    znu: np.ndarray
    dnw: np.ndarray
    rdnw: np.ndarray
    dn: np.ndarray
    rdn: np.ndarray
    fnp: np.ndarray
    fnm: np.ndarray
    hybrid_opt: int = 0
    etac: float = 0.2
    p_top: float | None = None
    c1f: np.ndarray | None = None
    c2f: np.ndarray | None = None
    c3f: np.ndarray | None = None
    c4f: np.ndarray | None = None
    c1h: np.ndarray | None = None
    c2h: np.ndarray | None = None
    c3h: np.ndarray | None = None
    c4h: np.ndarray | None = None


def hybrid_b_poly(etac: float) -> np.ndarray:
    """Cubic B(eta) = b0 + b1 eta + b2 eta^2 + b3 eta^3 with B(1)=1,
    B'(1)=1, B(etac)=0, B'(etac)=0 (WRF v4 hybrid coordinate)."""
    e = float(etac)
    m = np.array([[1.0, 1.0, 1.0, 1.0],
                  [0.0, 1.0, 2.0, 3.0],
                  [1.0, e, e * e, e ** 3],
                  [0.0, 1.0, 2.0 * e, 3.0 * e * e]])
    return np.linalg.solve(m, np.array([1.0, 1.0, 0.0, 0.0]))


def compute_hybrid_coeffs(znw, hybrid_opt: int, etac: float, p0: float,
                          pt: float) -> dict:
    """WRF v4 c1..c4 on full and half levels (reference grid.py:100-147).

    c3 = B(eta); c4 = (eta - B)(p0 - pt); c1 = dB/deta by WRF's finite
    differences; c2 = (1 - c1)(p0 - pt).
    """
    znw = np.asarray(znw, dtype=np.float64)
    if hybrid_opt in (0, 1):
        c3f = znw.copy()
    elif hybrid_opt == 2:
        b = hybrid_b_poly(etac)
        c3f = b[0] + b[1] * znw + b[2] * znw ** 2 + b[3] * znw ** 3
        c3f[znw < etac] = 0.0
        c3f[0] = 1.0
        c3f[-1] = 0.0
    else:
        raise ValueError(
            f"hybrid_opt must be 0/1 (B=eta) or 2 (Klemp cubic), got "
            f"{hybrid_opt}")
    dp = p0 - pt
    znu = 0.5 * (znw[:-1] + znw[1:])
    c3h = 0.5 * (c3f[:-1] + c3f[1:])
    c4f = (znw - c3f) * dp
    c4h = (znu - c3h) * dp
    c1f = np.empty_like(znw)
    c1f[1:-1] = np.diff(c3h) / np.diff(znu)
    c1f[0] = 1.0
    c1f[-1] = 1.0 if hybrid_opt in (0, 1) else 0.0
    c1h = np.diff(c3f) / np.diff(znw)
    return {"c1f": c1f, "c2f": (1.0 - c1f) * dp, "c3f": c3f, "c4f": c4f,
            "c1h": c1h, "c2h": (1.0 - c1h) * dp, "c3h": c3h, "c4h": c4h}


def _install_hybrid(coord: VerticalCoord, pt: float) -> None:
    for name, arr in compute_hybrid_coeffs(coord.znw, coord.hybrid_opt,
                                           coord.etac, c.P0, pt).items():
        setattr(coord, name, arr)


def make_vertical_coord(nz: int, stretch: float | None = None,
                        hybrid_opt: int = 0, etac: float = 0.2,
                        eta_levels=None) -> VerticalCoord:
    """Eta coordinate with nz half levels (reference grid.py:351-409):
    uniform by default, tanh-clustered toward the surface with ``stretch``,
    or explicit ``eta_levels``."""
    if eta_levels is not None:
        if stretch is not None:
            raise ValueError("stretch and eta_levels are mutually exclusive")
        znw = np.array(eta_levels, dtype=np.float64)
        if znw.shape != (nz + 1,):
            raise ValueError(f"eta_levels must have shape ({nz + 1},)")
        if znw[0] != 1.0 or znw[-1] != 0.0 or not np.all(np.diff(znw) < 0):
            raise ValueError("eta_levels must decrease strictly from 1 to 0")
    elif stretch is None:
        znw = np.linspace(1.0, 0.0, nz + 1)
    else:
        if stretch <= 0.0:
            raise ValueError(f"stretch must be positive, got {stretch}")
        zeta = np.linspace(0.0, 1.0, nz + 1)
        znw = np.tanh(stretch * (1.0 - zeta)) / np.tanh(stretch)
        znw[0], znw[-1] = 1.0, 0.0
    dnw = znw[1:] - znw[:-1]
    dn = np.zeros(nz)
    dn[1:] = 0.5 * (dnw[1:] + dnw[:-1])
    rdn = np.zeros(nz)
    rdn[1:] = 1.0 / dn[1:]
    fnp = np.zeros(nz)
    fnm = np.zeros(nz)
    fnp[1:] = 0.5 * dnw[1:] / dn[1:]
    fnm[1:] = 0.5 * dnw[:-1] / dn[1:]
    coord = VerticalCoord(znw=znw, znu=0.5 * (znw[:-1] + znw[1:]), dnw=dnw,
                          rdnw=1.0 / dnw, dn=dn, rdn=rdn, fnp=fnp, fnm=fnm,
                          hybrid_opt=hybrid_opt, etac=etac)
    _install_hybrid(coord, pt=c.P0)      # c2/c4 finalised by the base state
    return coord


@dataclass
class BaseState:
    mub: float | np.ndarray
    p_top: float
    pb: np.ndarray
    alb: np.ndarray
    thb: np.ndarray
    phb: np.ndarray
    terrain_z: np.ndarray | None = None


def _exner_profile(sounding: Sounding, p_surf: float, ztop: float):
    """Exner function on a 1 m grid (ztop appended): pi(z) integrated as
    pi[n+1] = pi[n] - g dz / (cp theta(z_mid))."""
    zf = np.append(np.arange(0.0, ztop, 1.0), ztop)
    dz = np.diff(zf)
    zmid = 0.5 * (zf[:-1] + zf[1:])
    pis = (p_surf / c.P0) ** c.RCP
    pi = np.empty(zf.size)
    pi[0] = pis
    pi[1:] = pis - np.cumsum(c.G * dz / (c.CP * sounding(zmid)))
    return zf, pi, c.P0 * pi ** (1.0 / c.RCP)


def make_base_state(coord: VerticalCoord, sounding: Sounding, p_surf: float,
                    ztop: float, terrain_z=None) -> BaseState:
    """Discretely hydrostatic base state (reference grid.py:435-549).

    p(z) from the fine-grid Exner integration of ``sounding``; p_top =
    p(ztop); dry mass mub = p_s - p_top (2-D with terrain); half-level dry
    pressure pb = c3h*mub + c4h + p_top; thb = sounding(z(pb)); alb =
    RD thb (pb/P0)^RCP / pb; and phb from the load-bearing recurrence
    phb[k+1] = phb[k] - dnw[k] (c1h[k] mub + c2h[k]) alb[k].
    """
    zf, pi, p_of_z = _exner_profile(sounding, p_surf, ztop)
    p_top = float(p_of_z[-1])
    if coord.p_top is None:
        _install_hybrid(coord, pt=p_top)
        coord.p_top = p_top
    elif coord.p_top != p_top:
        if coord.hybrid_opt not in (0, 1):
            raise ValueError(
                f"VerticalCoord already finalized for p_top={coord.p_top} Pa;"
                f" this base state has p_top={p_top} Pa")
        coord.p_top = p_top
    c1h, c2h, c3h, c4h = coord.c1h, coord.c2h, coord.c3h, coord.c4h
    if terrain_z is None:
        mub = float(p_surf - p_top)
        pb = c3h * mub + c4h + p_top
        phb = np.zeros(coord.znw.size)
        pd_full = coord.c3f * mub + coord.c4f + p_top
    else:
        terrain_z = np.array(terrain_z, dtype=np.float64)
        if terrain_z.ndim != 2:
            raise ValueError("terrain_z must be 2-D (ny, nx)")
        if terrain_z.min() < 0.0 or terrain_z.max() >= ztop:
            raise ValueError("terrain_z must lie in [0, ztop)")
        p_s = c.P0 * np.interp(terrain_z, zf, pi) ** (1.0 / c.RCP)
        mub = p_s - p_top
        pb = c3h[:, None, None] * mub[None] + c4h[:, None, None] + p_top
        phb = np.zeros((coord.znw.size,) + terrain_z.shape)
        phb[0] = c.G * terrain_z
        pd_full = (coord.c3f[:, None, None] * mub[None]
                   + coord.c4f[:, None, None] + p_top)
    if not np.all(np.diff(pd_full, axis=0) < 0.0):
        raise ValueError(
            "hybrid reference dry pressure is not monotonically decreasing "
            f"(hybrid_opt={coord.hybrid_opt}, etac={coord.etac}): terrain "
            "too high for this etac - reduce etac")
    z_of_pb = np.interp(pb, p_of_z[::-1], zf[::-1])
    thb = sounding(z_of_pb)
    alb = c.RD * thb * (pb / c.P0) ** c.RCP / pb
    for k in range(coord.dnw.size):
        phb[k + 1] = phb[k] - coord.dnw[k] * (c1h[k] * mub + c2h[k]) * alb[k]
    return BaseState(mub=mub, p_top=p_top, pb=pb, alb=alb, thb=thb, phb=phb,
                     terrain_z=terrain_z)


def rebalance_hydrostatic(th_total: np.ndarray, mub: float,
                          coord: VerticalCoord, p_surf: float) -> np.ndarray:
    """Balanced total geopotential (nz+1, ny, nx) for a perturbed theta on
    a flat base (reference grid.py:552-569): alpha from p = znu*mub +
    p_top, then the discrete recurrence with the column mass mub."""
    nz = th_total.shape[0]
    p_top = p_surf - mub
    p = (coord.znu * mub + p_top)[:, None, None]
    alpha = c.RD * th_total * (p / c.P0) ** c.RCP / p
    ph = np.zeros((nz + 1,) + th_total.shape[1:])
    for k in range(nz):
        ph[k + 1] = ph[k] - coord.dnw[k] * mub * alpha[k]
    return ph


def rebalance_hydrostatic_terrain(th_total: np.ndarray, base: BaseState,
                                  coord: VerticalCoord) -> np.ndarray:
    """Per-column balanced geopotential over terrain (reference
    state.py:922-958): pd = c3h*mub + c4h + p_top, increments c1h*mub+c2h,
    surface pinned at g*terrain_z."""
    nz = th_total.shape[0]
    mub = np.asarray(base.mub, dtype=np.float64)
    p = (coord.c3h[:, None, None] * mub[None] + coord.c4h[:, None, None]
         + base.p_top)
    alpha = c.RD * th_total * (p / c.P0) ** c.RCP / p
    ph = np.zeros((nz + 1,) + th_total.shape[1:])
    ph[0] = c.G * base.terrain_z
    for k in range(nz):
        ph[k + 1] = ph[k] - coord.dnw[k] * (coord.c1h[k] * mub
                                            + coord.c2h[k]) * alpha[k]
    return ph


def surface_extrapolation(coord: VerticalCoord) -> dict:
    """cf1..cf3 (surface) and cfn, cfn1 (top) extrapolation weights
    (reference state.py:757-771, WRF start_em)."""
    out = {"cf1": 0.0, "cf2": 0.0, "cf3": 0.0, "cfn": 0.0, "cfn1": 0.0}
    dn, dnw, fnp, fnm = coord.dn, coord.dnw, coord.fnp, coord.fnm
    if dnw.size >= 3:
        cof1 = (2.0 * dn[1] + dn[2]) / (dn[1] + dn[2]) * dnw[0] / dn[1]
        cof2 = dn[1] / (dn[1] + dn[2]) * dnw[0] / dn[2]
        out["cf1"] = fnp[1] + cof1
        out["cf2"] = fnm[1] - cof1 - cof2
        out["cf3"] = cof2
    if dnw.size >= 1:
        out["cfn"] = 1.0 + fnp[-1]
        out["cfn1"] = -fnp[-1]
    return out


def bell_hill(nx: int, ny: int, dx: float, height: float,
              halfwidth: float) -> np.ndarray:
    """Witch-of-Agnesi ridge on domain-centred cell centres, uniform in y
    (reference terrain.py:17-27)."""
    x = (np.arange(nx) + 0.5) * dx - 0.5 * nx * dx
    h = height / (1.0 + (x / halfwidth) ** 2)
    return np.broadcast_to(h, (ny, nx)).copy()


def gaussian_hill(nx: int, ny: int, dx: float, dy: float, height: float,
                  radius: float) -> np.ndarray:
    """Domain-centred Gaussian hill h exp(-r^2/a^2) on cell centres."""
    x = (np.arange(nx) + 0.5) * dx - 0.5 * nx * dx
    y = (np.arange(ny) + 0.5) * dy - 0.5 * ny * dy
    r2 = x[None, :] ** 2 + y[:, None] ** 2
    return height * np.exp(-r2 / radius ** 2)
