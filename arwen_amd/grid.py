"""Host (FP64) setup: the eta coordinate and a discretely hydrostatic base
state.

What this module produces is part of the drop-in contract: the WRF
coordinate arrays ``znw, znu, dnw, rdnw, dn, rdn, fnm, fnp``, the hybrid
coefficients ``c1..c4`` on full (``f``) and half (``h``) levels, and the base
profiles ``thb, pb, alb, phb`` with the column dry mass ``mub`` (reference
gpuwm/core/grid.py:100-147, :351-409, :435-549, :552-569).  Their
definitions are WRF's, so the arithmetic below follows WRF's recurrences
and evaluation order; everything is float64 and is rounded to float32
once, when a state loads it.

Conventions: ``znw[0] = 1`` at the surface and ``znw[nz] = 0`` at the top,
so ``dnw[k] = znw[k+1] - znw[k] < 0``; ``dn, rdn, fnm, fnp`` live on the
interior full levels k >= 1 and are zero at k = 0.

The pressure-weighted hybrid coefficients (``c2``, ``c4``) carry a factor
``P0 - p_top``, and p_top is only known once a sounding is integrated to
the model top.  A coordinate therefore starts *unbound* (those
coefficients zero) and is bound to a model top by the first base state
built on it; see :meth:`VerticalCoord.bind_model_top`.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable

import numpy as np

from . import constants as c

Sounding = Callable[[np.ndarray], np.ndarray]

# hybrid_opt values whose B(eta) is eta itself (c1 = 1, c2 = c4 = 0)
_TERRAIN_FOLLOWING = (0, 1)


# ---------------------------------------------------------------------------
# Hybrid coefficients
# ---------------------------------------------------------------------------

def hybrid_b_poly(etac: float) -> np.ndarray:
    """Coefficients (b0, b1, b2, b3) of WRF's cubic B(eta): B(1) = 1,
    B'(1) = 1, B(etac) = 0, B'(etac) = 0."""
    e = float(etac)
    conditions = np.array([[1.0, 1.0, 1.0, 1.0],          # B(1)
                           [0.0, 1.0, 2.0, 3.0],          # B'(1)
                           [1.0, e, e * e, e ** 3],       # B(etac)
                           [0.0, 1.0, 2.0 * e, 3.0 * e * e]])  # B'(etac)
    return np.linalg.solve(conditions, np.array([1.0, 1.0, 0.0, 0.0]))


def _b_of_eta(znw: np.ndarray, hybrid_opt: int, etac: float) -> np.ndarray:
    """c3 on full levels: B(eta) = eta, or the cubic clipped to 0 above
    etac with its end values pinned."""
    if hybrid_opt in _TERRAIN_FOLLOWING:
        return znw.copy()
    if hybrid_opt != 2:
        raise ValueError(f"unknown hybrid_opt {hybrid_opt!r}: this "
                         "coordinate is terrain-following (0 or 1) or WRF's "
                         "cubic hybrid (2)")
    b0, b1, b2, b3 = hybrid_b_poly(etac)
    b = b0 + b1 * znw + b2 * znw ** 2 + b3 * znw ** 3
    b[znw < etac] = 0.0
    b[0], b[-1] = 1.0, 0.0
    return b


def compute_hybrid_coeffs(znw, hybrid_opt: int, etac: float, p0: float,
                          pt: float) -> dict:
    """WRF v4 hybrid coefficients on full and half levels (reference
    grid.py:100-147): c3 = B(eta), c4 = (eta - B)(p0 - pt), c1 = dB/deta by
    WRF's differences (half levels from full-level B, interior full levels
    from half-level B, surface 1, top 1 or 0), c2 = (1 - c1)(p0 - pt)."""
    znw = np.asarray(znw, dtype=np.float64)
    c3f = _b_of_eta(znw, hybrid_opt, etac)
    span = p0 - pt
    znu = 0.5 * (znw[:-1] + znw[1:])
    c3h = 0.5 * (c3f[:-1] + c3f[1:])
    c1h = np.diff(c3f) / np.diff(znw)
    c1f = np.empty_like(znw)
    c1f[0] = 1.0
    c1f[1:-1] = np.diff(c3h) / np.diff(znu)
    c1f[-1] = 1.0 if hybrid_opt in _TERRAIN_FOLLOWING else 0.0
    return {"c1f": c1f, "c2f": (1.0 - c1f) * span, "c3f": c3f,
            "c4f": (znw - c3f) * span,
            "c1h": c1h, "c2h": (1.0 - c1h) * span, "c3h": c3h,
            "c4h": (znu - c3h) * span}


# ---------------------------------------------------------------------------
# The coordinate
# ---------------------------------------------------------------------------

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
    p_top: float | None = None         # model top the c2/c4 are bound to
    c1f: np.ndarray | None = None
    c2f: np.ndarray | None = None
    c3f: np.ndarray | None = None
    c4f: np.ndarray | None = None
    c1h: np.ndarray | None = None
    c2h: np.ndarray | None = None
    c3h: np.ndarray | None = None
    c4h: np.ndarray | None = None

    def _set_hybrid(self, pt: float) -> None:
        for name, arr in compute_hybrid_coeffs(
                self.znw, self.hybrid_opt, self.etac, c.P0, pt).items():
            setattr(self, name, arr)

    def bind_model_top(self, p_top: float) -> None:
        """Fix the pressure-weighted coefficients for model top ``p_top``.

        The first binding wins.  With B(eta) = eta those coefficients are
        zero whatever the top, so a later base state with another top may
        share the coordinate (the recorded top follows it).  With the cubic
        they are not: states already built on the coordinate carry the
        first top's c2/c4, so a second top is refused rather than silently
        changing them."""
        p_top = float(p_top)
        if self.p_top is None:
            self._set_hybrid(p_top)
            self.p_top = p_top
            return
        if self.p_top == p_top:
            return
        if self.hybrid_opt not in _TERRAIN_FOLLOWING:
            raise ValueError(
                f"this cubic hybrid coordinate is bound to a model top of "
                f"{self.p_top} Pa, and its c2/c4 already feed the states "
                f"built on it; a base state topping out at {p_top} Pa needs "
                "a coordinate of its own (make_vertical_coord)")
        self.p_top = p_top


def _full_eta_levels(nz: int, stretch, eta_levels) -> np.ndarray:
    """znw: explicit levels, tanh-clustered toward the surface, or
    uniform."""
    if eta_levels is not None and stretch is not None:
        raise ValueError("give either explicit eta_levels or a stretch "
                         "factor, not both")
    if eta_levels is not None:
        znw = np.array(eta_levels, dtype=np.float64)
        ok = (znw.shape == (nz + 1,) and np.all(np.isfinite(znw))
              and znw[0] == 1.0 and znw[-1] == 0.0
              and np.all(np.diff(znw) < 0.0))
        if not ok:
            raise ValueError(
                f"eta_levels must be {nz + 1} finite values running "
                "strictly downward from 1.0 (surface) to 0.0 (top); got "
                f"shape {znw.shape}")
        return znw
    if stretch is None:
        return np.linspace(1.0, 0.0, nz + 1)
    if not stretch > 0.0:
        raise ValueError(f"stretch must be a positive clustering factor, "
                         f"got {stretch!r}")
    # zeta: 0 at the surface, 1 at the top; thinnest layers at the surface
    zeta = np.linspace(0.0, 1.0, nz + 1)
    znw = np.tanh(stretch * (1.0 - zeta)) / np.tanh(stretch)
    znw[0], znw[-1] = 1.0, 0.0
    return znw


def make_vertical_coord(nz: int, stretch: float | None = None,
                        hybrid_opt: int = 0, etac: float = 0.2,
                        eta_levels=None) -> VerticalCoord:
    """Eta coordinate with nz half levels (reference grid.py:351-409),
    unbound to a model top (its c2/c4 are zero until a base state binds
    it)."""
    znw = _full_eta_levels(nz, stretch, eta_levels)
    dnw = znw[1:] - znw[:-1]
    dn, rdn = np.zeros(nz), np.zeros(nz)
    fnp, fnm = np.zeros(nz), np.zeros(nz)
    dn[1:] = 0.5 * (dnw[1:] + dnw[:-1])
    rdn[1:] = 1.0 / dn[1:]
    fnp[1:] = 0.5 * dnw[1:] / dn[1:]
    fnm[1:] = 0.5 * dnw[:-1] / dn[1:]
    coord = VerticalCoord(znw=znw, znu=0.5 * (znw[:-1] + znw[1:]), dnw=dnw,
                          rdnw=1.0 / dnw, dn=dn, rdn=rdn, fnp=fnp, fnm=fnm,
                          hybrid_opt=hybrid_opt, etac=etac)
    coord._set_hybrid(c.P0)          # P0 - pt = 0: c2/c4 start at zero
    return coord


# ---------------------------------------------------------------------------
# Base state
# ---------------------------------------------------------------------------

@dataclass
class BaseState:
    mub: float | np.ndarray           # column dry mass p_s - p_top
    p_top: float
    pb: np.ndarray                    # half-level dry pressure
    alb: np.ndarray                   # inverse density
    thb: np.ndarray                   # potential temperature
    phb: np.ndarray                   # full-level geopotential
    terrain_z: np.ndarray | None = None

    @property
    def t_init(self) -> np.ndarray:
        """WRF's name for the base theta profile."""
        return self.thb


@dataclass
class _FineColumn:
    """The sounding integrated hydrostatically on a 1 m grid."""
    z: np.ndarray        # heights, ztop appended
    exner: np.ndarray    # Exner function at z
    p: np.ndarray        # pressure at z

    def surface_pressure(self, height):
        """p at ``height``: the Exner function is piecewise linear between
        the fine-grid nodes, so interpolating it is exact."""
        return c.P0 * np.interp(height, self.z, self.exner) ** (1.0 / c.RCP)

    def height_of(self, p):
        return np.interp(p, self.p[::-1], self.z[::-1])


def _integrate_sounding(sounding: Sounding, p_surf: float,
                        ztop: float) -> _FineColumn:
    """exner[n+1] = exner[n] - g dz / (cp theta(z_mid)) up to ztop."""
    z = np.append(np.arange(0.0, ztop, 1.0), ztop)
    dz = np.diff(z)
    zmid = 0.5 * (z[:-1] + z[1:])
    ex0 = (p_surf / c.P0) ** c.RCP
    exner = np.empty(z.size)
    exner[0] = ex0
    exner[1:] = ex0 - np.cumsum(c.G * dz / (c.CP * sounding(zmid)))
    return _FineColumn(z, exner, c.P0 * exner ** (1.0 / c.RCP))


def _levels3(a, mub):
    """A coordinate profile against a scalar or (ny, nx) column mass."""
    return a if np.ndim(mub) == 0 else a[:, None, None]


def _terrain_field(terrain_z, ztop) -> np.ndarray:
    tz = np.array(terrain_z, dtype=np.float64)   # own copy
    if tz.ndim != 2:
        raise ValueError(f"terrain_z is a surface height field (ny, nx); "
                         f"got {tz.ndim} dimension(s)")
    lo, hi = float(tz.min()), float(tz.max())
    if lo < 0.0 or hi >= ztop:
        raise ValueError(f"terrain heights span [{lo}, {hi}] m; they must "
                         f"stay at or above 0 and below the model top "
                         f"{ztop} m")
    return tz


def make_base_state(coord: VerticalCoord, sounding: Sounding, p_surf: float,
                    ztop: float, terrain_z=None) -> BaseState:
    """Discretely hydrostatic base state (reference grid.py:435-549).

    The sounding is integrated on a 1 m grid to give p(z) and p_top =
    p(ztop), which binds ``coord``.  The column dry mass is mub = p_s -
    p_top (per column over terrain, p_s at the terrain height); half-level
    dry pressure pb = c3h mub + c4h + p_top; thb = sounding(z(pb)); alb =
    RD thb (pb/P0)^RCP / pb; and phb from phb[k+1] = phb[k] - dnw[k]
    (c1h[k] mub + c2h[k]) alb[k], starting from g * terrain height.
    """
    col = _integrate_sounding(sounding, p_surf, ztop)
    p_top = float(col.p[-1])
    coord.bind_model_top(p_top)
    if terrain_z is None:
        mub = float(p_surf - p_top)
        phb = np.zeros(coord.znw.size)
    else:
        terrain_z = _terrain_field(terrain_z, ztop)
        mub = col.surface_pressure(terrain_z) - p_top
        phb = np.zeros((coord.znw.size,) + terrain_z.shape)
        phb[0] = c.G * terrain_z
    mb = mub if np.ndim(mub) == 0 else mub[None]
    pb = _levels3(coord.c3h, mub) * mb + _levels3(coord.c4h, mub) + p_top
    dry_full = _levels3(coord.c3f, mub) * mb + _levels3(coord.c4f, mub) \
        + p_top
    if not np.all(np.diff(dry_full, axis=0) < 0.0):
        raise ValueError(
            f"the cubic hybrid coordinate folds over this terrain: the "
            f"reference dry pressure stops decreasing upward somewhere "
            f"(hybrid_opt={coord.hybrid_opt}, etac={coord.etac}); a smaller "
            "etac keeps the columns ordered")
    thb = sounding(col.height_of(pb))
    alb = c.RD * thb * (pb / c.P0) ** c.RCP / pb
    for k in range(coord.dnw.size):
        phb[k + 1] = phb[k] - coord.dnw[k] * (coord.c1h[k] * mub
                                              + coord.c2h[k]) * alb[k]
    return BaseState(mub=mub, p_top=p_top, pb=pb, alb=alb, thb=thb, phb=phb,
                     terrain_z=terrain_z)


def _balanced_geopotential(th_total, p, dmass, surface) -> np.ndarray:
    """phi[0] = surface, phi[k+1] = phi[k] - dnw-weighted mass * alpha with
    alpha = RD theta (p/P0)^RCP / p, per column; ``dmass[k]`` is the
    layer's dnw[k] * column-mass increment."""
    nz = th_total.shape[0]
    alpha = c.RD * th_total * (p / c.P0) ** c.RCP / p
    ph = np.zeros((nz + 1,) + th_total.shape[1:])
    ph[0] = surface
    for k in range(nz):
        ph[k + 1] = ph[k] - dmass[k] * alpha[k]
    return ph


def rebalance_hydrostatic(th_total: np.ndarray, mub: float,
                          coord: VerticalCoord, p_surf: float) -> np.ndarray:
    """Balanced total geopotential (nz+1, ny, nx) for a perturbed theta on
    a flat base (reference grid.py:552-569): pressure znu mub + p_top, layer
    mass dnw mub."""
    p_top = p_surf - mub
    p = (coord.znu * mub + p_top)[:, None, None]
    dmass = [coord.dnw[k] * mub for k in range(coord.dnw.size)]
    return _balanced_geopotential(th_total, p, dmass, 0.0)


def rebalance_hydrostatic_terrain(th_total: np.ndarray, base: BaseState,
                                  coord: VerticalCoord) -> np.ndarray:
    """Per-column balanced geopotential over terrain (reference
    state.py:936-958): pressure c3h mub + c4h + p_top, layer mass dnw (c1h
    mub + c2h), surface at g * terrain height."""
    mub = np.asarray(base.mub, dtype=np.float64)
    p = (coord.c3h[:, None, None] * mub[None] + coord.c4h[:, None, None]
         + base.p_top)
    dmass = [coord.dnw[k] * (coord.c1h[k] * mub + coord.c2h[k])
             for k in range(coord.dnw.size)]
    return _balanced_geopotential(th_total, p, dmass, c.G * base.terrain_z)


def surface_extrapolation(coord: VerticalCoord) -> dict:
    """cf1..cf3 (lowest three half levels to the surface) and cfn, cfn1
    (top) extrapolation weights (reference state.py:757-771, WRF
    start_em)."""
    w = dict.fromkeys(("cf1", "cf2", "cf3", "cfn", "cfn1"), 0.0)
    dn, dnw, fnp, fnm = coord.dn, coord.dnw, coord.fnp, coord.fnm
    if dnw.size >= 3:
        cof1 = (2.0 * dn[1] + dn[2]) / (dn[1] + dn[2]) * dnw[0] / dn[1]
        cof2 = dn[1] / (dn[1] + dn[2]) * dnw[0] / dn[2]
        w.update(cf1=fnp[1] + cof1, cf2=fnm[1] - cof1 - cof2, cf3=cof2)
    if dnw.size >= 1:
        w.update(cfn=1.0 + fnp[-1], cfn1=-fnp[-1])
    return w


# ---------------------------------------------------------------------------
# Terrain shapes
# ---------------------------------------------------------------------------

def _centred(n: int, d: float) -> np.ndarray:
    return (np.arange(n) + 0.5) * d - 0.5 * n * d


def bell_hill(nx: int, ny: int, dx: float, height: float,
              halfwidth: float) -> np.ndarray:
    """Witch-of-Agnesi ridge h / (1 + (x/a)^2), uniform in y (reference
    terrain.py:17-27)."""
    h = height / (1.0 + (_centred(nx, dx) / halfwidth) ** 2)
    return np.broadcast_to(h, (ny, nx)).copy()


def gaussian_hill(nx: int, ny: int, dx: float, dy: float, height: float,
                  radius: float) -> np.ndarray:
    """Domain-centred Gaussian hill h exp(-r^2/a^2) on cell centres."""
    r2 = _centred(nx, dx)[None, :] ** 2 + _centred(ny, dy)[:, None] ** 2
    return height * np.exp(-r2 / radius ** 2)
