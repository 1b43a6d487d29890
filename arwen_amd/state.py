"""Domain state on PyTorch-ROCm tensors, laid out like the reference.

``DomainState`` owns every array the step touches, with the reference's
names, shapes and dtype (gpuwm/core/state.py:359-695): layout ``(nz, ny,
nx)`` x-fastest, ``u (nz,ny,nx+1)``, ``v (nz,ny+1,nx)``, ``w, php
(nz+1,ny,nx)``, FP32.  The arrays live in HBM as plain contiguous tensors;
the HIP library receives raw pointers through :meth:`domain_struct`.
"""
from __future__ import annotations

import numpy as np
import torch

from . import constants as c
from ._lib import (DIFF_ARRAYS, DOMAIN_ARRAYS, MAX_EXTRA, MOIST_ARRAYS,
                   MOIST_SCRATCH, SCRATCH_ARRAYS, ArwDomain, ArwDomain64)
from .config import EXTRA_SPECIES, LOADING_EXTRA, RunConfig
from .grid import BaseState, VerticalCoord, surface_extrapolation

DTYPE = torch.float32          # the step's precision; FP64 is the reference build

_COORD = ("dnw", "rdnw", "dn", "rdn", "fnp", "fnm", "znu", "znw",
          "c1h", "c2h", "c1f", "c2f", "c3h", "c4h", "c3f", "c4f")


def mu_at_u_faces(mu: torch.Tensor) -> torch.Tensor:
    """(ny, nx) -> (ny, nx+1) periodic face average (face f between cells
    f-1 and f; face nx duplicates face 0) (reference state.py:339-346)."""
    mux = 0.5 * (mu + torch.roll(mu, 1, dims=1))
    return torch.cat([mux, mux[:, :1]], dim=1)


def mu_at_v_faces(mu: torch.Tensor) -> torch.Tensor:
    """(ny, nx) -> (ny+1, nx) (reference state.py:349-356)."""
    muy = 0.5 * (mu + torch.roll(mu, 1, dims=0))
    return torch.cat([muy, muy[:1]], dim=0)


class DomainState:
    """All prognostic, diagnostic, tendency, acoustic, base and coordinate
    arrays of one domain (reference state.py:359-695), on ``device``."""

    def __init__(self, cfg: RunConfig, device="cuda", dtype=DTYPE):
        nz, ny, nx = cfg.nz, cfg.ny, cfg.nx
        self.cfg = cfg
        self.device = torch.device(device)
        if dtype not in (torch.float32, torch.float64):
            raise ValueError("the step runs in float32 (or float64, the "
                             f"reference build), not {dtype}")
        self.dtype = dtype
        z = lambda *sh: torch.zeros(sh, dtype=dtype, device=self.device)
        self.u, self.v, self.w = z(nz, ny, nx + 1), z(nz, ny + 1, nx), \
            z(nz + 1, ny, nx)
        self.thp, self.php, self.mup = z(nz, ny, nx), z(nz + 1, ny, nx), \
            z(ny, nx)
        self.p, self.al, self.alt = z(nz, ny, nx), z(nz, ny, nx), \
            z(nz, ny, nx)
        if cfg.moist:
            for q in ("qv", "qc", "qr"):
                setattr(self, q, z(nz, ny, nx))
                setattr(self, q + "0", z(nz, ny, nx))
            self.h_diabatic = z(nz, ny, nx)
            # a mixed-phase scheme's transported species (reference
            # state.py:428-470; the scheme runs in the caller's physics)
            for q in EXTRA_SPECIES[cfg.mp_physics]:
                setattr(self, q, z(nz, ny, nx))
                setattr(self, q + "0", z(nz, ny, nx))
            if cfg.mp_physics == 18:
                # NSSL predicted CCN starts at nssl_cccn / 1.225 (reference
                # state.py:447-452)
                self.qnn.fill_(408163264.0)
        else:
            self.qv = self.qc = self.qr = None
            self.qv0 = self.qc0 = self.qr0 = None
            self.h_diabatic = None
        self.u0, self.v0, self.w0 = z(nz, ny, nx + 1), z(nz, ny + 1, nx), \
            z(nz + 1, ny, nx)
        self.thp0, self.php0, self.mup0 = z(nz, ny, nx), z(nz + 1, ny, nx), \
            z(ny, nx)
        self.ru_t, self.rv_t, self.rw_t = z(nz, ny, nx + 1), \
            z(nz, ny + 1, nx), z(nz + 1, ny, nx)
        self.rth_t, self.rph_t, self.rmu_t = z(nz, ny, nx), \
            z(nz + 1, ny, nx), z(ny, nx)
        self.u_pp, self.v_pp, self.w_pp = z(nz, ny, nx + 1), \
            z(nz, ny + 1, nx), z(nz + 1, ny, nx)
        self.th_pp, self.ph_pp, self.mu_pp = z(nz, ny, nx), \
            z(nz + 1, ny, nx), z(ny, nx)
        self.p_pp, self.p_pp_old = z(nz, ny, nx), z(nz, ny, nx)
        self.ww_pp, self.al_pp = z(nz + 1, ny, nx), z(nz, ny, nx)
        if cfg.terrain_opt == 0:
            self.thb, self.pb, self.alb = z(nz), z(nz), z(nz)
            self.phb = z(nz + 1)
        else:
            self.thb, self.pb, self.alb = z(nz, ny, nx), z(nz, ny, nx), \
                z(nz, ny, nx)
            self.phb = z(nz + 1, ny, nx)
        self.mub = 0.0
        self.p_top = None
        self.mub2d, self.ht = z(ny, nx), z(ny, nx)
        for name in _COORD:
            setattr(self, name, z(nz + 1 if name.endswith("f")
                                  or name == "znw" else nz))
        self.msft = torch.ones((ny, nx), dtype=dtype, device=self.device)
        self.msfu = torch.ones((ny, nx + 1), dtype=dtype, device=self.device)
        self.msfv = torch.ones((ny + 1, nx), dtype=dtype, device=self.device)
        self.f, self.e = z(ny, nx), z(ny, nx)
        self.has_msf = False
        self.rotational = False
        self.cf1 = self.cf2 = self.cf3 = self.cfn = self.cfn1 = 0.0
        self.elapsed_seconds = 0.0
        self._scratch: dict[str, torch.Tensor] = {}
        self._phb_host: np.ndarray | None = None

    # -- setup --------------------------------------------------------------

    def load_base(self, coord: VerticalCoord, base: BaseState) -> None:
        """Copy the FP64 host coordinate/base arrays to FP32 device arrays
        (reference state.py:722-771)."""
        put = lambda t, a: t.copy_(torch.as_tensor(
            np.asarray(a, dtype=self.np_dtype)))
        for name in _COORD:
            put(getattr(self, name), getattr(coord, name))
        for name in ("thb", "pb", "alb", "phb"):
            host = np.asarray(getattr(base, name), dtype=np.float64)
            dev = getattr(self, name)
            if host.ndim != dev.ndim:
                raise ValueError(
                    f"base state {name} is {host.ndim}-D but the state was "
                    f"allocated for {dev.ndim}-D profiles: cfg.terrain_opt "
                    "must match the terrain the base state was built with")
            put(dev, host)
        self.p_top = float(self.np_dtype.type(base.p_top))
        if np.ndim(base.mub) == 0:
            self.mub = float(self.np_dtype.type(base.mub))
            self.mub2d.fill_(self.mub)
        else:
            self.mub = None
            put(self.mub2d, base.mub)
        if base.terrain_z is not None:
            put(self.ht, base.terrain_z)
        else:
            self.ht.zero_()
        self._phb_host = np.array(base.phb, dtype=np.float64, copy=True)
        self._coord, self._base = coord, base
        for name, val in surface_extrapolation(coord).items():
            setattr(self, name, float(self.np_dtype.type(val)))

    @property
    def extra_species(self) -> tuple:
        """The transported species beyond qv/qc/qr this state carries."""
        if self.qv is None:
            return ()
        return EXTRA_SPECIES[self.cfg.mp_physics]

    @property
    def np_dtype(self) -> np.dtype:
        return np.dtype(np.float64 if self.dtype == torch.float64
                        else np.float32)

    @classmethod
    def from_host(cls, case, device="cuda", dtype=DTYPE) -> "DomainState":
        """Upload a host initial condition (``arwen_amd.cases.HostCase``),
        rounded once to the state's precision."""
        s = cls(case.cfg, device=device, dtype=dtype)
        s.load_base(case.coord, case.base)
        for name, arr in case.fields.items():
            getattr(s, name).copy_(torch.as_tensor(
                np.asarray(arr, dtype=s.np_dtype)))
        return s

    def host_case(self):
        """The state's FP64 coordinate and base objects with its current
        prognostic (and moist) fields downloaded, as an
        ``arwen_amd.cases.HostCase``."""
        from .cases import HostCase
        if getattr(self, "_coord", None) is None:
            raise RuntimeError("host_case() called before load_base()")
        names = ["u", "v", "w", "thp", "php", "mup"]
        if self.qv is not None:
            names += ["qv", "qc", "qr"] + list(self.extra_species)
            if self.cfg.mp_physics != 0:
                names.append("h_diabatic")
        fields = {n: getattr(self, n).detach().cpu().numpy().astype(
            np.float64) for n in names}
        return HostCase(self.cfg, self._coord, self._base, fields)

    # -- reference helpers ----------------------------------------------------

    def scratch(self, shape, slot: str, dtype=None) -> torch.Tensor:
        """Persistent named scratch buffer (reference state.py:805-830)."""
        shape = tuple(shape)
        buf = self._scratch.get(slot)
        if buf is None:
            buf = torch.zeros(shape, dtype=dtype or self.dtype,
                              device=self.device)
            self._scratch[slot] = buf
            self.__dict__.pop("_dom_memo", None)
        elif tuple(buf.shape) != shape:
            raise ValueError(f"scratch slot {slot!r} has shape "
                             f"{tuple(buf.shape)}, requested {shape}")
        return buf

    def total_theta(self) -> torch.Tensor:
        thb = self.thb
        return (thb if thb.dim() == 3 else thb[:, None, None]) + self.thp

    def total_mu(self) -> torch.Tensor:
        return self.mub2d + self.mup

    def height_half(self) -> np.ndarray:
        if self._phb_host is None:
            raise RuntimeError("height_half() called before load_base()")
        return 0.5 * (self._phb_host[:-1] + self._phb_host[1:]) / c.G

    @property
    def shape(self):
        return tuple(self.p.shape)

    # -- C-ABI view -----------------------------------------------------------

    def _diff_spacings(self):
        """Constant-K vertical spacings from the FP32 base geopotential,
        formed in double (reference diffusion.py:34-61, 117)."""
        cached = getattr(self, "_diff_rdz", None)
        if cached is not None:
            return cached
        if self.phb.dim() != 1:
            raise NotImplementedError(
                "constant-K diffusion assumes a flat base state")
        zf = self.phb.detach().cpu().numpy().astype(np.float64) / c.G
        zh = 0.5 * (zf[:-1] + zf[1:])
        rdzf_w = 1.0 / np.diff(zf)
        rdzc_w = np.zeros(zf.size)
        rdzc_w[1:-1] = 1.0 / (zh[1:] - zh[:-1])
        rdzf_h = 1.0 / np.diff(zh)
        rdzc_h = 1.0 / np.diff(zf)
        self._diff_rdz = [
            torch.as_tensor(np.asarray(a, dtype=self.np_dtype)).to(self.device)
            for a in (rdzf_h, rdzc_h, rdzf_w, rdzc_w)]
        return self._diff_rdz

    def __setattr__(self, name, value):
        # rebinding an array invalidates the cached raw-pointer views
        if isinstance(value, torch.Tensor) or value is None:
            self.__dict__.pop("_dom_memo", None)
        object.__setattr__(self, name, value)

    def domain_struct(self, diffusion: bool = False,
                      precision_form: bool = False,
                      diff6: bool = False,
                      hypsometric_opt: int | None = None,
                      moist_cq: bool = False,
                      smag: bool = False) -> ArwDomain:
        """Raw-pointer view of the state for the C-ABI (a fresh copy the
        caller may modify).  Built once per option set and cached until an
        array attribute is rebound or a scratch slot is created; arrays
        updated in place keep their pointers."""
        if hypsometric_opt is None:
            hypsometric_opt = getattr(self.cfg, "hypsometric_opt", 1)
        key = (bool(diffusion), bool(precision_form), bool(diff6),
               int(hypsometric_opt), bool(moist_cq), bool(smag))
        memo = self.__dict__.get("_dom_memo")
        if memo is not None and key in memo:
            m = memo[key]
            return type(m).from_buffer_copy(m)
        d = self._build_domain_struct(*key)
        self.__dict__.setdefault("_dom_memo", {})[key] = d
        return type(d).from_buffer_copy(d)

    def _build_domain_struct(self, diffusion: bool, precision_form: bool,
                             diff6: bool, hypsometric_opt: int,
                             moist_cq: bool, smag: bool) -> ArwDomain:
        nz, ny, nx = self.shape
        d = ArwDomain64() if self.dtype == torch.float64 else ArwDomain()
        d.nz, d.ny, d.nx = nz, ny, nx
        d.base3d = int(self.thb.dim() == 3)
        for name in DOMAIN_ARRAYS:
            t = getattr(self, name)
            if not t.is_contiguous() or t.dtype != self.dtype:
                raise ValueError(f"state.{name} must be contiguous "
                                 f"{self.dtype}")
            setattr(d, name, t.data_ptr())
        for name in ("cf1", "cf2", "cf3", "cfn", "cfn1"):
            setattr(d, name, float(getattr(self, name)))
        shapes = {"ww": (nz + 1, ny, nx), "gam": (nz + 1, ny, nx),
                  "dmdt_rf": (ny, nx), "mup_next": (ny, nx)}
        for name in SCRATCH_ARRAYS:
            setattr(d, name, self.scratch(shapes[name],
                                          "arw_" + name).data_ptr())
        if diffusion:
            for n, t in zip(DIFF_ARRAYS, self._diff_spacings()):
                setattr(d, n, t.data_ptr())
        if self.qv is not None:
            m3, w3 = (nz, ny, nx), (nz + 1, ny, nx)
            u3, v3 = (nz, ny, nx + 1), (nz, ny + 1, nx)
            for name in MOIST_ARRAYS:
                t = (getattr(self, name) if not name.startswith("qs")
                     else self.scratch(m3, "arw_" + name))
                if not t.is_contiguous() or t.dtype != self.dtype:
                    raise ValueError(f"state.{name} must be contiguous "
                                     f"{self.dtype}")
                setattr(d, name, t.data_ptr())
            shapes = {"ru_m": u3, "rv_m": v3, "ww_m": w3, "pd_fxl": u3,
                      "pd_fxc": u3, "pd_fyl": v3, "pd_fyc": v3, "pd_fzl": w3,
                      "pd_fzc": w3, "pd_sc": m3}
            for name in MOIST_SCRATCH:
                setattr(d, name, self.scratch(shapes[name],
                                              "arw_" + name).data_ptr())
        extra = self.extra_species
        if len(extra) > MAX_EXTRA:
            raise ValueError(f"{len(extra)} extra species exceed the "
                             f"ABI's {MAX_EXTRA}")
        d.n_extra = len(extra)
        d.n_load = LOADING_EXTRA[self.cfg.mp_physics] if extra else 0
        for k, name in enumerate(extra):
            d.qx[k] = getattr(self, name).data_ptr()
            d.qx0[k] = getattr(self, name + "0").data_ptr()
            d.qxs[k] = self.scratch((nz, ny, nx), "arw_qs_" + name).data_ptr()
        if self.qv is not None and self.cfg.mp_physics != 0:
            d.h_diabatic = self.h_diabatic.data_ptr()
        d.hypsometric_opt = int(hypsometric_opt)
        if hypsometric_opt == 2:
            if self.p_top is None:
                raise ValueError("hypsometric_opt=2 needs the base state's "
                                 "p_top: call load_base first")
            for name in ("c3h", "c4h", "c3f", "c4f"):
                d_arr = getattr(self, name)
                setattr(d, name, d_arr.data_ptr())
            d.p_top = float(self.p_top)
        if moist_cq and self.qv is not None:
            for name, shp in (("cqu", (nz, ny, nx + 1)),
                              ("cqv", (nz, ny + 1, nx)),
                              ("cqw", (nz + 1, ny, nx))):
                setattr(d, name, self.scratch(shp, "arw_" + name).data_ptr())
        if precision_form and self.dtype == torch.float32:
            d.tri64 = self.scratch((2, nz + 1, ny, nx), "arw_tri64",
                                   dtype=torch.float64).data_ptr()
            d.pp = self.scratch((nz, ny, nx), "arw_pp").data_ptr()
        if smag:
            # km_opt = 4: the deformation, K and stress work areas
            n, nw = nz * ny * nx, (nz + 1) * ny * nx
            d.sm_buf = self.scratch((9 * n + 5 * nw,),
                                    "arw_sm_buf").data_ptr()
        if (diff6 or smag) and self.qv is not None:
            # every transported species' held mixing tendency
            nq = 3 + len(self.extra_species)
            d.held_q = self.scratch((nq, nz, ny, nx),
                                    "arw_held_q").data_ptr()
        if diff6 or smag:
            shapes = {"d6_u": (nz, ny, nx + 1), "d6_v": (nz, ny + 1, nx),
                      "d6_w": (nz + 1, ny, nx), "d6_t": (nz, ny, nx)}
            for name, shp in shapes.items():
                setattr(d, name, self.scratch(shp, "arw_" + name).data_ptr())
        return d

    def numpy_fields(self, names=("u", "v", "w", "thp", "php", "mup")):
        return {n: getattr(self, n).detach().cpu().numpy().copy()
                for n in names}


def init_at_rest(cfg: RunConfig, coord: VerticalCoord, base: BaseState,
                 terrain_z=None, *, device="cuda",
                 dtype=DTYPE) -> DomainState:
    """Zero perturbations on ``base`` (reference state.py:922-932)."""
    from .cases import init_at_rest as _init
    return _init(cfg, coord, base, terrain_z, device=device, dtype=dtype)


def init_theta_perturbation(cfg: RunConfig, coord: VerticalCoord,
                            base: BaseState, thp_func, terrain_z=None, *,
                            device="cuda", dtype=DTYPE) -> DomainState:
    """At rest plus a hydrostatically rebalanced theta perturbation
    (reference state.py:961-993)."""
    from .cases import init_theta_perturbation as _init
    return _init(cfg, coord, base, thp_func, terrain_z, device=device,
                 dtype=dtype)
