"""Flux-form advection launchers and their adjoints.

``launch_flux_div_{scalar,u,v,w}`` keep the reference's launcher
signatures (gpuwm/core/advection.py:76-129: ``(field, ru, rv, rw, tend,
coord, dx, dy, ...)``, ADDING the periodic flux divergence into ``tend``)
and run the HIP kernels ``arw_flux_div_*``.  ``coord`` is anything carrying
``rdnw`` (and ``rdn`` for w), ``fnm`` and ``fnp``: a ``VerticalCoord`` or a
``DomainState``.  The open / specified lateral boundaries and map factors
of the reference's launchers are not part of this library and raise.

``flux_div_adjoint`` applies the transpose of an operator's Jacobian (with
respect to the advected field and the three mass fluxes) through the
hand-written kernels ``arw_flux_div_*_adjoint``; ``FluxDivergence`` is the
``torch.autograd.Function`` over an operator with that adjoint as its
backward.  The reference has no adjoint; these entry points are new.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib

OPS = ("scalar", "u", "v", "w")


def _extents(op: str, q: torch.Tensor) -> tuple[int, int, int]:
    a, b, c = q.shape
    return {"scalar": (a, b, c), "u": (a, b, c - 1), "v": (a, b - 1, c),
            "w": (a - 1, b, c)}[op]


def _shapes(nz, ny, nx):
    return {"ru": (nz, ny, nx + 1), "rv": (nz, ny + 1, nx),
            "rw": (nz + 1, ny, nx)}


def field_shape(op: str, nz: int, ny: int, nx: int) -> tuple[int, int, int]:
    """Shape of the advected field (and of the tendency) of ``op``."""
    return {"scalar": (nz, ny, nx), "u": (nz, ny, nx + 1),
            "v": (nz, ny + 1, nx), "w": (nz + 1, ny, nx)}[op]


def _coef(x, like: torch.Tensor, n: int, name: str) -> torch.Tensor:
    t = torch.as_tensor(np.asarray(x) if not torch.is_tensor(x) else x,
                        device=like.device).to(like.dtype).contiguous()
    if t.ndim != 1 or t.numel() < n:
        raise ValueError(f"coord.{name} needs at least {n} entries")
    return t


def _coords(op, coord, q, nz):
    """rdnw (rdn for w), fnm, fnp as device vectors of the field's dtype;
    the kernels read entries 0..nz-1 of each."""
    rdz = coord.rdn if op == "w" else coord.rdnw
    return (_coef(rdz, q, nz, "rdn" if op == "w" else "rdnw"),
            _coef(coord.fnm, q, nz, "fnm"), _coef(coord.fnp, q, nz, "fnp"))


def _check(op, q, ru, rv, rw, extra):
    if op not in OPS:
        raise ValueError(f"unknown advection operator {op!r}")
    nz, ny, nx = _extents(op, q)
    want = _shapes(nz, ny, nx)
    for name, t in (("ru", ru), ("rv", rv), ("rw", rw)) + tuple(extra):
        shp = want.get(name, field_shape(op, nz, ny, nx))
        if tuple(t.shape) != shp:
            raise ValueError(f"{name}: shape {tuple(t.shape)} != {shp}")
        if t.dtype != q.dtype or t.device != q.device:
            raise ValueError(f"{name}: dtype/device differ from the field")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
    if q.dtype not in (torch.float32, torch.float64) or q.device.type != "cuda":
        raise ValueError("fields must be float32/float64 device tensors")
    if not q.is_contiguous():
        raise ValueError("the advected field must be contiguous")
    return nz, ny, nx


def _prefix(q):
    return "arw64_" if q.dtype == torch.float64 else "arw_"


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _launch(op, q, ru, rv, rw, tend, coord, dx, dy, open_x, open_y, msf,
            has_msf, spec):
    if open_x or open_y or spec or msf is not None or has_msf:
        raise NotImplementedError("only the periodic, msf = 1 advection "
                                  "path is implemented")
    nz, ny, nx = _check(op, q, ru, rv, rw, (("tend", tend),))
    rdz, fnm, fnp = _coords(op, coord, q, nz)
    name = _prefix(q) + "flux_div_" + op
    fn = getattr(_lib.load(), name)
    _lib.check(fn(q.data_ptr(), ru.data_ptr(), rv.data_ptr(), rw.data_ptr(),
                  tend.data_ptr(), rdz.data_ptr(), fnm.data_ptr(),
                  fnp.data_ptr(), 1.0 / dx, 1.0 / dy, nz, ny, nx, _stream()),
               name)


def launch_flux_div_scalar(q, ru, rv, rw, tend, coord, dx, dy, open_x=False,
                           open_y=False, msf=None, has_msf=None, spec=False):
    """Add the flux divergence of scalar ``q`` (mass points) into ``tend``
    (reference advection.py:76-87)."""
    _launch("scalar", q, ru, rv, rw, tend, coord, dx, dy, open_x, open_y,
            msf, has_msf, spec)


def launch_flux_div_u(u, ru, rv, rw, tend, coord, dx, dy, open_x=False,
                      open_y=False, msf=None, has_msf=None, spec=False):
    """Add the flux divergence of ``u`` (u points) into ``tend``
    (reference advection.py:90-102)."""
    _launch("u", u, ru, rv, rw, tend, coord, dx, dy, open_x, open_y, msf,
            has_msf, spec)


def launch_flux_div_v(v, ru, rv, rw, tend, coord, dx, dy, open_x=False,
                      open_y=False, msf=None, has_msf=None, spec=False):
    """Add the flux divergence of ``v`` (v points) into ``tend``
    (reference advection.py:105-117)."""
    _launch("v", v, ru, rv, rw, tend, coord, dx, dy, open_x, open_y, msf,
            has_msf, spec)


def launch_flux_div_w(w, ru, rv, rw, tend, coord, dx, dy, open_x=False,
                      open_y=False, msf=None, has_msf=None, spec=False):
    """Add the flux divergence of ``w`` (w points, interior levels) into
    ``tend``; uses ``coord.rdn`` (reference advection.py:120-129)."""
    _launch("w", w, ru, rv, rw, tend, coord, dx, dy, open_x, open_y, msf,
            has_msf, spec)


def flux_div(op, q, ru, rv, rw, coord, dx, dy) -> torch.Tensor:
    """The tendency of ``op`` alone (a new tensor)."""
    tend = torch.zeros_like(q)
    _launch(op, q, ru, rv, rw, tend, coord, dx, dy, False, False, None,
            None, False)
    return tend


def flux_div_adjoint(op, lam, q, ru, rv, rw, coord, dx, dy,
                     need=("q", "ru", "rv", "rw")) -> dict:
    """{"q", "ru", "rv", "rw"}: the transposed Jacobian of ``op``'s
    tendency at (q, ru, rv, rw) applied to ``lam`` (shaped like the
    tendency), for the names in ``need``."""
    nz, ny, nx = _check(op, q, ru, rv, rw, (("lam", lam),))
    rdz, fnm, fnp = _coords(op, coord, q, nz)
    out = {"q": torch.zeros_like(q), "ru": torch.zeros_like(ru),
           "rv": torch.zeros_like(rv), "rw": torch.zeros_like(rw)}
    out = {k: v for k, v in out.items() if k in need}
    ptr = {k: (out[k].data_ptr() if k in out else None)
           for k in ("q", "ru", "rv", "rw")}
    name = _prefix(q) + "flux_div_" + op + "_adjoint"
    fn = getattr(_lib.load(), name)
    _lib.check(fn(q.data_ptr(), ru.data_ptr(), rv.data_ptr(), rw.data_ptr(),
                  lam.data_ptr(), ptr["q"], ptr["ru"], ptr["rv"], ptr["rw"],
                  rdz.data_ptr(), fnm.data_ptr(), fnp.data_ptr(), 1.0 / dx,
                  1.0 / dy, nz, ny, nx, _stream()), name)
    return out


class FluxDivergence(torch.autograd.Function):
    """(q, ru, rv, rw) -> the tendency of advection operator ``op``, with
    the hand-written adjoint as its backward.  ``apply(op, q, ru, rv, rw,
    coord, dx, dy)``."""

    @staticmethod
    def forward(ctx, op, q, ru, rv, rw, coord, dx, dy):
        q, ru, rv, rw = (x.detach().contiguous() for x in (q, ru, rv, rw))
        ctx.op, ctx.coord, ctx.dx, ctx.dy = op, coord, dx, dy
        ctx.save_for_backward(q, ru, rv, rw)
        return flux_div(op, q, ru, rv, rw, coord, dx, dy)

    @staticmethod
    def backward(ctx, grad):
        q, ru, rv, rw = ctx.saved_tensors
        need = tuple(n for n, f in zip(("q", "ru", "rv", "rw"),
                                       ctx.needs_input_grad[1:5]) if f)
        g = flux_div_adjoint(ctx.op, grad.contiguous(), q, ru, rv, rw,
                             ctx.coord, ctx.dx, ctx.dy, need=need)
        return (None, g.get("q"), g.get("ru"), g.get("rv"), g.get("rw"),
                None, None, None)
