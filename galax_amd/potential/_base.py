"""Potential classes: the host side of galax's potential plugin interface.

galax's plugin contract is `AbstractSinglePotential._potential(xyz, t)` with
gradients and Hessians from autodiff (potential/_src/base.py:113-239) and
composites summing their components (base_multi.py:39-82).  Here each model
flattens to rows of a component table (`gx_component`, include/galax_amd.h);
the HIP library evaluates Phi, its closed-form gradient and Hessian for the
whole table on the GPU.  Public methods mirror AbstractPotential's
(`potential`, `gradient`, `acceleration`, `hessian`, `density`,
`laplacian`, `d2potential_dr2`, `local_circular_velocity`,
`evaluate_orbit`).
"""

from __future__ import annotations

import ctypes
import math
import weakref
from dataclasses import dataclass

import numpy as np

from .. import _abi, _lib
from ..units import G, Q, ustrip

__all__ = ["AbstractPotential", "AbstractSinglePotential", "DenseTrack", "Track"]


@dataclass(frozen=True)
class Track:
    """Tabulated centre trajectory: knots t[n], positions q[n,3], velocities
    v[n,3] (galactic units), interpolated by cubic Hermite polynomials."""

    t: np.ndarray
    q: np.ndarray
    v: np.ndarray


@dataclass(frozen=True)
class DenseTrack:
    """Centre trajectory given as the dense solution of one integrated
    orbit: its accepted-step records rec[n, GX_DENSE_RECORD] in integration
    order, written by method `method` (GX_DOPRI8 or GX_DOPRI5); the position
    at t is the method's dense output over the step holding t."""

    rec: np.ndarray
    method: int


def fill_track(dst, tr) -> list:
    """Fill a gx_track from a Track or a DenseTrack; returns the arrays it
    points into, which must outlive the call that reads it."""
    if isinstance(tr, DenseTrack):
        rec = np.ascontiguousarray(tr.rec, dtype=np.float64)
        if rec.ndim != 2 or rec.shape[1] != _abi.GX_DENSE_RECORD or rec.shape[0] < 1:
            raise ValueError(f"dense track records must have shape (n >= 1, {_abi.GX_DENSE_RECORD})")
        dst.n = rec.shape[0]
        dst.rec = rec.ctypes.data
        dst.method = int(tr.method)
        return [rec]
    tt = np.ascontiguousarray(tr.t, dtype=np.float64)
    qq = np.ascontiguousarray(tr.q, dtype=np.float64)
    vv = np.ascontiguousarray(tr.v, dtype=np.float64)
    dst.n = tt.shape[0]
    dst.t = tt.ctypes.data
    dst.q = qq.ctypes.data
    dst.v = vv.ctypes.data
    return [tt, qq, vv]


def _destroy_handles(handles: dict) -> None:
    try:
        lib = _lib.lib()
    except Exception:  # library unloaded at interpreter exit
        return
    for h in handles.values():
        if h:
            lib.gx_potential_destroy(h)
    handles.clear()


class AbstractPotential:
    """Base of every potential.  Subclasses implement `_rows()`."""

    units = "galactic"

    # ---------------------------------------------------------- flattening
    def _rows(self) -> list[tuple[_abi.GxComponent, Track | None]]:
        """Component-table rows; a row with a Track uses GX_FRAME_TRACK."""
        raise NotImplementedError

    def _table(self) -> tuple[list[_abi.GxComponent], list[Track]]:
        comps, tracks = [], []
        for row, track in self._rows():
            if track is not None:
                row.frame = _abi.GX_FRAME_TRACK
                row.track = len(tracks)
                tracks.append(track)
            comps.append(row)
        if len(comps) > _abi.GX_MAX_COMPONENTS:
            raise ValueError(f"at most {_abi.GX_MAX_COMPONENTS} components")
        if len(tracks) > _abi.GX_MAX_TRACKS:
            raise ValueError(f"at most {_abi.GX_MAX_TRACKS} tracked centres")
        return comps, tracks

    # ------------------------------------------------------- device handle
    def _handle(self) -> int:
        t = _lib.require_gpu()
        dev = t.cuda.current_device()
        handles = self.__dict__.setdefault("_gx_handles", {})
        if dev in handles:
            return handles[dev]
        comps, tracks = self._table()
        carr = _abi.component_array(comps)
        keep = []
        tarr = (_abi.GxTrack * max(len(tracks), 1))()
        for i, tr in enumerate(tracks):
            keep += fill_track(tarr[i], tr)
        h = ctypes.c_void_p()
        rc = _lib.lib().gx_potential_create(
            ctypes.addressof(carr), len(comps), ctypes.addressof(tarr), len(tracks), ctypes.byref(h)
        )
        _lib.check(rc, "gx_potential_create")
        handles[dev] = h.value
        if "_gx_finalizer" not in self.__dict__:
            self.__dict__["_gx_finalizer"] = weakref.finalize(self, _destroy_handles, handles)
        return h.value

    # ------------------------------------------------------------- fields
    def _field(self, fname: str, width: int, q, t):
        """Evaluate a field kernel; returns output shaped batch + (width,)."""
        torch = _lib.require_gpu()
        is_q = isinstance(q, Q)
        is_torch = isinstance(q, torch.Tensor)
        qv = ustrip(q)
        if is_torch:
            xq = qv.to(device="cuda", dtype=torch.float64)
        else:
            xq = torch.as_tensor(np.ascontiguousarray(qv, dtype=np.float64), device="cuda")
        if xq.shape[-1:] != (3,):
            raise ValueError(f"positions must have shape (..., 3), got {tuple(xq.shape)}")
        batch = xq.shape[:-1]
        xq = xq.reshape(-1, 3).contiguous()
        n = xq.shape[0]
        tv = ustrip(t, "Myr") if isinstance(t, Q) else t
        if isinstance(tv, torch.Tensor):
            tt = tv.to(device="cuda", dtype=torch.float64)
        else:
            tt = torch.as_tensor(np.ascontiguousarray(tv, dtype=np.float64), device="cuda")
        if tt.numel() == 1:
            tt = tt.reshape(1).contiguous()
            stride = 0
        else:
            tt = torch.broadcast_to(tt, batch).reshape(-1).contiguous()
            stride = 1
        out = torch.empty((n, width), dtype=torch.float64, device="cuda")
        fn = getattr(_lib.lib(), fname)
        rc = fn(self._handle(), _lib.ptr(xq), _lib.ptr(tt), stride, n, _lib.ptr(out), _lib.stream_ptr())
        _lib.check(rc, fname)
        out = out.reshape(*batch, width) if width > 1 else out.reshape(batch)
        if is_torch:
            return out
        res = out.cpu().numpy()
        return res

    def potential(self, q, t=0.0):
        """Phi(q, t) [kpc^2/Myr^2] (galax: AbstractPotential.potential)."""
        out = self._field("gx_potential_eval", 1, q, t)
        return Q(out, "kpc2/Myr2") if isinstance(q, Q) else out

    __call__ = potential

    def gradient(self, q, t=0.0):
        """grad Phi [kpc/Myr^2] (galax: AbstractPotential.gradient)."""
        out = self._field("gx_gradient", 3, q, t)
        return Q(out, "kpc/Myr2") if isinstance(q, Q) else out

    def acceleration(self, q, t=0.0):
        """-grad Phi [kpc/Myr^2] (galax: AbstractPotential.acceleration,
        potential/_src/base.py:250-255)."""
        out = self._field("gx_acceleration", 3, q, t)
        return Q(out, "kpc/Myr2") if isinstance(q, Q) else out

    def hessian(self, q, t=0.0):
        """Hessian of Phi, shape (..., 3, 3) [1/Myr^2]."""
        out = self._field("gx_hessian", 9, q, t)
        return out.reshape(*out.shape[:-1], 3, 3)

    def laplacian(self, q, t=0.0):
        H = self.hessian(q, t)
        return H[..., 0, 0] + H[..., 1, 1] + H[..., 2, 2]

    def density(self, q, t=0.0):
        """rho = lap Phi / (4 pi G) [Msun/kpc^3] (galax: base.py _density)."""
        return self.laplacian(q, t) / (4.0 * math.pi * G)

    def d2potential_dr2(self, q, t=0.0):
        """r_hat . H . r_hat (galax: register_funcs.py:442-457)."""
        xyz = np.asarray(ustrip(q), dtype=np.float64)
        H = np.asarray(self.hessian(xyz, t))
        rhat = xyz / np.linalg.norm(xyz, axis=-1, keepdims=True)
        return np.einsum("...i,...ij,...j->...", rhat, H, rhat)

    def local_circular_velocity(self, q, t=0.0):
        """sqrt(r |dPhi/dr|) [kpc/Myr] (galax: register_funcs.py:384-412)."""
        xyz = np.asarray(ustrip(q), dtype=np.float64)
        r = np.linalg.norm(xyz, axis=-1)
        g = np.asarray(self.gradient(xyz, t))
        dphi_dr = np.sum(g * xyz, axis=-1) / r
        return np.sqrt(r * np.abs(dphi_dr))

    def evaluate_orbit(self, w0, t, **kw):
        """gd.evaluate_orbit(self, w0, t, ...) (potential/_src/base.py:384-456)."""
        from ..dynamics import evaluate_orbit

        return evaluate_orbit(self, w0, t, **kw)

    def __add__(self, other: AbstractPotential):
        from ._composite import CompositePotential

        if not isinstance(other, AbstractPotential):
            return NotImplemented
        parts = {}
        for p in (self, other):
            if isinstance(p, CompositePotential) and not getattr(p, "_preset", False):
                for k, v in p.items():
                    parts[f"c{len(parts)}_{k}"] = v
            else:
                parts[f"c{len(parts)}"] = p
        return CompositePotential(parts)


class AbstractSinglePotential(AbstractPotential):
    """One analytic model (galax: potential/_src/base_single.py)."""

    kind: int = _abi.GX_NULL

    def _params(self) -> tuple[float, ...]:
        raise NotImplementedError

    def _rows(self):
        return [(_abi.component(self.kind, self._params()), None)]
