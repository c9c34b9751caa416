"""N-body dynamics: galax's NBodyField (dynamics/_src/orbit/field_nbody.py
:22-131) and its integration as one ODE system.

The field is a direct sum over bodies with Plummer softening plus an
external potential; `solve_nbody` integrates all bodies with one shared
step sequence (the PID error norm is the RMS over the whole state), as
OrbitSolver.solve does for this field.  Both run in libgalaxhip.so
(gx_nbody_acceleration, gx_nbody_integrate).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from .. import _lib
from ..units import G, G_SOLARSYSTEM, Q, ustrip
from ._solver import OrbitSolver, _dev

__all__ = ["NBodyField", "NBodySolution", "solve_nbody"]

_G_UNITS = {"galactic": G, "solarsystem": G_SOLARSYSTEM}
_CONV = {  # unit system -> (length, speed, time, mass) unit names of Q
    "galactic": ("kpc", "kpc/Myr", "Myr", "Msun"),
    "solarsystem": ("AU", "AU/yr", "yr", "Msun"),
}


def _strip(x, unit, system):
    """Plain array in `system` units; Q converted via the galactic factors."""
    if isinstance(x, Q):
        v = x.to_value()  # galactic
        if system == "solarsystem":
            v = v / Q(1.0, unit).to_value()
        return np.asarray(v, dtype=np.float64)
    return x


class NBodyField:
    """dq/dt = p, dp/dt = sum_i G m_i (q_i - q) / (|q_i - q|^2 + eps^2)^(3/2)
    - grad Phi_ext(q, t).  `units` is "galactic" (kpc, Myr, Msun) or
    "solarsystem" (AU, yr, Msun; external potential NullPotential only)."""

    def __init__(self, masses, eps, external_potential=None, *, units=None):
        from ..potential import NullPotential

        ext = external_potential if external_potential is not None else NullPotential()
        if units is None:
            units = getattr(ext, "units", "galactic")
        if units not in _G_UNITS:
            raise ValueError(f"unsupported unit system {units!r}")
        if units != "galactic" and any(r.kind != 0 for r in ext._table()[0]):
            raise ValueError("a non-galactic unit system needs a NullPotential external potential")
        self.units = units
        L, V, T, M = _CONV[units]
        self.masses = np.asarray(_strip(masses, M, units), dtype=np.float64).reshape(-1)
        self.eps = float(np.asarray(_strip(eps, L, units), dtype=np.float64))
        self.external_potential = ext

    @property
    def _G(self) -> float:
        return _G_UNITS[self.units]

    def _gm(self, torch):
        return torch.as_tensor(self._G * self.masses, dtype=torch.float64, device="cuda")

    def __call__(self, t, qp, args=None):
        """(dq/dt, dp/dt) at time t for q, p of shape (N, 3)."""
        torch = _lib.require_gpu()
        L, V, T, _ = _CONV[self.units]
        q, p = qp
        qd = _dev(torch, _strip(q, L, self.units), torch.float64).reshape(-1, 3).contiguous()
        n = int(qd.shape[0])
        if n != self.masses.size:
            raise ValueError(f"{n} positions for {self.masses.size} masses")
        acc = torch.empty_like(qd)
        gm = self._gm(torch)
        tt = float(np.asarray(_strip(t, T, self.units)))
        rc = _lib.lib().gx_nbody_acceleration(self.external_potential._handle(), _lib.ptr(qd), _lib.ptr(gm),
                                              self.eps, tt, n, _lib.ptr(acc), _lib.stream_ptr())
        _lib.check(rc, "gx_nbody_acceleration")
        return _dev(torch, _strip(p, V, self.units), torch.float64).reshape(-1, 3), acc


@dataclass
class NBodySolution:
    """ts (T,); ys = (q, p) each (T, N, 3) as diffrax returns them; status
    (diffrax RESULTS code), attempted and accepted steps of the system."""

    ts: np.ndarray
    ys: tuple
    status: int
    nsteps: int
    naccept: int


def solve_nbody(field: NBodyField, qp, t0, t1, *, saveat=None, solver: OrbitSolver | None = None, dt0=None,
                max_steps="default") -> NBodySolution:
    """Integrate the bodies from t0 to t1 as one system; saves at `saveat`
    (default: t1 only)."""
    torch = _lib.require_gpu()
    solver = solver or OrbitSolver()
    L, V, T, _ = _CONV[field.units]
    q, p = qp
    q = np.asarray(_strip(q, L, field.units), dtype=np.float64).reshape(-1, 3)
    p = np.asarray(_strip(p, V, field.units), dtype=np.float64).reshape(-1, 3)
    n = q.shape[0]
    if n != field.masses.size:
        raise ValueError(f"{n} bodies for {field.masses.size} masses")
    t0 = float(np.asarray(_strip(t0, T, field.units)))
    t1 = float(np.asarray(_strip(t1, T, field.units)))
    ts = np.array([t1]) if saveat is None else np.asarray(_strip(saveat, T, field.units), dtype=np.float64).reshape(-1)
    cfg = solver.cfg(dt0=dt0, max_steps=max_steps)
    w0 = torch.as_tensor(np.concatenate([q, p], axis=1), device="cuda")
    tsd = torch.as_tensor(ts, device="cuda")
    out = torch.empty((n, ts.size, 6), dtype=torch.float64, device="cuda")
    st = torch.empty(1, dtype=torch.int32, device="cuda")
    ns = torch.empty(1, dtype=torch.int64, device="cuda")
    na = torch.empty(1, dtype=torch.int64, device="cuda")
    lib = _lib.lib()
    ws = torch.empty((int(lib.gx_nbody_workspace_bytes(n)) + 7) // 8, dtype=torch.float64, device="cuda")
    rc = lib.gx_nbody_integrate(field.external_potential._handle(), _lib.ptr(field._gm(torch)), field.eps,
                                ctypes.addressof(cfg), _lib.ptr(w0), t0, t1, _lib.ptr(tsd), int(ts.size), n,
                                _lib.ptr(out), _lib.ptr(st), _lib.ptr(ns), _lib.ptr(na), _lib.ptr(ws),
                                _lib.stream_ptr())
    _lib.check(rc, "gx_nbody_integrate")
    o = out.cpu().numpy()
    ys = (np.transpose(o[..., :3], (1, 0, 2)), np.transpose(o[..., 3:], (1, 0, 2)))
    return NBodySolution(ts=ts, ys=ys, status=int(st.item()), nsteps=int(ns.item()), naccept=int(na.item()))
