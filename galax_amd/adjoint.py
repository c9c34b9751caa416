"""Gradients through fixed-step orbit integration (discrete adjoint).

`adjoint_orbits` calls gx_adjoint: given cotangents of the saved states it
returns cotangents of the initial states and of the potential's table
parameters.  `OrbitFlow` wraps forward integration and adjoint as a
torch.autograd.Function so that losses built with torch operations on the
outputs (and on the inputs, e.g. stream release conditions) differentiate
end to end.  `MilkyWayParams` maps the six Milky Way parameters (halo m,
r_s; disk m_tot, a, b; bulge m_tot) to a potential and to the table-parameter
gradient layout.

galax obtains such derivatives with jax.jacfwd through diffrax
(tests/functional/test_mockstreamgenerator.py:24-116; ForwardMode adjoint,
dynamics/_src/orbit/solver.py:133).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _abi, _lib
from .dynamics._solver import ConstantStepSize, Dopri8, OrbitSolver, _dev, integrate_orbits
from .potential import HernquistPotential, MilkyWayPotential, MiyamotoNagaiPotential, NFWPotential
from .units import G

__all__ = ["adjoint_orbits", "OrbitFlow", "FlowSpec", "MilkyWayParams", "steps_bound"]


def steps_bound(t0, t1, dt0: float) -> int:
    """An upper bound on the fixed steps of any orbit."""
    span = np.max(np.abs(np.asarray(t1, dtype=np.float64) - np.asarray(t0, dtype=np.float64)))
    return int(np.ceil(span / abs(dt0))) + 2


def adjoint_orbits(pot, w0, t0, t1, cot, *, dt0: float, ts=None, method=None, K: int = 64,
                   ntheta: int = 0, max_steps: int | None = None):
    """Returns (w0_bar (n,6), theta_bar (n, ntheta), status (n,)) as device
    tensors.  `cot` (n, T, 6) are the cotangents of the states saved at `ts`
    ((T,) shared, (n,T) per orbit, or None for the final state only)."""
    torch = _lib.require_gpu()
    method = method or Dopri8()
    cfg = OrbitSolver(method, stepsize_controller=ConstantStepSize(), max_steps=None).cfg(dt0=dt0)
    w0d = _dev(torch, w0, torch.float64)
    n = int(w0d.shape[0])
    t0d = _dev(torch, t0, torch.float64).reshape(-1)
    t1d = _dev(torch, t1, torch.float64).reshape(-1)
    if t0d.numel() == 1 and t1d.numel() == 1:
        stride = 0
    else:
        t0d = torch.broadcast_to(t0d, (n,)).contiguous()
        t1d = torch.broadcast_to(t1d, (n,)).contiguous()
        stride = 1
    tsd = None
    if ts is None:
        mode, T = _abi.GX_TS_FINAL, 1
    else:
        tsd = _dev(torch, ts, torch.float64)
        mode = _abi.GX_TS_SHARED if tsd.ndim == 1 else _abi.GX_TS_PER_ORBIT
        T = int(tsd.shape[-1])
    cotd = _dev(torch, cot, torch.float64).reshape(n, T, 6).contiguous()
    if max_steps is None:
        max_steps = steps_bound(t0d.cpu().numpy(), t1d.cpu().numpy(), dt0)
    lib = _lib.lib()
    nbytes = int(lib.gx_adjoint_workspace_bytes(n, int(max_steps), int(K)))
    ws = torch.empty((nbytes + 7) // 8, dtype=torch.float64, device="cuda")
    w0_bar = torch.empty((n, 6), dtype=torch.float64, device="cuda")
    th = torch.zeros((n, max(ntheta, 1)), dtype=torch.float64, device="cuda")
    status = torch.empty((n,), dtype=torch.int32, device="cuda")
    rc = lib.gx_adjoint(
        pot._handle(), ctypes.addressof(cfg), _lib.ptr(w0d), _lib.ptr(t0d), _lib.ptr(t1d), stride,
        _lib.ptr(tsd), T, mode, n, _lib.ptr(cotd), _lib.ptr(w0_bar), _lib.ptr(th), int(ntheta),
        int(max_steps), int(K), _lib.ptr(status), _lib.ptr(ws), _lib.stream_ptr(),
    )
    _lib.check(rc, "gx_adjoint")
    return w0_bar, th[:, :ntheta], status


class MilkyWayParams:
    """theta = (halo m, halo r_s, disk m_tot, disk a, disk b, bulge m_tot)
    over galax's MilkyWayPotential defaults (nucleus fixed)."""

    names = ("halo_m", "halo_r_s", "disk_m_tot", "disk_a", "disk_b", "bulge_m_tot")
    default = np.array([5.4e11, 15.62, 6.8e10, 3.0, 0.28, 5e9])
    ntheta = 12  # table parameters: 3 per component x 4 components

    @staticmethod
    def potential(theta) -> MilkyWayPotential:
        t = np.asarray(theta, dtype=np.float64)
        return MilkyWayPotential(
            disk=MiyamotoNagaiPotential(m_tot=t[2], a=t[3], b=t[4]),
            halo=NFWPotential(m=t[0], r_s=t[1]),
            bulge=HernquistPotential(m_tot=t[5], r_s=1.0),
        )

    @staticmethod
    def grad_from_table(raw):
        """dL/dtheta from dL/d(table parameters) (..., 12): table order is
        disk (G m, a, b), halo (G m, r_s, -), bulge (G m, r_s, -), nucleus."""
        raw = raw.reshape(*raw.shape[:-1], 4, 3)
        return _stack(
            [G * raw[..., 1, 0], raw[..., 1, 1], G * raw[..., 0, 0], raw[..., 0, 1], raw[..., 0, 2], G * raw[..., 2, 0]]
        )


def _stack(cols):
    torch = _lib.torch()
    return torch.stack(cols, dim=-1)


@dataclass
class FlowSpec:
    """Fixed-step integration settings for OrbitFlow."""

    t0: object
    t1: object
    dt0: float
    ts: object = None
    method: object = None
    K: int = 64
    params: type = MilkyWayParams


class OrbitFlow:
    """torch.autograd.Function: (theta, w0) -> saved states (n, T, 6)."""

    @staticmethod
    def apply(theta, w0, spec: FlowSpec):
        return _orbit_flow_fn().apply(theta, w0, spec)


_FN = None


def _orbit_flow_fn():
    global _FN
    if _FN is not None:
        return _FN
    torch = _lib.torch()

    class _Flow(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, w0, spec):
            pot = spec.params.potential(theta.detach().cpu().numpy())
            solver = OrbitSolver(spec.method or Dopri8(), stepsize_controller=ConstantStepSize(), max_steps=None)
            res = integrate_orbits(pot, w0.detach(), spec.t0, spec.t1, ts=spec.ts, solver=solver, dt0=spec.dt0)
            if int(res.status.abs().sum()) != 0:
                raise RuntimeError("orbit integration failed for some orbits")
            ctx.spec = spec
            ctx.pot = pot
            ctx.save_for_backward(theta, w0)
            return res.out

        @staticmethod
        def backward(ctx, grad_out):
            theta, w0 = ctx.saved_tensors
            spec = ctx.spec
            w0_bar, raw, status = adjoint_orbits(
                ctx.pot, w0.detach(), spec.t0, spec.t1, grad_out.contiguous(), dt0=spec.dt0, ts=spec.ts,
                method=spec.method, K=spec.K, ntheta=spec.params.ntheta,
            )
            if int(status.abs().sum()) != 0:
                raise RuntimeError("adjoint sweep ran past its step bound")
            g_theta = spec.params.grad_from_table(raw.sum(0)).to(theta.dtype)
            return g_theta.to(theta.device), w0_bar.to(w0.dtype), None

    _FN = _Flow
    return _FN
