"""Solver configuration objects and the integration core.

Mirrors the configuration surface galax passes to diffrax:
  OrbitSolver(solver, stepsize_controller, max_steps)
      dynamics/_src/orbit/solver.py:121-141
  Integrator(dynamics_solver, diffeq_kw)   legacy/integrator.py:43-245
  Dopri8 / Dopri5, PIDController, ConstantStepSize   (diffrax 0.7.0 names
  and defaults: PIDController(pcoeff=0, icoeff=1, dcoeff=0, dtmin=None,
  dtmax=None, force_dtmin=True, factormin=0.2, factormax=10, safety=0.9))
and runs the HIP integrator (gx_integrate) on device buffers.
"""

from __future__ import annotations

import ctypes

from dataclasses import dataclass, field, replace

import numpy as np

from .. import _abi, _lib

__all__ = [
    "Dopri8",
    "Dopri5",
    "PIDController",
    "ConstantStepSize",
    "OrbitSolver",
    "integrate_orbits",
    "IntegrationResult",
]


@dataclass(frozen=True)
class Dopri8:
    """Prince & Dormand RK8(7)13M (diffrax Dopri8)."""

    scan_kind: str | None = None
    method = _abi.GX_DOPRI8
    order = 8


@dataclass(frozen=True)
class Dopri5:
    """Dormand & Prince RK5(4)7M (diffrax Dopri5)."""

    scan_kind: str | None = None
    method = _abi.GX_DOPRI5
    order = 5


@dataclass(frozen=True)
class PIDController:
    rtol: float
    atol: float
    pcoeff: float = 0.0
    icoeff: float = 1.0
    dcoeff: float = 0.0
    dtmin: float | None = None
    dtmax: float | None = None
    force_dtmin: bool = True
    factormin: float = 0.2
    factormax: float = 10.0
    safety: float = 0.9


@dataclass(frozen=True)
class ConstantStepSize:
    pass


@dataclass(frozen=True)
class OrbitSolver:
    """Solver settings; `max_steps=None` means unbounded."""

    solver: object = field(default_factory=Dopri8)
    stepsize_controller: object = field(default_factory=lambda: PIDController(rtol=1e-8, atol=1e-8))
    max_steps: int | None = 2**16

    @classmethod
    def from_(cls, obj):
        if isinstance(obj, OrbitSolver):
            return obj
        if isinstance(obj, dict):
            return cls(**obj)
        raise TypeError(f"cannot build an OrbitSolver from {type(obj).__name__}")

    def cfg(self, *, dt0: float | None = None, max_steps="default") -> _abi.GxSolverCfg:
        c = _abi.GxSolverCfg()
        c.method = self.solver.method
        ms = self.max_steps if max_steps == "default" else max_steps
        c.max_steps = 0 if ms is None else int(ms)
        ctl = self.stepsize_controller
        if isinstance(ctl, ConstantStepSize):
            if dt0 is None or not abs(dt0) > 0:
                raise ValueError("ConstantStepSize needs a nonzero dt0")
            c.controller = _abi.GX_CONSTANT
            c.dt0 = float(abs(dt0))
            c.rtol = c.atol = 0.0
            c.safety, c.factormin, c.factormax = 0.9, 0.2, 10.0
            c.icoeff = 1.0
        elif isinstance(ctl, PIDController):
            c.controller = _abi.GX_PID
            c.dt0 = 0.0 if dt0 is None else float(abs(dt0))
            c.rtol, c.atol = float(ctl.rtol), float(ctl.atol)
            c.pcoeff, c.icoeff, c.dcoeff = float(ctl.pcoeff), float(ctl.icoeff), float(ctl.dcoeff)
            c.safety, c.factormin, c.factormax = float(ctl.safety), float(ctl.factormin), float(ctl.factormax)
            c.dtmin = 0.0 if ctl.dtmin is None else float(ctl.dtmin)
            c.dtmax = 0.0 if ctl.dtmax is None else float(ctl.dtmax)
            c.force_dtmin = int(bool(ctl.force_dtmin))
        else:
            raise TypeError(f"unsupported stepsize controller {type(ctl).__name__}")
        return c

    def with_(self, **kw) -> OrbitSolver:
        return replace(self, **kw)

    def solve(self, field, qp, t0, t1, *, saveat=None, dt0=None, max_steps="default"):
        """OrbitSolver.solve (orbit/solver.py:121-141, 774-1039).

        For an NBodyField the bodies integrate as one system.  For a
        HamiltonianField (or a potential) over a batch of orbits, scalar t0
        and t1 solve the batch as one ODE with a shared step sequence, as
        diffeqsolve does over a (B, 6) state; per-orbit t0/t1 give every
        orbit its own steps (the vectorized path, orbit/solver.py:854-1039).
        Returns a solution with ts (T,) and ys = (q, p) of shape (T, B, 3)
        (the saved times first, as diffrax returns them); without `saveat`
        only t1 is saved."""
        from ._nbody import NBodyField, NBodySolution, solve_nbody
        from .fields import HamiltonianField

        if isinstance(field, NBodyField):
            return solve_nbody(field, qp, t0, t1, saveat=saveat, solver=self, dt0=dt0, max_steps=max_steps)
        pot = field.potential if isinstance(field, HamiltonianField) else field
        if not hasattr(pot, "_handle"):
            raise TypeError(f"OrbitSolver.solve needs a potential, HamiltonianField or NBodyField, "
                            f"not {type(field).__name__}")
        from ..units import ustrip

        q, p = qp
        q = np.asarray(ustrip(q, "kpc"), dtype=np.float64).reshape(-1, 3)
        p = np.asarray(ustrip(p, "kpc/Myr"), dtype=np.float64).reshape(-1, 3)
        t0v = np.asarray(ustrip(t0, "Myr"), dtype=np.float64)
        t1v = np.asarray(ustrip(t1, "Myr"), dtype=np.float64)
        shared = t0v.ndim == 0 and t1v.ndim == 0
        ts = None if saveat is None else np.atleast_1d(np.asarray(ustrip(saveat, "Myr"), dtype=np.float64))
        if ts is None and shared:
            ts = np.atleast_1d(t1v)
        res = integrate_orbits(pot, np.concatenate([q, p], 1), t0v, t1v, ts=ts, solver=self, dt0=dt0,
                               max_steps=max_steps, shared=shared)
        out = res.out.cpu().numpy()  # (B, T, 6)
        ys = (np.transpose(out[..., :3], (1, 0, 2)), np.transpose(out[..., 3:], (1, 0, 2)))
        st = res.status.cpu().numpy()
        ns = res.nsteps.cpu().numpy()
        na = res.naccept.cpu().numpy()
        if shared:
            return NBodySolution(ts=ts, ys=ys, status=int(st[0]), nsteps=int(ns[0]), naccept=int(na[0]))
        return NBodySolution(ts=ts if ts is not None else t1v, ys=ys, status=st, nsteps=ns, naccept=na)


class DenseSolution:
    """The dense solution of a solve on the device: every accepted step's
    record (start time, signed step, q, v, stage accelerations; device
    tensors) and the method's dense output over it (gx_dense_eval), as
    diffrax's DenseInterpolation keeps and evaluates them."""

    def __init__(self, method: int, rec, nrec, cap: int):
        self.method = method
        self.rec = rec  # (n, cap, GX_DENSE_RECORD) float64
        self.nrec = nrec  # (n,) int64
        self.cap = cap

    @property
    def n(self) -> int:
        return int(self.nrec.shape[0])

    def evaluate(self, t):
        """States (n, T, 6) at times t: (T,) for every orbit or (n, T) per
        orbit; NaN outside an orbit's integrated interval."""
        torch = _lib.require_gpu()
        td = _dev(torch, t, torch.float64)
        scalar = td.ndim == 0
        td = td.reshape(1) if scalar else td
        per_orbit = 1 if td.ndim == 2 else 0
        if per_orbit and td.shape[0] != self.n:
            raise ValueError("per-orbit times must have shape (n, T)")
        T = int(td.shape[-1])
        out = torch.empty((self.n, T, 6), dtype=torch.float64, device="cuda")
        rc = _lib.lib().gx_dense_eval(self.method, _lib.ptr(self.rec), self.cap, _lib.ptr(self.nrec), self.n,
                                      _lib.ptr(td), per_orbit, T, _lib.ptr(out), _lib.stream_ptr())
        _lib.check(rc, "gx_dense_eval")
        return out


@dataclass
class IntegrationResult:
    """Device-side result of one gx_integrate call (torch tensors)."""

    out: object  # (n, T, 6)
    status: object  # (n,) int32
    nsteps: object  # (n,) int64 attempted steps
    naccept: object  # (n,) int64 accepted steps
    dense: object = None  # DenseSolution when requested

    @property
    def q(self):
        return self.out[..., :3]

    @property
    def p(self):
        return self.out[..., 3:]


def _dev(torch, x, dtype):
    if isinstance(x, torch.Tensor):
        return x.to(device="cuda", dtype=dtype).contiguous()
    return torch.as_tensor(np.array(x, copy=True), dtype=dtype, device="cuda").contiguous()


def integrate_orbits(
    pot,
    w0,
    t0,
    t1,
    *,
    ts=None,
    T: int | None = None,
    ts_mode: int | None = None,
    solver: OrbitSolver | None = None,
    dt0: float | None = None,
    max_steps="default",
    order=None,
    validate_order: bool = True,
    out=None,
    workspace=None,
    stream=None,
    shared: bool = False,
    dense: bool | int = False,
    _records=None,
) -> IntegrationResult:
    """Integrate n orbits on the GPU.

    w0: (n, 6); t0, t1: scalars or (n,); ts: (T,) shared save times, (n, T)
    per-orbit save times, or None (with `T`: per-orbit linspace(t0, t1, T);
    without: the final state only).  Every orbit takes its own step sequence
    unless `shared`: then the batch is one system with one step sequence
    (scalar t0, t1 and shared or no save times), as galax's Integrator
    solves a batch with scalar times.  `dense` (True, or a record capacity
    per orbit) also keeps every accepted step's dense-output record
    (`result.dense`, a DenseSolution); a capacity too small for an orbit is
    raised until it fits.  Returns device tensors."""
    if dense:
        return _integrate_dense(pot, w0, t0, t1, ts=ts, T=T, solver=solver, dt0=dt0, max_steps=max_steps,
                                ts_mode=ts_mode, order=order, validate_order=validate_order, out=out,
                                workspace=workspace, stream=stream, shared=shared,
                                cap=None if dense is True else int(dense))
    torch = _lib.require_gpu()
    solver = solver or OrbitSolver()
    cfg = solver.cfg(dt0=dt0, max_steps=max_steps)
    w0d = _dev(torch, w0, torch.float64)
    if w0d.ndim != 2 or w0d.shape[1] != 6:
        raise ValueError(f"w0 must have shape (n, 6), got {tuple(w0d.shape)}")
    n = int(w0d.shape[0])
    if shared:
        return _integrate_shared(torch, pot, cfg, w0d, n, t0, t1, ts, T, out, workspace, stream, _records)
    t0d = _dev(torch, t0, torch.float64).reshape(-1)
    t1d = _dev(torch, t1, torch.float64).reshape(-1)
    if t0d.numel() == 1 and t1d.numel() == 1:
        t_stride = 0
    else:
        t0d = torch.broadcast_to(t0d, (n,)).contiguous()
        t1d = torch.broadcast_to(t1d, (n,)).contiguous()
        t_stride = 1
    tsd = None
    if ts is not None:
        tsd = _dev(torch, ts, torch.float64)
        if tsd.ndim == 1:
            mode, T = _abi.GX_TS_SHARED, int(tsd.shape[0])
        elif tsd.ndim == 2 and tsd.shape[0] == n:
            mode, T = _abi.GX_TS_PER_ORBIT, int(tsd.shape[1])
        else:
            raise ValueError("ts must have shape (T,) or (n, T)")
    elif T is not None:
        mode = _abi.GX_TS_LINSPACE
    else:
        mode, T = _abi.GX_TS_FINAL, 1
    if tsd is not None and tsd.shape[-1] > 1:
        # saves are served in order along the integration (diffrax requires
        # ordered SaveAt(ts) too)
        step = tsd[..., 1:] - tsd[..., :-1]
        sgn = torch.sign(t1d - t0d).reshape(-1, 1)
        if bool(torch.any(step.reshape(-1 if tsd.ndim == 2 else 1, T - 1) * sgn < 0)):
            raise ValueError("save times must be ordered in the direction of integration")
    if ts_mode is not None:
        mode = ts_mode
    if T < 1:
        raise ValueError("need at least one save time")
    if out is None:
        out = torch.empty((n, T, 6), dtype=torch.float64, device="cuda")
    else:
        _check_out(torch, out, n, T)
    status = torch.empty((n,), dtype=torch.int32, device="cuda")
    nsteps = torch.empty((n,), dtype=torch.int64, device="cuda")
    naccept = torch.empty((n,), dtype=torch.int64, device="cuda")
    lib = _lib.lib()
    nbytes = int(lib.gx_integrate_workspace_bytes(n))
    if workspace is None:
        workspace = torch.empty((nbytes + 7) // 8, dtype=torch.float64, device="cuda")
    else:
        _check_workspace(torch, workspace, nbytes)
    orderd = None
    if order is not None:
        orderd = _dev(torch, order, torch.int64)
        if orderd.shape != (n,):
            raise ValueError("order must be a permutation of range(n)")
        if validate_order and not bool(
            torch.equal(torch.sort(orderd).values, torch.arange(n, device="cuda", dtype=torch.int64))
        ):
            raise ValueError("order must be a permutation of range(n)")
    args = [pot._handle(), __import__("ctypes").addressof(cfg), _lib.ptr(w0d), _lib.ptr(t0d), _lib.ptr(t1d),
            t_stride, _lib.ptr(tsd), T, mode, n, _lib.ptr(orderd), _lib.ptr(out), _lib.ptr(status), _lib.ptr(nsteps),
            _lib.ptr(naccept)]
    tail = [_lib.ptr(workspace), stream if stream is not None else _lib.stream_ptr()]
    if _records is None:
        rc = lib.gx_integrate(*args, *tail)
    else:
        rec, cap, nrec = _records
        rc = lib.gx_integrate_dense(*args, _lib.ptr(rec), cap, _lib.ptr(nrec), *tail)
    _lib.check(rc, "gx_integrate")
    return IntegrationResult(out=out, status=status, nsteps=nsteps, naccept=naccept)


def _integrate_dense(pot, w0, t0, t1, *, cap, shared, **kw) -> IntegrationResult:
    """integrate_orbits with dense records: the solve runs with `cap`
    records per orbit (default 1024) and again with eight times more while
    an orbit's accepted steps do not fit."""
    torch = _lib.require_gpu()
    solver = kw.get("solver") or OrbitSolver()
    fixed_cap = cap is not None
    if cap is None:
        # a first capacity within about 256 MB of records
        n = int(np.shape(w0)[0]) if not hasattr(w0, "shape") else int(w0.shape[0])
        cap = int(min(65536, max(1024, 256e6 / (8 * _abi.GX_DENSE_RECORD * max(n, 1)))))
    cap = int(cap)
    if cap < 1:
        raise ValueError("dense record capacity must be positive")
    cfg = solver.cfg(dt0=kw.get("dt0"), max_steps=kw.get("max_steps", "default"))
    limit = cfg.max_steps if cfg.max_steps > 0 else 10_000_000
    while True:
        res = _integrate_records(torch, pot, w0, t0, t1, cap=cap, shared=shared, **kw)
        overflow = (res.status == _abi.GX_MAX_STEPS_REACHED) & (res.naccept > cap)
        if fixed_cap or not bool(overflow.any()) or cap >= limit:
            return res
        cap = min(8 * cap, limit)


def _integrate_records(torch, pot, w0, t0, t1, *, cap, shared, ts=None, T=None, solver=None, dt0=None,
                       max_steps="default", ts_mode=None, order=None, validate_order=True, out=None,
                       workspace=None, stream=None) -> IntegrationResult:
    solver = solver or OrbitSolver()
    w0d = _dev(torch, w0, torch.float64)
    n = int(w0d.shape[0])
    rec = torch.empty((n, cap, _abi.GX_DENSE_RECORD), dtype=torch.float64, device="cuda")
    nrec = torch.zeros((n,), dtype=torch.int64, device="cuda")
    res = integrate_orbits(pot, w0d, t0, t1, ts=ts, T=T, solver=solver, dt0=dt0, max_steps=max_steps,
                           ts_mode=ts_mode, order=order, validate_order=validate_order, out=out,
                           workspace=workspace, stream=stream, shared=shared, _records=(rec, cap, nrec))
    res.dense = DenseSolution(solver.cfg(dt0=dt0).method, rec, nrec, cap)
    return res


def _check_out(torch, out, n: int, T: int):
    """A caller's output buffer must be exactly what the kernels write:
    float64 (n, T, 6), contiguous, on the GPU."""
    if not isinstance(out, torch.Tensor):
        raise ValueError("out must be a torch tensor")
    if tuple(out.shape) != (n, T, 6):
        raise ValueError(f"out must have shape {(n, T, 6)}, got {tuple(out.shape)}")
    if out.dtype != torch.float64 or out.device.type != "cuda" or not out.is_contiguous():
        raise ValueError("out must be a contiguous float64 CUDA tensor")


def _check_workspace(torch, ws, nbytes: int):
    """A caller's workspace must be a contiguous CUDA buffer of at least
    nbytes, 256-byte aligned (the library's counters and records sit at its
    start)."""
    if not isinstance(ws, torch.Tensor):
        raise ValueError("workspace must be a torch tensor")
    if ws.numel() * ws.element_size() < nbytes:
        raise ValueError(f"workspace holds {ws.numel() * ws.element_size()} bytes; this call needs {nbytes}")
    if ws.device.type != "cuda" or not ws.is_contiguous():
        raise ValueError("workspace must be a contiguous CUDA tensor")
    if ws.data_ptr() % 256:
        raise ValueError("workspace must be 256-byte aligned")


def _integrate_shared(torch, pot, cfg, w0d, n, t0, t1, ts, T, out, workspace, stream,
                      records=None) -> IntegrationResult:
    t0v = np.asarray(t0.cpu() if isinstance(t0, torch.Tensor) else t0, dtype=np.float64)
    t1v = np.asarray(t1.cpu() if isinstance(t1, torch.Tensor) else t1, dtype=np.float64)
    if t0v.size != 1 or t1v.size != 1:
        raise ValueError("a shared step sequence needs scalar t0 and t1")
    t0f, t1f = float(t0v.reshape(())), float(t1v.reshape(()))
    if ts is None:
        ts = np.linspace(t0f, t1f, T) if T is not None else np.array([t1f])
    tsd = _dev(torch, ts, torch.float64)
    if tsd.ndim != 1:
        raise ValueError("a shared step sequence needs shared save times ts of shape (T,)")
    T = int(tsd.shape[0])
    if T < 1:
        raise ValueError("need at least one save time")
    if T > 1 and bool(torch.any((tsd[1:] - tsd[:-1]) * float(np.sign(t1f - t0f)) < 0)):
        raise ValueError("save times must be ordered in the direction of integration")
    if out is None:
        out = torch.empty((n, T, 6), dtype=torch.float64, device="cuda")
    else:
        _check_out(torch, out, n, T)
    status = torch.empty((n,), dtype=torch.int32, device="cuda")
    nsteps = torch.empty((n,), dtype=torch.int64, device="cuda")
    naccept = torch.empty((n,), dtype=torch.int64, device="cuda")
    lib = _lib.lib()
    nbytes = int(lib.gx_integrate_shared_workspace_bytes(n))
    if workspace is None:
        workspace = torch.empty((nbytes + 7) // 8, dtype=torch.float64, device="cuda")
    else:
        _check_workspace(torch, workspace, nbytes)
    args = [pot._handle(), __import__("ctypes").addressof(cfg), _lib.ptr(w0d), t0f, t1f, _lib.ptr(tsd), T, n,
            _lib.ptr(out), _lib.ptr(status), _lib.ptr(nsteps), _lib.ptr(naccept)]
    tail = [_lib.ptr(workspace), stream if stream is not None else _lib.stream_ptr()]
    if records is None:
        rc = lib.gx_integrate_shared(*args, *tail)
    else:
        rec, cap, nrec = records
        rc = lib.gx_integrate_shared_dense(*args, _lib.ptr(rec), cap, _lib.ptr(nrec), *tail)
    _lib.check(rc, "gx_integrate_shared")
    return IntegrationResult(out=out, status=status, nsteps=nsteps, naccept=naccept)
