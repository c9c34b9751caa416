"""integrate_orbit of galax's experimental API (experimental/integrate.py).

The default solver is the reference's: Dopri8 with PIDController(rtol=1e-7,
atol=1e-7, dtmin=0.3, force_dtmin=True) and max_steps=10,000
(integrate.py:28-35).  One orbit, ((3,), (3,)), returns ys of shape
((T, 3), (T, 3)); a batch ((B, 3), (B, 3)) returns ((B, T, 3), (B, T, 3)),
each orbit with its own steps (the VMap/Scan loop strategies,
integrate.py:446-570).  `saveat` is shared (T,) or per orbit (B, T).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ...units import ustrip
from .._orbit import HamiltonianField
from .._solver import Dopri8, OrbitSolver, PIDController, integrate_orbits

__all__ = ["Solution", "default_solver", "integrate_orbit", "parse_t0_t1_saveat"]

default_solver = OrbitSolver(
    solver=Dopri8(scan_kind="bounded"),
    stepsize_controller=PIDController(rtol=1e-7, atol=1e-7, dtmin=0.3, dtmax=None, force_dtmin=True),
    max_steps=10_000,
)


@dataclass
class Solution:
    """The fields of diffrax's Solution that galax reads: t0, t1, ts, ys =
    (q, p), the per-orbit result code and step count, and the dense
    interpolation when requested (ts and ys are then None)."""

    t0: object
    t1: object
    ts: np.ndarray | None
    ys: tuple | None
    result: np.ndarray | int
    num_steps: np.ndarray | int
    interpolation: object = None

    def evaluate(self, t):
        """The dense solution at time(s) t (dense=True only)."""
        if self.interpolation is None:
            raise ValueError("this solution was computed with dense=False")
        return self.interpolation(t)


def parse_t0_t1_saveat(t0, t1, saveat, *, dense: bool):
    """(t0, t1, save times) as the reference parses them (integrate.py:38-142):
    no saveat saves t1 only; a scalar saveat is the end time, saving t0 and
    it; an array saves at its times, with t0 = min and t1 = max where they
    are not given."""
    ts = None if saveat is None else np.asarray(ustrip(saveat, "Myr"), dtype=np.float64)
    t0 = None if t0 is None else np.asarray(ustrip(t0, "Myr"), dtype=np.float64)
    t1 = None if t1 is None else np.asarray(ustrip(t1, "Myr"), dtype=np.float64)
    if ts is None:
        if t0 is None or t1 is None:
            raise ValueError("t0, t1 must be specified if saveat is None")
        return t0, t1, (None if dense else t1[..., None] if t1.ndim else np.atleast_1d(t1))
    if ts.ndim == 0:
        if t0 is None:
            raise ValueError("t0 must be specified if saveat is a scalar")
        return t0, ts, (None if dense else np.stack([np.broadcast_to(t0, ts.shape), ts], -1))
    if ts.shape[-1] < 2 and (t0 is None or t1 is None):
        raise ValueError("if t0 or t1 are None, saveat must be [t0, ..., t1]")
    t0 = ts.min(axis=-1) if t0 is None else t0
    t1 = ts.max(axis=-1) if t1 is None else t1
    return t0, t1, (None if dense else ts)


def integrate_orbit(pot, qp0, t0=None, t1=None, *, saveat=None, solver: OrbitSolver | None = None,
                    solver_kwargs: dict | None = None, dense: bool = False, dense_vectorize: bool = False):
    """Integrate one orbit or a batch (experimental/integrate.py:149-570)."""
    pot = pot.potential if isinstance(pot, HamiltonianField) else pot
    solver = default_solver if solver is None else OrbitSolver.from_(solver)
    kw = dict(solver_kwargs or {})
    q, p = qp0
    q = np.asarray(ustrip(q, "kpc"), dtype=np.float64)
    p = np.asarray(ustrip(p, "kpc/Myr"), dtype=np.float64)
    single = q.ndim == 1
    w0 = np.concatenate([np.atleast_2d(q), np.atleast_2d(p)], -1)
    n = w0.shape[0]
    t0v, t1v, ts = parse_t0_t1_saveat(t0, t1, saveat, dense=dense)
    if dense:
        res = integrate_orbits(pot, w0, np.broadcast_to(t0v, (n,)), np.broadcast_to(t1v, (n,)), solver=solver,
                               dt0=kw.get("dt0"), max_steps=kw.get("max_steps", "default"), dense=True)
        status = res.status.cpu().numpy()
        nsteps = res.nsteps.cpu().numpy()
        sol_dense = res.dense

        def interpolation(t, _d=sol_dense):
            o = _d.evaluate(np.atleast_1d(np.asarray(ustrip(t, "Myr"), dtype=np.float64))).cpu().numpy()
            qs, ps = o[..., :3], o[..., 3:]
            return (qs[0], ps[0]) if single else (qs, ps)

        return Solution(t0=t0v, t1=t1v, ts=None, ys=None, result=status[0] if single else status,
                        num_steps=nsteps[0] if single else nsteps, interpolation=interpolation)
    res = integrate_orbits(pot, w0, np.broadcast_to(t0v, (n,)), np.broadcast_to(t1v, (n,)),
                           ts=np.broadcast_to(ts, (n, ts.shape[-1])) if ts.ndim == 2 else ts,
                           solver=solver, dt0=kw.get("dt0"), max_steps=kw.get("max_steps", "default"))
    out = res.out.cpu().numpy()  # (n, T, 6)
    status = res.status.cpu().numpy()
    nsteps = res.nsteps.cpu().numpy()
    ys = (out[..., :3], out[..., 3:])
    if single:
        ys = (ys[0][0], ys[1][0])
        return Solution(t0=t0v, t1=t1v, ts=ts, ys=ys, result=int(status[0]), num_steps=int(nsteps[0]))
    return Solution(t0=t0v, t1=t1v, ts=ts, ys=ys, result=status, num_steps=nsteps)
