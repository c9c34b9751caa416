"""Phase-space containers, `Orbit`, `Integrator` and `evaluate_orbit`.

  PhaseSpacePosition / PhaseSpaceCoordinate  coordinates/_src/psps, pscs
  Orbit(q, p, t, frame, interpolant)         dynamics/_src/orbit/orbit.py:17-75
  Integrator                                 dynamics/_src/legacy/integrator.py:43-245
  evaluate_orbit                             dynamics/_src/legacy/funcs.py:42-213
  HamiltonianField                           dynamics/_src/orbit/field_hamiltonian.py:22-333

Arrays are numpy float64 in galactic units unless the caller passed torch
tensors, in which case results stay on the device.
"""

from __future__ import annotations

import numpy as np

from .. import _lib
from ..units import Q, ustrip
from ._solver import ConstantStepSize, Dopri8, OrbitSolver, PIDController, integrate_orbits

__all__ = [
    "PhaseSpacePosition",
    "PhaseSpaceCoordinate",
    "Orbit",
    "Integrator",
    "HamiltonianField",
    "evaluate_orbit",
    "parse_w0",
]


def _arr(x, unit):
    if isinstance(x, Q):
        return np.asarray(x.to_value(unit), dtype=np.float64)
    return x


class PhaseSpacePosition:
    """Position and velocity without a time (galax gc.PhaseSpacePosition)."""

    def __init__(self, q, p):
        self.q = _arr(q, "kpc")
        self.p = _arr(p, "kpc/Myr")
        qs, ps = np.shape(self.q), np.shape(self.p)
        bshape = np.broadcast_shapes(qs[:-1], ps[:-1])
        self.q = np.broadcast_to(self.q, bshape + (3,))
        self.p = np.broadcast_to(self.p, bshape + (3,))

    @property
    def shape(self):
        return np.shape(self.q)[:-1]

    def w(self, units="galactic"):
        return np.concatenate([np.asarray(self.q), np.asarray(self.p)], axis=-1)


class PhaseSpaceCoordinate(PhaseSpacePosition):
    """Position, velocity and time (galax gc.PhaseSpaceCoordinate)."""

    def __init__(self, q, p, t, frame="simulation"):
        super().__init__(q, p)
        self.t = np.asarray(_arr(t, "Myr"), dtype=np.float64)
        self.frame = frame

    def __repr__(self):
        return f"PhaseSpaceCoordinate(q={self.q!r}, p={self.p!r}, t={self.t!r})"


def parse_w0(w0):
    """(w (n,6) array, batch shape, time or None) from galax's w0 forms:
    PhaseSpaceCoordinate/Position, (q, p) tuple, or (..., 6) array."""
    torch = None
    t = None
    if isinstance(w0, PhaseSpacePosition):
        w = w0.w()
        if isinstance(w0, PhaseSpaceCoordinate):
            t = w0.t
    elif isinstance(w0, tuple) and len(w0) == 2:
        q, p = _arr(w0[0], "kpc"), _arr(w0[1], "kpc/Myr")
        q, p = np.broadcast_arrays(np.asarray(q, dtype=np.float64), np.asarray(p, dtype=np.float64))
        w = np.concatenate([q, p], axis=-1)
    else:
        try:
            import torch as _t

            torch = _t
        except ImportError:  # pragma: no cover
            pass
        if torch is not None and isinstance(w0, torch.Tensor):
            w = w0
        else:
            w = np.asarray(w0, dtype=np.float64)
    if w.shape[-1] != 6:
        raise ValueError(f"w0 must have a trailing axis of 6, got {tuple(w.shape)}")
    batch = tuple(w.shape[:-1])
    w = w.reshape(-1, 6)
    return w, batch, t


class HamiltonianField:
    """dq/dt = p, dp/dt = -grad Phi(q, t) for a potential."""

    def __init__(self, potential):
        self.potential = potential

    @property
    def units(self):
        return "galactic"

    def __call__(self, t, qp, args=None):
        q, p = qp
        return p, self.potential.acceleration(q, t)


class Orbit:
    """q, p with shape (*batch, T, 3) and t with shape (T,)."""

    def __init__(self, q, p, t, frame="simulation", interpolant=None, status=None, nsteps=None):
        self.q = q
        self.p = p
        self.t = t
        self.frame = frame
        self.interpolant = interpolant
        self.status = status
        self.nsteps = nsteps

    @property
    def shape(self):
        return tuple(self.q.shape[:-1])

    def w(self, units="galactic"):
        if isinstance(self.q, np.ndarray):
            return np.concatenate([self.q, self.p], axis=-1)
        torch = _lib.torch()
        return torch.cat([self.q, self.p], dim=-1)

    def __call__(self, t):
        if self.interpolant is None:
            raise ValueError("this orbit was computed with dense=False")
        return self.interpolant(t)

    def __getitem__(self, idx):
        return Orbit(self.q[idx], self.p[idx], self.t, self.frame, self.interpolant)

    def __repr__(self):
        return f"Orbit(q={tuple(self.q.shape)}, p={tuple(self.p.shape)}, t={tuple(np.shape(self.t))}, frame={self.frame!r})"


class OrbitInterpolant:
    """Dense orbit (galax's Orbit interpolant, orbit/interp.py:42-145, over
    diffrax's DenseInterpolation): the solve's accepted-step records stay on
    the device and every query is the method's dense output over the step
    holding it, one GPU lane per (orbit, time) -- the same polynomial the
    solve's own saves come from."""

    def __init__(self, dense, batch=None):
        self.dense = dense  # DenseSolution
        self.batch = batch

    def __call__(self, t):
        t = np.atleast_1d(np.asarray(ustrip(t, "Myr"), dtype=np.float64))
        out = self.dense.evaluate(t).cpu().numpy()
        if self.batch is not None:
            out = out.reshape(*self.batch, t.size, 6)
        return PhaseSpaceCoordinate(q=out[..., :3], p=out[..., 3:], t=t)


class Integrator:
    """galax's legacy Integrator: Dopri8 with PIDController(rtol=1e-7,
    atol=1e-7) and unbounded max_steps by default.

    As in galax (legacy/integrator.py:269-355 and 449-527), a batch with
    scalar start and end times is solved as one system with a shared step
    sequence, and a batch with per-orbit times as independent orbits.
    `shared_steps=False` gives every orbit its own steps in both cases (the
    fast path for large batches of independent orbits); fixed-step runs are
    identical either way."""

    def __init__(self, dynamics_solver=None, *, diffeq_kw=None, shared_steps: bool | None = None):
        self.shared_steps = shared_steps
        if dynamics_solver is None:
            dynamics_solver = OrbitSolver(
                solver=Dopri8(scan_kind="bounded"),
                stepsize_controller=PIDController(rtol=1e-7, atol=1e-7),
                max_steps=2**16,
            )
        self.dynamics_solver = OrbitSolver.from_(dynamics_solver)
        kw = {"max_steps": None}
        kw.update(diffeq_kw or {})
        self.diffeq_kw = kw

    def _shared(self, w, t0, t1) -> bool:
        """galax's rule: scalar times couple the batch.  A single orbit, or a
        fixed step, gives the same steps either way, so those take the
        independent path."""
        if self.shared_steps is False:
            return False
        scalar = np.ndim(t0) == 0 and np.ndim(t1) == 0
        adaptive = self.dynamics_solver.stepsize_controller.__class__.__name__ == "PIDController"
        return scalar and adaptive and int(np.shape(w)[0]) > 1

    def _run(self, pot, w, t0, t1, ts=None, dense=False):
        kw = self.diffeq_kw
        return integrate_orbits(
            pot,
            w,
            t0,
            t1,
            ts=ts,
            solver=self.dynamics_solver,
            dt0=kw.get("dt0"),
            max_steps=kw.get("max_steps", "default"),
            shared=self._shared(w, t0, t1) and (ts is None or np.ndim(ts) == 1),
            dense=dense,
        )

    def __call__(self, field, w0, t0, t1, *, saveat=None, dense=False):
        pot = field.potential if isinstance(field, HamiltonianField) else field
        w, batch, _ = parse_w0(w0)
        t0v = np.asarray(ustrip(t0, "Myr"), dtype=np.float64)
        t1v = np.asarray(ustrip(t1, "Myr"), dtype=np.float64)
        ts = None if saveat is None else np.atleast_1d(np.asarray(ustrip(saveat, "Myr"), dtype=np.float64))
        res = self._run(pot, w, t0v, t1v, ts)
        out = res.out.cpu().numpy()
        T = out.shape[1]
        if saveat is None:
            out = out[:, 0, :].reshape(*batch, 6)
            tt = t1v
        else:
            out = out.reshape(*batch, T, 6)
            tt = ts
        return PhaseSpaceCoordinate(q=out[..., :3], p=out[..., 3:], t=tt)


default_integrator = Integrator()


# evaluate_orbit serves more than _MANY_SAVES save times of at most
# _FEW_ORBITS orbits from dense records
_MANY_SAVES = 256
_FEW_ORBITS = 64


def evaluate_orbit(pot, w0, t, *, integrator: Integrator | None = None, dense: bool = False) -> Orbit:
    """Integrate orbits in `pot`, returning them at the times `t`.

    As in galax (legacy/funcs.py:42-213): a PhaseSpaceCoordinate w0 is first
    integrated from w0.t to t[0] (per-orbit times: independent orbits); then
    the batch is integrated from t[0] to t[-1] and saved at each t, as one
    system with a shared adaptive step sequence, as galax does (scalar
    times).  `Integrator(shared_steps=False)` gives each orbit its own steps
    instead."""
    integrator = integrator or default_integrator
    t = np.atleast_1d(np.asarray(ustrip(t, "Myr"), dtype=np.float64))
    w, batch, tw0 = parse_w0(w0)
    torch = _lib.require_gpu()
    is_torch = isinstance(w, torch.Tensor)
    n = w.shape[0]
    # stage 1: w0.t -> t[0]
    if tw0 is not None:
        tw0 = np.broadcast_to(np.asarray(tw0, dtype=np.float64), batch).reshape(-1)
        if np.any(tw0 != t[0]):
            r1 = integrator._run(pot, w, tw0, np.full(n, t[0]))
            w = r1.out[:, 0, :]
    # stage 2: t[0] -> t[-1], saving at t (and keeping the dense records)
    T = t.shape[0]
    if T > _MANY_SAVES and n <= _FEW_ORBITS:
        # many save times on few orbits (a stream's progenitor): rather than
        # one lane serving every save in turn, keep the step records and
        # evaluate the same dense output at all the times in parallel
        r2 = integrator._run(pot, w, t[0], t[-1], dense=True)
        out = r2.dense.evaluate(t)
        out[:, -1, :] = r2.out[:, 0, :]  # the end point as the solve left it
    else:
        r2 = integrator._run(pot, w, t[0], t[-1], ts=t, dense=dense)
        out = r2.out
    if not is_torch:
        out = out.cpu().numpy()
    out = out.reshape(*batch, T, 6)
    interp = OrbitInterpolant(r2.dense, batch) if dense else None
    status = r2.status if is_torch else r2.status.cpu().numpy().reshape(batch)
    nsteps = r2.nsteps if is_torch else r2.nsteps.cpu().numpy().reshape(batch)
    return Orbit(out[..., :3], out[..., 3:], t, interpolant=interp, status=status, nsteps=nsteps)
