"""StreamSimulator of galax's experimental API (experimental/stream.py).

`init` integrates the progenitor to the sorted release times and draws the
release states from a kinematic DF; `run` integrates every released particle
from its release time to t1.  The reference's default solver here is Dopri5
with PIDController(rtol=1e-7, atol=1e-7, dtmin=0.3, force_dtmin=True) and
max_steps=10,000 (stream.py:32-39).  Both arms of all particles go to the GPU
as one batch of independent orbits (the reference vmaps over them,
stream.py:384-446).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ...units import ustrip
from .._solver import Dopri5, OrbitSolver, PIDController
from .df import AbstractKinematicDF, Fardal2015DF
from .integrate import integrate_orbit

__all__ = ["StreamICs", "StreamSimulator", "default_solver", "default_kinematic_df"]

default_solver = OrbitSolver(
    solver=Dopri5(scan_kind="bounded"),
    stepsize_controller=PIDController(rtol=1e-7, atol=1e-7, dtmin=0.3, dtmax=None, force_dtmin=True),
    max_steps=10_000,
)

default_kinematic_df = Fardal2015DF()


@dataclass
class StreamICs:
    """Release times, progenitor mass at each, and the leading and trailing
    release states (x (N, 3), v (N, 3)) (stream.py:46-56)."""

    release_times: np.ndarray
    prog_mass: np.ndarray
    qp_lead: tuple
    qp_trail: tuple


class StreamSimulator:
    """Simulate a stellar stream (stream.py:66-446)."""

    def init(self, pot, prog_w0, prog_t0, /, release_times, Msat, kinematic_df: AbstractKinematicDF | None = None,
             *, key=0, solver: OrbitSolver = default_solver, solver_kwargs: dict | None = None) -> StreamICs:
        """Release states along the progenitor's orbit (stream.py:110-236).

        `key` seeds NumPy's PCG64 (or is a numpy Generator) in place of a JAX
        PRNG key."""
        release_times = np.sort(np.asarray(ustrip(release_times, "Myr"), dtype=np.float64).reshape(-1))
        t0 = float(np.asarray(ustrip(prog_t0, "Myr")))
        # to the first release time, then saved at every release time
        w = integrate_orbit(pot, prog_w0, t0, release_times[0], solver=solver, solver_kwargs=solver_kwargs).ys
        w = (w[0][-1], w[1][-1])
        sol = integrate_orbit(pot, w, saveat=release_times, solver=solver, solver_kwargs=solver_kwargs)
        xs, vs = sol.ys
        df = default_kinematic_df if kinematic_df is None else kinematic_df
        m = np.broadcast_to(np.asarray(ustrip(Msat, "Msun"), dtype=np.float64), release_times.shape).copy()
        xl, vl, xt, vt = df.sample(key, pot, t=release_times, x=xs, v=vs, Msat=m)
        return StreamICs(release_times=release_times, prog_mass=m, qp_lead=(xl, vl), qp_trail=(xt, vt))

    def run(self, pot, prog: StreamICs, /, *, t1, dense: bool = False, solver: OrbitSolver = default_solver,
            solver_kwargs: dict | None = None):
        """Every particle from its release time to t1 (stream.py:384-446):
        returns ((x_lead, v_lead), (x_trail, v_trail)), each (N, 3), or with
        dense=True the dense solutions of the two arms."""
        t1 = float(np.asarray(ustrip(t1, "Myr")))
        n = prog.release_times.size
        q = np.concatenate([prog.qp_lead[0], prog.qp_trail[0]])
        p = np.concatenate([prog.qp_lead[1], prog.qp_trail[1]])
        t0 = np.concatenate([prog.release_times, prog.release_times])
        sol = integrate_orbit(pot, (q, p), t0, np.full(2 * n, t1), solver=solver, solver_kwargs=solver_kwargs,
                              dense=dense)
        if dense:
            return sol
        xs, vs = sol.ys[0][:, -1], sol.ys[1][:, -1]
        return (xs[:n], vs[:n]), (xs[n:], vs[n:])
