"""Gradient of a stream-track loss with respect to Milky Way parameters.

The pipeline follows MockStreamGenerator.run
(legacy/mockstream/mockstream_generator.py:161-275) with fixed-step Dopri8:

1. the progenitor is integrated backwards from its present-day state,
   saving at the release times;
2. leading and trailing particles are released there with the Fardal (2015)
   offsets (df/fardal15.py:50-94) for fixed k-draws;
3. every particle is integrated forwards from its release time to t = 0;
4. the loss is the mean squared distance of the particles to the nearest
   point of a target track.

Steps 1 and 3 are `OrbitFlow` (forward in gx_integrate, reverse in the
checkpointed discrete adjoint gx_adjoint); step 2 and the loss are torch
operations, so one backward pass gives dL/dtheta through the progenitor,
the release conditions and the particles.  The tidal radius is evaluated at
the reference parameters theta0 and held fixed (it enters only through the
release offsets).  galax obtains such derivatives with jax.jacfwd through
the whole generator (tests/functional/test_mockstreamgenerator.py:24-116).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from .adjoint import FlowSpec, MilkyWayParams, OrbitFlow

__all__ = ["StreamFitProblem", "make_problem", "stream_loss", "release_torch"]


def _cross(a, b):
    torch = _lib.torch()
    return torch.linalg.cross(a, b, dim=-1)


def release_torch(x, v, r_t, k):
    """Fardal (2015) leading/trailing release states, as
    FardalStreamDF.release, in differentiable torch operations."""
    torch = _lib.torch()
    r = x.norm(dim=-1, keepdim=True)
    L = _cross(x, v)
    om = L.norm(dim=-1, keepdim=True) / r**2
    r_hat = x / r
    z_hat = L / L.norm(dim=-1, keepdim=True)
    phi_vec = v - (v * r_hat).sum(-1, keepdim=True) * r_hat
    phi_hat = phi_vec / phi_vec.norm(dim=-1, keepdim=True)
    rt = r_t[:, None]
    vc = om * rt
    kr, kvphi, kz, kvz = (k[:, i : i + 1] for i in range(4))
    x_trail = x + rt * (kr * r_hat + kz * z_hat)
    v_trail = v + vc * (kvphi * phi_hat + kvz * z_hat)
    x_lead = x - rt * (kr * r_hat - kz * z_hat)
    v_lead = v - vc * (kvphi * phi_hat - kvz * z_hat)
    return torch.cat([torch.cat([x_lead, v_lead], -1), torch.cat([x_trail, v_trail], -1)], 0)


@dataclass
class StreamFitProblem:
    """Inputs of the stream-track loss (galactic units)."""

    w_prog: np.ndarray  # (6,) present-day progenitor state
    ts: np.ndarray  # (N,) ascending release times, all < 0
    k: np.ndarray  # (N, 4) Fardal draws
    r_t: np.ndarray  # (N,) tidal radii at theta0
    target: np.ndarray  # (M, 3) track points
    dt: float = 1.0
    K: int = 64

    @property
    def n_particles(self) -> int:
        return 2 * len(self.ts)


def _progenitor_states(theta, prob: StreamFitProblem):
    """Progenitor states at the release times, (N, 6), ascending in time."""
    torch = _lib.require_gpu()
    w = torch.as_tensor(prob.w_prog, dtype=torch.float64, device="cuda").reshape(1, 6)
    ts_desc = np.ascontiguousarray(prob.ts[::-1])
    spec = FlowSpec(t0=0.0, t1=float(prob.ts[0]), dt0=prob.dt, ts=ts_desc, K=prob.K)
    out = OrbitFlow.apply(theta, w, spec)  # (1, N, 6), descending time
    return torch.flip(out[0], dims=[0])


def stream_loss(theta, prob: StreamFitProblem, return_positions: bool = False):
    """L(theta) = mean_i min_j |x_i(0) - y_j|^2 over the 2N particles."""
    torch = _lib.require_gpu()
    prog = _progenitor_states(theta, prob)
    r_t = torch.as_tensor(prob.r_t, dtype=torch.float64, device="cuda")
    k = torch.as_tensor(prob.k, dtype=torch.float64, device="cuda")
    w0 = release_torch(prog[:, :3], prog[:, 3:], r_t, k)
    t0 = np.concatenate([prob.ts, prob.ts])
    spec = FlowSpec(t0=t0, t1=0.0, dt0=prob.dt, K=prob.K)
    out = OrbitFlow.apply(theta, w0, spec)  # (2N, 1, 6)
    pos = out[:, 0, :3]
    y = torch.as_tensor(prob.target, dtype=torch.float64, device="cuda")
    d2 = ((pos[:, None, :] - y[None, :, :]) ** 2).sum(-1)
    loss = d2.min(dim=1).values.mean()
    return (loss, pos) if return_positions else loss


def make_problem(n_release: int = 10_000, span: float = 3000.0, seed: int = 6, n_target: int = 200,
                 target_scale=(1.05, 1.0, 1.0, 1.0, 1.0, 1.0), prog_mass: float = 1e5, dt: float = 1.0,
                 K: int = 64) -> StreamFitProblem:
    """A stream from a progenitor on a mildly eccentric orbit at R = 15 kpc,
    released at n_release times over `span` Myr; the target track is 200
    seeded particles of the same stream grown in a Milky Way whose
    parameters are theta0 * target_scale."""
    torch = _lib.require_gpu()
    from .dynamics import ConstantStepSize, Dopri8, FardalStreamDF, OrbitSolver, integrate_orbits
    from .dynamics._mockstream import tidal_radius

    theta0 = MilkyWayParams.default.copy()
    pot0 = MilkyWayParams.potential(theta0)
    x = np.array([15.0, 0.0, 0.5])
    vc = float(np.asarray(pot0.local_circular_velocity(x, 0.0)).reshape(-1)[0])
    w_prog = np.concatenate([x, [0.0, 0.85 * vc, 0.12 * vc]])
    ts = np.linspace(-span, 0.0, n_release + 1)[:-1]
    k = FardalStreamDF().draw_k(seed, n_release)
    solver = OrbitSolver(Dopri8(), stepsize_controller=ConstantStepSize(), max_steps=None)
    res = integrate_orbits(pot0, w_prog[None], 0.0, float(ts[0]), ts=ts[::-1].copy(), solver=solver, dt0=dt)
    prog = res.out[0].cpu().numpy()[::-1]
    r_t = tidal_radius(pot0, prog[:, :3], prog[:, 3:], mass=np.full(n_release, prog_mass), t=ts)
    prob = StreamFitProblem(w_prog=w_prog, ts=ts, k=k, r_t=np.asarray(r_t), target=np.zeros((1, 3)), dt=dt, K=K)
    theta_t = torch.as_tensor(theta0 * np.asarray(target_scale), dtype=torch.float64)
    with torch.no_grad():
        _, pos = stream_loss(theta_t, prob, return_positions=True)
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    sel = rng.choice(pos.shape[0], size=n_target, replace=False)
    prob.target = pos[torch.as_tensor(sel, device=pos.device)].cpu().numpy()
    return prob
