"""Gradient of a stream-track loss with respect to Milky Way parameters.

The pipeline follows MockStreamGenerator.run
(legacy/mockstream/mockstream_generator.py:161-275) with fixed-step Dopri8:

1. the progenitor is integrated backwards from its present-day state,
   saving at the release times;
2. leading and trailing particles are released there with the Fardal (2015)
   offsets (df/fardal15.py:50-94) for fixed k-draws, at the King (1962)
   tidal radius r_t = cbrt(G M / (Omega^2 - d2Phi/dr2)) of the progenitor
   (cluster/radius.py:200-215, used at fardal15.py:68), which depends on
   theta directly and through the progenitor's orbit;
3. every particle is integrated forwards from its release time to t = 0;
4. the loss is the mean squared distance of the particles to the nearest
   point of a target track.

Steps 1 and 3 are `OrbitFlow` (forward in gx_integrate, reverse in the
checkpointed discrete adjoint gx_adjoint); step 2 and the loss are torch
operations (the radial second derivative of the Milky Way model is written
out in closed form, so its theta- and position-derivatives come from torch's
autograd), so one backward pass gives dL/dtheta through the progenitor, the
tidal radius, the release conditions and the particles.  galax obtains such
derivatives with jax.jacfwd through the whole generator
(tests/functional/test_mockstreamgenerator.py:24-116).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from .adjoint import FlowSpec, MilkyWayParams, OrbitFlow
from .dynamics._mockstream import release_torch

__all__ = ["StreamFitProblem", "make_problem", "stream_loss", "release_torch", "mw_d2phi_dr2", "tidal_radius_torch"]

# galax's MilkyWayPotential nucleus (Hernquist), not among the fitted
# parameters
NUCLEUS_M, NUCLEUS_C = 1.71e9, 0.07


def mw_d2phi_dr2(x, theta):
    """r_hat^T H r_hat of the Milky Way model at x (..., 3) for theta = (halo
    m, r_s; disk m_tot, a, b; bulge m_tot) (nucleus fixed), in torch.  The
    spherical components contribute their radial second derivative,
      Hernquist  -2 G M / (r + c)^3,
      NFW        G m [1 / (r^2 (r + r_s)) - 2 ln(1 + r/r_s) / r^3
                      + (2 r + r_s) / (r^2 (r + r_s)^2)],
    and the Miyamoto-Nagai disk (zeta = sqrt(z^2 + b^2), A = a + zeta,
    D = R^2 + A^2, w = A z / zeta) its Hessian
      H_ij = G M [delta_ij' D^-3/2 - 3 u_i u_j D^-5/2],  u = (x, y, w),
    with the zz entry's first term (z^2 / zeta^2 + A b^2 / zeta^3) D^-3/2,
    contracted with r_hat."""
    from .units import G

    torch = _lib.torch()
    th = theta.to(x.dtype).to(x.device)
    hm, rs, dm, a, b, bm = (th[i] for i in range(6))
    r = x.norm(dim=-1)
    rh = x / r[..., None]
    # spherical components
    d2 = -2.0 * G * bm / (r + 1.0) ** 3 - 2.0 * G * NUCLEUS_M / (r + NUCLEUS_C) ** 3
    d2 = d2 + G * hm * (1.0 / (r**2 * (r + rs)) - 2.0 * torch.log1p(r / rs) / r**3
                        + (2.0 * r + rs) / (r**2 * (r + rs) ** 2))
    # disk
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    zeta = torch.sqrt(Z * Z + b * b)
    A = a + zeta
    D = X * X + Y * Y + A * A
    D32, D52 = D ** -1.5, D ** -2.5
    w = A * Z / zeta
    u = torch.stack([X, Y, w], dim=-1)
    ur = (u * rh).sum(-1)
    diag = torch.stack([D32, D32, (Z * Z / zeta**2 + A * b * b / zeta**3) * D32], dim=-1)
    d2 = d2 + G * dm * ((diag * rh * rh).sum(-1) - 3.0 * ur * ur * D52)
    return d2


def tidal_radius_torch(x, v, theta, mass):
    """King (1962) tidal radius cbrt(G M / (Omega^2 - d2Phi/dr2)) of a
    cluster at (x, v) in the Milky Way model theta (radius.py:200-215)."""
    from .units import G

    torch = _lib.torch()
    L = torch.linalg.cross(x, v, dim=-1)
    om = L.norm(dim=-1) / (x * x).sum(-1)
    a = G * mass / (om * om - mw_d2phi_dr2(x, theta))
    return torch.sign(a) * torch.abs(a) ** (1.0 / 3.0)  # a real cube root, as np.cbrt


@dataclass
class StreamFitProblem:
    """Inputs of the stream-track loss (galactic units)."""

    w_prog: np.ndarray  # (6,) present-day progenitor state
    ts: np.ndarray  # (N,) ascending release times, all < 0
    k: np.ndarray  # (N, 4) Fardal draws
    mass: float  # progenitor mass (Msun) for the tidal radius
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
    r_t = tidal_radius_torch(prog[:, :3], prog[:, 3:], theta, prob.mass)
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
    from .dynamics import FardalStreamDF

    theta0 = MilkyWayParams.default.copy()
    pot0 = MilkyWayParams.potential(theta0)
    x = np.array([15.0, 0.0, 0.5])
    vc = float(np.asarray(pot0.local_circular_velocity(x, 0.0)).reshape(-1)[0])
    w_prog = np.concatenate([x, [0.0, 0.85 * vc, 0.12 * vc]])
    ts = np.linspace(-span, 0.0, n_release + 1)[:-1]
    k = FardalStreamDF().draw_k(seed, n_release)
    prob = StreamFitProblem(w_prog=w_prog, ts=ts, k=k, mass=float(prog_mass), target=np.zeros((1, 3)), dt=dt, K=K)
    theta_t = torch.as_tensor(theta0 * np.asarray(target_scale), dtype=torch.float64)
    with torch.no_grad():
        _, pos = stream_loss(theta_t, prob, return_positions=True)
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    sel = rng.choice(pos.shape[0], size=n_target, replace=False)
    prob.target = pos[torch.as_tensor(sel, device=pos.device)].cpu().numpy()
    return prob
