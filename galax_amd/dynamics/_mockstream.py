"""Mock stellar streams.

  MockStreamGenerator.run      dynamics/_src/legacy/mockstream/mockstream_generator.py:161-275
  FardalStreamDF._sample       legacy/mockstream/df/fardal15.py:50-94
  AbstractStreamDF.sample      legacy/mockstream/df/base.py:32-131
  tidal_radius (King 1962)     dynamics/_src/cluster/radius.py:200-215
  omega                        dynamics/_src/register_api.py:82-88
  MockStream / MockStreamArm   dynamics/_src/mockstream/arm.py:20-98, core.py:20

The progenitor orbit is integrated at the stripping times; the DF places a
leading and a trailing particle at each stripping time; every particle is
then integrated from its release time to ts[-1] + 1e-3 (the reference's
bump, mockstream_generator.py:88,139).  All particles go in ONE launch of the
persistent integrator, whose work queue absorbs their uneven lengths.

Random numbers: the reference draws the four k-parameters with JAX's
threefry `jr.normal`; here they come from NumPy's PCG64 seeded by `rng`
(an int seed or a numpy Generator), so the samples differ from JAX's while
following the same distributions.

With `progenitor_potential`, the particles also feel the progenitor: that
potential is translated along the progenitor's own dense solution (the
method's dense output over each of its accepted steps, the polynomial its
saves come from); an orbit without one (computed with dense=False) gives a
cubic Hermite track through its saved positions and velocities instead.

`MockStreamGenerator.run` samples the release on the device (the King radius
from the device Hessian, Fardal's formula in torch); `FardalStreamDF.sample`
is the same computation in NumPy, for callers that want host arrays.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ..potential import CompositePotential, TimeDependentTranslationParameter, TranslatedPotential
from ..units import G, ustrip
from ._orbit import Integrator, PhaseSpaceCoordinate, evaluate_orbit, parse_w0
from ._solver import integrate_orbits

__all__ = [
    "omega",
    "tidal_radius",
    "FardalStreamDF",
    "MockStreamArm",
    "MockStream",
    "MockStreamGenerator",
]

# Fardal et al. (2015) means; the dispersions are the reference's placeholder
# values (df/fardal15.py:26-35).
KR_BAR, KVPHI_BAR, KZ_BAR, KVZ_BAR = 2.0, 0.3, 0.0, 0.0
SIGMA_KR = SIGMA_KVPHI = SIGMA_KZ = SIGMA_KVZ = 0.5


def omega(x, v):
    """|x cross v| / |x|^2 (register_api.py:82-88)."""
    x = np.asarray(x, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    r = np.linalg.norm(x, axis=-1, keepdims=True)
    return np.linalg.norm(np.cross(x, v) / r**2, axis=-1)


def tidal_radius(pot, x, v, *, mass, t=0.0):
    """King (1962): r_t = cbrt(G M / (Omega^2 - d2Phi/dr2)), with the radial
    second derivative from the potential's Hessian (radius.py:200-215)."""
    x = np.asarray(x, dtype=np.float64)
    d2 = np.asarray(pot.d2potential_dr2(x, t))
    return np.cbrt(G * np.asarray(mass, dtype=np.float64) / (omega(x, v) ** 2 - d2))


def _normalize(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _generator(rng):
    if isinstance(rng, np.random.Generator):
        return rng
    return np.random.Generator(np.random.PCG64(int(rng)))


def release_torch(x, v, r_t, k):
    """Fardal (2015) leading/trailing release states, as
    FardalStreamDF.release, in torch operations (differentiable, and on the
    device when the inputs are): rows [lead; trail], each (x, v)."""
    import torch

    r = x.norm(dim=-1, keepdim=True)
    L = torch.linalg.cross(x, v, dim=-1)
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


def tidal_radius_device(pot, x, v, mass, t):
    """King (1962) tidal radius at device positions/velocities (torch, FP64),
    the radial second derivative from the device Hessian (radius.py:200-215)."""
    import torch

    H = pot.hessian(x, t)
    r_hat = x / x.norm(dim=-1, keepdim=True)
    d2 = (r_hat[:, :, None] * H * r_hat[:, None, :]).sum(dim=(1, 2))
    L = torch.linalg.cross(x, v, dim=-1)
    om = L.norm(dim=-1) / (x * x).sum(-1)
    a = G * mass / (om * om - d2)
    return torch.sign(a) * torch.abs(a) ** (1.0 / 3.0)  # a real cube root, as np.cbrt


class FardalStreamDF:
    """Fardal+2015 stream distribution function."""

    def draw_k(self, rng, n: int) -> np.ndarray:
        """(n, 4) samples of (k_r, k_vphi, k_z, k_vz)."""
        g = _generator(rng)
        kr = KR_BAR + g.normal(size=n) * SIGMA_KR
        kvphi = kr * (KVPHI_BAR + g.normal(size=n) * SIGMA_KVPHI)
        kz = KZ_BAR + g.normal(size=n) * SIGMA_KZ
        kvz = KVZ_BAR + g.normal(size=n) * SIGMA_KVZ
        return np.stack([kr, kvphi, kz, kvz], axis=-1)

    @staticmethod
    def release(x, v, r_tidal, k):
        """Leading and trailing ICs for given k samples (fardal15.py:50-94)."""
        om = omega(x, v)[..., None]
        r_hat = _normalize(x)
        r_t = np.asarray(r_tidal)[..., None]
        v_circ = om * r_t
        z_hat = _normalize(np.cross(x, v))
        phi_vec = v - np.sum(v * r_hat, axis=-1, keepdims=True) * r_hat
        phi_hat = _normalize(phi_vec)
        kr, kvphi, kz, kvz = (k[..., i : i + 1] for i in range(4))
        x_trail = x + r_t * (kr * r_hat + kz * z_hat)
        v_trail = v + v_circ * (kvphi * phi_hat + kvz * z_hat)
        x_lead = x - r_t * (kr * r_hat - kz * z_hat)
        v_lead = v - v_circ * (kvphi * phi_hat - kvz * z_hat)
        return x_lead, v_lead, x_trail, v_trail

    def sample(self, rng, pot, prog_orbit, prog_mass):
        """Stream ICs along `prog_orbit` (an Orbit of one progenitor saved at
        the release times).  Returns {'lead': MockStreamArm, 'trail': ...}."""
        x = np.asarray(prog_orbit.q).reshape(-1, 3)
        v = np.asarray(prog_orbit.p).reshape(-1, 3)
        ts = np.asarray(prog_orbit.t, dtype=np.float64)
        mass = prog_mass(ts) if callable(prog_mass) else np.full(ts.shape, float(ustrip(prog_mass, "Msun")))
        r_t = tidal_radius(pot, x, v, mass=mass, t=ts)
        k = self.draw_k(rng, ts.size)
        xl, vl, xt, vt = self.release(x, v, r_t, k)
        return {
            "lead": MockStreamArm(q=xl, p=vl, t=ts, release_time=ts),
            "trail": MockStreamArm(q=xt, p=vt, t=ts, release_time=ts),
        }


@dataclass
class MockStreamArm:
    q: np.ndarray
    p: np.ndarray
    t: np.ndarray
    release_time: np.ndarray
    frame: str = "simulation"

    def w(self, units="galactic"):
        return np.concatenate([self.q, self.p], axis=-1)

    @property
    def shape(self):
        return self.q.shape[:-1]


class MockStream(dict):
    """{'lead': MockStreamArm, 'trail': MockStreamArm}."""

    @property
    def q(self):
        return np.concatenate([self["lead"].q, self["trail"].q])

    @property
    def p(self):
        return np.concatenate([self["lead"].p, self["trail"].p])


def release_order(t0, t1):
    """Work-queue order for a stream's particles: the longest integrations
    first, so that both arms' earliest releases start on the queue head
    (ties keep the lead arm first).  Results do not depend on the order."""
    return np.argsort(-(np.asarray(t1) - np.asarray(t0)), kind="stable")


class MockStreamGenerator:
    """Generate a mock stream in `potential` (default integrators: galax's
    legacy Integrator, Dopri8 with PIDController(1e-7, 1e-7))."""

    def __init__(self, df, potential, *, progenitor_integrator=None, stream_integrator=None,
                 progenitor_potential=None):
        self.df = df
        self.potential = potential
        self.progenitor_integrator = progenitor_integrator or Integrator()
        self.stream_integrator = stream_integrator or Integrator()
        self.progenitor_potential = progenitor_potential

    @property
    def units(self):
        return "galactic"

    def stream_potential(self, prog_o):
        """The field the particles move in: the host, plus the progenitor's
        own potential riding its track when one is given (the progenitor's
        dense solution when `prog_o` carries one, else its saved states)."""
        if self.progenitor_potential is None:
            return self.potential
        if getattr(prog_o, "interpolant", None) is not None:
            track = TimeDependentTranslationParameter.from_dense(prog_o)
        else:
            track = TimeDependentTranslationParameter(
                np.asarray(prog_o.t), np.asarray(prog_o.q).reshape(-1, 3), np.asarray(prog_o.p).reshape(-1, 3)
            )
        return CompositePotential(
            host=self.potential, progenitor=TranslatedPotential(self.progenitor_potential, track)
        )

    def run(self, rng, ts, prog_w0, prog_mass):
        """Returns (MockStream, final progenitor PhaseSpaceCoordinate)."""
        ts = np.asarray(ustrip(ts, "Myr"), dtype=np.float64)
        if ts.size > 1 and ts[1] < ts[0]:
            ts = ts[::-1].copy()  # cond_reverse (mockstream_generator.py:236)
        if isinstance(prog_w0, PhaseSpaceCoordinate):
            w0 = prog_w0
        else:
            w, _, t = parse_w0(prog_w0)
            w0 = PhaseSpaceCoordinate(q=w[0, :3], p=w[0, 3:], t=ts[0] if t is None else t)
        prog_o = evaluate_orbit(self.potential, w0, ts, integrator=self.progenitor_integrator,
                                dense=self.progenitor_potential is not None)
        prog_o.q = np.asarray(prog_o.q).reshape(-1, 3)
        prog_o.p = np.asarray(prog_o.p).reshape(-1, 3)
        t_f = ts[-1] + 1e-3
        n = ts.size
        if type(self.df) is FardalStreamDF:
            # the release on the device: King radius from the device Hessian,
            # Fardal's formula in torch; only the k draws come from the host
            # generator (the same draws FardalStreamDF.sample makes)
            import torch

            dev = dict(device="cuda", dtype=torch.float64)
            x = torch.as_tensor(prog_o.q, **dev)
            v = torch.as_tensor(prog_o.p, **dev)
            tt = torch.as_tensor(ts, **dev)
            mass = prog_mass(ts) if callable(prog_mass) else np.full(ts.shape, float(ustrip(prog_mass, "Msun")))
            r_t = tidal_radius_device(self.potential, x, v, torch.as_tensor(np.asarray(mass, np.float64), **dev), tt)
            k = torch.as_tensor(self.df.draw_k(rng, n), **dev)
            w_all = release_torch(x, v, r_t, k)
        else:
            mock0 = self.df.sample(rng, self.potential, prog_o, prog_mass)
            w_all = np.concatenate([mock0["lead"].w(), mock0["trail"].w()])
        t0_all = np.concatenate([ts, ts])
        pot = self.stream_potential(prog_o)
        integ = self.stream_integrator
        res = integrate_orbits(
            pot, w_all, t0_all, np.full(2 * n, t_f), solver=integ.dynamics_solver,
            dt0=integ.diffeq_kw.get("dt0"), max_steps=integ.diffeq_kw.get("max_steps", "default"),
            order=release_order(t0_all, t_f), validate_order=False,  # argsort: a permutation
        )
        final = res.out[:, 0, :].cpu().numpy()
        tt = np.full(n, ts[-1])
        stream = MockStream(
            lead=MockStreamArm(q=final[:n, :3], p=final[:n, 3:], t=tt, release_time=ts),
            trail=MockStreamArm(q=final[n:, :3], p=final[n:, 3:], t=tt, release_time=ts),
        )
        stream.status = res.status.cpu().numpy()
        stream.nsteps = res.nsteps.cpu().numpy()
        prog_final = PhaseSpaceCoordinate(q=prog_o.q[-1], p=prog_o.p[-1], t=ts[-1])
        return stream, prog_final
