"""Kinematic distribution functions of galax's experimental stream API
(experimental/df.py)."""

from __future__ import annotations

import abc
from dataclasses import dataclass

import numpy as np

from .._mockstream import FardalStreamDF, _generator, tidal_radius

__all__ = ["AbstractKinematicDF", "Fardal2015DF"]


class AbstractKinematicDF(metaclass=abc.ABCMeta):
    """Samples leading and trailing release states (experimental/df.py:23-41)."""

    @abc.abstractmethod
    def sample(self, rng, pot, /, t, x, v, Msat):
        """(x_lead, v_lead, x_trail, v_trail) for the progenitor state (t, x,
        v) of mass Msat; x, v of shape (3,) or (N, 3), t and Msat scalars or
        (N,)."""


@dataclass
class Fardal2015DF(AbstractKinematicDF):
    """Fardal et al. (2015) release model with the reference's default means
    and dispersions (experimental/df.py:44-166): k_r ~ N(2, 0.5),
    k_vphi = k_r N(0.3, 0.5), k_z ~ N(0, 0.5), k_vz ~ N(0, 0.5);
    x_trail = x + r_t (k_r r^ + k_z z^), v_trail = v + Omega r_t (k_vphi phi^ +
    k_vz z^), the leading arm with the radial terms flipped; r_t is King's
    tidal radius and z^ the orbit normal.

    `rng` (an int seed or a numpy Generator) replaces JAX's PRNG key: the
    draws follow the same distributions, not JAX's numbers."""

    kr_bar: float = 2.0
    kvphi_bar: float = 0.3
    kz_bar: float = 0.0
    kvz_bar: float = 0.0
    sigma_kr: float = 0.5
    sigma_kvphi: float = 0.5
    sigma_kz: float = 0.5
    sigma_kvz: float = 0.5

    def draw_k(self, rng, n: int) -> np.ndarray:
        """(n, 4) samples of (k_r, k_vphi, k_z, k_vz)."""
        g = _generator(rng)
        kr = self.kr_bar + g.normal(size=n) * self.sigma_kr
        kvphi = kr * (self.kvphi_bar + g.normal(size=n) * self.sigma_kvphi)
        kz = self.kz_bar + g.normal(size=n) * self.sigma_kz
        kvz = self.kvz_bar + g.normal(size=n) * self.sigma_kvz
        return np.stack([kr, kvphi, kz, kvz], axis=-1)

    def sample(self, rng, pot, /, t, x, v, Msat):
        x = np.asarray(x, dtype=np.float64)
        v = np.asarray(v, dtype=np.float64)
        single = x.ndim == 1
        x2, v2 = np.atleast_2d(x), np.atleast_2d(v)
        n = x2.shape[0]
        t = np.broadcast_to(np.asarray(t, dtype=np.float64), (n,))
        m = np.broadcast_to(np.asarray(Msat, dtype=np.float64), (n,))
        r_t = tidal_radius(pot, x2, v2, mass=m, t=t)
        k = self.draw_k(rng, n)
        xl, vl, xt, vt = FardalStreamDF.release(x2, v2, r_t, k)
        if single:
            return xl[0], vl[0], xt[0], vt[0]
        return xl, vl, xt, vt
