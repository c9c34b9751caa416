"""Analytic potential models on galax's orbit hot path.

Each class keeps the reference's constructor parameters and units and maps
them to a component-table row (include/galax_amd.h):

  KeplerPotential(m_tot)              potential/_src/builtin/kepler.py:34-94
  HernquistPotential(m_tot, r_s)      builtin/hernquist.py:27-82
  MiyamotoNagaiPotential(m_tot, a, b) builtin/miyamotonagai.py:30-78
  NFWPotential(m, r_s)                builtin/nfw/base.py:36-100,327-338
  PlummerPotential(m_tot, r_s)        builtin/plummer.py:33-132
  MonariEtAl2016BarPotential(...)     builtin/monari2016.py:29-138
  NullPotential()                     builtin/null.py

and the further models below, each citing its reference file.

Parameters are numbers in galactic units (kpc, Myr, Msun, rad) or `Q`
quantities.  Mass parameters enter the table as G*m.
"""

from __future__ import annotations

import math

from .. import _abi
from ..units import G, Q, to_float
from ._base import AbstractSinglePotential

__all__ = [
    "KeplerPotential",
    "HernquistPotential",
    "MiyamotoNagaiPotential",
    "NFWPotential",
    "PlummerPotential",
    "MonariEtAl2016BarPotential",
    "NullPotential",
    "IsochronePotential",
    "JaffePotential",
    "LogarithmicPotential",
    "LMJ09LogarithmicPotential",
    "LongMuraliBarPotential",
    "MN3ExponentialPotential",
    "MN3Sech2Potential",
    "PowerLawCutoffPotential",
    "BurkertPotential",
    "KuzminPotential",
    "SatohPotential",
    "StoneOstriker15Potential",
    "TriaxialHernquistPotential",
    "TriaxialNFWPotential",
    "LeeSutoTriaxialNFWPotential",
    "Vogelsberger08TriaxialNFWPotential",
]


def _p(x, unit: str | None = None) -> float:
    if isinstance(x, Q):
        return float(x.to_value(unit))
    return to_float(x)


class _MassScale(AbstractSinglePotential):
    _mass_name = "m_tot"

    def __init__(self, m_tot, r_s, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m_tot = _p(m_tot, "Msun")
        self.r_s = _p(r_s, "kpc")

    def _params(self):
        return (G * self.m_tot, self.r_s)

    def __repr__(self):
        return f"{type(self).__name__}(m_tot={self.m_tot!r}, r_s={self.r_s!r})"


class KeplerPotential(AbstractSinglePotential):
    """Point mass: Phi = -G m_tot / r."""

    kind = _abi.GX_KEPLER

    def __init__(self, m_tot, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m_tot = _p(m_tot, "Msun")

    def _params(self):
        return (G * self.m_tot,)

    def __repr__(self):
        return f"KeplerPotential(m_tot={self.m_tot!r})"


class HernquistPotential(_MassScale):
    """Phi = -G m_tot / (r + r_s)."""

    kind = _abi.GX_HERNQUIST


class PlummerPotential(_MassScale):
    """Phi = -G m_tot / sqrt(r^2 + r_s^2)."""

    kind = _abi.GX_PLUMMER


class NFWPotential(AbstractSinglePotential):
    """Phi = -(G m / r_s) log(1 + r/r_s) / (r/r_s)."""

    kind = _abi.GX_NFW

    def __init__(self, m, r_s, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m = _p(m, "Msun")
        self.r_s = _p(r_s, "kpc")

    def _params(self):
        return (G * self.m, self.r_s)

    def __repr__(self):
        return f"NFWPotential(m={self.m!r}, r_s={self.r_s!r})"


class MiyamotoNagaiPotential(AbstractSinglePotential):
    """Phi = -G m_tot / sqrt(R^2 + (a + sqrt(z^2 + b^2))^2)."""

    kind = _abi.GX_MIYAMOTO_NAGAI

    def __init__(self, m_tot, a, b, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m_tot = _p(m_tot, "Msun")
        self.a = _p(a, "kpc")
        self.b = _p(b, "kpc")

    def _params(self):
        return (G * self.m_tot, self.a, self.b)

    def __repr__(self):
        return f"MiyamotoNagaiPotential(m_tot={self.m_tot!r}, a={self.a!r}, b={self.b!r})"


class MonariEtAl2016BarPotential(AbstractSinglePotential):
    """Rotating bar of Monari et al. (2016):
    Phi = alpha (v0^2/3) (R0/Rb)^3 U(r/Rb) (R^2/r^2) cos 2(phi - phi_b - Omega t).

    Defaults follow the reference: alpha = 0.01, Rb = 3.5 kpc,
    phi_b = 25 deg, Omega = 52.2 km/s/kpc; R0 and v0 have no default."""

    kind = _abi.GX_MONARI_BAR

    def __init__(
        self,
        *,
        alpha=0.01,
        R0,
        v0,
        Rb=3.5,
        phi_b=Q(25.0, "deg"),
        Omega=Q(52.2, "km/(s kpc)"),
        units="galactic",
    ):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.alpha = _p(alpha)
        self.R0 = _p(R0, "kpc")
        self.v0 = _p(v0, "kpc/Myr")
        self.Rb = _p(Rb, "kpc")
        self.phi_b = _p(phi_b, "rad")
        self.Omega = _p(Omega, "1/Myr")

    def _params(self):
        return (self.alpha, self.R0, self.v0, self.Rb, self.phi_b, self.Omega)

    def __repr__(self):
        return (
            f"MonariEtAl2016BarPotential(alpha={self.alpha!r}, R0={self.R0!r}, v0={self.v0!r}, "
            f"Rb={self.Rb!r}, phi_b={self.phi_b!r}, Omega={self.Omega!r})"
        )


class NullPotential(AbstractSinglePotential):
    """Phi = 0."""

    kind = _abi.GX_NULL

    def __init__(self, *, units="galactic"):
        # Phi = 0 in any unit system; the label tells fields which G to use
        self.units = units

    def _params(self):
        return ()

    def __repr__(self):
        return "NullPotential()"


# angle helper used by callers that build bars from degrees
def deg(x: float) -> float:
    return math.radians(x)


class IsochronePotential(_MassScale):
    """Phi = -G m_tot / (r_s + sqrt(r^2 + r_s^2)) (isochrone.py:61-90)."""

    kind = _abi.GX_ISOCHRONE


class JaffePotential(_MassScale):
    """Phi = -(G m_tot / r_s) log(1 + r_s / r) (jaffe.py:34-60)."""

    kind = _abi.GX_JAFFE


class LogarithmicPotential(AbstractSinglePotential):
    """Phi = v_c^2 / 2 log(r_s^2 + r^2) (logarithmic.py:19-53)."""

    kind = _abi.GX_LOGARITHMIC

    def __init__(self, v_c, r_s, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.v_c = _p(v_c, "kpc/Myr")
        self.r_s = _p(r_s, "kpc")

    def _params(self):
        return (self.v_c, self.r_s)

    def __repr__(self):
        return f"LogarithmicPotential(v_c={self.v_c!r}, r_s={self.r_s!r})"


class LMJ09LogarithmicPotential(AbstractSinglePotential):
    """Triaxial logarithmic halo of Law, Majewski & Johnston (2009):
    Phi = v_c^2 / 2 log(r_s^2 + (x'/q1)^2 + (y'/q2)^2 + (z/q3)^2), with
    (x', y') the position rotated by -phi about z (logarithmic.py:56-108)."""

    kind = _abi.GX_LMJ09_LOG

    def __init__(self, v_c, r_s, q1, q2, q3, phi, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.v_c = _p(v_c, "kpc/Myr")
        self.r_s = _p(r_s, "kpc")
        self.q1, self.q2, self.q3 = _p(q1), _p(q2), _p(q3)
        self.phi = _p(phi, "rad")

    def _params(self):
        return (self.v_c, self.r_s, self.q1, self.q2, self.q3, self.phi)

    def __repr__(self):
        return (f"LMJ09LogarithmicPotential(v_c={self.v_c!r}, r_s={self.r_s!r}, q1={self.q1!r}, "
                f"q2={self.q2!r}, q3={self.q3!r}, phi={self.phi!r})")


class LongMuraliBarPotential(AbstractSinglePotential):
    """Bar of Long & Murali (1992): Phi = G m / (2a) log((x' - a + T-) /
    (x' + a + T+)), T+- = sqrt((a +- x')^2 + y'^2 + (b + sqrt(c^2 + z^2))^2),
    (x', y') rotated by -alpha about z (longmurali.py:19-91)."""

    kind = _abi.GX_LONGMURALI_BAR

    def __init__(self, m_tot, a, b, c, alpha, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m_tot = _p(m_tot, "Msun")
        self.a, self.b, self.c = _p(a, "kpc"), _p(b, "kpc"), _p(c, "kpc")
        self.alpha = _p(alpha, "rad")

    def _params(self):
        return (G * self.m_tot, self.a, self.b, self.c, self.alpha)

    def __repr__(self):
        return (f"LongMuraliBarPotential(m_tot={self.m_tot!r}, a={self.a!r}, b={self.b!r}, c={self.c!r}, "
                f"alpha={self.alpha!r})")


# Smith et al. (2015, MNRAS 448, 2934) fits for three Miyamoto-Nagai disks
# approximating an exponential disk (mn3.py:31-52): rows give the three mass
# fractions and three a / h_R as quartics in b / h_R, for the density kept
# positive everywhere or allowed negative far out (the better fit to 7-8
# scale lengths); b / h_R is a cubic in h_z / h_R.
_MN3_K_POS = (
    (0.0036, -0.0330, 0.1117, -0.1335, 0.1749),
    (-0.0131, 0.1090, -0.3035, 0.2921, -5.7976),
    (-0.0048, 0.0454, -0.1425, 0.1012, 6.7120),
    (-0.0158, 0.0993, -0.2070, -0.7089, 0.6445),
    (-0.0319, 0.1514, -0.1279, -0.9325, 2.6836),
    (-0.0326, 0.1816, -0.2943, -0.6329, 2.3193),
)
_MN3_K_NEG = (
    (-0.0090, 0.0640, -0.1653, 0.1164, 1.9487),
    (0.0173, -0.0903, 0.0877, 0.2029, -1.3077),
    (-0.0051, 0.0287, -0.0361, -0.0544, 0.2242),
    (-0.0358, 0.2610, -0.6987, -0.1193, 2.0074),
    (-0.0830, 0.4992, -0.7967, -1.2966, 4.4441),
    (-0.0247, 0.1718, -0.4124, -0.5944, 0.7333),
)
_MN3_B_EXP = (-0.269, 1.08, 1.092)
_MN3_B_SECH2 = (-0.033, 0.262, 0.659)


class _MN3(AbstractSinglePotential):
    """Sum of three Miyamoto-Nagai disks approximating an exponential disk of
    mass m_tot, radial scale h_R and vertical scale h_z; flattened into three
    Miyamoto-Nagai rows of the component table (mn3.py:90-120)."""

    _b_coeffs = _MN3_B_EXP

    def __init__(self, m_tot, h_R, h_z, positive_density=False, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m_tot = _p(m_tot, "Msun")
        self.h_R = _p(h_R, "kpc")
        self.h_z = _p(h_z, "kpc")
        self.positive_density = bool(positive_density)

    def components(self):
        """(masses, a's, b) of the three Miyamoto-Nagai disks."""
        q = self.h_z / self.h_R
        c3, c2, c1 = self._b_coeffs
        b_hR = c3 * q**3 + c2 * q**2 + c1 * q
        powers = (b_hR**4, b_hR**3, b_hR**2, b_hR, 1.0)
        K = _MN3_K_POS if self.positive_density else _MN3_K_NEG
        v = [sum(k * x for k, x in zip(row, powers)) for row in K]
        return [f * self.m_tot for f in v[:3]], [f * self.h_R for f in v[3:]], b_hR * self.h_R

    def _rows(self):
        ms, As, b = self.components()
        return [(_abi.component(_abi.GX_MIYAMOTO_NAGAI, (G * m, a, b)), None) for m, a in zip(ms, As)]

    def __repr__(self):
        return (f"{type(self).__name__}(m_tot={self.m_tot!r}, h_R={self.h_R!r}, h_z={self.h_z!r}, "
                f"positive_density={self.positive_density!r})")


class MN3ExponentialPotential(_MN3):
    """Three Miyamoto-Nagai disks for a double-exponential disk."""

    _b_coeffs = _MN3_B_EXP


class MN3Sech2Potential(_MN3):
    """Three Miyamoto-Nagai disks for an exponential disk with a sech^2
    vertical profile."""

    _b_coeffs = _MN3_B_SECH2


class PowerLawCutoffPotential(AbstractSinglePotential):
    """Spherical power law with an exponential cut-off, rho ~ r^-alpha
    exp(-(r/r_c)^2), normalised to m_tot (powerlawcutoff.py:31-117);
    0 <= alpha < 2."""

    kind = _abi.GX_POWERLAW_CUTOFF

    def __init__(self, m_tot, alpha, r_c, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m_tot = _p(m_tot, "Msun")
        self.alpha = _p(alpha)
        self.r_c = _p(r_c, "kpc")
        if not 0.0 <= self.alpha < 2.0:
            raise ValueError("alpha must lie in [0, 2)")

    def _params(self):
        return (G * self.m_tot, self.alpha, self.r_c)

    def __repr__(self):
        return f"PowerLawCutoffPotential(m_tot={self.m_tot!r}, alpha={self.alpha!r}, r_c={self.r_c!r})"


class BurkertPotential(AbstractSinglePotential):
    """Burkert (1995) cored halo, rho ~ 1 / ((r + r_s)(r^2 + r_s^2)), with
    characteristic mass m (burkert.py:150-230)."""

    kind = _abi.GX_BURKERT

    def __init__(self, m, r_s, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m = _p(m, "Msun")
        self.r_s = _p(r_s, "kpc")

    def _params(self):
        return (G * self.m, self.r_s)

    def __repr__(self):
        return f"BurkertPotential(m={self.m!r}, r_s={self.r_s!r})"


class KuzminPotential(_MassScale):
    """Kuzmin (1956) razor-thin disk: Phi = -G m_tot / sqrt(R^2 + (|z| +
    r_s)^2) (kuzmin.py:60-85)."""

    kind = _abi.GX_KUZMIN


class SatohPotential(AbstractSinglePotential):
    """Satoh (1980) disk: Phi = -G m_tot / sqrt(R^2 + z^2 + a (a + 2
    sqrt(z^2 + b^2))) (satoh.py:50-69)."""

    kind = _abi.GX_SATOH

    def __init__(self, m_tot, a, b, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m_tot = _p(m_tot, "Msun")
        self.a = _p(a, "kpc")
        self.b = _p(b, "kpc")

    def _params(self):
        return (G * self.m_tot, self.a, self.b)

    def __repr__(self):
        return f"SatohPotential(m_tot={self.m_tot!r}, a={self.a!r}, b={self.b!r})"


class StoneOstriker15Potential(AbstractSinglePotential):
    """Stone & Ostriker (2015) cored halo, rho ~ 1 / ((1 + r^2/r_c^2)(1 +
    r^2/r_h^2)), total mass m_tot (stoneostriker15.py:90-147)."""

    kind = _abi.GX_STONE

    def __init__(self, m_tot, r_c, r_h, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m_tot = _p(m_tot, "Msun")
        self.r_c = _p(r_c, "kpc")
        self.r_h = _p(r_h, "kpc")

    def _params(self):
        return (G * self.m_tot, self.r_c, self.r_h)

    def __repr__(self):
        return f"StoneOstriker15Potential(m_tot={self.m_tot!r}, r_c={self.r_c!r}, r_h={self.r_h!r})"


class TriaxialHernquistPotential(AbstractSinglePotential):
    """Hernquist profile of the ellipsoidal radius r' = sqrt(x^2 + (y/q1)^2 +
    (z/q2)^2): Phi = -G m_tot / (r' + r_s) (hernquist.py:89-182)."""

    kind = _abi.GX_TRIAXIAL_HERNQUIST

    def __init__(self, m_tot, r_s, q1=1.0, q2=1.0, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m_tot = _p(m_tot, "Msun")
        self.r_s = _p(r_s, "kpc")
        self.q1, self.q2 = _p(q1), _p(q2)
        if not (self.q1 > 0 and self.q2 > 0):
            raise ValueError("q1 and q2 must be positive")

    def _params(self):
        return (G * self.m_tot, self.r_s, self.q1, self.q2)

    def __repr__(self):
        return (f"TriaxialHernquistPotential(m_tot={self.m_tot!r}, r_s={self.r_s!r}, "
                f"q1={self.q1!r}, q2={self.q2!r})")


class TriaxialNFWPotential(AbstractSinglePotential):
    """NFW density of the ellipsoidal radius (x^2 + y^2/q1^2 + z^2/q2^2)^(1/2):
    its potential by Chandrasekhar's integral over ellipsoidal shells, taken
    with an `integration_order`-point Gauss-Legendre rule on [0, 1]
    (nfw/triaxial.py:29-230; the rule is the reference's
    GaussLegendreIntegrator, potential/_src/utils.py:131-162)."""

    kind = _abi.GX_TRIAXIAL_NFW

    def __init__(self, m, r_s, q1=1.0, q2=1.0, *, integration_order=50, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m = _p(m, "Msun")
        self.r_s = _p(r_s, "kpc")
        self.q1, self.q2 = _p(q1), _p(q2)
        self.integration_order = int(integration_order)
        if not (self.q1 > 0 and self.q2 > 0 and self.r_s > 0):
            raise ValueError("q1, q2 and r_s must be positive")
        if not 1 <= self.integration_order <= _abi.GX_MAX_GL_ORDER:
            raise ValueError(f"integration_order must be in [1, {_abi.GX_MAX_GL_ORDER}]")

    def _params(self):
        return (G * self.m, self.r_s, self.q1, self.q2, float(self.integration_order))

    def __repr__(self):
        return (f"TriaxialNFWPotential(m={self.m!r}, r_s={self.r_s!r}, q1={self.q1!r}, "
                f"q2={self.q2!r}, integration_order={self.integration_order})")


class LeeSutoTriaxialNFWPotential(AbstractSinglePotential):
    """Lee & Suto (2003) approximation to the potential of a triaxial NFW
    density, first order in the eccentricities e_b^2 = 1 - (a2/a1)^2,
    e_c^2 = 1 - (a3/a1)^2; `m` sets v_c^2 = G m / r_s (nfw/leesuto.py:29-164).
    Requires a1 >= a2 >= a3."""

    kind = _abi.GX_LEESUTO_NFW

    def __init__(self, m, r_s, a1=1.0, a2=1.0, a3=1.0, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m = _p(m, "Msun")
        self.r_s = _p(r_s, "kpc")
        self.a1, self.a2, self.a3 = _p(a1), _p(a2), _p(a3)
        if self.a1 < self.a2 or self.a2 < self.a3:
            raise ValueError(f"a1 {self.a1} >= a2 {self.a2} >= a3 {self.a3} is required")
        if not (self.a3 > 0 and self.r_s > 0):
            raise ValueError("axes and r_s must be positive")

    def _params(self):
        return (G * self.m, self.r_s, self.a1, self.a2, self.a3)

    def __repr__(self):
        return (f"LeeSutoTriaxialNFWPotential(m={self.m!r}, r_s={self.r_s!r}, a1={self.a1!r}, "
                f"a2={self.a2!r}, a3={self.a3!r})")


class Vogelsberger08TriaxialNFWPotential(AbstractSinglePotential):
    """NFW potential of the radius r~ = (r_a + r) r_e / (r_a + r_e), which is
    ellipsoidal (r_e^2 = x^2 + y^2/q1^2 + z^2/q2^2, q2^2 = 3 - q1^2) inside
    r_a = a_r r_s and spherical outside: Phi = -G m log(1 + r~/r_s) / r~
    (nfw/vogelsburger08.py:27-110)."""

    kind = _abi.GX_VOGELSBERGER_NFW

    def __init__(self, m, r_s, q1=1.0, a_r=1.0, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m = _p(m, "Msun")
        self.r_s = _p(r_s, "kpc")
        self.q1, self.a_r = _p(q1), _p(a_r)
        if not (0 < self.q1 * self.q1 < 3 and self.r_s > 0):
            raise ValueError("need 0 < q1^2 < 3 and r_s > 0")

    def _params(self):
        return (G * self.m, self.r_s, self.q1, self.a_r)

    def __repr__(self):
        return (f"Vogelsberger08TriaxialNFWPotential(m={self.m!r}, r_s={self.r_s!r}, "
                f"q1={self.q1!r}, a_r={self.a_r!r})")


class HarmonicOscillatorPotential(AbstractSinglePotential):
    """Phi = omega^2 |q|^2 / 2 (example.py:17-100); omega in 1/Myr."""

    kind = _abi.GX_HARMONIC_OSCILLATOR

    def __init__(self, omega, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.omega = _p(omega, "1 / Myr")

    def _params(self):
        return (self.omega,)

    def __repr__(self):
        return f"HarmonicOscillatorPotential(omega={self.omega!r})"


class HenonHeilesPotential(AbstractSinglePotential):
    """Henon & Heiles (1964) in specific-energy units: Phi = (R^2/2 +
    coeff (x^2 y - y^3/3)) / timescale^2 (example.py:103-176); coeff in
    1/kpc, timescale in Myr."""

    kind = _abi.GX_HENON_HEILES

    def __init__(self, coeff, timescale, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.coeff = _p(coeff, "1 / kpc")
        self.timescale = _p(timescale, "Myr")
        if self.timescale == 0:
            raise ValueError("timescale must be nonzero")

    def _params(self):
        return (self.coeff, self.timescale)

    def __repr__(self):
        return f"HenonHeilesPotential(coeff={self.coeff!r}, timescale={self.timescale!r})"


class HardCutoffNFWPotential(AbstractSinglePotential):
    """NFW sharply truncated at r_t: Phi_NFW + G m / (r_s + r_t) inside,
    -G M_NFW(<r_t) / r outside (nfw/truncated.py:37-330)."""

    kind = _abi.GX_HARDCUTOFF_NFW

    def __init__(self, m, r_s, r_t, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m = _p(m, "Msun")
        self.r_s = _p(r_s, "kpc")
        self.r_t = _p(r_t, "kpc")
        if not (self.r_s > 0 and self.r_t > 0):
            raise ValueError("need r_s > 0 and r_t > 0")

    def _params(self):
        return (G * self.m, self.r_s, self.r_t)

    def __repr__(self):
        return f"HardCutoffNFWPotential(m={self.m!r}, r_s={self.r_s!r}, r_t={self.r_t!r})"


class gNFWPotential(AbstractSinglePotential):  # noqa: N801 (the reference's name)
    """Generalized NFW, rho ~ x^-gamma (1 + x)^(gamma - 3), gamma in [0, 2):
    M(<r) = m B_z(3 - gamma, 0), z = x/(1 + x), and
    Phi = -G (M(<r)/r + m/r_s B_{1/(1+x)}(1, 2 - gamma))
    (nfw/generalized.py:30-330)."""

    kind = _abi.GX_GNFW

    def __init__(self, m, r_s, gamma, *, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m = _p(m, "Msun")
        self.r_s = _p(r_s, "kpc")
        self.gamma = _p(gamma)
        if not (0.0 <= self.gamma < 2.0 and self.r_s > 0):
            raise ValueError("need 0 <= gamma < 2 and r_s > 0")

    def _params(self):
        return (G * self.m, self.r_s, self.gamma)

    def __repr__(self):
        return f"gNFWPotential(m={self.m!r}, r_s={self.r_s!r}, gamma={self.gamma!r})"


class GaussianPotential(_MassScale):
    """Spherical Gaussian density: Phi = -G m_tot erf(r / (sqrt2 r_s)) / r
    (gaussian/base.py:40-259)."""

    kind = _abi.GX_GAUSSIAN


class TriaxialGaussianPotential(AbstractSinglePotential):
    """Gaussian density on ellipsoids, rho = m_tot / (q1 q2 (2 pi)^(3/2)
    r_s^3) exp(-xi^2/2), r_s^2 xi^2 = x^2 + y^2/q1^2 + z^2/q2^2; Phi by the
    reference's Gauss-Legendre shell integral (gaussian/triaxial.py:26-240)."""

    kind = _abi.GX_TRIAXIAL_GAUSSIAN

    def __init__(self, m_tot, r_s, q1=1.0, q2=1.0, *, integration_order=50, units="galactic"):
        if units != "galactic":
            raise ValueError("only the galactic unit system is supported")
        self.m_tot = _p(m_tot, "Msun")
        self.r_s = _p(r_s, "kpc")
        self.q1, self.q2 = _p(q1), _p(q2)
        self.integration_order = int(integration_order)
        if not (1 <= self.integration_order <= _abi.GX_MAX_GL_ORDER):
            raise ValueError(f"integration_order must lie in 1..{_abi.GX_MAX_GL_ORDER}")
        if not (self.r_s > 0 and self.q1 > 0 and self.q2 > 0):
            raise ValueError("need r_s, q1, q2 > 0")

    def _params(self):
        return (G * self.m_tot, self.r_s, self.q1, self.q2, float(self.integration_order))

    def __repr__(self):
        return (f"TriaxialGaussianPotential(m_tot={self.m_tot!r}, r_s={self.r_s!r}, q1={self.q1!r}, "
                f"q2={self.q2!r}, integration_order={self.integration_order})")


class AxisymmetricGaussianPotential(TriaxialGaussianPotential):
    """The triaxial Gaussian with q1 = 1: rho ~ exp(-(R^2 + z^2/q2^2) /
    (2 r_s^2)) (gaussian/axisymmetric.py:29-206)."""

    def __init__(self, m_tot, r_s, q2=1.0, *, integration_order=50, units="galactic"):
        super().__init__(m_tot, r_s, 1.0, q2, integration_order=integration_order, units=units)

    def __repr__(self):
        return (f"AxisymmetricGaussianPotential(m_tot={self.m_tot!r}, r_s={self.r_s!r}, q2={self.q2!r}, "
                f"integration_order={self.integration_order})")
