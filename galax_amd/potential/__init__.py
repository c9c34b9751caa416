"""galax.potential-shaped API backed by the HIP library."""

from ._base import AbstractPotential, AbstractSinglePotential, Track
from ._builtin import (
    HernquistPotential,
    IsochronePotential,
    JaffePotential,
    LMJ09LogarithmicPotential,
    LogarithmicPotential,
    LongMuraliBarPotential,
    MN3ExponentialPotential,
    MN3Sech2Potential,
    PowerLawCutoffPotential,
    BurkertPotential,
    KuzminPotential,
    SatohPotential,
    StoneOstriker15Potential,
    TriaxialHernquistPotential,
    TriaxialNFWPotential,
    LeeSutoTriaxialNFWPotential,
    Vogelsberger08TriaxialNFWPotential,
    HarmonicOscillatorPotential,
    HenonHeilesPotential,
    HardCutoffNFWPotential,
    gNFWPotential,
    GaussianPotential,
    TriaxialGaussianPotential,
    AxisymmetricGaussianPotential,
    KeplerPotential,
    MiyamotoNagaiPotential,
    MonariEtAl2016BarPotential,
    NFWPotential,
    NullPotential,
    PlummerPotential,
)
from ._composite import (
    CompositePotential,
    BovyMWPotential2014,
    LM10Potential,
    MilkyWayPotential,
    MilkyWayPotential2022,
    RotatingTranslation,
    TimeDependentTranslationParameter,
    TranslatedPotential,
)


def potential(pot: AbstractPotential, q, t=0.0):
    """gp.potential (potential/_src/api.py)."""
    return pot.potential(q, t)


def gradient(pot: AbstractPotential, q, t=0.0):
    """gp.gradient (potential/_src/register_funcs.py:86-160)."""
    return pot.gradient(q, t)


def acceleration(pot: AbstractPotential, q, t=0.0):
    """gp.acceleration (potential/_src/api.py:779-848)."""
    return pot.acceleration(q, t)


def hessian(pot: AbstractPotential, q, t=0.0):
    return pot.hessian(q, t)


def density(pot: AbstractPotential, q, t=0.0):
    return pot.density(q, t)


def laplacian(pot: AbstractPotential, q, t=0.0):
    return pot.laplacian(q, t)


def d2potential_dr2(pot: AbstractPotential, q, t=0.0):
    return pot.d2potential_dr2(q, t)


def local_circular_velocity(pot: AbstractPotential, q, t=0.0):
    return pot.local_circular_velocity(q, t)


__all__ = [
    "BurkertPotential",
    "KuzminPotential",
    "SatohPotential",
    "StoneOstriker15Potential",
    "TriaxialHernquistPotential",
    "TriaxialNFWPotential",
    "LeeSutoTriaxialNFWPotential",
    "Vogelsberger08TriaxialNFWPotential",
    "HarmonicOscillatorPotential",
    "HenonHeilesPotential",
    "HardCutoffNFWPotential",
    "gNFWPotential",
    "GaussianPotential",
    "TriaxialGaussianPotential",
    "AxisymmetricGaussianPotential",
    "PowerLawCutoffPotential",
    "BovyMWPotential2014",
    "MN3ExponentialPotential",
    "MN3Sech2Potential",
    "MilkyWayPotential2022",
    "LM10Potential",
    "IsochronePotential",
    "JaffePotential",
    "LMJ09LogarithmicPotential",
    "LogarithmicPotential",
    "LongMuraliBarPotential",
    "AbstractPotential",
    "AbstractSinglePotential",
    "Track",
    "KeplerPotential",
    "HernquistPotential",
    "MiyamotoNagaiPotential",
    "NFWPotential",
    "PlummerPotential",
    "MonariEtAl2016BarPotential",
    "NullPotential",
    "CompositePotential",
    "MilkyWayPotential",
    "TranslatedPotential",
    "TimeDependentTranslationParameter",
    "RotatingTranslation",
    "potential",
    "gradient",
    "acceleration",
    "hessian",
    "density",
    "laplacian",
    "d2potential_dr2",
    "local_circular_velocity",
]
