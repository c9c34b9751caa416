"""galax.dynamics-shaped API backed by the HIP library."""

from ._orbit import (
    HamiltonianField,
    Integrator,
    Orbit,
    OrbitInterpolant,
    PhaseSpaceCoordinate,
    PhaseSpacePosition,
    evaluate_orbit,
    parse_w0,
)
from ._solver import (
    ConstantStepSize,
    Dopri5,
    Dopri8,
    IntegrationResult,
    OrbitSolver,
    PIDController,
    integrate_orbits,
)

__all__ = [
    "HamiltonianField",
    "Integrator",
    "Orbit",
    "OrbitInterpolant",
    "PhaseSpaceCoordinate",
    "PhaseSpacePosition",
    "evaluate_orbit",
    "parse_w0",
    "ConstantStepSize",
    "Dopri5",
    "Dopri8",
    "IntegrationResult",
    "OrbitSolver",
    "PIDController",
    "integrate_orbits",
]
