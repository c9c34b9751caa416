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
from ._mockstream import (
    FardalStreamDF,
    MockStream,
    MockStreamArm,
    MockStreamGenerator,
    omega,
    tidal_radius,
)
from . import experimental, fields
from ._nbody import NBodyField, NBodySolution, solve_nbody
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
    "experimental",
    "fields",
    "NBodyField",
    "NBodySolution",
    "solve_nbody",
    "FardalStreamDF",
    "MockStream",
    "MockStreamArm",
    "MockStreamGenerator",
    "omega",
    "tidal_radius",
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
