"""galax_amd: galax's orbit hot path (composite-potential accelerations inside
Dopri8/Dopri5 orbit integration, mock-stream release integration) on AMD
Instinct MI355X, through a C ABI over hand-written HIP kernels.

    import galax_amd.potential as gp
    import galax_amd.dynamics as gd
    orbit = gd.evaluate_orbit(gp.MilkyWayPotential(), w0, t)
"""

from . import dynamics, potential, units  # noqa: F401
from .units import G, Q  # noqa: F401

__version__ = "0.1.0"
