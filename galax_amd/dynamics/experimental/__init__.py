"""galax's experimental dynamics API (dynamics/_src/experimental/), backed by
the HIP integrator.

  integrate_orbit, parse_t0_t1_saveat   experimental/integrate.py:38-575
  StreamSimulator, StreamICs             experimental/stream.py:54-446
  Fardal2015DF                           experimental/df.py:54-166

Batches are integrated as independent orbits, as the reference's vmapped
loop strategies do; all particles of a stream go to the GPU in one launch.
"""

from . import stream
from .df import AbstractKinematicDF, Fardal2015DF
from .integrate import Solution, default_solver, integrate_orbit, parse_t0_t1_saveat
from .stream import StreamICs, StreamSimulator

__all__ = [
    "stream",
    "AbstractKinematicDF",
    "Fardal2015DF",
    "Solution",
    "default_solver",
    "integrate_orbit",
    "parse_t0_t1_saveat",
    "StreamICs",
    "StreamSimulator",
]
