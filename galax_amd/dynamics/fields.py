"""Orbit fields (galax.dynamics.fields): the Hamiltonian field of a
potential and the N-body field."""

from ._nbody import NBodyField
from ._orbit import HamiltonianField

__all__ = ["HamiltonianField", "NBodyField"]
