"""Units: galax's "galactic" system (kpc, Myr, Msun, rad) and a minimal
quantity type.

galax takes values either as plain arrays in the potential's unit system or
as `unxt.Quantity` objects (potential/_src/utils.py:171-626,
dynamics/_src/utils.py:124-536).  `Q` here carries a value and a unit string
and converts to the galactic system; plain arrays are taken to be in it
already.  The constants follow astropy 7.1 (CODATA 2018 G, IAU 2015 solar
mass, the Julian year) as galax does (potential/_src/base.py:14,35).
"""

from __future__ import annotations

import math

import numpy as np

# G in kpc^3 Msun^-1 Myr^-2 (astropy 7.1 G = 6.6743e-11 m^3 kg^-1 s^-2,
# Msun = 1.988409870698051e30 kg, kpc = 3.0856775814913673e19 m,
# Myr = 3.15576e13 s).
G = 4.498502151469553e-12
# 1 km/s in kpc/Myr
KMS = 1.022712165045695e-3
# the "solarsystem" unit system (AU, yr, Msun): 1 AU in kpc (IAU 2012 AU),
# G in AU^3 Msun^-1 yr^-2, and 1 km/s in AU/yr
AU = 4.848136811095360e-9
G_SOLARSYSTEM = 39.47692640889764
KMS_SOLARSYSTEM = 0.2109495265696987

_FACTORS = {
    # length -> kpc
    "kpc": 1.0,
    "pc": 1e-3,
    "Mpc": 1e3,
    "AU": AU,
    # time -> Myr
    "Myr": 1.0,
    "Gyr": 1e3,
    "kyr": 1e-3,
    "yr": 1e-6,
    # speed -> kpc/Myr
    "kpc/Myr": 1.0,
    "kpc / Myr": 1.0,
    "km/s": KMS,
    "km / s": KMS,
    "AU/yr": AU * 1e6,
    # mass -> Msun
    "Msun": 1.0,
    "solMass": 1.0,
    # angle -> rad
    "rad": 1.0,
    "deg": math.pi / 180.0,
    # frequency -> 1/Myr
    "1/Myr": 1.0,
    "1 / Myr": 1.0,
    "km/(s kpc)": KMS,
    "km / (kpc s)": KMS,
    "km/s/kpc": KMS,
    "1/Gyr": 1e-3,
    # wavenumber -> 1/kpc
    "1/kpc": 1.0,
    "1 / kpc": 1.0,
    "1/pc": 1e3,
    # acceleration -> kpc/Myr^2
    "kpc/Myr2": 1.0,
    "kpc / Myr2": 1.0,
    # specific energy -> kpc^2/Myr^2
    "kpc2/Myr2": 1.0,
    "kpc2 / Myr2": 1.0,
    "km2/s2": KMS * KMS,
    # dimensionless
    "": 1.0,
    "one": 1.0,
}


class Q:
    """A value with a unit, convertible to galax's galactic system."""

    __slots__ = ("value", "unit")

    def __init__(self, value, unit: str = ""):
        if unit not in _FACTORS:
            raise ValueError(f"unknown unit {unit!r}; known: {sorted(_FACTORS)}")
        self.value = np.asarray(value, dtype=np.float64)
        self.unit = unit

    def to_value(self, unit: str | None = None) -> np.ndarray:
        """Value in `unit` (default: the galactic unit of its dimension)."""
        v = self.value * _FACTORS[self.unit]
        if unit is not None:
            v = v / _FACTORS[unit]
        return v

    @property
    def shape(self):
        return self.value.shape

    def __repr__(self) -> str:
        return f"Q({np.array2string(self.value, precision=8)}, {self.unit!r})"


def ustrip(x, unit: str | None = None):
    """Plain float64 array in galactic units (Q converted; arrays passed)."""
    if isinstance(x, Q):
        return x.to_value(unit)
    return x


def to_float(x) -> float:
    return float(np.asarray(ustrip(x), dtype=np.float64))
