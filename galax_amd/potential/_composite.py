"""Composite, preset and translated potentials.

  CompositePotential      potential/_src/composite.py:33-111 (sum of
                          components, base_multi.py:39-82)
  MilkyWayPotential       builtin/milkyway.py:173-236 (disk, halo, bulge,
                          nucleus, summed in that order)
  TranslatedPotential     xfm/translate.py:29-110: Phi(x - Delta(t))
  TimeDependentTranslationParameter  xfm/translate.py:117-199
"""

from __future__ import annotations

from collections.abc import Mapping

import numpy as np

from .. import _abi
from ..units import Q, ustrip
from ._base import AbstractPotential, DenseTrack, Track
from ._builtin import (
    HernquistPotential,
    LMJ09LogarithmicPotential,
    MiyamotoNagaiPotential,
    MN3Sech2Potential,
    NFWPotential,
    PowerLawCutoffPotential,
)

__all__ = [
    "CompositePotential",
    "MilkyWayPotential",
    "MilkyWayPotential2022",
    "LM10Potential",
    "BovyMWPotential2014",
    "TranslatedPotential",
    "TimeDependentTranslationParameter",
    "RotatingTranslation",
]


class CompositePotential(AbstractPotential, Mapping):
    """Ordered sum of named potentials."""

    _preset = False

    def __init__(self, _data: Mapping | None = None, /, **kwargs):
        data = dict(_data or {})
        data.update(kwargs)
        for k, v in data.items():
            if not isinstance(v, AbstractPotential):
                raise TypeError(f"component {k!r} is not a potential")
        self._data = data

    def __getitem__(self, k):
        return self._data[k]

    def __iter__(self):
        return iter(self._data)

    def __len__(self):
        return len(self._data)

    def _rows(self):
        rows = []
        for v in self._data.values():
            rows.extend(v._rows())
        return rows

    def __repr__(self):
        inner = ", ".join(f"{k}={v!r}" for k, v in self._data.items())
        return f"{type(self).__name__}({inner})"


def _as(cls, v, defaults: dict):
    if isinstance(v, cls):
        return v
    if v is None:
        return cls(**defaults)
    kw = dict(defaults)
    kw.update(v)
    return cls(**kw)


class MilkyWayPotential(CompositePotential):
    """Bovy (2015) disk + NFW halo + Hernquist bulge + Hernquist nucleus with
    galax's default parameters."""

    _preset = True
    DISK = {"m_tot": 6.8e10, "a": 3.0, "b": 0.28}
    HALO = {"m": 5.4e11, "r_s": 15.62}
    BULGE = {"m_tot": 5e9, "r_s": 1.0}
    NUCLEUS = {"m_tot": 1.71e9, "r_s": 0.07}

    def __init__(self, *, disk=None, halo=None, bulge=None, nucleus=None, units="galactic"):
        super().__init__(
            disk=_as(MiyamotoNagaiPotential, disk, self.DISK),
            halo=_as(NFWPotential, halo, self.HALO),
            bulge=_as(HernquistPotential, bulge, self.BULGE),
            nucleus=_as(HernquistPotential, nucleus, self.NUCLEUS),
        )


class MilkyWayPotential2022(CompositePotential):
    """galax's 2022 Milky Way model (milkyway.py:240-313): a three-disk
    Miyamoto-Nagai sum for an exponential sech^2 disk, NFW halo, Hernquist
    bulge and nucleus."""

    _preset = True
    DISK = {"m_tot": 4.7717e10, "h_R": 2.6, "h_z": 0.3, "positive_density": True}
    HALO = {"m": 5.5427e11, "r_s": 15.626}
    BULGE = {"m_tot": 5e9, "r_s": 1.0}
    NUCLEUS = {"m_tot": 1.8142e9, "r_s": 0.0688867}

    def __init__(self, *, disk=None, halo=None, bulge=None, nucleus=None, units="galactic"):
        super().__init__(
            disk=_as(MN3Sech2Potential, disk, self.DISK),
            halo=_as(NFWPotential, halo, self.HALO),
            bulge=_as(HernquistPotential, bulge, self.BULGE),
            nucleus=_as(HernquistPotential, nucleus, self.NUCLEUS),
        )


class LM10Potential(CompositePotential):
    """Law & Majewski (2010) Milky Way (milkyway.py:101-170): Miyamoto-Nagai
    disk, Hernquist bulge, triaxial LMJ09 logarithmic halo."""

    _preset = True
    DISK = {"m_tot": 1e11, "a": 6.5, "b": 0.26}
    BULGE = {"m_tot": 3.4e10, "r_s": 0.7}
    HALO = {"v_c": Q(2.0**0.5 * 121.858, "km/s"), "r_s": 12.0, "q1": 1.38, "q2": 1.0, "q3": 1.36,
            "phi": Q(97.0, "deg")}

    def __init__(self, *, disk=None, bulge=None, halo=None, units="galactic"):
        super().__init__(
            disk=_as(MiyamotoNagaiPotential, disk, self.DISK),
            bulge=_as(HernquistPotential, bulge, self.BULGE),
            halo=_as(LMJ09LogarithmicPotential, halo, self.HALO),
        )


class BovyMWPotential2014(CompositePotential):
    """Bovy (2015) MWPotential2014 (milkyway.py:34-99): Miyamoto-Nagai disk,
    power-law bulge with an exponential cut-off, NFW halo."""

    _preset = True
    DISK = {"m_tot": 68_193_902_782.346756, "a": 3.0, "b": 0.28}
    BULGE = {"m_tot": 4501365375.06545, "alpha": 1.8, "r_c": 1.9}
    HALO = {"m": 4.3683325e11, "r_s": 16.0}

    def __init__(self, *, disk=None, bulge=None, halo=None, units="galactic"):
        super().__init__(
            disk=_as(MiyamotoNagaiPotential, disk, self.DISK),
            bulge=_as(PowerLawCutoffPotential, bulge, self.BULGE),
            halo=_as(NFWPotential, halo, self.HALO),
        )


class RotatingTranslation:
    """Delta(t) = Rz(Omega t) . xyz0: a centre fixed in a frame that rotates
    rigidly about z (a circular orbit of radius |xyz0| when z0 = 0)."""

    def __init__(self, xyz0, Omega):
        self.xyz0 = np.asarray(ustrip(xyz0, "kpc") if isinstance(xyz0, Q) else xyz0, dtype=np.float64)
        self.Omega = float(ustrip(Omega, "1/Myr") if isinstance(Omega, Q) else Omega)

    def __call__(self, t):
        t = np.asarray(t, dtype=np.float64)
        c, s = np.cos(self.Omega * t), np.sin(self.Omega * t)
        x0, y0, z0 = self.xyz0
        return np.stack([x0 * c - y0 * s, x0 * s + y0 * c, np.broadcast_to(z0, np.shape(t))], axis=-1)


def not_a_knot_slopes(t, y):
    """First derivatives at the knots of the not-a-knot cubic spline through
    (t, y) (y: (n, ...) values), the spline interpax's and SciPy's
    CubicSpline build by default.  On each interval that spline is the cubic
    Hermite interpolant of the end values and these slopes.

    Slopes s_i satisfy, for the interior knots,
      h_i s_(i-1) + 2 (h_(i-1) + h_i) s_i + h_(i-1) s_(i+1)
          = 3 (h_i d_(i-1) + h_(i-1) d_i),
    with h_i = t_(i+1) - t_i and d_i = (y_(i+1) - y_i) / h_i, and at each end
    the not-a-knot condition (one cubic over the first two and over the last
    two intervals).  Each end row is combined with its neighbour so that the
    remaining system is strictly diagonally dominant and solved without
    pivoting (Thomas algorithm); n = 2 gives the line and n = 3 the parabola
    through the three points."""
    t = np.asarray(t, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = t.size
    if n < 2:
        return np.zeros_like(y)
    h = np.diff(t)
    d = np.diff(y, axis=0) / h.reshape((-1,) + (1,) * (y.ndim - 1))
    if n == 2:
        return np.stack([d[0], d[0]])
    if n == 3:
        # the parabola through the three points
        c2 = (d[1] - d[0]) / (t[2] - t[0])
        s0 = d[0] - c2 * h[0]
        return np.stack([s0, s0 + 2 * c2 * h[0], s0 + 2 * c2 * (t[2] - t[0])])
    bshape = (-1,) + (1,) * (y.ndim - 1)
    hb = lambda a: a.reshape(bshape)
    # rows 1 .. n-2 of the tridiagonal system: lower a, diagonal b, upper c, rhs r
    a = hb(h[1:].copy())                  # coefficient of s_(i-1), i = 1 .. n-2
    b = hb(2.0 * (h[:-1] + h[1:]))
    c = hb(h[:-1].copy())                 # coefficient of s_(i+1)
    r = 3.0 * (hb(h[1:]) * d[:-1] + hb(h[:-1]) * d[1:])
    a, b, c = (np.broadcast_to(x, r.shape).copy() for x in (a, b, c))
    # left end: h1 s0 + (h0 + h1) s1 = r0; row 1 minus it eliminates s0
    h0, h1 = h[0], h[1]
    r0 = ((h0 + 2.0 * (h0 + h1)) * h1 * d[0] + h0 * h0 * d[1]) / (h0 + h1)
    b[0] -= (h0 + h1)
    r[0] -= r0
    a[0] = 0.0
    # right end: (hm + hn) s_(n-2) + hn s_(n-1) = rn with hn = h_(n-3),
    # hm = h_(n-2); row n-2 minus (h_(n-3)/hn) times it eliminates s_(n-1)
    hn, hm = h[-2], h[-1]
    rn = ((hm + 2.0 * (hm + hn)) * hn * d[-1] + hm * hm * d[-2]) / (hm + hn)
    b[-1] -= (hm + hn)
    r[-1] -= rn
    c[-1] = 0.0
    # Thomas algorithm over rows 1 .. n-2
    m = r.shape[0]
    for i in range(1, m):
        w = a[i] / b[i - 1]
        b[i] = b[i] - w * c[i - 1]
        r[i] = r[i] - w * r[i - 1]
    s_in = np.empty_like(r)
    s_in[-1] = r[-1] / b[-1]
    for i in range(m - 2, -1, -1):
        s_in[i] = (r[i] - c[i] * s_in[i + 1]) / b[i]
    s0 = (r0 - (h0 + h1) * s_in[0]) / h1
    sn = (rn - (hm + hn) * s_in[-1]) / hn
    return np.concatenate([s0[None], s_in, sn[None]])


class TimeDependentTranslationParameter:
    """Tabulated centre track.  Given positions only, the track is the
    not-a-knot cubic spline through them, as the reference's interpax
    CubicSpline (xfm/translate.py:185-199); given velocities as well, it is
    the cubic Hermite interpolant of positions and velocities at the knots.
    `from_dense` takes an integrated orbit's own dense solution instead
    (what the mock-stream generator passes for the progenitor).  Outside the
    covered range the track holds its end positions, where the reference's
    spline (extrapolate=False) gives NaN."""

    def __init__(self, t, q, v=None):
        t = np.asarray(ustrip(t, "Myr") if isinstance(t, Q) else t, dtype=np.float64)
        q = np.asarray(ustrip(q, "kpc") if isinstance(q, Q) else q, dtype=np.float64)
        order = np.argsort(t)
        t, q = t[order], q[order]
        if v is None:
            v = not_a_knot_slopes(t, q)
        else:
            v = np.asarray(ustrip(v, "kpc/Myr") if isinstance(v, Q) else v, dtype=np.float64)[order]
        self.track = Track(t=t, q=q, v=v)

    @classmethod
    def from_(cls, t, q, v=None):
        return cls(t, q, v)

    @classmethod
    def from_dense(cls, solution, i: int = 0):
        """The dense solution of orbit i of a solve (a DenseSolution, or an
        Orbit computed with dense=True) as the track: between the solve's
        accepted steps the position is the method's dense output, the same
        polynomial the solve's saves and Orbit.interpolant come from (no
        knots, no interpolation error beyond the solve's own).  Outside the
        integrated interval the end positions are held."""
        dense = getattr(getattr(solution, "interpolant", None), "dense", solution)
        if not hasattr(dense, "rec") or not hasattr(dense, "nrec"):
            raise TypeError("from_dense needs a DenseSolution or an Orbit computed with dense=True")
        nr = int(dense.nrec[i])
        if nr < 1:
            raise ValueError(f"orbit {i} of the dense solution has no accepted steps")
        rec = dense.rec[i, :nr]
        rec = rec.cpu().numpy() if hasattr(rec, "cpu") else np.asarray(rec)
        obj = cls.__new__(cls)
        obj.track = DenseTrack(rec=np.ascontiguousarray(rec, dtype=np.float64), method=int(dense.method))
        return obj

    @classmethod
    def from_callable(cls, fn, t):
        """A callable translation t -> (x, y, z) (the reference accepts any
        callable, translate.py:165-176) tabulated at the times t and
        interpolated by the not-a-knot cubic spline through those values:
        the device evaluates tracks, not Python callables, so the caller
        chooses the knots that resolve the motion."""
        t = np.asarray(ustrip(t, "Myr") if isinstance(t, Q) else t, dtype=np.float64).reshape(-1)
        q = np.asarray(fn(t), dtype=np.float64)
        if q.shape != (t.size, 3):
            raise ValueError(f"the translation callable must map times (n,) to positions (n, 3); got {q.shape}")
        return cls(t, q)


class TranslatedPotential(AbstractPotential):
    """Phi_base(x - Delta(t)).  `translation` is a constant offset (3,), a
    RotatingTranslation, or a TimeDependentTranslationParameter."""

    def __init__(self, original_potential: AbstractPotential, translation, *, units="galactic"):
        self.original_potential = original_potential
        self.translation = translation

    def _rows(self):
        rows = []
        for row, track in self.original_potential._rows():
            if row.frame != _abi.GX_FRAME_NONE or track is not None:
                raise ValueError("nested translations are not supported")
            tr = self.translation
            if isinstance(tr, RotatingTranslation):
                row.frame = _abi.GX_FRAME_ROTATING
                for i, v in enumerate((*tr.xyz0, tr.Omega)):
                    row.f[i] = float(v)
                rows.append((row, None))
            elif isinstance(tr, TimeDependentTranslationParameter):
                rows.append((row, tr.track))
            elif callable(tr):
                raise TypeError(
                    "a callable translation runs in Python; tabulate it on the times that resolve "
                    "it with TimeDependentTranslationParameter.from_callable(fn, t)")
            else:
                off = np.asarray(ustrip(tr, "kpc") if isinstance(tr, Q) else tr, dtype=np.float64)
                if off.shape != (3,):
                    raise ValueError("a constant translation has shape (3,)")
                row.frame = _abi.GX_FRAME_OFFSET
                for i in range(3):
                    row.f[i] = float(off[i])
                rows.append((row, None))
        return rows

    def __repr__(self):
        return f"TranslatedPotential({self.original_potential!r}, {type(self.translation).__name__})"
