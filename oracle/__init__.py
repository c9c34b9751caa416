"""TEST INFRASTRUCTURE ONLY -- the CPU oracle.

ctypes wrapper around oracle/_build/liboracle.so (built from
oracle/gx_oracle.cpp by oracle/Makefile), a CPU FP64 restatement of galax's
orbit hot path: potentials differentiated by forward-mode AD (as galax's
jax.grad / jax.hessian), diffrax-style Dopri8/Dopri5 steps, PID controller,
initial step and clip-to-end rule, independent or shared batch semantics.
See the header of gx_oracle.cpp for the reference file:line each part
follows.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg use it, always as the checker.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"
_lib = None

_vp = ctypes.c_void_p


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        L.orc_eval.restype = ctypes.c_int
        L.orc_eval.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int64, _vp]
        L.orc_gauss_legendre.restype = ctypes.c_int
        L.orc_gauss_legendre.argtypes = [ctypes.c_int, _vp, _vp]
        L.orc_tableau.restype = ctypes.c_int
        L.orc_tableau.argtypes = [ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_int)]
        L.orc_dense_weights.restype = ctypes.c_int
        L.orc_dense_weights.argtypes = [ctypes.c_int, ctypes.c_double, _vp]
        L.orc_rk_step.restype = ctypes.c_int
        L.orc_rk_step.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _vp, ctypes.c_double, ctypes.c_int64, _vp, _vp]
        L.orc_integrate.restype = ctypes.c_int
        L.orc_integrate.argtypes = [
            _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_int, _vp,
            ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp,
        ]
        L.orc_nbody_integrate.restype = ctypes.c_int
        L.orc_nbody_integrate.argtypes = [
            _vp, ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_double, _vp, _vp, ctypes.c_double,
            ctypes.c_double, _vp, ctypes.c_int, ctypes.c_int64, _vp, _vp, _vp, _vp,
        ]
        _lib = L
    return _lib


class _Table:
    """Component and track arrays of a galax_amd potential, kept alive."""

    def __init__(self, pot):
        from galax_amd import _abi

        comps, tracks = pot._table()
        self.comps = _abi.component_array(comps)
        self.n = len(comps)
        self.keep = []
        self.tracks = (_abi.GxTrack * max(len(tracks), 1))()
        from galax_amd.potential._base import fill_track

        for i, tr in enumerate(tracks):
            self.keep += fill_track(self.tracks[i], tr)
        self.ntracks = len(tracks)

    def args(self):
        return (ctypes.addressof(self.comps), self.n, ctypes.addressof(self.tracks), self.ntracks)


def _c(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def evaluate(pot, what: str, xyz, t=0.0) -> np.ndarray:
    """what in {'potential', 'gradient', 'hessian', 'acceleration'}."""
    code = {"potential": 0, "gradient": 1, "hessian": 2, "acceleration": 3}[what]
    width = {0: 1, 1: 3, 2: 9, 3: 3}[code]
    x = _c(xyz)
    batch = x.shape[:-1]
    x = x.reshape(-1, 3)
    n = x.shape[0]
    tt = _c(t).reshape(-1)
    stride = 0 if tt.size == 1 else 1
    if stride:
        tt = _c(np.broadcast_to(tt, (n,)))
    out = np.empty((n, width))
    tab = _Table(pot)
    rc = lib().orc_eval(*tab.args(), code, x.ctypes.data, tt.ctypes.data, stride, n, out.ctypes.data)
    assert rc == 0, rc
    if code == 0:
        return out.reshape(batch)
    if code == 2:
        return out.reshape(*batch, 3, 3)
    return out.reshape(*batch, 3)


def gauss_legendre(n: int):
    """The oracle's Gauss-Legendre rule on [0, 1]: nodes and weights."""
    s = np.zeros(n)
    w = np.zeros(n)
    if lib().orc_gauss_legendre(int(n), s.ctypes.data, w.ctypes.data) != 0:
        raise ValueError(f"bad order {n}")
    return s, w


def tableau(method: int):
    c = np.zeros(16)
    a = np.zeros(256)
    b = np.zeros(16)
    be = np.zeros(16)
    order = ctypes.c_int()
    s = lib().orc_tableau(method, c.ctypes.data, a.ctypes.data, b.ctypes.data, be.ctypes.data, ctypes.byref(order))
    return c[:s], a[: s * s].reshape(s, s), b[:s], be[:s], order.value


def dense_weights(method: int, th: float) -> np.ndarray:
    """Weights w_i(th) of the dense output, y(t0 + th h) = y0 + sum_i w_i k_i,
    over the step's increments (Dopri8: 13 stages and the first-same-as-last
    increment h f(t1, y1); Dopri5: its 7 stages)."""
    w = np.zeros(16)
    ns = lib().orc_dense_weights(int(method), float(th), w.ctypes.data)
    return w[:ns]


def rk_step(pot, method: int, t: float, y, h: float):
    y = _c(y).reshape(-1, 6)
    B = y.shape[0]
    y1 = np.empty_like(y)
    err = np.empty_like(y)
    tab = _Table(pot)
    lib().orc_rk_step(*tab.args(), method, float(t), y.ctypes.data, float(h), B, y1.ctypes.data, err.ctypes.data)
    return y1, err


def integrate(pot, solver, w0, t0, t1, *, ts=None, T=None, dt0=None, max_steps="default",
              shared=False, nthreads=0):
    """Returns (out (n,T,6), status (n,), nsteps (n,), naccept (n,)).
    `solver` is a galax_amd.dynamics.OrbitSolver (configuration only)."""
    from galax_amd import _abi

    cfg = solver.cfg(dt0=dt0, max_steps=max_steps)
    w = _c(w0).reshape(-1, 6)
    n = w.shape[0]
    t0a, t1a = _c(t0).reshape(-1), _c(t1).reshape(-1)
    stride = 0
    if t0a.size > 1 or t1a.size > 1:
        t0a = _c(np.broadcast_to(t0a, (n,)))
        t1a = _c(np.broadcast_to(t1a, (n,)))
        stride = 1
    tsa = None
    if ts is not None:
        tsa = _c(ts)
        if tsa.ndim == 1:
            mode, T = _abi.GX_TS_SHARED, tsa.shape[0]
        else:
            mode, T = _abi.GX_TS_PER_ORBIT, tsa.shape[1]
    elif T is not None:
        mode = _abi.GX_TS_LINSPACE
    else:
        mode, T = _abi.GX_TS_FINAL, 1
    out = np.empty((n, T, 6))
    status = np.empty(n, dtype=np.int32)
    nsteps = np.empty(n, dtype=np.int64)
    nacc = np.empty(n, dtype=np.int64)
    tab = _Table(pot)
    rc = lib().orc_integrate(
        *tab.args(), ctypes.addressof(cfg), w.ctypes.data, t0a.ctypes.data, t1a.ctypes.data, stride,
        None if tsa is None else tsa.ctypes.data, T, mode, n, int(bool(shared)), int(nthreads),
        out.ctypes.data, status.ctypes.data, nsteps.ctypes.data, nacc.ctypes.data,
    )
    assert rc == 0, rc
    return out, status, nsteps, nacc


def nbody_integrate(ext_pot, gm, eps, solver, w0, t0, t1, *, ts=None, dt0=None, max_steps="default"):
    """The N-body system (NBodyField) as one ODE system.  Returns
    (out (n, T, 6), status, nsteps, naccept)."""
    cfg = solver.cfg(dt0=dt0, max_steps=max_steps)
    w = _c(w0).reshape(-1, 6)
    n = w.shape[0]
    tsa = _c(np.array([t1], dtype=np.float64) if ts is None else ts)
    T = tsa.shape[0]
    g = _c(gm).reshape(n)
    out = np.empty((n, T, 6))
    status = ctypes.c_int32()
    nsteps = ctypes.c_int64()
    nacc = ctypes.c_int64()
    tab = _Table(ext_pot)
    rc = lib().orc_nbody_integrate(
        *tab.args(), g.ctypes.data, float(eps), ctypes.addressof(cfg), w.ctypes.data, float(t0), float(t1),
        tsa.ctypes.data, T, n, out.ctypes.data, ctypes.byref(status), ctypes.byref(nsteps), ctypes.byref(nacc),
    )
    assert rc == 0, rc
    return out, status.value, nsteps.value, nacc.value
