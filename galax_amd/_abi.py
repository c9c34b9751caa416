"""ctypes mirror of include/galax_amd.h (structures and constants)."""

from __future__ import annotations

import ctypes

c_double_p = ctypes.c_void_p

GX_OK, GX_EINVAL, GX_ELAUNCH, GX_ENOMEM, GX_ETOOBIG = 0, -1, -2, -3, -4

GX_NULL = 0
GX_KEPLER = 1
GX_HERNQUIST = 2
GX_MIYAMOTO_NAGAI = 3
GX_NFW = 4
GX_PLUMMER = 5
GX_MONARI_BAR = 6
GX_ISOCHRONE = 7
GX_JAFFE = 8
GX_LOGARITHMIC = 9
GX_LMJ09_LOG = 10
GX_LONGMURALI_BAR = 11
GX_POWERLAW_CUTOFF = 12
GX_BURKERT = 13
GX_KUZMIN = 14
GX_SATOH = 15
GX_STONE = 16
GX_TRIAXIAL_HERNQUIST = 17
GX_TRIAXIAL_NFW = 18
GX_LEESUTO_NFW = 19
GX_VOGELSBERGER_NFW = 20
GX_HARMONIC_OSCILLATOR = 21
GX_HENON_HEILES = 22
GX_HARDCUTOFF_NFW = 23
GX_GNFW = 24
GX_GAUSSIAN = 25
GX_TRIAXIAL_GAUSSIAN = 26

GX_MAX_GL_ORDER = 128
GX_MATH_CHECK_OUTPUTS = 9

GX_FRAME_NONE = 0
GX_FRAME_OFFSET = 1
GX_FRAME_ROTATING = 2
GX_FRAME_TRACK = 3

GX_DOPRI8 = 0
GX_DOPRI5 = 1
GX_CONSTANT = 0
GX_PID = 1

GX_SUCCESS = 0
GX_MAX_STEPS_REACHED = 1
GX_DT_MIN_REACHED = 2
GX_NONFINITE = 3
STATUS_NAMES = {
    GX_SUCCESS: "successful",
    GX_MAX_STEPS_REACHED: "max_steps_reached",
    GX_DT_MIN_REACHED: "dt_min_reached",
    GX_NONFINITE: "nonfinite",
}

GX_TS_SHARED = 0
GX_TS_PER_ORBIT = 1
GX_TS_LINSPACE = 2
GX_TS_FINAL = 3

GX_DENSE_RECORD = 50

GX_MAX_COMPONENTS = 256
GX_MAX_TRACKS = 8


class GxComponent(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("frame", ctypes.c_int32),
        ("track", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("p", ctypes.c_double * 8),
        ("f", ctypes.c_double * 6),
    ]


class GxTrack(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("t", ctypes.c_void_p),
        ("q", ctypes.c_void_p),
        ("v", ctypes.c_void_p),
        ("rec", ctypes.c_void_p),
        ("method", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class GxSolverCfg(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int32),
        ("controller", ctypes.c_int32),
        ("max_steps", ctypes.c_int64),
        ("dt0", ctypes.c_double),
        ("rtol", ctypes.c_double),
        ("atol", ctypes.c_double),
        ("pcoeff", ctypes.c_double),
        ("icoeff", ctypes.c_double),
        ("dcoeff", ctypes.c_double),
        ("safety", ctypes.c_double),
        ("factormin", ctypes.c_double),
        ("factormax", ctypes.c_double),
        ("dtmin", ctypes.c_double),
        ("dtmax", ctypes.c_double),
        ("force_dtmin", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


def component(kind: int, p=(), frame: int = GX_FRAME_NONE, f=(), track: int = 0) -> GxComponent:
    c = GxComponent()
    c.kind = kind
    c.frame = frame
    c.track = track
    c.reserved = 0
    for i, v in enumerate(p):
        c.p[i] = float(v)
    for i, v in enumerate(f):
        c.f[i] = float(v)
    return c


def component_array(comps) -> ctypes.Array:
    arr = (GxComponent * max(len(comps), 1))()
    for i, c in enumerate(comps):
        arr[i] = c
    return arr

# surrogate GEMM epilogues (gx_mlp_gemm)
GX_EPI_STORE = 0
GX_EPI_BIAS = 1
GX_EPI_BIAS_SILU = 2
GX_EPI_DSILU = 3
