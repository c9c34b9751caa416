"""Loader for libgalaxhip.so and the device-memory plumbing around it.

PyTorch-ROCm supplies device memory and the current HIP stream; every
computation runs in the HIP library.  There is no CPU fallback: if the
library or a GPU is missing, calls raise.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

from . import _abi

# GALAX_AMD_LIB names another build of the same library (e.g. one compiled
# with different options, for timing comparisons); the default is the in-tree
# build next to this file.
LIB_PATH = Path(os.environ.get("GALAX_AMD_LIB") or Path(__file__).resolve().parent / "libgalaxhip.so")
_lib = None

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f32 = ctypes.c_float
_f64 = ctypes.c_double


class GalaxHipError(RuntimeError):
    """Error returned by the HIP library."""


SIGNATURES = {
    "gx_potential_create": (ctypes.c_int, [_vp, _i32, _vp, _i32, ctypes.POINTER(_vp)]),
    "gx_potential_destroy": (ctypes.c_int, [_vp]),
    "gx_acceleration": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i64, _vp, _vp]),
    "gx_gradient": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i64, _vp, _vp]),
    "gx_potential_eval": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i64, _vp, _vp]),
    "gx_hessian": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i64, _vp, _vp]),
    "gx_integrate_workspace_bytes": (_i64, [_i64]),
    "gx_integrate": (
        ctypes.c_int,
        [_vp, _vp, _vp, _vp, _vp, _i32, _vp, _i32, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    ),
    "gx_integrate_noalloc_max_orbits": (_i64, []),
    "gx_integrate_noalloc": (
        ctypes.c_int,
        [_vp, _vp, _vp, _vp, _vp, _i32, _vp, _i32, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    ),
    "gx_integrate_dense": (
        ctypes.c_int,
        [_vp, _vp, _vp, _vp, _vp, _i32, _vp, _i32, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp],
    ),
    "gx_dense_eval": (ctypes.c_int, [_i32, _vp, _i64, _vp, _i64, _vp, _i32, _i32, _vp, _vp]),
    "gx_integrate_shared_dense": (
        ctypes.c_int,
        [_vp, _vp, _vp, _f64, _f64, _vp, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp],
    ),
    "gx_integrate_shared_workspace_bytes": (_i64, [_i64]),
    "gx_integrate_shared": (
        ctypes.c_int, [_vp, _vp, _vp, _f64, _f64, _vp, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    ),
    "gx_set_launch_shape": (ctypes.c_int, [_i32, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "gx_set_launch_fraction": (ctypes.c_int, [_i32, ctypes.POINTER(_i32)]),
    "gx_set_queue_policy": (ctypes.c_int, [_i64, _i32, ctypes.POINTER(_i64), ctypes.POINTER(_i32)]),
    "gx_device_info": (ctypes.c_int, [ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "gx_math_check": (ctypes.c_int, [_vp, _i64, _vp, _vp]),
    "gx_nbody_acceleration": (ctypes.c_int, [_vp, _vp, _vp, _f64, _f64, _i64, _vp, _vp]),
    "gx_nbody_workspace_bytes": (_i64, [_i64]),
    "gx_nbody_integrate": (
        ctypes.c_int, [_vp, _vp, _f64, _vp, _vp, _f64, _f64, _vp, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    ),
    "gx_adjoint_workspace_bytes": (_i64, [_i64, _i64, _i32]),
    "gx_adjoint": (
        ctypes.c_int,
        [_vp, _vp, _vp, _vp, _vp, _i32, _vp, _i32, _i32, _i64, _vp, _vp, _vp, _i32, _i64, _i32, _vp, _vp, _vp],
    ),
    # surrogate kernels
    "gx_mlp_gemm": (
        ctypes.c_int,
        [_vp, _i64, _i64, _vp, _i64, _i64, _i64, _i64, _i64, _i32, _i32, _i32, _vp, _vp, _i64, _vp, _i64,
         _vp, _i64, _vp],
    ),
    "gx_mlp_forward_fused": (
        ctypes.c_int,
        [_vp, _i32, _i32, _i64, _i32, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _i32, _vp, _i32, _vp],
    ),
    "gx_mlp_colsum": (ctypes.c_int, [_vp, _i64, _i64, _i64, _vp, _vp]),
    "gx_mlp_adam": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i64, _f32, _f32, _f32, _f32, _i64, _vp]),
    "gx_mlp_huber": (ctypes.c_int, [_vp, _i64, _vp, _i64, _f32, _f32, _vp, _i64, _vp, _vp]),
    "gx_sur_inputs": (ctypes.c_int, [_vp, _vp, _i32, _f64, _i64, _f64, _f64, _vp, _i64, _vp]),
    "gx_sur_targets": (ctypes.c_int, [_vp, _i32, _i64, _vp, _vp, _f64, _f64, _vp, _vp]),
    "gx_sur_combine": (
        ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, _f64, _i64, _vp, _i64, _f64, _f64, _f64, _vp, _vp, _vp],
    ),
    "gx_sur_check": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i64, _vp, _f64, _vp, _vp, _vp]),
    "gx_method_nodes": (ctypes.c_int, [_i32, _vp, _i32]),
    "gx_lockstep_stage": (ctypes.c_int, [_i32, _vp, _vp, _vp, _i64, _f64, _vp, _vp]),
    "gx_lockstep_update": (ctypes.c_int, [_vp, _vp, _vp, _i64, _f64, _vp, _vp]),
}


def _declare(lib: ctypes.CDLL) -> None:
    missing = [name for name in SIGNATURES if not hasattr(lib, name)]
    if missing:
        raise GalaxHipError(f"{LIB_PATH} lacks {', '.join(missing)}: it is older than this package; "
                            "rebuild it with `python -m galax_amd._build` (or __graft_entry__.build())")
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib() -> ctypes.CDLL:
    """The loaded HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise GalaxHipError(
                f"{LIB_PATH} is missing: build it with `python -m galax_amd._build` "
                "(or __graft_entry__.build())"
            )
        # torch ships its own libamdhip64 under the same soname as the system
        # ROCm's.  Loading this library first would map the system runtime
        # and torch would then map its own beside it: two HIP runtimes in one
        # process, whose device memory and streams are not each other's
        # (calls fail with runtime errors).  Importing torch first makes the
        # library bind to the runtime torch already holds.
        import torch  # noqa: F401

        handle = ctypes.CDLL(str(LIB_PATH))
        _declare(handle)  # every signature, or none: never a half-declared handle
        _lib = handle
    return _lib


def check(rc: int, what: str) -> None:
    if rc != _abi.GX_OK:
        names = {-1: "invalid argument", -2: "HIP launch/runtime error", -3: "out of device memory",
                 -4: "more than the fixed-capacity entry point holds"}
        raise GalaxHipError(f"{what} failed: {names.get(rc, rc)} ({rc})")


def torch():
    import torch as _torch

    return _torch


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise GalaxHipError("galax_amd computes on a ROCm GPU; none is visible to this process")
    return t


def stream_ptr() -> int:
    t = torch()
    return int(t.cuda.current_stream().cuda_stream)


def ptr(x) -> int | None:
    return None if x is None else int(x.data_ptr())
